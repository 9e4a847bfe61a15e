import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "nuppn_modular2-ininet3_amd")
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def load_package():
    """Import the package directory (its name is not a Python identifier) as `nuppn_amd`."""
    if "nuppn_amd" in sys.modules:
        return sys.modules["nuppn_amd"]
    spec = importlib.util.spec_from_file_location(
        "nuppn_amd", os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["nuppn_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long CPU-oracle run")


@pytest.fixture(scope="session")
def pkg():
    return load_package()


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    oracle_lib.lib()
    return oracle_lib
