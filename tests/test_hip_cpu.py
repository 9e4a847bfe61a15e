"""CPU-side checks of the HIP library: it loads, exports every entry point of include/*.h, and its
host symbolic phase (COLAMD + column-etree post-order) reproduces the reference SuperLU's column
order bit for bit.  No GPU calls."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        for m in re.finditer(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b([A-Za-z_]\w*)\s*\(", txt, flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof", "defined"):
                names.append(name)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(pkg):
    hip = pkg.hip
    lib = hip.load()
    names = _declared_functions()
    assert {"nuppn_ctx_create", "nuppn_integrate_batch", "slu_solve", "c_fortran_dgssv_"} <= set(names)
    for n in names:
        assert hasattr(lib, n), n
        assert isinstance(getattr(lib, n), C._CFuncPtr)


def _network_patterns(pkg, oracle):
    out = []
    for cfg in (0, 1, 3):
        w = pkg.network.config(cfg, nz=1)
        o = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0))
        v = o.rates(w.t9[0, 0], w.rho[0, 0])
        n, rp, ci, va, mask = o.reduced_matrix(w.dt[0], w.x0[0], v)
        cp, ri, cv = oracle.csr_to_csc(n, rp, ci, va)
        out.append((f"net{cfg}", n, cp, ri))
    return out


def test_column_order_matches_reference_superlu(pkg, oracle):
    if not oracle.ref_available():
        pytest.skip("reference SuperLU (oracle/_ref) not built")
    hip = pkg.hip
    g = np.load(os.path.join(ROOT, "tests", "golden", "g20.npz"))
    cases = [("g20", int(g["n"]), g["colptr"], g["rowind"])] + _network_patterns(pkg, oracle)
    import scipy.sparse as sp
    rng = np.random.default_rng(7)
    for t in range(25):
        n = int(rng.integers(3, 300))
        A = sp.random(n, n, density=rng.uniform(0.003, 0.1), random_state=int(rng.integers(1 << 30)),
                      format="csc") + sp.eye(n, format="csc")
        if t % 4 == 1:   # dense rows and columns exercise COLAMD's dense handling
            d = np.zeros((n, n))
            d[0, :] = 1
            d[:, n // 2] = 1
            A = A + sp.csc_matrix(d)
        if t % 5 == 2:   # an empty column
            A = A.tolil()
            A[:, n - 1] = 0
            A = A.tocsc()
        A = A.tocsc()
        A.eliminate_zeros()
        A.sort_indices()
        cases.append((f"rand{t}", n, A.indptr, A.indices))
    for name, n, cp, ri in cases:
        pc, et = hip.column_order(n, cp, ri, with_etree=True)
        p0, p1, et_ref = oracle.ref_colperm(n, cp, ri)
        np.testing.assert_array_equal(pc, p1, err_msg=name)
        np.testing.assert_array_equal(et, et_ref, err_msg=name)


def test_no_cpu_fallback_without_library(pkg, monkeypatch):
    """The product path has no CPU fallback: a missing library raises."""
    hip = pkg.hip
    monkeypatch.setattr(hip, "_lib", None)
    monkeypatch.setattr(hip, "LIB_PATH", "/nonexistent/libnuppn_hip.so")
    with pytest.raises(RuntimeError):
        hip.load()
