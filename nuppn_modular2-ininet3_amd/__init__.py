"""MI355X-native network-integration path of NuPPN: REACLIB rates, sparse Jacobian assembly and
sparse LU inside the backward-Euler Newton step and the Bader-Deuflhard step, behind a C ABI
(include/nuppn_hip.h) that mirrors the reference's solver entry points."""
from . import network  # noqa: F401
from . import hip  # noqa: F401
