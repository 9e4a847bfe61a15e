"""Golden output of one very stiff Bader-Deuflhard zone (integration_method 1,
bader_deuflhard.F90:139-488) of the 4,096-zone x 1,100-species workload: zone 7 of the seeded
generator (T9 1.65, rho 3.4e6 g/cc, dt 2.3e-3 s: 43 sub-steps, 176 factorisations), integrated
by the oracle with its own mid-point rates from the initial BD state.  The oracle takes about a
minute and a half for it on one core, too long for a test run, so the result is stored.
Run:  python tests/golden/make_bd_stiff_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib as oracle  # noqa: E402
from __graft_entry__ import _load_package  # noqa: E402

ZONE = 7


def main():
    pkg = _load_package()
    w = pkg.network.config(4, nz=16)
    o = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0, integration_method=1))
    z = ZONE
    ie, y, ye, st = o.integrate_zone(w.x0[z], w.t9[z, 0], w.t9[z, 1], w.rho[z, 0], w.rho[z, 1],
                                     w.dt[z])
    np.savez_compressed(os.path.join(HERE, "bd_stiff_c4_zone7.npz"), zone=z, nz=16, ierr=ie, y=y,
                        ye=ye, nsubt=st.nsubt, iter=st.iter, nsucc=st.nsucc, nfail=st.nfail,
                        n_factor=st.n_factor, n_solve=st.n_solve)
    print(f"zone {z}: ierr {ie} nsubt {st.nsubt} iter {st.iter} succ/fail {st.nsucc}/{st.nfail} "
          f"factor {st.n_factor} solve {st.n_solve}")


if __name__ == "__main__":
    main()
