"""GPU parity of the dynamic network (cyminformal >= 0, the reference's default: jac_rhs.F90:
308-436) against the oracle's restatement (pinned in tests/test_oracle_dynamic.py).

The device keeps the reference's reduced system embedded in the static one (species outside the
mask solve to an exact zero, entries left out are explicit zeros), so:

* plain path: the oracle's step decisions and nvar1, <= 1e-10 on every X > 1e-20 (the GEPP
  eliminates in the static column order restricted to the mask, the reference's SuperLU in
  COLAMD's order of the reduced pattern: the solves agree to rounding);
* batched path: the exact-solve oracle's decisions and nvar1, <= 1e-8;
* a species the mask leaves out keeps its abundance exactly.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DYN = dict(cyminformal=1.0e-30, min_rate_ov_dt=1.0e-99)   # solver_knobs.F90:47-48


def _maxrel(a, b):
    m = (a > 1e-20) | (b > 1e-20)
    return float(np.max(np.abs(a[m] - b[m]) / np.maximum(a[m], b[m]))) if m.any() else 0.0


def test_dynamic_trajectory_plain_path(pkg, oracle):
    """the 271-species network over 20 ppn cycles: the mask grows from 13 to about 30 species"""
    w = pkg.network.config(0)
    ctx = pkg.hip.Context(w.net, pkg.hip.default_knobs(**DYN))
    o = oracle.Oracle(w.net, oracle.default_knobs(**DYN))
    try:
        y_g = w.x0[0].copy()
        y_o = w.x0[0].copy()
        seen = set()
        for s in range(20):
            args = (w.t9[s, 0], w.t9[s, 1], w.rho[s, 0], w.rho[s, 1], w.dt[s])
            ie_g, y_g, ye_g, st_g = ctx.integrate_network(y_g, *args)
            ie_o, y_o, ye_o, st_o = o.integrate_zone(y_o, *args)
            assert ie_g == ie_o == 0
            assert (st_g.nsubt, st_g.n_newton, st_g.nfail, st_g.nvar1) == \
                (st_o.nsubt, st_o.n_newton, st_o.nfail, st_o.nvar1), s
            e = _maxrel(y_g, y_o)
            assert e <= 1e-10, (s, e)
            assert abs(ye_g - ye_o) <= 1e-12
            seen.add(st_g.nvar1)
        print(f"nvar1 over the cycles: {sorted(seen)}; last rel {e:.1e}")
        assert min(seen) < 40 < w.net.nsp   # the mask is far smaller than the network
    finally:
        ctx.close()


def test_dynamic_batched_path_matches_exact_oracle(pkg, oracle):
    w = pkg.network.config(1, nz=16)
    ctx = pkg.hip.Context(w.net, pkg.hip.default_knobs(**DYN))
    o = oracle.Oracle(w.net, oracle.default_knobs(**DYN), oracle.LS_EXACT)
    try:
        sel = np.array([0, 2, 6, 9])
        nv = []
        args = [a[sel] for a in (w.t9[:, 0], w.t9[:, 1], w.rho[:, 0], w.rho[:, 1], w.dt)]
        ie_b, y_b, ye_b, st_b = ctx.integrate_batch_mixed(w.x0[sel], *args)
        v = ctx.rates(0.5 * (args[0] + args[1]), 0.5 * (args[2] + args[3]))
        for j in range(len(sel)):
            ie_o, y_o, ye_o, st_o = o.integrate_zone(w.x0[sel[j]], *(a[j] for a in args),
                                                     v_mid=v[j])
            e = _maxrel(y_b[j], y_o)
            print(f"zone {sel[j]} T9 {args[0][j]:.3f}: nvar1 {st_b['nvar1'][j]}/{st_o.nvar1} "
                  f"nsubt {st_b['nsubt'][j]}/{st_o.nsubt} newton "
                  f"{st_b['n_newton'][j]}/{st_o.n_newton} rel {e:.1e}")
            assert ie_b[j] == ie_o == 0
            assert (st_b["nsubt"][j], st_b["n_newton"][j], st_b["nvar1"][j]) == \
                (st_o.nsubt, st_o.n_newton, st_o.nvar1)
            assert e <= 1e-8
            nv.append(st_o.nvar1)
        # the seed abundances put nearly every species of these zones in the solve; the cold
        # zones still leave some out
        assert min(nv) < w.net.nsp - 1
    finally:
        ctx.close()


def test_species_outside_the_mask_keep_their_abundance(pkg, oracle):
    """a threshold that leaves out every reaction of some species with a nonzero dX/dt: one
    ppn cycle on the plain path leaves them exactly where they were, as the oracle does"""
    w = pkg.network.config(0)
    knobs = dict(cyminformal=1.0e-12, min_rate_ov_dt=1.0e-99)
    ctx = pkg.hip.Context(w.net, pkg.hip.default_knobs(**knobs))
    o = oracle.Oracle(w.net, oracle.default_knobs(**knobs))
    try:
        args = (w.t9[0, 0], w.t9[0, 1], w.rho[0, 0], w.rho[0, 1], w.dt[0])
        y0 = w.x0[0]
        ie_o, y_o, _, st_o = o.integrate_zone(y0.copy(), *args)
        ie_g, y_g, _, st_g = ctx.integrate_network(y0.copy(), *args)
        assert ie_o == ie_g == 0
        frozen = (y_o == y0) & (o.dxdt(y0, o.rates(args[0], args[2])) != 0.0)
        assert frozen.any()
        np.testing.assert_array_equal(y_g[frozen], y0[frozen])
        assert _maxrel(y_g, y_o) <= 1e-10
    finally:
        ctx.close()
