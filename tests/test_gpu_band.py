"""GPU tests of the batched path's band form (csrc/band_lu.hpp): the pivot-free LU of the
arrowhead matrix (a band of half-width ~100 plus the border of hub species) that serves the
5,000-species network, whose interior is too wide for the block-tridiagonal form.

* the 5,000-species network: the band form is chosen, and its zones take the exact-solve
  oracle's step decisions and end within 1e-8 of it (the refined solves are exactly rounded, so
  the form of the factorisation does not enter the result), with |sum X - 1| <= 1e-12;
* the 1,100-species network run in both forms: the band form gives the block form's results bit
  for bit and the same decisions.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _maxrel(a, b):
    m = (a > 1e-20) | (b > 1e-20)
    return float(np.max(np.abs(a[m] - b[m]) / np.maximum(a[m], b[m]))) if m.any() else 0.0


def test_c3_band_form_matches_exact_oracle(pkg, oracle):
    w = pkg.network.config(3, nz=8)
    ctx = pkg.hip.Context(w.net)
    try:
        info = ctx.block_info()
        print(info)
        assert info["available"] and info["NB"] == -1 and info["bandwidth"] > 64
        sel = np.array([0, 5, 6, 7])   # one zone of 15 sub-steps, three of one
        args = [a[sel] for a in (w.t9[:, 0], w.t9[:, 1], w.rho[:, 0], w.rho[:, 1], w.dt)]
        ie_b, y_b, ye_b, st_b = ctx.integrate_batch_mixed(w.x0[sel], *args)
        v = ctx.rates(0.5 * (args[0] + args[1]), 0.5 * (args[2] + args[3]))
        o = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0), oracle.LS_EXACT)
        for j in range(len(sel)):
            ie_o, y_o, ye_o, st_o = o.integrate_zone(w.x0[sel[j]], *(a[j] for a in args),
                                                     v_mid=v[j])
            e = _maxrel(y_b[j], y_o)
            print(f"zone {sel[j]} T9 {args[0][j]:.3f}: nsubt {st_b['nsubt'][j]}/{st_o.nsubt} "
                  f"newton {st_b['n_newton'][j]}/{st_o.n_newton} fp64 factors "
                  f"{st_b['n_factor_fp64'][j]}/{st_b['n_factor'][j]} rel {e:.1e}")
            assert ie_b[j] == ie_o == 0
            assert (st_b["nsubt"][j], st_b["n_newton"][j]) == (st_o.nsubt, st_o.n_newton)
            assert e <= 1e-8
        assert np.all(np.abs(y_b.sum(axis=1) - 1.0) <= 1e-12)
    finally:
        ctx.close()


def test_band_form_equals_block_form(pkg):
    hip = pkg.hip
    w = pkg.network.config(1, nz=16)
    ctx = hip.Context(w.net)
    try:
        sel = np.arange(16)
        args = [a[sel] for a in (w.t9[:, 0], w.t9[:, 1], w.rho[:, 0], w.rho[:, 1], w.dt)]
        assert ctx.block_info()["NB"] == 32
        ie_a, y_a, _, st_a = ctx.integrate_batch_mixed(w.x0[sel], *args)
        assert ctx.set_block_form(-1) == hip.OK
        info = ctx.block_info()
        assert info["NB"] == -1
        ie_b, y_b, _, st_b = ctx.integrate_batch_mixed(w.x0[sel], *args)
        print(f"newton {st_a['n_newton'].sum()} / {st_b['n_newton'].sum()}, fp64 factors "
              f"{st_a['n_factor_fp64'].sum()} / {st_b['n_factor_fp64'].sum()}")
        np.testing.assert_array_equal(ie_a, ie_b)
        np.testing.assert_array_equal(st_a["nsubt"], st_b["nsubt"])
        np.testing.assert_array_equal(st_a["n_newton"], st_b["n_newton"])
        np.testing.assert_array_equal(y_a, y_b)
        assert ctx.set_block_form(32) == hip.OK and ctx.block_info()["NB"] == 32
    finally:
        ctx.close()
