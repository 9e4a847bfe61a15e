"""GPU parity of the rate physics (nuppn_set_rate_physics: partition-function factors of the
REACLIB reverse rates, detailed balance, Chugunov screening; evaluate_rates.F90:14-172) with the
oracle's restatement (tests/test_oracle_rate_physics.py pins it).

Bars: the rates within the conditioning bound of the REACLIB exponent plus 2e-12 (the device's
and the host's exp / log / pow may differ in the last place, and exp(-Q / kT) and the screening
exponent amplify that by their arguments, <= 200 and <= 69); the same pairing as the oracle's;
integrations with the rate physics on (rates re-evaluated at the current abundances whenever
T9 or rho change) take the oracle's step decisions and end within 1e-10 (plain path, each side
its own rates) or 1e-8 of the exact-solve oracle (batched path, the device's rate vector).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def _rate_bound(coef, t9):
    t = max(t9, 0.01)
    terms = np.abs(coef * np.array([1, 1 / t, t ** (-1 / 3), t ** (1 / 3), t, t ** (5 / 3),
                                    abs(math.log(t))])[:, None]).sum(0)
    return (12 * terms + 8) * EPS


@pytest.fixture(scope="module")
def c1(pkg):
    w = pkg.network.config(1, nz=16)
    rp = pkg.network.rate_physics(w.net)
    ctx = pkg.hip.Context(w.net)
    ctx.set_rate_physics(rp)
    yield w, rp, ctx
    ctx.close()


def _oracle(oracle, w, rp, ls=None):
    o = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0),
                      oracle.LS_GEPP if ls is None else ls)
    o.set_rate_physics(rp)
    return o


def test_reverse_pairs_equal_oracle(c1, oracle):
    w, rp, ctx = c1
    np.testing.assert_array_equal(ctx.reverse_pairs(), _oracle(oracle, w, rp).reverse_pairs())


def test_rates_with_rate_physics(c1, oracle):
    w, rp, ctx = c1
    o = _oracle(oracle, w, rp)
    t9s = np.array([0.005, 0.05, 0.3, 1.2, 3.0, 7.0])
    rhos = np.array([1e3, 10.0, 1e5, 3e7, 1e8, 1e10])
    x = w.x0[:6]
    v = ctx.evaluate_all_rates(t9s, rhos, x)
    vr = ctx.rates(t9s, rhos)   # reaclib_calculate_rates: the partition-function factors only
    irev = o.reverse_pairs()
    fwd = np.arange(w.net.nre)
    fwd[irev[irev >= 0]] = np.where(irev >= 0)[0]   # a reverse rate inherits its forward bound
    for i in range(len(t9s)):
        vo = o.rates_full(t9s[i], rhos[i], x[i])
        b = _rate_bound(w.net.coef, t9s[i])
        rel = np.abs(v[i] - vo) / vo
        print(f"T9 {t9s[i]}: max rel {rel.max():.1e}")
        assert np.all(rel <= b[fwd] + 2e-12)
        vro = o.rates(t9s[i], rhos[i])
        assert np.all(np.abs(vr[i] - vro) / vro <= b + 2e-12)
        # the physics does something: screening raises and balance sets some rates
        assert np.any(v[i] > 1.01 * vr[i]) or t9s[i] < 0.01


def test_cache_with_partition_functions(pkg, oracle, c1):
    w, rp, _ = c1
    ctx = pkg.hip.Context(w.net)
    try:
        ctx.set_rate_cache(True)
        ctx.set_rate_physics(rp)   # rebuilds the cache with the reverse factors
        tab = ctx.rate_cache_table()
        o = _oracle(oracle, w, rp)
        o.set_rate_cache(True)
        to = o.rate_cache()
        for i in (0, 1000, 5000):
            t9 = 0.01 + i * (10.0 - 0.01) / 5000
            rel = np.abs(tab[i] - to[i]) / to[i]
            assert np.all(rel <= _rate_bound(w.net.coef, t9) + 2e-12)
    finally:
        ctx.close()


def test_ppn_trajectory_with_rate_physics(pkg, oracle):
    """The 271-species network over 20 ppn cycles whose T9 and rho change within each cycle, so
    the controller re-evaluates the rates (with screening at the current abundances) on every
    sub-step: the oracle's decisions, <= 1e-10."""
    w = pkg.network.config(0)
    rp = pkg.network.rate_physics(w.net)
    ctx = pkg.hip.Context(w.net)
    ctx.set_rate_physics(rp)
    o = _oracle(oracle, w, rp)
    try:
        y_g = w.x0[0].copy()
        y_o = w.x0[0].copy()
        for s in range(20):
            args = (w.t9[s, 0], w.t9[s, 1], w.rho[s, 0], w.rho[s, 1], w.dt[s])
            ie_g, y_g, ye_g, st_g = ctx.integrate_network(y_g, *args)
            ie_o, y_o, ye_o, st_o = o.integrate_zone(y_o, *args)
            assert ie_g == ie_o == 0
            assert (st_g.nsubt, st_g.n_newton, st_g.nfail) == \
                (st_o.nsubt, st_o.n_newton, st_o.nfail)
            m = y_o > 1e-20
            e = float(np.max(np.abs(y_g[m] - y_o[m]) / y_o[m]))
            assert e <= 1e-10, (s, e)
            assert abs(ye_g - ye_o) <= 1e-12
        print(f"20 cycles, last: nsubt {st_g.nsubt} newton {st_g.n_newton} rel {e:.1e}")
    finally:
        ctx.close()


def test_batched_path_with_rate_physics(c1, oracle):
    w, rp, ctx = c1
    sel = np.array([0, 2, 6])   # 1, 8 and 1 sub-steps with the rate physics on
    args = [a[sel] for a in (w.t9[:, 0], w.t9[:, 1], w.rho[:, 0], w.rho[:, 1], w.dt)]
    ie_b, y_b, ye_b, st_b = ctx.integrate_batch_mixed(w.x0[sel], *args)
    v = ctx.evaluate_all_rates(0.5 * (args[0] + args[1]), 0.5 * (args[2] + args[3]), w.x0[sel])
    o = _oracle(oracle, w, rp, oracle.LS_EXACT)
    for j in range(len(sel)):
        ie_o, y_o, ye_o, st_o = o.integrate_zone(w.x0[sel[j]], *(a[j] for a in args), v_mid=v[j])
        m = (y_b[j] > 1e-20) | (y_o > 1e-20)
        e = float(np.max(np.abs(y_b[j][m] - y_o[m]) / np.maximum(y_b[j][m], y_o[m])))
        print(f"zone {sel[j]} T9 {args[0][j]:.3f}: nsubt {st_b['nsubt'][j]}/{st_o.nsubt} newton "
              f"{st_b['n_newton'][j]}/{st_o.n_newton} rel {e:.1e}")
        assert ie_b[j] == ie_o == 0
        assert (st_b["nsubt"][j], st_b["n_newton"][j]) == (st_o.nsubt, st_o.n_newton)
        assert e <= 1e-8
    assert np.all(np.abs(y_b.sum(axis=1) - 1.0) <= 1e-12)
