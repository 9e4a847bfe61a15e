"""The REACLIB T9 rate cache (physics knob use_cache; reaclib_preprocessor, physics/source/
reaclib.F90:609-646, and the interpolating branch of reaclib_calculate_rates, 446-470).

CPU: the oracle's table and lookup against what the reference's formulas fix - the table rows are
the direct rates at rho = 1 on T9 = 0.01 + i (10 - 0.01) / 5000; at a grid point the lookup returns
the row, between two points the straight line through them (times rho for two-body rates), below
T9 = 0.01 the first row, above 10 the last interval extended; and the interpolation error stays
within the second difference of the table (|f - p| <= h^2 / 8 max|f''|).
GPU: the device table against the oracle's within exp's last-place differences, the device lookup
bit for bit against the oracle's on the device's own table, and a whole integration with the
cache on against the oracle with the cache on."""
import numpy as np
import pytest

NT9 = 5001
DT9 = (10.0 - 0.01) / (NT9 - 1)


@pytest.fixture(scope="module")
def c1net(pkg):
    return pkg.network.config(1, nz=8)


def test_cache_table_rows_are_direct_rates(oracle, c1net):
    w = c1net
    o = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0))
    o.set_rate_cache(True)
    tab = o.rate_cache()
    assert tab.shape == (NT9, w.net.nre)
    d = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0))
    for i in (0, 1, 777, 2500, 5000):
        np.testing.assert_array_equal(tab[i], d.rates(0.01 + i * DT9, 1.0))
    assert np.all(tab >= 1e-99)


def test_cache_lookup_is_linear_interpolation(oracle, c1net):
    w = c1net
    o = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0))
    o.set_rate_cache(True)
    tab = o.rate_cache()
    two = w.net.rtype == 0
    rho = 3.7e3
    scale = np.where(two, rho, 1.0)
    rng = np.random.default_rng(11)
    for i in rng.integers(0, NT9 - 1, 12):
        for f in (0.0, 0.25, 0.5, 0.9):
            t9 = 0.01 + i * DT9 + f * DT9
            v = o.rates(t9, rho)
            j = int(np.floor((t9 - 0.01) / DT9))
            j = min(max(j, 0), NT9 - 2)
            g = (t9 - (0.01 + j * DT9)) / DT9
            expect = np.maximum((tab[j] + g * (tab[j + 1] - tab[j])) * scale, 1e-99)
            np.testing.assert_allclose(v, expect, rtol=1e-14, atol=0)
    # below the grid: the clamp to T9 = 0.01 (reaclib.F90:443) gives the first row
    np.testing.assert_allclose(o.rates(0.001, rho), np.maximum(tab[0] * scale, 1e-99), rtol=1e-15)
    # above the grid: the last interval, extended
    t9 = 12.0
    g = (t9 - (0.01 + (NT9 - 2) * DT9)) / DT9
    expect = np.maximum((tab[NT9 - 2] + g * (tab[NT9 - 1] - tab[NT9 - 2])) * scale, 1e-99)
    np.testing.assert_allclose(o.rates(t9, rho), expect, rtol=1e-13)


def test_cache_interpolation_error_bound(oracle, c1net):
    """|lerp - direct| <= h^2/8 max|f''| on the interval, f'' taken from the table's second
    differences around it (the rates are smooth and convex on one grid interval)."""
    w = c1net
    o = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0))
    o.set_rate_cache(True)
    tab = o.rate_cache()
    d = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0))
    rng = np.random.default_rng(5)
    for i in rng.integers(2, NT9 - 3, 20):
        t9 = 0.01 + (i + 0.5) * DT9
        vc, vd = o.rates(t9, 1.0), d.rates(t9, 1.0)
        d2 = np.max(np.abs(tab[i - 1:i + 3][:-2] - 2 * tab[i:i + 2] + tab[i + 1:i + 3]), axis=0)
        bound = 0.25 * d2 + 1e-13 * vd + 1e-99
        big = vd > 1e-90
        assert np.all(np.abs(vc - vd)[big] <= bound[big])


@pytest.mark.gpu
def test_device_cache_table_and_lookup(pkg, oracle, c1net):
    hip = pkg.hip
    w = c1net
    ctx = hip.Context(w.net)
    o = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0))
    o.set_rate_cache(True)
    try:
        ctx.set_rate_cache(True)
        tab_d = ctx.rate_cache_table()
        tab_o = o.rate_cache()
        # both tables evaluate the same formula at the same T9; each is within the bound that
        # the parity test of the direct rates uses, so they differ by at most twice it
        t = 0.01 + np.arange(NT9) * DT9
        basis = np.stack([np.ones_like(t), 1 / t, t ** (-1 / 3), t ** (1 / 3), t, t ** (5 / 3),
                          np.abs(np.log(t))], 1)
        terms = basis @ np.abs(w.net.coef)
        bound = 2 * (12 * terms + 8) * np.finfo(float).eps
        live = (tab_o > 1e-99) & (tab_d > 1e-99)
        rel = np.abs(tab_d - tab_o)[live] / tab_o[live]
        print(f"table: max rel {rel.max():.2e}, max rel/bound {(rel / bound[live]).max():.2f}")
        assert np.all(rel <= bound[live])
        assert np.array_equal(tab_d <= 1e-99, tab_o <= 1e-99)
        # the lookup, bit for bit, on the device's table: interpolate it on the host the same way
        t9 = np.array([0.005, 0.0123, 0.3, 1.7777, 9.99, 11.0])
        rho = np.array([1.0, 1e3, 5e5, 2.0, 1e8, 10.0])
        v = ctx.rates(t9, rho)
        two = w.net.rtype == 0
        for z in range(len(t9)):
            t = max(t9[z], 0.01)
            j = min(max(int(np.floor((t - 0.01) / DT9)), 0), NT9 - 2)
            g = (t - (0.01 + j * DT9)) / DT9
            e = tab_d[j] + g * (tab_d[j + 1] - tab_d[j])
            e = np.maximum(np.where(two, e * rho[z], e), 1e-99)
            np.testing.assert_array_equal(v[z], e)
        # off again: the direct formula (reaclib_calculate_rates without the cache)
        ctx.set_rate_cache(False)
        o.set_rate_cache(False)
        vd = ctx.rates(t9[2:3], rho[2:3])[0]
        vo = o.rates(t9[2], rho[2])
        assert not np.array_equal(vd, v[2])
        np.testing.assert_allclose(vd, vo, rtol=2e-12)
    finally:
        ctx.close()


@pytest.mark.gpu
def test_integration_with_cache_matches_oracle(pkg, oracle):
    """A ppn-style trajectory of the 271-species network with the cache on (the in-loop rate
    re-evaluations interpolate too): identical decisions and <= 1e-10 against the oracle with
    the cache on (the tables differ in exp's last place)."""
    hip = pkg.hip
    w = pkg.network.config(0)
    ctx = hip.Context(w.net)
    o = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0))
    ctx.set_rate_cache(True)
    o.set_rate_cache(True)
    y_g = w.x0[0].copy()
    y_o = w.x0[0].copy()
    try:
        for s in range(6):
            args = (w.t9[s, 0], w.t9[s, 1], w.rho[s, 0], w.rho[s, 1], w.dt[s])
            ie_g, y_g, ye_g, st_g = ctx.integrate_network(y_g, *args)
            ie_o, y_o, ye_o, st_o = o.integrate_zone(y_o, *args)
            m = y_o > 1e-20
            e = float(np.max(np.abs(y_g[m] - y_o[m]) / y_o[m]))
            print(f"cycle {s}: T9 {args[0]:.3f}..{args[1]:.3f} ierr {ie_g}/{ie_o} nsubt "
                  f"{st_g.nsubt}/{st_o.nsubt} newton {st_g.n_newton}/{st_o.n_newton} rel {e:.2e}")
            assert ie_g == ie_o == 0
            assert (st_g.nsubt, st_g.n_newton) == (st_o.nsubt, st_o.n_newton)
            assert e <= 1e-10
    finally:
        ctx.close()
