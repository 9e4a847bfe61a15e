"""GPU tests of the mppnp cycle on the device (nuppn_mppnp_cycle: zoneselect, the selected zones
as one batch, renorm_1d, mixing over all zones, renorm_2d, clip) against the oracle's
restatement (orc_mppnp_cycle, frames/mppnp/source/mppnp.F90:637-781 and mixing.F90).

Bars:
* mixing alone (no zone hot enough to integrate): bit for bit where the Lagrangian coefficients
  avoid libm (rho = rad = 1: both logarithms are exactly 0), both matrix forms, with and without
  the 2-d renormalisation; within 1e-13 on a stellar profile (log / exp of device and host may
  differ in the last place);
* the whole cycle on the plain FP64 path: the oracle's step decisions zone by zone and <= 1e-10
  on every X > 1e-20; on the batched path: the exact-solve oracle's decisions and <= 1e-8;
* an integration error ends the cycle before mixing, as mppnp stops there; a zero shell mass
  returns NUPPN_ERR_MIXING where mixing.F90 stops.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip(pkg):
    return pkg.hip


@pytest.fixture(scope="module")
def c1(pkg, hip):
    w = pkg.network.config(1, nz=64)
    ctx = hip.Context(w.net)
    yield w, ctx
    ctx.close()


def _maxrel(a, b):
    m = (a > 1e-20) | (b > 1e-20)
    if not m.any():
        return 0.0
    return float(np.max(np.abs(a[m] - b[m]) / np.maximum(a[m], b[m])))


@pytest.mark.parametrize("mesa", [0, 1])
@pytest.mark.parametrize("renorm", [False, True])
def test_mixing_bitwise(c1, hip, oracle, pkg, mesa, renorm):
    w, ctx = c1
    m = 64
    s = pkg.network.stellar_model(m)
    one = np.ones(m)
    y0 = w.x0[:m].copy()
    if renorm:
        # the last species (never mixed) uneven across zones: the zone sums drift by ~1e-3 and
        # the 2-d renormalisation runs
        y0[:, -1] = np.linspace(0.0, 2e-3, m)
    o = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0))
    rc_o, ie_o, y_o, _ = oracle.mppnp_cycle(o, y0, s.t9, one, s.xm, s.dq, one, s.d_eul, 1e6,
                                            t9threshold=100.0, mesa=bool(mesa))
    rc, ie, y, st = ctx.mppnp_cycle(y0, s.t9, one, s.xm, s.dq, one, s.d_eul, 1e6,
                                    opts=hip.mppnp_opts(t9threshold=100.0, code_source=mesa))
    assert rc == rc_o == 0
    assert not ie.any() and not ie_o.any()
    assert np.max(np.abs(y - y0)) > 1e-6   # it mixed
    sums = np.abs(y.sum(axis=1) - 1.0)
    print(f"mesa {mesa} renorm {renorm}: max |sum X - 1| after {sums.max():.1e}")
    np.testing.assert_array_equal(y, y_o)


def test_mixing_stellar_profile(c1, hip, oracle, pkg):
    w, ctx = c1
    m = 64
    s = pkg.network.stellar_model(m)
    y0 = w.x0[:m].copy()
    o = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0))
    for mesa in (0, 1):
        _, _, y_o, _ = oracle.mppnp_cycle(o, y0, s.t9, s.rho, s.xm, s.dq, s.rad, s.d_eul, 1e6,
                                          t9threshold=100.0, mesa=bool(mesa))
        rc, ie, y, _ = ctx.mppnp_cycle(y0, s.t9, s.rho, s.xm, s.dq, s.rad, s.d_eul, 1e6,
                                       opts=hip.mppnp_opts(t9threshold=100.0, code_source=mesa))
        assert rc == 0
        e = _maxrel(y, y_o)
        print(f"mesa {mesa}: max rel {e:.1e}, max change {np.max(np.abs(y - y0)):.1e}")
        assert e <= 1e-13


def _cycle_case(pkg, w, m=10):
    # a cool star: the inner zones (T9 0.3 .. 0.08) are integrated, the outer ones only mixed
    s = pkg.network.stellar_model(m, t9_center=0.3, t9_surface=0.01)
    return s, w.x0[:m].copy(), 0.08


def test_cycle_plain_path_matches_oracle(c1, hip, oracle, pkg):
    w, ctx = c1
    s, y0, thr = _cycle_case(pkg, w)
    o = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0))
    rc_o, ie_o, y_o, st_o = oracle.mppnp_cycle(o, y0, s.t9, s.rho, s.xm, s.dq, s.rad, s.d_eul,
                                               1e4, t9threshold=thr)
    rc, ie, y, st = ctx.mppnp_cycle(y0, s.t9, s.rho, s.xm, s.dq, s.rad, s.d_eul, 1e4,
                                    opts=hip.mppnp_opts(t9threshold=thr, path=0))
    sel = s.t9 > thr
    print(f"{int(sel.sum())} of {s.m} zones integrated; newton {st['n_newton'].tolist()}")
    assert rc == rc_o == 0
    np.testing.assert_array_equal(ie, ie_o)
    for j in range(s.m):
        assert (st["nsubt"][j], st["n_newton"][j]) == (st_o[j].nsubt, st_o[j].n_newton)
    assert 2 <= sel.sum() < s.m
    e = _maxrel(y, y_o)
    print(f"max rel vs oracle {e:.1e}")
    assert e <= 1e-10


def test_cycle_batched_path_matches_exact_oracle(c1, hip, oracle, pkg):
    w, ctx = c1
    s, y0, thr = _cycle_case(pkg, w)
    o = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0), oracle.LS_EXACT)
    rc_o, ie_o, y_o, st_o = oracle.mppnp_cycle(o, y0, s.t9, s.rho, s.xm, s.dq, s.rad, s.d_eul,
                                               1e4, t9threshold=thr)
    rc, ie, y, st = ctx.mppnp_cycle(y0, s.t9, s.rho, s.xm, s.dq, s.rad, s.d_eul, 1e4,
                                    opts=hip.mppnp_opts(t9threshold=thr))
    assert rc == rc_o == 0
    np.testing.assert_array_equal(ie, ie_o)
    for j in range(s.m):
        assert (st["nsubt"][j], st["n_newton"][j]) == (st_o[j].nsubt, st_o[j].n_newton)
    e = _maxrel(y, y_o)
    print(f"max rel vs exact oracle {e:.1e}")
    assert e <= 1e-8
    assert np.all(np.abs(y.sum(axis=1) - 1.0) <= 1e-12)


def test_cycle_stops_before_mixing_on_an_integration_error(pkg, hip, oracle):
    w = pkg.network.config(1, nz=10)
    # a hotter star: its second zone (T9 0.32) needs more than the one sub-step allowed here
    s = pkg.network.stellar_model(10, t9_center=1.0, t9_surface=0.01)
    y0, thr = w.x0[:10].copy(), 0.08
    ctx = hip.Context(w.net, hip.default_knobs(max_sub_steps=1))
    o = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0, max_sub_steps=1),
                      oracle.LS_EXACT)
    try:
        rc, ie, y, st = ctx.mppnp_cycle(y0, s.t9, s.rho, s.xm, s.dq, s.rad, s.d_eul, 1e4,
                                        opts=hip.mppnp_opts(t9threshold=thr))
        rc_o, ie_o, y_o, _ = oracle.mppnp_cycle(o, y0, s.t9, s.rho, s.xm, s.dq, s.rad, s.d_eul,
                                                1e4, t9threshold=thr)
        print("ierr", ie.tolist())
        assert rc == rc_o == 0
        np.testing.assert_array_equal(ie, ie_o)
        assert (ie == 2).any()                       # MAX_SUBSTEPS
        cold = s.t9 <= thr
        np.testing.assert_array_equal(y[cold], y0[cold])   # neither mixed nor clipped
        assert _maxrel(y, y_o) <= 1e-8
        # a zero shell mass stops the MESA-form mixing
        bad = s.dq.copy()
        bad[0] = 0.0
        ctx2 = hip.Context(w.net)
        try:
            rc, _, _, _ = ctx2.mppnp_cycle(y0, s.t9, s.rho, s.xm, bad, s.rad, s.d_eul, 1e4,
                                           opts=hip.mppnp_opts(t9threshold=100.0,
                                                               code_source=1))
            assert rc == hip.ERR_MIXING
        finally:
            ctx2.close()
    finally:
        ctx.close()
