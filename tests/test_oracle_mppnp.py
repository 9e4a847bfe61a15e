"""Pins of the oracle's mppnp-cycle pieces (oracle/nuppn_oracle.c: zoneselect, ksum, renorm_1d,
mix_calculate_D_Lagrangian, matrixdiffus, matrixdiffus_mesa, tridag, do_mixing, mppnp_cycle)
by what the mathematics fixes, each chosen so that a plausible slip fails it:

* tridag against a dense solve (a swapped sub/super-diagonal or an off-by-one in gam fails);
* every row of matrixdiffus sums to one, so a uniform composition is a fixed point of mixing
  (a sign or xl / xr swap fails), in both matrix forms;
* matrixdiffus mixing conserves sum_j w_j X_j with the finite-volume weights w_1 = xm_2 - xm_1,
  w_j = (xm_{j+1} - xm_{j-1}) / 2, w_m = xm_m - xm_{m-1} (any wrong interface width fails);
* D_lagr against its closed form D_eul (4 pi Rsun^2 / Msun)^2 r^4 rho^2 with r and rho the
  geometric means over the zone and its inner neighbour;
* the coefficient limit: with sig_term_limit below every D h / (xmm xr), the limited rows are
  exactly those of D = sig xmm xr / h;
* ksum recovers what plain summation loses; renorm_1d gives a unit Kahan sum;
* zoneselect keeps exactly the zones above the threshold and reports the last one;
* the whole cycle with no mixing (D = 0) and nothing selected leaves the composition as it is
  (clip aside), and with every zone selected equals the integrator zone by zone.
"""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def L(oracle):
    return oracle.lib()


def _tri_dense(a, b, c):
    m = b.size
    A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    return A


def test_tridag_against_dense_solve(L, oracle):
    rng = np.random.default_rng(5)
    for m in (2, 3, 17, 200):
        a = -rng.uniform(0.0, 1.0, m)
        c = -rng.uniform(0.0, 1.0, m)
        a[0] = 0.0
        c[-1] = 0.0
        b = 1.0 + np.abs(a) + np.abs(c) + rng.uniform(0.0, 1.0, m)
        r = rng.uniform(-1.0, 1.0, m)
        u = np.zeros(m)
        assert L.orc_tridag(m, a, b, c, r, u) == 0
        x = np.linalg.solve(_tri_dense(a, b, c), r)
        np.testing.assert_allclose(u, x, rtol=1e-12, atol=1e-14)
    # a zero pivot is where the reference stops
    assert L.orc_tridag(2, np.zeros(2), np.zeros(2), np.zeros(2), np.ones(2), np.zeros(2)) == 1


def _model(pkg, m=40):
    return pkg.network.stellar_model(m)


def test_matrix_rows_sum_to_one_and_uniform_is_fixed(L, oracle, pkg):
    s = _model(pkg)
    m = s.m
    dl = np.zeros(m)
    L.orc_mix_d_lagrangian(m, s.rho, s.rad, s.d_eul, dl)
    for mesa in (0, 1):
        a, b, c = np.zeros(m), np.zeros(m), np.zeros(m)
        D = dl.copy()
        if mesa:
            L.orc_matrixdiffus_mesa(m, 1e4, s.dq, D, 1e10, a, b, c)
        else:
            assert L.orc_matrixdiffus(m, 1e4, s.xm, D, 1e10, a, b, c) == 0
        scale = np.abs(a) + np.abs(b) + np.abs(c)
        assert np.all(np.abs(a + b + c - 1.0) <= 4e-16 * scale)
        assert a[0] == 0.0 and c[-1] == 0.0 and np.all(a <= 0) and np.all(c <= 0)
        y = np.full((m, 5), 0.2)
        rc, y2 = oracle.do_mixing(y, np.array([1, 1, 1, 1, 1]), s.xm, s.dq, s.rho, s.rad,
                                  s.d_eul, 1e4, mesa=bool(mesa))
        assert rc == 0
        np.testing.assert_allclose(y2[:, :4], 0.2, rtol=1e-13)
        assert np.array_equal(y2[:, 4], y[:, 4])   # the last species is not mixed


def test_mixing_conserves_weighted_mass(oracle, pkg):
    s = _model(pkg)
    m = s.m
    rng = np.random.default_rng(9)
    y = rng.uniform(0.0, 1.0, (m, 3))
    rc, y2 = oracle.do_mixing(y, np.array([1, 1, 1]), s.xm, s.dq, s.rho, s.rad, s.d_eul, 1e4)
    assert rc == 0
    xm = s.xm
    w = np.empty(m)
    w[0] = xm[1] - xm[0]
    w[1:-1] = 0.5 * (xm[2:] - xm[:-2])
    w[-1] = xm[-1] - xm[-2]
    for k in range(2):
        assert abs(w @ y2[:, k] - w @ y[:, k]) <= 1e-13 * (w @ y[:, k])
        assert np.max(np.abs(y2[:, k] - y[:, k])) > 1e-3   # it did mix


def test_d_lagrangian_closed_form(L, pkg):
    s = _model(pkg)
    m = s.m
    dl = np.zeros(m)
    L.orc_mix_d_lagrangian(m, s.rho, s.rad, s.d_eul, dl)
    dc = 3.0589703419947736e-11
    r = np.sqrt(s.rad * np.concatenate([[s.rad[0]], s.rad[:-1]]))
    rho = np.sqrt(s.rho * np.concatenate([[s.rho[0]], s.rho[:-1]]))
    ref = np.where(s.d_eul > 0, s.d_eul * (dc * r ** 2 * rho) ** 2, 0.0)
    np.testing.assert_allclose(dl, ref, rtol=1e-13)
    assert np.all(dl[s.d_eul <= 0] == 0.0)


def test_coefficient_limit(L, pkg):
    s = _model(pkg)
    m = s.m
    h = 1e4
    D = np.full(m, 1e3)   # far above any limit below
    a, b, c = np.zeros(m), np.zeros(m), np.zeros(m)
    sig = 1e-3
    assert L.orc_matrixdiffus(m, h, s.xm, D, sig, a, b, c) == 0
    xm = s.xm
    for j in range(1, m - 1):   # rows 2..m-1: D(j+1) limited to sig xmm xr / h
        xr = xm[j + 1] - xm[j]
        xmm = 0.5 * (xm[j + 1] - xm[j - 1])
        alpha = h / xmm
        np.testing.assert_allclose(-c[j], sig, rtol=1e-12)   # D(j+1) alpha / xr == sig
        assert D[j + 1] == pytest.approx(sig / alpha * xr, rel=1e-15)
    assert b[0] == pytest.approx(1.0 + sig, rel=1e-15)   # row 1 limits alpha itself


def test_ksum_and_renorm(L):
    x = np.array([1.0] + [1e-16] * 1000)
    assert L.orc_ksum(x.size, x) == pytest.approx(1.0 + 1e-13, rel=1e-15)
    plain = 0.0
    for v in x:
        plain = plain + v
    assert plain == 1.0   # sequential FP64 summation loses every small term
    y = np.random.default_rng(1).uniform(0, 1, 300)
    L.orc_renorm_1d(y.size, y)
    assert abs(L.orc_ksum(y.size, y) - 1.0) <= 4e-16
    assert L.orc_mass_residual(3, np.array([0.5, 0.25, 0.25])) == 0.0


def test_zoneselect(L):
    t9 = np.array([0.5, 0.05, 0.2, 0.1, 0.01])
    sel = np.zeros(5, np.uint8)
    assert L.orc_zoneselect(0.1, t9, 5, sel) == 3
    assert sel.tolist() == [1, 0, 1, 0, 0]
    assert L.orc_zoneselect(1.0, t9, 5, sel) == 0 and not sel.any()


def test_cycle_without_mixing(oracle, pkg):
    w = pkg.network.config(1, nz=6)
    o = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0))
    s = pkg.network.stellar_model(6, t9_center=0.3, t9_surface=0.06)
    zero = np.zeros(6)
    y0 = w.x0.copy()
    # nothing hot enough, no diffusion: the composition only gets clipped
    rc, ierr, y, _ = oracle.mppnp_cycle(o, y0, s.t9, s.rho, s.xm, s.dq, s.rad, zero, 1e4,
                                        t9threshold=10.0)
    assert rc == 0 and not ierr.any()
    np.testing.assert_array_equal(y, np.maximum(y0, 1e-99))
    # every zone selected: each is integrate_network at constant (T9, rho)
    rc, ierr, y, st = oracle.mppnp_cycle(o, y0, s.t9, s.rho, s.xm, s.dq, s.rad, zero, 1e4)
    assert rc == 0 and not ierr.any()
    for j in range(6):
        yj = y0[j].copy()
        ye = np.zeros(1)
        ie, yj, _, stj = o.integrate_zone(yj, s.t9[j], s.t9[j], s.rho[j], s.rho[j], 1e4)
        assert ie == 0 and stj.nsubt == st[j].nsubt
        if abs(1.0 - np.sum(yj)) > 5e-7:
            yj = yj / oracle.lib().orc_ksum(yj.size, yj)
        np.testing.assert_array_equal(y[j], np.maximum(yj, 1e-99))
