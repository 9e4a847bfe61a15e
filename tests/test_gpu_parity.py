"""GPU parity: the HIP path (through the C ABI) against the CPU oracle and the reference SuperLU.

Bars: bit-exact for every integer result (patterns, column order, row pivots) and for every
floating-point result computed from equal inputs (dX/dt, Jacobian, LU solution, whole
integrations given the same rate vector); rates within the conditioning bound of the REACLIB
formula (device and host libm may differ in the last place of exp/log/cbrt); full integrations
with each side's own rates within 1e-10 relative on every X > 1e-20 with identical decisions.
"""
import json
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
EPS = 2.0 ** -53


@pytest.fixture(scope="module")
def hip(pkg):
    return pkg.hip


@pytest.fixture(scope="module")
def c0(pkg, oracle, hip):
    w = pkg.network.config(0, nz=40)
    ctx = hip.Context(w.net)
    o = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0))
    yield w, ctx, o
    ctx.close()


@pytest.fixture(scope="module")
def c1(pkg, oracle, hip):
    w = pkg.network.config(1, nz=8)
    ctx = hip.Context(w.net)
    o = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0))
    yield w, ctx, o
    ctx.close()


@pytest.fixture(scope="module")
def c3(pkg, oracle, hip):
    """The 5,000-species network (34,055 reactions, n = 4,999 reduced species)."""
    w = pkg.network.config(3, nz=4)
    ctx = hip.Context(w.net)
    o = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0))
    yield w, ctx, o
    ctx.close()


def _rate_bound(coef, t9):
    t = max(t9, 0.01)
    terms = np.abs(coef * np.array([1, 1 / t, t ** (-1 / 3), t ** (1 / 3), t, t ** (5 / 3),
                                    abs(math.log(t))])[:, None]).sum(0)
    return (12 * terms + 8) * EPS


def test_rates_parity(c1, hip, pkg):
    w, ctx, o = c1
    t9s = np.array([0.005, 0.05, 0.3, 1.2, 3.0])
    rhos = np.array([1e3, 10.0, 1e5, 3e7, 1e8])
    v = ctx.rates(t9s, rhos)
    for i in range(len(t9s)):
        vo = o.rates(t9s[i], rhos[i])
        rel = np.abs(v[i] - vo) / vo
        assert np.all(rel <= _rate_bound(w.net.coef, t9s[i]))
    g = np.load(os.path.join(GOLD, "rates_mpmath.npz"))
    nre = g["coef"].shape[1]
    net = pkg.network.Network("golden", 3, nre, np.array([0.0, 1.0, 2.0]), np.zeros(3),
                              np.zeros((8, nre), np.int32), g["rtype"], g["coef"],
                              np.zeros((3, 2), np.int32))
    net.k[0] = 1
    net.k[2] = 2
    net.k[4] = 0
    net.k[6] = 2
    net.k[1] = 1
    net.k[7] = 1
    cg = hip.Context(net)
    vg = cg.rates(g["t9"], g["rho"])
    for i in range(len(g["t9"])):
        exact = g["v"][i]
        ok = exact > 1e-99
        assert np.all(np.abs(vg[i][ok] - exact[ok]) / exact[ok] <= _rate_bound(g["coef"], g["t9"][i])[ok])
        assert np.all(vg[i][~ok] == 1e-99)
    cg.close()


@pytest.mark.parametrize("which", ["c0", "c1", "c3"])
def test_dxdt_and_jacobian_bitwise(which, request, oracle):
    w, ctx, o = request.getfixturevalue(which)
    rng = np.random.default_rng(11)
    cp_g, ri_g, r2s = ctx.reduced_pattern()
    for z in range(3):
        x = w.x0[z % w.x0.shape[0]].copy()
        x[1:] = np.maximum(x[1:], 10.0 ** rng.uniform(-30, -8, x.size - 1))
        x /= x.sum()
        x[0] = 0.0
        v = o.rates(w.t9[z, 0] * (1 + 0.3 * z), w.rho[z, 0])
        np.testing.assert_array_equal(ctx.dxdt(x, v), o.dxdt(x, v))
        for dt in (1e-2, 3e3, 1e8):
            rc, jv = ctx.jacobian_reduced(dt, x, v)
            n, rp, ci, va, mask = o.reduced_matrix(dt, x, v)
            cp, ri, cv = oracle.csr_to_csc(n, rp, ci, va)
            assert rc == 0 and n == ctx.n
            np.testing.assert_array_equal(cp, cp_g)
            np.testing.assert_array_equal(ri, ri_g)
            np.testing.assert_array_equal(jv, cv)


@pytest.mark.parametrize("which,nzones", [("c1", 6), ("c3", 2)])
def test_lu_pivots_and_solution(which, nzones, request, oracle):
    """Device GEPP on I - dt J of stiff and mild zones: row pivots equal the oracle's and the
    reference SuperLU's, nnz(L)/nnz(U) equal the oracle's, solution bit-equal to the oracle's.
    On the 5,000-species network COLAMD leaves ~2.6 M entries in L + U."""
    w, ctx, o = request.getfixturevalue(which)
    perm_c = ctx.column_order()
    ref = oracle.ref_available()
    for z in range(nzones):
        v = o.rates(w.t9[z, 0], w.rho[z, 0])
        n, rp, ci, va, mask = o.reduced_matrix(w.dt[z], w.x0[z], v)
        cp, ri, cv = oracle.csr_to_csc(n, rp, ci, va)
        b = np.random.default_rng(z).standard_normal((2, n))
        rc, xg, prg, nlu = ctx.lu_solve(cv, b)
        assert rc == 0
        info, pro, nl, nu, solve = oracle.gepp(n, cp, ri, cv, perm_c)
        np.testing.assert_array_equal(prg, pro)
        assert (nlu[0], nlu[1]) == (nl, nu)
        for r in range(2):
            np.testing.assert_array_equal(xg[r], solve(b[r]))
        if ref:
            rr = oracle.ref_factor(n, cp, ri, cv)
            np.testing.assert_array_equal(rr["perm_c"], perm_c)
            np.testing.assert_array_equal(prg, rr["perm_r"])


def test_column_order_override(c0, oracle):
    w, ctx, o = c0
    v = o.rates(0.3, 1e3)
    n, rp, ci, va, mask = o.reduced_matrix(1e5, w.x0[0], v)
    cp, ri, cv = oracle.csr_to_csc(n, rp, ci, va)
    saved = ctx.column_order()
    try:
        natural = np.arange(n, dtype=np.int32)
        ctx.set_column_order(natural)
        b = np.ones(n)
        rc, xg, prg, nlu = ctx.lu_solve(cv, b)
        info, pro, nl, nu, solve = oracle.gepp(n, cp, ri, cv, natural)
        np.testing.assert_array_equal(prg, pro)
        np.testing.assert_array_equal(xg, solve(b))
    finally:
        ctx.set_column_order(saved)


def test_integrate_trajectory_c0(c0):
    """ppn-style trajectory: each side with its own rates; identical decisions, <= 1e-10."""
    w, ctx, o = c0
    y_g = w.x0[0].copy()
    y_o = w.x0[0].copy()
    for s in range(40):
        args = (w.t9[s, 0], w.t9[s, 1], w.rho[s, 0], w.rho[s, 1], w.dt[s])
        ie_g, y_g, ye_g, st_g = ctx.integrate_network(y_g, *args)
        ie_o, y_o, ye_o, st_o = o.integrate_zone(y_o, *args)
        assert ie_g == ie_o == 0
        assert (st_g.nsubt, st_g.n_newton, st_g.nfail) == (st_o.nsubt, st_o.n_newton, st_o.nfail)
        m = y_o > 1e-20
        assert np.max(np.abs(y_g[m] - y_o[m]) / y_o[m]) <= 1e-10
        assert abs(ye_g - ye_o) <= 1e-12


def test_integrate_batch_bitwise_given_rates(c1, c0):
    """With the same mid-point rate vector, the device integration is bit-for-bit the oracle's:
    abundances, Ye, error codes, sub-step and Newton counts (1,100-species zones, one stiff)."""
    for w, ctx, o, sel in ((c1[0], c1[1], c1[2], [0, 2, 5, 6]), (c0[0], c0[1], c0[2], [0, 1, 2])):
        x0 = w.x0[sel] if w.x0.shape[0] > 1 else np.repeat(w.x0, len(sel), axis=0)
        t9_0, t9_1, r0, r1, dt = (a[sel] for a in (w.t9[:, 0], w.t9[:, 1], w.rho[:, 0],
                                                   w.rho[:, 1], w.dt))
        v = ctx.rates(0.5 * (t9_0 + t9_1), 0.5 * (r0 + r1))
        ie, y, ye, st = ctx.integrate_batch(x0, t9_0, t9_1, r0, r1, dt)
        for j in range(len(sel)):
            ie_o, y_o, ye_o, st_o = o.integrate_zone(x0[j], t9_0[j], t9_1[j], r0[j], r1[j], dt[j],
                                                     v_mid=v[j])
            assert ie[j] == ie_o
            assert (st["nsubt"][j], st["n_newton"][j], st["nfail"][j]) == \
                (st_o.nsubt, st_o.n_newton, st_o.nfail)
            np.testing.assert_array_equal(y[j], y_o)
            assert ye[j] == ye_o


def test_batch_equals_single_zone_calls(c0):
    w, ctx, o = c0
    sel = [3, 7, 11, 19]
    x0 = np.repeat(w.x0, len(sel), axis=0)
    args = [a[sel] for a in (w.t9[:, 0], w.t9[:, 1], w.rho[:, 0], w.rho[:, 1], w.dt)]
    ie, y, ye, st = ctx.integrate_batch(x0, *args)
    for j in range(len(sel)):
        ie1, y1, ye1, st1 = ctx.integrate_network(x0[j], *(a[j] for a in args))
        assert ie1 == ie[j]
        np.testing.assert_array_equal(y1, y[j])
        assert ye1 == ye[j]


def test_error_codes_match_oracle(pkg, oracle, hip):
    """Forced failures: max_sub_steps=2 with a too-large step gives MAX_SUBSTEPS on both sides;
    ittd=0 gives the same sequence of rejected steps."""
    w = pkg.network.config(1, nz=8)
    z = 5
    for ms, it in ((2, 6), (30, 0)):
        ctx = hip.Context(w.net, hip.default_knobs(max_sub_steps=ms, ittd=it))
        o = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0, max_sub_steps=ms, ittd=it))
        v = ctx.rates([w.t9[z, 0]], [w.rho[z, 0]])[0]
        args = (w.t9[z, 0], w.t9[z, 1], w.rho[z, 0], w.rho[z, 1], w.dt[z])
        ie_g, y_g, ye_g, st_g = ctx.integrate_network(w.x0[z], *args)
        ie_o, y_o, ye_o, st_o = o.integrate_zone(w.x0[z], *args, v_mid=v)
        assert ie_g == ie_o
        assert (st_g.nsubt, st_g.n_newton, st_g.nfail) == (st_o.nsubt, st_o.n_newton, st_o.nfail)
        np.testing.assert_array_equal(y_g, y_o)
        ctx.close()
    assert ie_g in (0, 2)


def test_slu_solve_dropin(hip, oracle):
    """slu_solve: matrixtest.f90's system and g20 (via CSR) against a dense solve and the
    reference; the index arrays come back shifted to 0-based, as the reference leaves them."""
    m = json.load(open(os.path.join(GOLD, "matrixtest.json")))
    n = m["n"]
    val = np.array(m["val"])
    ci1 = np.array(m["colind0"], np.int32) + 1
    rp1 = np.array(m["rowptr0"], np.int32) + 1
    rhs = np.array(m["rhs"])
    dense = np.zeros((n, n))
    for r in range(n):
        for p in range(rp1[r] - 1, rp1[r + 1] - 1):
            dense[r, ci1[p] - 1] = val[p]
    b = rhs.copy()
    ierr = hip.slu_solve(val.copy(), ci1, rp1, b)
    assert ierr == 0
    np.testing.assert_allclose(b, np.linalg.solve(dense, rhs), rtol=1e-13)
    np.testing.assert_array_equal(ci1, np.array(m["colind0"]))
    np.testing.assert_array_equal(rp1, np.array(m["rowptr0"]))
    # a structurally singular column: SuperLU's info = column index + 1
    val2 = val.copy()
    b2 = rhs.copy()
    ci2 = np.array(m["colind0"], np.int32) + 1
    rp2 = np.array(m["rowptr0"], np.int32) + 1
    val2[:] = 0.0
    assert hip.slu_solve(val2, ci2, rp2, b2) > 0


def test_slu_solve_reuses_its_buffers_across_sizes(hip, oracle, c1):
    """slu_solve keeps its device buffers from call to call and grows them for a larger system:
    g20 (via CSR), a 1,099-species network matrix, then g20 again and the network matrix again.
    Every solution is bit-equal to the oracle GEPP in SuperLU's column order, whatever buffers
    the call found."""
    g = np.load(os.path.join(GOLD, "g20.npz"))
    w, ctx, o = c1
    v = o.rates(w.t9[2, 0], w.rho[2, 0])
    n1, rp, ci, va, mask = o.reduced_matrix(w.dt[2], w.x0[2], v)
    cp1, ri1, cv1 = oracle.csr_to_csc(n1, rp, ci, va)
    cases = [(int(g["n"]), g["colptr"], g["rowind"], g["values"]), (n1, cp1, ri1, cv1)]
    for n, cp, ri, cv in cases + cases:
        cp = np.asarray(cp)
        ri = np.asarray(ri)
        cv = np.ascontiguousarray(cv, dtype=np.float64)
        rhs = np.random.default_rng(n).standard_normal(n)
        perm_c = hip.column_order(n, cp, ri)
        solve = oracle.gepp(n, cp, ri, cv, perm_c)[4]
        # the CSC matrix handed over as the CSR pattern of its transpose: isrow = 0 takes
        # (rowptr, colind) as the column pointers and row indices
        b = rhs.copy()
        ierr = hip.slu_solve(cv.copy(), (ri + 1).astype(np.int32), (cp + 1).astype(np.int32), b,
                             isrow=0)
        assert ierr == 0, (n, ierr)
        np.testing.assert_array_equal(b, solve(rhs), err_msg=f"n = {n}")


def test_c_fortran_dgssv_dropin(hip, oracle, c1):
    """iopt=1 factor / iopt=2 solve (two right-hand sides, ldb > n) / iopt=3 free on g20 and on
    a network matrix; solutions bit-equal to the oracle GEPP in SuperLU's column order."""
    g = np.load(os.path.join(GOLD, "g20.npz"))
    w, ctx, o = c1
    v = o.rates(w.t9[1, 0], w.rho[1, 0])
    n1, rp, ci, va, mask = o.reduced_matrix(w.dt[1], w.x0[1], v)
    cp1, ri1, cv1 = oracle.csr_to_csc(n1, rp, ci, va)
    for n, cp, ri, cv in ((int(g["n"]), g["colptr"], g["rowind"], g["values"]), (n1, cp1, ri1, cv1)):
        cpf = (np.asarray(cp) + 1).astype(np.int32)
        rif = (np.asarray(ri) + 1).astype(np.int32)
        vals = np.ascontiguousarray(cv, dtype=np.float64)
        ldb = n + 3
        b = np.zeros(2 * ldb)
        rng = np.random.default_rng(n)
        b[:n] = rng.standard_normal(n)
        b[ldb:ldb + n] = rng.standard_normal(n)
        b0 = b.copy()
        info, h = hip.c_fortran_dgssv(1, n, int(cp[n]), 2, vals, rif, cpf, b, ldb, 0)
        assert info == 0 and h != 0
        info, h2 = hip.c_fortran_dgssv(2, n, int(cp[n]), 2, vals, rif, cpf, b, ldb, h)
        assert info == 0
        perm_c = hip.column_order(n, cp, ri)
        info_o, pr, nl, nu, solve = oracle.gepp(n, cp, ri, cv, perm_c)
        np.testing.assert_array_equal(b[:n], solve(b0[:n]))
        np.testing.assert_array_equal(b[ldb:ldb + n], solve(b0[ldb:ldb + n]))
        assert np.all(b[n:ldb] == 0)
        info, h3 = hip.c_fortran_dgssv(3, n, int(cp[n]), 2, vals, rif, cpf, b, ldb, h)
        assert info == 0 and h3 == 0
