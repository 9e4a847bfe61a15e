"""Pins of the oracle's dynamic network (oracle/nuppn_oracle.c: orc_jacobian's mask and row /
column assignment, orc_reduce_jacobian's packing, newton_iteration's ddecompac; jac_rhs.F90:
308-436,134-197, backward_euler.F90:120-176) on a network small enough to work by hand:

    gamma, p, B (A 10), C (A 11), D (A 12)
    R_BC: B + p -> C + gamma     R_CD: C + p -> D + gamma
    R_DC: D -> C + p (photo)     R_CB: C -> B + p (photo)

* loop order: with no B, R_BC carries no flux; placed first, the reactions that mark its species
  come after it and its entries stay out of the reduced matrix, placed last they are in (the
  entries worked out by hand from the statement list ice1..ice8 of 365-392);
* the threshold: |dfdN| = fact v Y1^k2 Y3 > cyminformal / dt marks a reaction's species, each
  flux worked out by hand; a species no reaction marks leaves the solve and keeps its abundance
  exactly, although its dX/dt is not zero;
* min_rate_ov_dt: a reaction with |v| / dt at or below it adds no Jacobian term (the entry's
  value is the 1e-99 seed plus the other reaction's term, by hand) and marks nothing;
* with every reaction carrying flux (cyminformal 0) the dynamic network is the static one: the
  same reduced matrix and the same integration, bit for bit.
"""
import math

import numpy as np
import pytest

G, P, B, CC, D = 0, 1, 2, 3, 4
AW = [0.0, 1.0, 10.0, 11.0, 12.0]
RATES = {"BC": 1.0e-2, "CD": 2.0e-2, "DC": 3.0e-3, "CB": 4.0e-4}


def _net(pkg, order):
    net = pkg.network
    topo = {
        "BC": ((B, 1, P, 1, G, 0, CC, 1), net.RTYPE_CAPTURE),
        "CD": ((CC, 1, P, 1, G, 0, D, 1), net.RTYPE_CAPTURE),
        "DC": ((D, 1, G, 0, P, 1, CC, 1), net.RTYPE_PHOTO),
        "CB": ((CC, 1, G, 0, P, 1, B, 1), net.RTYPE_PHOTO),
    }
    nre = len(order)
    k = np.zeros((8, nre), np.int32)
    rtype = np.zeros(nre, np.int32)
    coef = np.zeros((7, nre))
    for l, name in enumerate(order):
        k[:, l] = topo[name][0]
        rtype[l] = topo[name][1]
        coef[0, l] = math.log(RATES[name])
    return net.Network("dyn", 5, nre, np.array(AW), np.array([0.0, 1, 5, 5, 6]), k, rtype, coef,
                       np.array([[0, 0], [1, 0], [5, 5], [5, 6], [6, 6]], np.int32))


def _oracle(oracle, net, **kn):
    return oracle.Oracle(net, oracle.default_knobs(**kn))


def _entries(o, dt, x, v):
    """(row species, col species) of the reduced matrix, and the masked species"""
    n, rp, ci, va, mask = o.reduced_matrix(dt, x, v)
    sp = np.where(mask)[0]
    ent = {(int(sp[r]), int(sp[ci[e]])) for r in range(n) for e in range(rp[r], rp[r + 1])}
    return ent, set(sp.tolist())


X0 = np.array([0.0, 0.3, 0.0, 0.5, 0.2])
DIAG = {(s, s) for s in (P, B, CC, D)}
# off-diagonal entries of each reaction, (row, col), from ice2..ice8 with gamma's slots left out
E_CD = {(P, CC), (CC, P), (D, CC), (D, P)}
E_DC = {(P, D), (CC, D)}
E_CB = {(P, CC), (B, CC)}
E_BC = {(P, B), (B, P), (CC, B), (CC, P)}


def test_loop_order_decides_the_entries(oracle, pkg):
    for order, want in ((("BC", "CD", "DC", "CB"), E_CD | E_DC | E_CB),
                        (("CD", "DC", "CB", "BC"), E_CD | E_DC | E_CB | E_BC)):
        net = _net(pkg, order)
        o = _oracle(oracle, net, cyminformal=1.0e-30)
        v = o.rates(0.3, 1.0)
        ent, masked = _entries(o, 1.0, X0, v)
        # R_BC has no flux (no B); the other three mark p, B, C, D between them
        assert masked == {P, B, CC, D}
        assert ent == DIAG | want, order
    # the reaction left out contributes three entries no other reaction has
    assert len(E_BC - (E_CD | E_DC | E_CB)) == 3


def test_threshold_marks_by_flux(oracle, pkg):
    order = ("BC", "CD", "DC", "CB")
    net = _net(pkg, order)
    x = X0.copy()
    x[B] = 1.0e-3
    x[CC] -= 1.0e-3
    y = x / np.where(np.array(AW) > 0, AW, 1.0)
    dfdn = {"BC": RATES["BC"] * y[B] * y[P], "CD": RATES["CD"] * y[CC] * y[P],
            "DC": RATES["DC"] * y[D], "CB": RATES["CB"] * y[CC]}
    # 3.0e-7, 2.7e-4, 5.0e-5, 1.8e-5: a threshold between C -> B + p and D -> C + p
    thr = 3.0e-5
    assert dfdn["CB"] < thr < dfdn["DC"] and dfdn["BC"] < thr
    # (short enough that the mass C -> B + p moves into the frozen B, dt v Y_C A_B ~ 2e-7,
    # stays below SUM_TOL)
    dt = 1.0e-3
    o = _oracle(oracle, net, cyminformal=thr * dt)
    v = o.rates(0.3, 1.0)
    ent, masked = _entries(o, dt, x, v)
    assert masked == {P, CC, D}
    assert ent == {(P, P), (CC, CC), (D, D)} | E_CD | E_DC
    # B is out of the solve: one backward-Euler step leaves it exactly where it was, although
    # C -> B + p feeds it
    assert o.dxdt(x, v)[B] > 0.0
    ierr, ynew, dtr, st = o.backward_euler(x, dt, v)
    assert ierr == 0 and st.nvar1 == 3
    assert ynew[B] == x[B]
    assert np.all(ynew[[P, CC, D]] != x[[P, CC, D]])
    # below the smallest flux every species is in
    o2 = _oracle(oracle, net, cyminformal=0.5 * dfdn["BC"] * dt)
    assert _entries(o2, dt, x, v)[1] == {P, B, CC, D}


def test_min_rate_skips_a_reaction(oracle, pkg):
    order = ("BC", "CD", "DC", "CB")
    net = _net(pkg, order)
    dt = 2.0
    # |v| / dt = 2e-4 for C -> B + p: at the bound it is skipped, just below kept
    v = _oracle(oracle, net).rates(0.3, 1.0)
    o = _oracle(oracle, net, cyminformal=1.0e-30, min_rate_ov_dt=v[3] / dt)
    rc, jac, row, col, mask = o.jacobian(dt, X0, v)
    assert rc == 0
    assert not mask[B]            # only C -> B + p could mark B (no B for R_BC)
    rp, ci, _ = o.full_pattern()
    q = rp[P] + int(np.where(ci[rp[P]:rp[P + 1]] == CC)[0][0])   # the (p, C) entry
    y = X0 / np.where(np.array(AW) > 0, AW, 1.0)
    # d(dY_p/dt)/dY_C from C + p -> D alone: -v Y_p, times A_p / A_C; the seed added first
    assert jac[q] == 1.0e-99 + (-(v[1] * y[P])) * 1.0 / 11.0
    o2 = _oracle(oracle, net, cyminformal=1.0e-30, min_rate_ov_dt=np.nextafter(v[3] / dt, 0.0))
    rc, jac2, _, _, mask2 = o2.jacobian(dt, X0, v)
    assert mask2[B]
    assert jac2[q] == jac[q] + (v[3] * 1.0) * 1.0 / 11.0


def test_all_marking_reactions_give_the_static_network(oracle, pkg):
    order = ("BC", "CD", "DC", "CB")
    net = _net(pkg, order)
    x = np.array([0.0, 0.3, 0.1, 0.4, 0.2])
    dyn = _oracle(oracle, net, cyminformal=0.0)
    sta = _oracle(oracle, net, cyminformal=-1.0)
    v = dyn.rates(0.3, 1.0)
    for dt in (1.0, 1.0e3):
        a, b = dyn.reduced_matrix(dt, x, v), sta.reduced_matrix(dt, x, v)
        assert a[0] == b[0] == 4
        for u, w in zip(a[1:4], b[1:4]):
            np.testing.assert_array_equal(u, w)
    r1 = dyn.integrate_zone(x.copy(), 0.3, 0.3, 1.0, 1.0, 1.0e4)
    r2 = sta.integrate_zone(x.copy(), 0.3, 0.3, 1.0, 1.0, 1.0e4)
    assert r1[0] == r2[0] == 0
    np.testing.assert_array_equal(r1[1], r2[1])
    assert r1[3].n_newton == r2[3].n_newton
