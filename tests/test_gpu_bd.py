"""GPU tests of the Bader-Deuflhard step (integration_method = 1, bader_deuflhard.F90:139-488).

The device runs one whole bader_deuflhard_integration per zone and controller pass: the
Jacobian once, then per level I - h J factored once and nseq(k)+1 stage solves, Neville
extrapolation and the order / step-size decisions.  The reference keeps kopt, frst, reduce and
the error terms between calls (module variables and SAVE'd locals); the context carries them the
same way, so a ppn-style sequence of one-zone calls follows the oracle's sequence.

Bars: the plain path (device GEPP, the reference's pivot order) with the oracle's rate vector
against the oracle (GEPP restatement): identical error codes, sub-step counts, stage counts,
accepted and rejected steps, and <= 1e-10 relative on every X > 1e-20 (the last-place
difference of pow() between device and host libm moves dtrcmd, not the decisions).  The batched
path (pivot-free block factors, FP64 refinement) against the plain path within the FP64
rounding spread of the solve, as the backward-Euler batched tests.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def hip(pkg):
    return pkg.hip


def _args(w, sel):
    return [a[sel] for a in (w.t9[:, 0], w.t9[:, 1], w.rho[:, 0], w.rho[:, 1], w.dt)]


def _maxrel(a, b):
    m = (a > 1e-20) | (b > 1e-20)
    if not m.any():
        return 0.0
    return float(np.max(np.abs(a[m] - b[m]) / np.maximum(a[m], b[m])))


def test_bd_trajectory_c0(pkg, oracle, hip):
    """ppn-style: 12 consecutive cycles of the 271-species trajectory, one zone per call, the BD
    state carried between calls on both sides; each call with the oracle's mid-point rates."""
    w = pkg.network.config(0)
    ctx = hip.Context(w.net, hip.default_knobs(integration_method=1))
    o = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0, integration_method=1))
    y_g = w.x0[0].copy()
    y_o = w.x0[0].copy()
    try:
        for s in range(12):
            args = (w.t9[s, 0], w.t9[s, 1], w.rho[s, 0], w.rho[s, 1], w.dt[s])
            v = ctx.rates(np.array([0.5 * (args[0] + args[1])]), np.array([0.5 * (args[2] + args[3])]))
            ie_g, y_g, ye_g, st_g = ctx.integrate_network(y_g, *args)
            ie_o, y_o, ye_o, st_o = o.integrate_zone(y_o, *args, v_mid=v[0])
            e = _maxrel(y_g, y_o)
            print(f"cycle {s}: ierr {ie_g}/{ie_o} nsubt {st_g.nsubt}/{st_o.nsubt} iter "
                  f"{st_g.iter}/{st_o.iter} succ/fail {st_g.nsucc},{st_g.nfail}/{st_o.nsucc},"
                  f"{st_o.nfail} factor {st_g.n_factor}/{st_o.n_factor} relerr {e:.2e}", flush=True)
            assert ie_g == ie_o == 0
            assert (st_g.nsubt, st_g.iter, st_g.nsucc, st_g.nfail) == \
                (st_o.nsubt, st_o.iter, st_o.nsucc, st_o.nfail)
            assert st_g.n_factor == st_o.n_factor and st_g.n_solve == st_o.n_solve
            assert e <= 1e-10
            # Ye of the last step's raw result, before the DEUFL renormalisation by the plain
            # sum (solver.F90:234, 272)
            print(f"   ye {ye_g:.15f} / {ye_o:.15f}")
            assert abs(ye_g - ye_o) <= 1e-12
            assert abs(y_g.sum() - 1.0) <= 1e-14
    finally:
        ctx.close()


def test_bd_stiff_zone_c4_golden(pkg, hip):
    """The very stiff regime of the 4,096-zone Bader-Deuflhard workload (T9 1..4, rho 1e5..1e9
    g/cc, dt 1e-3..1e2 s): zone 7 (T9 1.65, 43 sub-steps, 176 factorisations) from a fresh
    context, each side with its own mid-point rates, against the oracle's stored result
    (tests/golden/make_bd_stiff_golden.py; the oracle needs about 90 s for it): identical error
    code, sub-steps, stage and step counts, and <= 1e-10 relative on every X > 1e-20."""
    g = np.load(os.path.join(GOLD, "bd_stiff_c4_zone7.npz"))
    w = pkg.network.config(4, nz=int(g["nz"]))
    z = int(g["zone"])
    ctx = hip.Context(w.net, hip.default_knobs(integration_method=1))
    try:
        ie, y, ye, st = ctx.integrate_network(w.x0[z].copy(), w.t9[z, 0], w.t9[z, 1], w.rho[z, 0],
                                              w.rho[z, 1], w.dt[z])
        e = _maxrel(y, g["y"])
        print(f"zone {z}: ierr {ie}/{int(g['ierr'])} nsubt {st.nsubt}/{int(g['nsubt'])} iter "
              f"{st.iter}/{int(g['iter'])} succ/fail {st.nsucc},{st.nfail}/{int(g['nsucc'])},"
              f"{int(g['nfail'])} factor {st.n_factor}/{int(g['n_factor'])} relerr {e:.2e}",
              flush=True)
        assert ie == int(g["ierr"]) == 0
        assert (st.nsubt, st.iter, st.nsucc, st.nfail, st.n_factor, st.n_solve) == \
            tuple(int(g[k]) for k in ("nsubt", "iter", "nsucc", "nfail", "n_factor", "n_solve"))
        assert e <= 1e-10
        print(f"ye {ye:.15f} / {float(g['ye']):.15f}")
        assert abs(ye - float(g["ye"])) <= 1e-12
        assert abs(y.sum() - 1.0) <= 1e-14
    finally:
        ctx.close()


def test_bd_batched_path_unsupported(pkg, hip):
    """The batched path refuses integration_method 1 (switched off) with NUPPN_ERR_UNSUPPORTED and
    leaves the abundances untouched."""
    w = pkg.network.config(1, nz=8)
    sel = np.array([0])
    ctx = hip.Context(w.net, hip.default_knobs(integration_method=1))
    try:
        with pytest.raises(RuntimeError, match=str(hip.ERR_UNSUPPORTED)):
            ctx.integrate_batch_mixed(w.x0[sel], *_args(w, sel))
    finally:
        ctx.close()


def test_bd_batch_matches_oracle_c1(pkg, oracle, hip):
    """1,100-species zones (T9 0.08 .. 0.43) in one batch call, every zone from the entry BD
    state; the oracle per zone from the same state (reset before each zone).  (Zones with more
    BD steps take the CPU oracle minutes.)"""
    w = pkg.network.config(1, nz=8)
    sel = np.array([0, 2, 5])
    args = _args(w, sel)
    ctx = hip.Context(w.net, hip.default_knobs(integration_method=1))
    o = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0, integration_method=1))
    try:
        v = ctx.rates(0.5 * (args[0] + args[1]), 0.5 * (args[2] + args[3]))
        ie, y, ye, st = ctx.integrate_batch(w.x0[sel], *args)
        for j in range(len(sel)):
            o.reset_state()
            ie_o, y_o, ye_o, st_o = o.integrate_zone(w.x0[sel[j]], *(a[j] for a in args),
                                                     v_mid=v[j])
            e = _maxrel(y[j], y_o)
            print(f"zone {sel[j]} T9 {args[0][j]:.3f}: ierr {ie[j]}/{ie_o} nsubt "
                  f"{st['nsubt'][j]}/{st_o.nsubt} iter {st['iter'][j]}/{st_o.iter} fail "
                  f"{st['nfail'][j]}/{st_o.nfail} relerr {e:.2e}", flush=True)
            assert ie[j] == ie_o
            assert (st["nsubt"][j], st["iter"][j], st["nsucc"][j], st["nfail"][j]) == \
                (st_o.nsubt, st_o.iter, st_o.nsucc, st_o.nfail)
            assert e <= 1e-10
            assert abs(ye[j] - ye_o) <= 1e-12, (ye[j], ye_o)
    finally:
        ctx.close()


@pytest.mark.skip(reason="batched-path Bader-Deuflhard switched off: its kernel took a GPU memory "
                         "aperture violation after the streamed segmented sums, cause not found; "
                         "nuppn_integrate_batch_mixed returns NUPPN_ERR_UNSUPPORTED for "
                         "integration_method 1")
@pytest.mark.parametrize("prec", [32, 64])
def test_bd_batched_matches_plain(pkg, hip, prec):
    """The batched path's BD (block factors in FP32 or FP64, every stage solve refined to
    FP64 backward stability) against the plain path's BD, each from a fresh context."""
    w = pkg.network.config(1, nz=8)
    sel = np.array([0, 2, 5, 6])
    args = _args(w, sel)
    kn = hip.default_knobs(integration_method=1)
    cp = hip.Context(w.net, kn)
    cb = hip.Context(w.net, kn)
    try:
        ie_p, y_p, ye_p, st_p = cp.integrate_batch(w.x0[sel], *args)
        ie_b, y_b, ye_b, st_b = cb.integrate_batch_mixed(w.x0[sel], *args,
                                                         opts=hip.batch_opts(factor_precision=prec))
        for j in range(len(sel)):
            e = _maxrel(y_b[j], y_p[j])
            print(f"zone {sel[j]}: ierr {ie_p[j]}/{ie_b[j]} nsubt {st_p['nsubt'][j]}/"
                  f"{st_b['nsubt'][j]} iter {st_p['iter'][j]}/{st_b['iter'][j]} factor "
                  f"{st_b['n_factor'][j]} fp64 {st_b['n_factor_fp64'][j]} solves "
                  f"{st_b['n_solve'][j]} relerr {e:.2e}", flush=True)
            assert ie_b[j] == ie_p[j] == 0
            assert (st_b["nsubt"][j], st_b["iter"][j], st_b["nfail"][j]) == \
                (st_p["nsubt"][j], st_p["iter"][j], st_p["nfail"][j])
            assert abs(y_b[j].sum() - 1.0) <= 1e-14
            assert e <= 1e-8
    finally:
        cp.close()
        cb.close()
