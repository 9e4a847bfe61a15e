"""GPU tests of the batched path (nuppn_integrate_batch_mixed): all zones of a cycle per call,
pivot-free block factorisation in FP32 or FP64, refinement against the FP64 matrix with an
extra-precise residual, dX/dt and the Newton rhs summed in twice the FP64 precision.

Bars (no bar depends on a measured spread):
* FP32 factors ("mixed") against FP64 factors: the refinement converges every new abundance below
  half an ulp of the FP64 system's own solution, so both give the same bits, the same error codes
  and the same step decisions on every zone, cold or hot;
* |sum X - 1| <= 1e-12 on every zone (the doubled-precision dX/dt conserves mass to the rounding
  of each species' net rate);
* against the exact-arithmetic oracle (the reference algorithm with J, dX/dt and the Newton rhs
  rounded once from their real sums and every Newton system solved exactly, ORC_LS_EXACT) -
  which is what the batched path computes: identical step decisions and <= 1e-8 relative on
  every X > 1e-20, cold zones and hot ones.  Against the reference's own FP64 arithmetic the bar
  holds on the cold zones only: over a hot-zone cycle the reference algorithm does not repeat
  under a change of rounding (its SuperLU and the oracle's GEPP, same pivots, part by up to 28 %
  there: tests/test_oracle_pins.py).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip(pkg):
    return pkg.hip


@pytest.fixture(scope="module")
def c1(pkg, hip):
    w = pkg.network.config(1, nz=8)
    ctx = hip.Context(w.net)
    yield w, ctx
    ctx.close()


@pytest.fixture(scope="module")
def c1_full(pkg, hip):
    w = pkg.network.config(1)
    ctx = hip.Context(w.net)
    yield w, ctx
    ctx.close()


def _args(w, sel):
    return [a[sel] for a in (w.t9[:, 0], w.t9[:, 1], w.rho[:, 0], w.rho[:, 1], w.dt)]


def _maxrel(a, b):
    m = (a > 1e-20) | (b > 1e-20)
    if not m.any():
        return 0.0
    return float(np.max(np.abs(a[m] - b[m]) / np.maximum(a[m], b[m])))


def test_block_structure(c1):
    """The 1,100-species capture network: n, p, alpha form the border; the interior half-bandwidth
    (alpha captures, (Z,N) -> (Z+2,N+2)) fits 32-species blocks."""
    w, ctx = c1
    info = ctx.block_info()
    print(info)
    assert info["available"]
    assert info["NB"] == 32 and info["nbd"] == 3
    assert info["bandwidth"] <= 32
    assert info["ni"] + info["nbd"] == ctx.n


# cold and hot zones of the 2,048-zone workload, T9 0.05 .. 2.1 (the hot ones take 100+ sub-steps)
MIXED_ZONES = [0, 2, 5, 6, 329, 439, 554, 1498, 1090, 1735]


def test_mixed_equals_fp64_factors(c1_full, hip):
    """FP32 factors + refinement and FP64 factors + refinement: the same bits, codes, sub-steps
    and Newton iterations on every zone, both precisions exercised, mass conserved to 1e-12."""
    w, ctx = c1_full
    sel = np.array(MIXED_ZONES)
    args = _args(w, sel)
    out = {}
    for prec in (32, 64):
        out[prec] = ctx.integrate_batch_mixed(w.x0[sel], *args,
                                              opts=hip.batch_opts(factor_precision=prec))
    ie32, y32, ye32, st32 = out[32]
    ie64, y64, ye64, st64 = out[64]
    for j, z in enumerate(sel):
        print(f"zone {z} T9 {w.t9[z, 0]:.3f}: ierr {ie32[j]}/{ie64[j]} nsubt {st32['nsubt'][j]}/"
              f"{st64['nsubt'][j]} newton {st32['n_newton'][j]} fp64 factors "
              f"{st32['n_factor_fp64'][j]}/{st32['n_factor'][j]} solves {st32['n_solve'][j]}/"
              f"{st64['n_solve'][j]} relerr {_maxrel(y32[j], y64[j]):.1e} sum-1 "
              f"{y32[j].sum() - 1:.1e}", flush=True)
    np.testing.assert_array_equal(ie32, 0)
    np.testing.assert_array_equal(ie64, 0)
    for f in ("nsubt", "n_newton", "nsucc", "nfail", "titers"):
        np.testing.assert_array_equal(st32[f], st64[f])
    np.testing.assert_array_equal(y32, y64)
    np.testing.assert_array_equal(ye32, ye64)
    assert np.all(np.abs(y32.sum(axis=1) - 1.0) <= 1e-12)
    # some zones stay on FP32 factors throughout, some fall back to FP64
    assert np.any(st32["n_factor_fp64"] == 0) and np.any(st32["n_factor_fp64"] > 0)


def test_team_kernel_equals_single_wavefront(c1_full, hip):
    """The multi-wavefront kernels (passes whose work list is short) and the one-wavefront
    kernels give the same bits: the cycle of these cold and hot zones once with every pass on one
    wavefront per zone (team_zones 0), then with every pass on four and on two (a bound above
    the list), the same abundances, Ye, codes, sub-steps and Newton iterations.  The segmented
    sums cut their terms into the same pieces for all; the factorisations differ (block Thomas on
    one wavefront, the twisted two-ended form on two or four), and so do the factors' rounding
    errors and the refinement steps they need - with FP32 factors the twisted form contracts
    better and keeps some zones off FP64 factors longer - but the refinement to below half an ulp
    of the FP64 system's solution gives every Newton iterate the same bits."""
    import time
    w, ctx = c1_full
    sel = np.array(MIXED_ZONES)
    args = _args(w, sel)
    out = {}
    for tz, tw in ((0, 0), (1 << 20, 4), (1 << 20, 2)):
        t = time.perf_counter()
        out[tw] = ctx.integrate_batch_mixed(w.x0[sel], *args,
                                            opts=hip.batch_opts(team_zones=tz, team_waves=tw))
        print(f"team_zones {tz} team_waves {tw}: {time.perf_counter() - t:.2f} s", flush=True)
    ie1, y1, ye1, st1 = out[0]
    np.testing.assert_array_equal(ie1, 0)
    for tw in (4, 2):
        ie4, y4, ye4, st4 = out[tw]
        np.testing.assert_array_equal(ie4, ie1)
        for f in ("n_factor", "n_factor_fp64", "n_solve", "n_refine"):
            print(tw, f, st1[f], st4[f])
        for f in ("nsubt", "n_newton", "nsucc", "nfail", "titers"):
            np.testing.assert_array_equal(st4[f], st1[f], err_msg=f)
        np.testing.assert_array_equal(y4, y1)
        np.testing.assert_array_equal(ye4, ye1)


def test_queued_and_graph_passes_equal_one_pass_at_a_time(c1_full, hip):
    """Passes queued ahead of the host (the counts read on the device) and chunks of 16 passes
    replayed from a captured hipGraph give the results of one pass per host round trip, bit for
    bit; the graph path runs twice in a row (replays of one call's graphs after the previous
    call's), with the program's own synchronising work in between."""
    w, ctx = c1_full
    sel = np.array(MIXED_ZONES)
    args = _args(w, sel)
    ref = ctx.integrate_batch_mixed(w.x0[sel], *args,
                                    opts=hip.batch_opts(queue_passes=1, use_graph=0))
    t_ref = ctx.last_timing()
    assert t_ref["graph_launches"] == 0
    runs = [("queued", hip.batch_opts(use_graph=0)), ("graph", hip.batch_opts()),
            ("graph again", hip.batch_opts())]
    for name, o in runs:
        ie, y, ye, st = ctx.integrate_batch_mixed(w.x0[sel], *args, opts=o)
        t = ctx.last_timing()
        print(f"{name}: passes {t['passes']:.0f} newton launches {t['newton_launches']:.0f} graph "
              f"launches {t['graph_launches']:.0f} total {t['total_ms']:.1f} ms "
              f"(one pass at a time: {t_ref['total_ms']:.1f} ms)", flush=True)
        np.testing.assert_array_equal(ie, ref[0])
        np.testing.assert_array_equal(y, ref[1])
        np.testing.assert_array_equal(ye, ref[2])
        for f in ("nsubt", "n_newton", "nsucc", "nfail", "titers", "n_factor", "n_solve"):
            np.testing.assert_array_equal(st[f], ref[3][f], err_msg=f)
        assert t["passes"] == t_ref["passes"]
        assert t["newton_launches"] == t_ref["newton_launches"]
        if name.startswith("graph"):
            assert t["graph_launches"] > 0
        # the program's own eager work between the calls
        _ = ctx.rates(args[0], args[2])


def _exact_oracle(oracle, net):
    return oracle.Oracle(net, oracle.default_knobs(cyminformal=-1.0), oracle.LS_EXACT)


@pytest.mark.parametrize("prec", [32, 64])
def test_batched_matches_exact_oracle_cold(c1, hip, oracle, prec):
    """Zones with T9 0.06 .. 0.43: the same error codes, sub-steps and Newton iterations as the
    oracle with exact linear solves, <= 1e-8 relative on every X > 1e-20, |sum X - 1| <= 1e-12;
    and the plain path (the reference's GEPP) on the same zones within the same bar."""
    w, ctx = c1
    sel = np.array([0, 2, 5, 6])
    args = _args(w, sel)
    ie_b, y_b, ye_b, st_b = ctx.integrate_batch_mixed(w.x0[sel], *args,
                                                      opts=hip.batch_opts(factor_precision=prec))
    ie_p, y_p, ye_p, st_p = ctx.integrate_batch(w.x0[sel], *args)
    o = _exact_oracle(oracle, w.net)
    v = ctx.rates(0.5 * (args[0] + args[1]), 0.5 * (args[2] + args[3]))
    for j in range(len(sel)):
        ie_o, y_o, ye_o, st_o = o.integrate_zone(w.x0[sel[j]], *(a[j] for a in args), v_mid=v[j])
        e, ep = _maxrel(y_b[j], y_o), _maxrel(y_p[j], y_o)
        print(f"zone {sel[j]} T9 {args[0][j]:.3f}: ierr {ie_b[j]}/{ie_o} nsubt {st_b['nsubt'][j]}/"
              f"{st_o.nsubt} newton {st_b['n_newton'][j]}/{st_o.n_newton} fp64 "
              f"{st_b['n_factor_fp64'][j]}/{st_b['n_factor'][j]} relerr {e:.2e} (plain {ep:.2e}) "
              f"sum-1 {y_b[j].sum() - 1:.1e}", flush=True)
        assert ie_b[j] == ie_o == ie_p[j] == 0
        assert (st_b["nsubt"][j], st_b["n_newton"][j]) == (st_o.nsubt, st_o.n_newton)
        assert (st_p["nsubt"][j], st_p["n_newton"][j]) == (st_o.nsubt, st_o.n_newton)
        assert e <= 1e-8 and ep <= 1e-8
        assert abs(y_b[j].sum() - 1.0) <= 1e-12
        assert abs(ye_b[j] - ye_o) <= 1e-12


def test_hot_zones_first_substeps_match_exact_oracle(c1_full, hip, oracle):
    """Hot zones (T9 0.8 .. 2.1) over their first sub-steps (max_sub_steps = 6: the call ends
    with MAX_SUBSTEPS): the same accepted and failed steps as the oracle with exact linear
    solves and <= 1e-8 relative on every X > 1e-20."""
    w, _ = c1_full
    sel = np.array([439, 1498, 1090, 1735])
    args = _args(w, sel)
    kn = hip.default_knobs(max_sub_steps=6)
    ctx = hip.Context(w.net, kn)
    o = oracle.Oracle(w.net, oracle.default_knobs(cyminformal=-1.0, max_sub_steps=6),
                      oracle.LS_EXACT)
    try:
        ie_b, y_b, ye_b, st_b = ctx.integrate_batch_mixed(w.x0[sel], *args)
        v = ctx.rates(0.5 * (args[0] + args[1]), 0.5 * (args[2] + args[3]))
        for j in range(len(sel)):
            # the device's rate vector: the last-place differences of exp / pow between the
            # device and host libm would otherwise blur the comparison of the integrators
            ie_o, y_o, ye_o, st_o = o.integrate_zone(w.x0[sel[j]], *(a[j] for a in args),
                                                     v_mid=v[j])
            e = _maxrel(y_b[j], y_o)
            print(f"zone {sel[j]} T9 {args[0][j]:.3f}: ierr {ie_b[j]}/{ie_o} succ/fail "
                  f"{st_b['nsucc'][j]},{st_b['nfail'][j]}/{st_o.nsucc},{st_o.nfail} newton "
                  f"{st_b['n_newton'][j]}/{st_o.n_newton} relerr {e:.2e}", flush=True)
            assert ie_b[j] == ie_o
            assert (st_b["nsucc"][j], st_b["nfail"][j], st_b["n_newton"][j]) == \
                (st_o.nsucc, st_o.nfail, st_o.n_newton)
            assert e <= 1e-8
    finally:
        ctx.close()


def test_hot_zones_full_cycle_match_exact_oracle(c1_full, hip, oracle):
    """Whole cycles of hot zones (T9 0.87 and 2.11; 21 and about 110 sub-steps): the batched
    path assembles J, dX/dt and the rhs as once-rounded values of their real sums and solves
    every Newton system exactly, which is what the exact-solve oracle computes, so it takes the
    oracle's decisions and ends within 1e-8 of it on every X > 1e-20."""
    w, ctx = c1_full
    sel = np.array([1498, 1090])
    args = _args(w, sel)
    ie_b, y_b, ye_b, st_b = ctx.integrate_batch_mixed(w.x0[sel], *args)
    v = ctx.rates(0.5 * (args[0] + args[1]), 0.5 * (args[2] + args[3]))
    o = _exact_oracle(oracle, w.net)
    for j in range(len(sel)):
        ie_o, y_o, ye_o, st_o = o.integrate_zone(w.x0[sel[j]], *(a[j] for a in args), v_mid=v[j])
        e = _maxrel(y_b[j], y_o)
        print(f"zone {sel[j]} T9 {args[0][j]:.3f}: ierr {ie_b[j]}/{ie_o} nsubt {st_b['nsubt'][j]}/"
              f"{st_o.nsubt} newton {st_b['n_newton'][j]}/{st_o.n_newton} relerr {e:.2e} sum-1 "
              f"{y_b[j].sum() - 1:.1e}/{y_o.sum() - 1:.1e}", flush=True)
        assert ie_b[j] == ie_o == 0
        assert (st_b["nsubt"][j], st_b["n_newton"][j]) == (st_o.nsubt, st_o.n_newton)
        assert e <= 1e-8
        assert abs(ye_b[j] - ye_o) <= 1e-12


def test_device_pointer_entry_matches_host_entry(c1, hip):
    """nuppn_integrate_batch_mixed_dev on device arrays gives the host entry's results bit for
    bit (the same kernels on the same data)."""
    import torch
    w, ctx = c1
    sel = np.array([0, 2, 6])
    args = _args(w, sel)
    ie1, y1, ye1, st1 = ctx.integrate_batch_mixed(w.x0[sel], *args)
    dev = torch.device("cuda:0")
    t = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    nz = len(sel)
    yps = t(w.x0[sel])
    tt = [t(a) for a in args]
    ye = torch.full((nz,), 0.5, dtype=torch.float64, device=dev)
    ierr = torch.zeros(nz, dtype=torch.int32, device=dev)
    stats = torch.zeros((nz, hip.STATS_DTYPE.itemsize // 8), dtype=torch.int64, device=dev)
    ctx.integrate_batch_mixed_dev(nz, yps.data_ptr(), tt[0].data_ptr(), tt[1].data_ptr(),
                                  tt[2].data_ptr(), tt[3].data_ptr(), ye.data_ptr(),
                                  tt[4].data_ptr(), ierr.data_ptr(), stats.data_ptr())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(yps.cpu().numpy(), y1)
    np.testing.assert_array_equal(ierr.cpu().numpy(), ie1)


def test_single_zone_equals_batch(c1):
    """A zone's result does not depend on the other zones of the batch or on its slot."""
    w, ctx = c1
    sel = np.array([0, 2, 5, 6])
    args = _args(w, sel)
    ie, y, ye, st = ctx.integrate_batch_mixed(w.x0[sel], *args)
    for j in (0, 3):
        ie1, y1, ye1, st1 = ctx.integrate_batch_mixed(w.x0[sel][j:j + 1], *(a[j:j + 1] for a in args))
        np.testing.assert_array_equal(y1[0], y[j])
        assert ie1[0] == ie[j]


def test_uneven_batch_c2(pkg, hip, oracle):
    """The uneven workload (1,100 species, T9 0.02 .. 4, dt 1 .. 1e8 s): 16 zones in one batched
    call with per-zone sub-step counts from 1 to hundreds.  Every zone converges with finite
    X >= 1e-99 and |sum X - 1| <= 1e-12; FP32 factors give the FP64 factors' bits; the zones with
    dt < 100 s take the exact-solve oracle's decisions and match it within 1e-8."""
    w = pkg.network.config(2, nz=16)
    ctx = hip.Context(w.net)
    try:
        allz = np.arange(16)
        ie_b, y_b, ye_b, st_b = ctx.integrate_batch_mixed(w.x0, *_args(w, allz))
        ie_6, y_6, ye_6, st_6 = ctx.integrate_batch_mixed(
            w.x0, *_args(w, allz), opts=hip.batch_opts(factor_precision=64))
        for j in allz:
            print(f"zone {j} T9 {w.t9[j, 0]:.3f} dt {w.dt[j]:.2e}: ierr {ie_b[j]} nsubt "
                  f"{st_b['nsubt'][j]} newton {st_b['n_newton'][j]} fp64 "
                  f"{st_b['n_factor_fp64'][j]}/{st_b['n_factor'][j]} sum-1 {y_b[j].sum() - 1:.1e}")
        np.testing.assert_array_equal(ie_b, 0)
        assert np.all(np.isfinite(y_b)) and np.all(y_b >= 1e-99)
        assert np.all(np.abs(y_b.sum(axis=1) - 1.0) <= 1e-12)
        np.testing.assert_array_equal(y_b, y_6)
        sel = np.where(w.dt < 100.0)[0]
        assert len(sel) >= 3
        o = _exact_oracle(oracle, w.net)
        v = ctx.rates(0.5 * (w.t9[sel, 0] + w.t9[sel, 1]), 0.5 * (w.rho[sel, 0] + w.rho[sel, 1]))
        for i, j in enumerate(sel):
            ie_o, y_o, ye_o, st_o = o.integrate_zone(w.x0[j], *(a[j] for a in _args(w, allz)),
                                                     v_mid=v[i])
            e = _maxrel(y_b[j], y_o)
            print(f"zone {j}: oracle nsubt {st_o.nsubt} newton {st_o.n_newton} relerr {e:.2e}")
            assert ie_o == 0
            assert (st_b["nsubt"][j], st_b["n_newton"][j]) == (st_o.nsubt, st_o.n_newton)
            assert e <= 1e-8
    finally:
        ctx.close()


def test_uneven_workload_full_size(pkg, hip):
    """All 8,192 zones of the uneven workload, one cycle: every zone ends with error code 0,
    finite abundances >= 1e-99 and |sum X - 1| <= 1e-12."""
    w = pkg.network.config(2)
    ctx = hip.Context(w.net)
    try:
        allz = np.arange(w.nz)
        ie, y, ye, st = ctx.integrate_batch_mixed(w.x0, *_args(w, allz))
        s = np.abs(y.sum(axis=1) - 1.0)
        print(f"{w.nz} zones: ierr != 0: {int((ie != 0).sum())}, max sub-steps "
              f"{int(st['nsubt'].max())}, max Newton {int(st['n_newton'].max())}, zones on FP64 "
              f"factors {int((st['n_factor_fp64'] > 0).sum())}, max |sum X - 1| {s.max():.1e}",
              flush=True)
        np.testing.assert_array_equal(ie, 0)
        assert np.all(np.isfinite(y)) and np.all(y >= 1e-99)
        assert np.all(s <= 1e-12)
    finally:
        ctx.close()


def test_autotune_picks_a_shape_and_caches_it(pkg, hip, tmp_path):
    """nuppn_autotune times one cycle of a sample under block sizes 32 and 64 and the team-kernel
    settings, keeps the fastest (every candidate gives the same bits), rebuilds the context with
    it and uses it for calls without options; a second context reads the choice from the cache
    file without timing and integrates to the same bits."""
    w = pkg.network.config(1)
    sel = np.array([0, 2, 5, 6, 439, 554, 1735, 1090])
    args = _args(w, sel)
    cache = tmp_path / "tune.txt"
    ctx = hip.Context(w.net)
    try:
        r = ctx.autotune(w.x0[sel], *args, cache_path=cache)
        for nb, tz, tw, ms in r["candidates"]:
            print(f"NB {nb} team_zones {tz} team_waves {tw}: {ms:.1f} ms")
        print("chosen", r["block_size"], r["team_zones"], r["team_waves"])
        assert r["consistent"] and not r["from_cache"]
        assert {c[0] for c in r["candidates"]} == {32, 64}
        best = min(r["candidates"], key=lambda c: c[3])
        assert (r["block_size"], r["team_zones"], r["team_waves"]) == best[:3]
        assert ctx.block_info()["NB"] == r["block_size"]
        ie, y, ye, st = ctx.integrate_batch_mixed(w.x0[sel], *args)
        np.testing.assert_array_equal(ie, 0)
    finally:
        ctx.close()
    ctx2 = hip.Context(w.net)
    try:
        r2 = ctx2.autotune(w.x0[sel], *args, cache_path=cache)
        assert r2["from_cache"] and r2["block_size"] == r["block_size"]
        ie2, y2, ye2, st2 = ctx2.integrate_batch_mixed(w.x0[sel], *args)
        np.testing.assert_array_equal(y2, y)
        ctx3 = hip.Context(w.net)   # untuned: block size 32, default options
        ie3, y3, _, _ = ctx3.integrate_batch_mixed(w.x0[sel], *args)
        ctx3.close()
        np.testing.assert_array_equal(y3, y)
    finally:
        ctx2.close()
