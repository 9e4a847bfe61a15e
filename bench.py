"""Benchmark: zone-timesteps/s of one mppnp-style cycle on one MI355X.

Workload: 2,048 independent zones sharing the synthetic 1,100-species network (7,128 reactions),
T9 log-uniform 0.05..3, rho 1e0..1e8, one cycle of dt = 1e4 s (backward-Euler Newton).  A step is
one cycle over all zones: the rate kernel, then the sub-step / Newton pipeline until every zone
is done.  Inputs are resident in HBM when the timed region starts; each step restores the initial
abundances with a device-to-device copy.

`value` is the batched mixed-precision path (block factorisation in FP32, FP64 refinement, FP64
fallback per zone).  Beside it: the same path with FP64 factors, and the plain FP64 path (device
GEPP in the reference's pivot order) on a bounded sample of zones, timed together with the
batched path on the same sample.  `roofline` is for the dominant kernels (the block Newton
kernels), algorithmic bytes per launch over the HIP-event launch time; `cpu_baseline` times the
oracle integrator with the reference SuperLU 5.0 built from source on a bounded sample of the same
zones on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def _args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--zones", type=int, default=2048)
    p.add_argument("--plain-sample", type=int, default=32,
                   help="zones (stratified in T9 up to 1) on which the plain FP64 path is timed")
    p.add_argument("--cpu-sample", type=int, default=128, help="zones in the CPU-baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-plain", action="store_true")
    p.add_argument("--no-uneven", action="store_true",
                   help="skip the one-cycle run of the uneven 8,192-zone workload")
    p.add_argument("--no-extra", action="store_true",
                   help="skip the mppnp-cycle and rate-physics runs")
    p.add_argument("--uneven-cycles", type=int, default=1,
                   help="cycles of the uneven 8,192-zone workload (its definition: 50; the "
                        "abundances carry over from cycle to cycle); the FP64-factor "
                        "comparison runs on the first cycle only")
    p.add_argument("--save", default="", help="write the step's abundances and counters (npz)")
    p.add_argument("--bd-only", type=int, default=0, metavar="ZONES",
                   help="run only the Bader-Deuflhard workload (4,096 zones x 1,100 species, "
                        "T9 1..4) on the plain FP64 path for its first ZONES zones and print its "
                        "figures (the batched path's BD step is switched off)")
    p.add_argument("--large-only", type=int, default=0, metavar="ZONES",
                   help="run only the 5,000-species network's batched cycle on its first ZONES "
                        "zones and print its figures (a profiling workload)")
    return p.parse_args()


TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "newton_kernels_2048x1100_backward_euler_traffic.json")


def measured_traffic():
    """HBM bytes per launch of the Newton kernel from the committed rocprofv3 counter passes
    (FETCH_SIZE and WRITE_SIZE, calibrated against a copy of known size), or None."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            return float(json.load(f)["hbm_bytes_per_dispatch"])
    except (OSError, KeyError, ValueError):
        return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def block_bytes(net, ctx, st):
    """Algorithmic HBM bytes of the block Newton kernels over one step, from the per-zone
    counters: per Newton iteration the Jacobian assembly (8R + 8N + 8 nnz(A)) and the rhs
    (8R + 16N); per factorisation the stored factors (3 nblk NB^2 + 2 nbd nip entries, 4 or 8
    bytes) plus reading A (8 nnz(A)); per solve the factors again plus 16N; per refinement
    residual 8 nnz(A) + 16N."""
    info = ctx.block_info()
    R, N, nnzA = net.nre, ctx.n, ctx.nnzA
    fac_entries = 3 * info["nblk"] * info["NB"] ** 2 + 2 * info["nbd"] * (info["nblk"] * info["NB"])
    nf = st["n_factor"].astype(np.int64)
    nf64 = st["n_factor_fp64"].astype(np.int64)
    nf32 = nf - nf64
    # solves are attributed to the precision of the zone's factorisations
    frac64 = np.where(nf > 0, nf64 / np.maximum(nf, 1), 0.0)
    ns = st["n_solve"].astype(np.float64)
    per_newton = (8 * R + 8 * N + 8 * nnzA) + (8 * R + 16 * N)
    total = (st["n_newton"].sum() * per_newton
             + (nf32.sum() * 4 + nf64.sum() * 8) * fac_entries + nf.sum() * 8 * nnzA
             + (ns * (frac64 * 8 + (1 - frac64) * 4)).sum() * fac_entries + ns.sum() * 16 * N
             + st["n_refine"].sum() * (8 * nnzA + 16 * N))
    return float(total)


def cpu_baseline(sample_idx, cores):
    """Oracle integrate_network + reference SuperLU on the sampled zones, one process per core
    (the reference's one-zone-per-worker model).  Run before anything touches the GPU, so the
    forked workers hold no device state; the pool is shut down before the GPU part starts."""
    import multiprocessing
    from concurrent.futures import ProcessPoolExecutor
    t = time.time()
    ex = ProcessPoolExecutor(max_workers=cores, initializer=_cpu_init,
                             mp_context=multiprocessing.get_context("fork"))
    try:
        results = list(ex.map(_cpu_zone, [int(i) for i in sample_idx]))
    finally:
        ex.shutdown(wait=True)
    dt = time.time() - t
    ok = sum(1 for r, _ in results if r == 0)
    return {"value": len(sample_idx) / dt, "unit": "zone-timesteps/s", "cores": cores,
            "kind": "reference",
            "sample": f"{len(sample_idx)} of the 2,048 zones (every ~{2048 // len(sample_idx)}th), "
                      f"one cycle each; oracle integrate_network driving the reference SuperLU 5.0 "
                      f"built from source (slu_solve per Newton iteration, as linalg.F90:115); "
                      f"{ok} ok; {dt:.1f} s wall",
            "_zone_seconds": sum(e for _, e in results)}


_CPU = {}


def _cpu_init():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    from __graft_entry__ import _load_package
    pkg = _load_package()
    w = pkg.network.config(1)
    ls = oracle_lib.LS_REF if oracle_lib.ref_available() else oracle_lib.LS_GEPP
    _CPU["w"] = w
    _CPU["o"] = oracle_lib.Oracle(w.net, oracle_lib.default_knobs(cyminformal=-1.0), ls)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_zone(z):
    w, o = _CPU["w"], _CPU["o"]
    t = time.perf_counter()
    ierr, y, ye, st = o.integrate_zone(w.x0[z], w.t9[z, 0], w.t9[z, 1], w.rho[z, 0], w.rho[z, 1],
                                       w.dt[z])
    return ierr, time.perf_counter() - t


def bd_workload(pkg, hip, zones):
    """The Bader-Deuflhard workload (semi-implicit extrapolation, integration_method = 1): its
    first `zones` zones of 4,096 (1,100 species, T9 log-uniform 1..4, rho 1e5..1e9, dt
    1e-3..1e2 s) in one call on the plain FP64 path (device GEPP in the reference's pivot order,
    one zone per wavefront, all zones concurrently)."""
    w = pkg.network.config(4)
    ctx = hip.Context(w.net, hip.default_knobs(integration_method=1))
    try:
        sel = np.arange(zones)
        args = [v[sel] for v in (w.t9[:, 0], w.t9[:, 1], w.rho[:, 0], w.rho[:, 1], w.dt)]
        ctx.integrate_batch(w.x0[:1], *[v[:1] for v in args])   # warm-up: one zone
        t0 = time.perf_counter()
        ie, y, _, st = ctx.integrate_batch(w.x0[sel], *args)
        el = time.perf_counter() - t0
    finally:
        ctx.close()
    out = {"zones": zones, "path": "plain FP64 (device GEPP), Bader-Deuflhard",
           "seconds": el, "zone_timesteps_per_s": zones / el,
           "ierr_nonzero": int((ie != 0).sum()),
           "sub_steps_max": int(st["nsubt"].max()),
           "factorisations": int(st["n_factor"].sum()),
           "factorisations_max_per_zone": int(st["n_factor"].max()),
           "stage_solves": int(st["n_solve"].sum()),
           "max_abs_sumX_minus_1": float(np.abs(y.sum(axis=1) - 1.0).max()),
           "what": "4,096 zones x 1,100 species, T9 1..4, rho 1e5..1e9, dt 1e-3..1e2 s: the "
                   "first zones of the workload"}
    log(f"BD workload: {el:.1f} s for {zones} zones")
    return out


def large_network(pkg, hip, plain=True, zones=None):
    """The large network: 256 zones x 5,000 species (34,055 reactions) on the batched path's band
    form, one cycle, and the plain FP64 path on four of its cool zones."""
    w3 = pkg.network.config(3)
    ctx3 = hip.Context(w3.net)
    info3 = ctx3.block_info()
    n3 = w3.x0.shape[0] if zones is None else zones
    a3 = [v[:n3] for v in (w3.t9[:, 0], w3.t9[:, 1], w3.rho[:, 0], w3.rho[:, 1], w3.dt)]
    ctx3.integrate_batch_mixed(w3.x0[:8], *[v[:8] for v in a3])   # warm-up
    t0 = time.perf_counter()
    ie3, y3, _, st3 = ctx3.integrate_batch_mixed(w3.x0[:n3], *a3)
    el3 = time.perf_counter() - t0
    tm3 = ctx3.last_timing()
    out = {
        "zones": n3, "species": w3.net.nsp, "reactions": w3.net.nre,
        "form": "band (half-bandwidth %d, border %d)" % (info3["bandwidth"], info3["nbd"]),
        "zone_timesteps_per_s_batched": n3 / el3, "cycle_ms": 1e3 * el3,
        "newton_ms": tm3["newton_ms"], "passes": tm3["passes"],
        "newton_iterations": int(st3["n_newton"].sum()),
        "max_newton_per_zone": int(st3["n_newton"].max()),
        "fp64_factorisations": int(st3["n_factor_fp64"].sum()),
        "factorisations": int(st3["n_factor"].sum()),
        "solves": int(st3["n_solve"].sum()),
        "ierr_nonzero": int((ie3 != 0).sum()),
        "max_abs_sumX_minus_1": float(np.abs(y3.sum(axis=1) - 1.0).max()),
        "what": "256 zones x 5,000 species, T9 0.1..2, rho 1e2..1e6, dt 1e4 s"}
    log(f"large network: {el3 * 1e3:.1f} ms for {n3} zones, {tm3['passes']:.0f} passes")
    if plain:
        sel3 = np.array([5, 6, 7, 10])
        t0 = time.perf_counter()
        ie3p, y3p, _, st3p = ctx3.integrate_batch(w3.x0[sel3], *[v[sel3] for v in a3])
        el3p = time.perf_counter() - t0
        out["plain_sample"] = {"zones": sel3.tolist(), "seconds": el3p,
                               "zone_timesteps_per_s": len(sel3) / el3p,
                               "newton_iterations": int(st3p["n_newton"].sum()),
                               "ierr_nonzero": int((ie3p != 0).sum())}
        log(f"large network plain sample: {el3p:.1f} s")
    ctx3.close()
    return out


def main():
    a = _args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported: the path shards "
              f"zones across GPUs (one contiguous block of zones per GPU), which this one-GPU "
              f"build does not run", file=sys.stderr)
        sys.exit(2)
    if a.large_only or a.bd_only:
        from __graft_entry__ import _load_package
        pkg = _load_package()
        if a.large_only:
            print(json.dumps(large_network(pkg, pkg.hip, plain=False, zones=a.large_only)))
        else:
            print(json.dumps(bd_workload(pkg, pkg.hip, a.bd_only)))
        return
    # the CPU baseline first: no GPU state exists yet in this process or its workers
    cpu = None
    cpu_one = None
    if not a.no_cpu:
        # the GPU box's CPU share is 16 cores; never more workers than that
        cores = min(16, os.cpu_count() or 1)
        idx = np.linspace(0, a.zones - 1, min(a.cpu_sample, a.zones)).round().astype(int)
        cpu = cpu_baseline(idx, cores)
        cpu["sample"] += f"; host CPU: {_cpu_model()}, {os.cpu_count()} logical CPUs visible"
        log(f"cpu baseline: {cpu['value']:.2f} zone-timesteps/s on {cores} cores")
        # one core: the same zones back to back, from each zone's own time in the run above
        cpu_one = {"value": len(idx) / cpu.pop("_zone_seconds"), "unit": "zone-timesteps/s",
                   "sample": "the same zones, summed per-zone times of the run above"}
        log(f"cpu baseline, one core: {cpu_one['value']:.2f} zone-timesteps/s")

    import torch
    from __graft_entry__ import _load_package
    pkg = _load_package()
    hip = pkg.hip
    w = pkg.network.config(1, nz=a.zones)
    net = w.net
    ctx = hip.Context(net)
    dev = torch.device("cuda:0")
    nz = w.nz
    tens = lambda x: torch.tensor(np.ascontiguousarray(x), dtype=torch.float64, device=dev)
    x0 = tens(w.x0)
    yps = x0.clone()
    t9_0, t9_1, r0, r1, dt = (tens(x) for x in (w.t9[:, 0], w.t9[:, 1], w.rho[:, 0], w.rho[:, 1],
                                                 w.dt))
    ye = torch.full((nz,), 0.5, dtype=torch.float64, device=dev)
    ierr = torch.zeros(nz, dtype=torch.int32, device=dev)
    nstat = hip.STATS_DTYPE.itemsize // 8
    stats = torch.zeros((nz, nstat), dtype=torch.int64, device=dev)

    def run(path, sel=None):
        """one step (cycle) of `path` over all zones, or over the zone subset `sel`"""
        if sel is None:
            y, a0, a1, b0, b1, d, e, ie, s, n = yps, t9_0, t9_1, r0, r1, dt, ye, ierr, stats, nz
            y.copy_(x0)
        else:
            n = len(sel)
            idx = torch.tensor(sel, device=dev)
            y = x0[idx].clone()
            a0, a1, b0, b1, d = (t[idx].contiguous() for t in (t9_0, t9_1, r0, r1, dt))
            e = torch.full((n,), 0.5, dtype=torch.float64, device=dev)
            ie = torch.zeros(n, dtype=torch.int32, device=dev)
            s = torch.zeros((n, nstat), dtype=torch.int64, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ptrs = [t.data_ptr() for t in (y, a0, a1, b0, b1, e, d, ie, s)]
        if path == "plain":
            ctx.integrate_batch_dev(n, *ptrs, 0)
        else:
            ctx.integrate_batch_mixed_dev(n, *ptrs, opts=hip.batch_opts(
                factor_precision=32 if path == "mixed" else 64))
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        st = np.frombuffer(s.cpu().numpy().tobytes(), dtype=hip.STATS_DTYPE)
        return el, y.cpu().numpy(), ie.cpu().numpy(), st, ctx.last_timing()

    # --- timed: the batched mixed path over all zones (the workload's metric)
    for i in range(a.warmup):
        el, *_ = run("mixed")
        log(f"warmup step {i + 1}/{a.warmup}: {el * 1e3:.1f} ms")
    elapsed = 0.0
    timings = []
    for i in range(a.steps):
        el, y_mixed, ie_mixed, st_mixed, tm = run("mixed")
        elapsed += el
        timings.append(tm)
        log(f"step {i + 1}/{a.steps}: {el * 1e3:.1f} ms")
    value = nz * a.steps / elapsed
    tm = timings[-1]
    bytes_step = block_bytes(net, ctx, st_mixed)
    launches = max(1, int(tm["newton_launches"]))
    newton_s = tm["newton_ms"] / 1e3
    avg_launch_s = newton_s / launches
    achieved = (bytes_step / launches) / avg_launch_s / 1e9 if newton_s > 0 else 0.0

    # --- the same path with FP64 factors (accuracy reference and throughput)
    el64, y64, ie64, st64, _ = run("fp64")
    log(f"fp64-factor step: {el64 * 1e3:.1f} ms")
    m = (y64 > 1e-20) | (y_mixed > 1e-20)
    rel = np.where(m, np.abs(y_mixed - y64) / np.maximum(np.maximum(y_mixed, y64), 1e-300), 0.0)
    err_mixed_vs_fp64 = float(rel.max())
    same = ((st_mixed["nsubt"] == st64["nsubt"]) & (st_mixed["n_newton"] == st64["n_newton"]))
    err_same = float(rel[same].max()) if same.any() else None

    # --- plain FP64 path (reference pivot order) on a bounded sample of light zones
    plain = None
    if not a.no_plain:
        # stratified: zones from the cold end up to T9 1 (the plain path's one-wavefront GEPP
        # takes about 0.1 s per Newton iteration, so the hottest zones' 1,300 iterations take
        # minutes; T9 up to 1 keeps zones of up to ~30 sub-steps)
        order = np.argsort(w.t9[:, 0])
        warm = order[w.t9[order, 0] < 1.0]
        sel = np.sort(warm[np.linspace(0, warm.size - 1, min(a.plain_sample, warm.size)).round()
                           .astype(int)])
        elp, yp, iep, stp, _ = run("plain", sel)
        elb, yb, ieb, stb, _ = run("mixed", sel)
        mm = (yp > 1e-20) | (yb > 1e-20)
        plain = {"zones": int(len(sel)), "selection": "stratified in T9 over the zones with "
                 f"T9 < 1 (T9 {w.t9[sel, 0].min():.3f} .. {w.t9[sel, 0].max():.3f}; the plain "
                 "path's per-iteration latency makes the hottest zones take minutes)",
                 "plain_zone_timesteps_per_s": len(sel) / elp,
                 "mixed_zone_timesteps_per_s": len(sel) / elb, "mixed_over_plain": elp / elb,
                 "max_rel_err_mixed_vs_plain": float(np.max(np.abs(yp[mm] - yb[mm]) /
                                                            np.maximum(yp[mm], yb[mm]))),
                 "same_decisions": bool(np.array_equal(stp["nsubt"], stb["nsubt"]) and
                                        np.array_equal(stp["n_newton"], stb["n_newton"]))}
        log(f"plain sample: {elp * 1e3:.1f} ms vs mixed {elb * 1e3:.1f} ms on {len(sel)} zones")

    # --- the uneven 8,192-zone workload (same network; T9 0.02..4, dt 1..1e8 s), one cycle
    uneven = None
    if not a.no_uneven:
        w2 = pkg.network.config(2)
        n2 = w2.nz
        y2 = tens(w2.x0)
        a2 = [tens(x) for x in (w2.t9[:, 0], w2.t9[:, 1], w2.rho[:, 0], w2.rho[:, 1], w2.dt)]
        e2 = torch.full((n2,), 0.5, dtype=torch.float64, device=dev)
        i2 = torch.zeros(n2, dtype=torch.int32, device=dev)
        s2 = torch.zeros((n2, nstat), dtype=torch.int64, device=dev)
        res = {}
        for path in ("mixed", "fp64"):
            y2.copy_(tens(w2.x0))
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.integrate_batch_mixed_dev(n2, y2.data_ptr(), *(t.data_ptr() for t in a2[:4]),
                                          e2.data_ptr(), a2[4].data_ptr(), i2.data_ptr(),
                                          s2.data_ptr(), opts=hip.batch_opts(
                                              factor_precision=32 if path == "mixed" else 64))
            torch.cuda.synchronize()
            el2 = time.perf_counter() - t0
            st2 = np.frombuffer(s2.cpu().numpy().tobytes(), dtype=hip.STATS_DTYPE)
            res[path] = (el2, y2.cpu().numpy(), i2.cpu().numpy(), st2)
            log(f"uneven 8,192-zone cycle ({path}): {el2 * 1e3:.1f} ms")
        el2, yy, ii, st2 = res["mixed"]
        multi = None
        if a.uneven_cycles > 1:
            # the workload's 50 cycles: each cycle continues from the previous one's abundances
            y2.copy_(tens(w2.x0))
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            per = []
            for cyc in range(a.uneven_cycles):
                tc = time.perf_counter()
                ctx.integrate_batch_mixed_dev(n2, y2.data_ptr(), *(t.data_ptr() for t in a2[:4]),
                                              e2.data_ptr(), a2[4].data_ptr(), i2.data_ptr(),
                                              s2.data_ptr())
                torch.cuda.synchronize()
                per.append(time.perf_counter() - tc)
                log(f"uneven cycle {cyc + 1}/{a.uneven_cycles}: {per[-1] * 1e3:.1f} ms")
                if int((i2 != 0).sum().item()) != 0:
                    break
            elm = time.perf_counter() - t0
            ym = y2.cpu().numpy()
            multi = {"cycles": len(per), "zone_timesteps_per_s": n2 * len(per) / elm,
                     "ms_per_cycle_min_max": [1e3 * min(per), 1e3 * max(per)],
                     "ierr_nonzero_last": int((i2.cpu().numpy() != 0).sum()),
                     "max_abs_sumX_minus_1_end": float(np.abs(ym.sum(axis=1) - 1.0).max()),
                     "finite": bool(np.isfinite(ym).all())}
        # the plain FP64 path on a sample of the same workload: zones stratified by the batched
        # run's Newton count up to 120 (the plain path's one-wavefront GEPP takes ~0.1 s per
        # Newton iteration, so the workload's hottest zones would take minutes), and the batched
        # path on the same zones
        uplain = None
        if not a.no_plain:
            nn = st2["n_newton"].astype(np.int64)
            cand = np.flatnonzero(nn <= 120)
            cand = cand[np.argsort(nn[cand], kind="stable")]
            sel2 = np.sort(cand[np.linspace(0, cand.size - 1, min(a.plain_sample, cand.size))
                                .round().astype(int)])
            out = {}
            for path in ("plain", "mixed"):
                idx = torch.tensor(sel2, device=dev)
                ys = tens(w2.x0)[idx].contiguous()
                av = [t[idx].contiguous() for t in a2]
                es = torch.full((len(sel2),), 0.5, dtype=torch.float64, device=dev)
                is_ = torch.zeros(len(sel2), dtype=torch.int32, device=dev)
                ss = torch.zeros((len(sel2), nstat), dtype=torch.int64, device=dev)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                ptrs = [t.data_ptr() for t in (ys, av[0], av[1], av[2], av[3], es, av[4], is_, ss)]
                if path == "plain":
                    ctx.integrate_batch_dev(len(sel2), *ptrs, 0)
                else:
                    ctx.integrate_batch_mixed_dev(len(sel2), *ptrs)
                torch.cuda.synchronize()
                out[path] = (time.perf_counter() - t0, ys.cpu().numpy(), is_.cpu().numpy(),
                             np.frombuffer(ss.cpu().numpy().tobytes(), dtype=hip.STATS_DTYPE))
                log(f"uneven plain sample ({path}): {out[path][0] * 1e3:.1f} ms on {len(sel2)} zones")
            (elp2, yp2, iep2, stp2), (elb2, yb2, ieb2, stb2) = out["plain"], out["mixed"]
            mp = (yp2 > 1e-20) | (yb2 > 1e-20)
            uplain = {"zones": int(len(sel2)),
                      "selection": "stratified in the batched run's Newton count over the zones "
                                   f"with at most 120 ({int(nn[sel2].min())} .. "
                                   f"{int(nn[sel2].max())} Newton iterations, T9 "
                                   f"{w2.t9[sel2, 0].min():.3f} .. {w2.t9[sel2, 0].max():.3f})",
                      "plain_zone_timesteps_per_s": len(sel2) / elp2,
                      "mixed_zone_timesteps_per_s": len(sel2) / elb2,
                      "mixed_over_plain": elp2 / elb2,
                      "ierr_nonzero_plain": int((iep2 != 0).sum()),
                      "ierr_nonzero_mixed": int((ieb2 != 0).sum()),
                      "max_rel_err_mixed_vs_plain": float(np.max(np.abs(yp2[mp] - yb2[mp]) /
                                                                 np.maximum(yp2[mp], yb2[mp]))),
                      "zones_same_decisions": int(((stp2["nsubt"] == stb2["nsubt"]) &
                                                   (stp2["n_newton"] == stb2["n_newton"])).sum())}
        y64u, st64u = res["fp64"][1], res["fp64"][3]
        mu = (yy > 1e-20) | (y64u > 1e-20)
        relu = np.where(mu, np.abs(yy - y64u) / np.maximum(np.maximum(yy, y64u), 1e-300), 0.0)
        sameu = (st2["nsubt"] == st64u["nsubt"]) & (st2["n_newton"] == st64u["n_newton"])
        uneven = {"zones": n2, "zone_timesteps_per_s_mixed": n2 / el2,
                  "plain_sample": uplain,
                  "zone_timesteps_per_s_fp64_factors": n2 / res["fp64"][0],
                  "mixed_equals_fp64_bitwise": bool(np.array_equal(yy, y64u)),
                  "zones_not_bitwise_equal": int((yy != y64u).any(axis=1).sum()),
                  "max_rel_err_mixed_vs_fp64_factors": float(relu.max()),
                  "zones_rel_err_above_1e-8": int((relu.max(axis=1) > 1e-8).sum()),
                  "zones_same_decisions_as_fp64": int(sameu.sum()),
                  "ierr_nonzero": int((ii != 0).sum()),
                  "max_abs_sumX_minus_1": float(np.abs(yy.sum(axis=1) - 1.0).max()),
                  "max_sub_steps": int(st2["nsubt"].max()),
                  "max_newton_per_zone": int(st2["n_newton"].max()),
                  "zones_on_fp64": int((st2["n_factor_fp64"] > 0).sum()),
                  "all_cycles": multi}
        del y2, a2, e2, i2, s2

    # --- the mppnp cycle on the device (zoneselect, batch, renormalisation, mixing) and the
    # same cycle with the rate physics on (partition functions, detailed balance, screening)
    extra = None
    if not a.no_extra:
        sm = pkg.network.stellar_model(nz, t9_center=3.0, t9_surface=0.01)
        yc = x0.clone()
        tt = [tens(v) for v in (sm.t9, sm.rho, sm.xm, sm.dq, sm.rad, sm.d_eul)]
        ie_m = torch.zeros(nz, dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rc = ctx.mppnp_cycle_dev(nz, yc.data_ptr(), *(t.data_ptr() for t in tt), 1e4,
                                 ie_m.data_ptr(), opts=hip.mppnp_opts(t9threshold=0.05))
        torch.cuda.synchronize()
        el_m = time.perf_counter() - t0
        tm_m = ctx.last_timing()
        nsel = int((sm.t9 > 0.05).sum())
        ctx_p = hip.Context(net)
        ctx_p.set_rate_physics(pkg.network.rate_physics(net))
        yp = x0.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx_p.integrate_batch_mixed_dev(nz, yp.data_ptr(), *(t.data_ptr() for t in (t9_0, t9_1,
                                        r0, r1)), ye.data_ptr(), dt.data_ptr(), ierr.data_ptr(),
                                        stats.data_ptr())
        torch.cuda.synchronize()
        el_p = time.perf_counter() - t0
        tm_p = ctx_p.last_timing()
        ie_p = ierr.cpu().numpy()
        ctx_p.close()
        extra = {"mppnp_cycle": {"zones": nz, "integrated": nsel, "rc": int(rc),
                                 "ierr_nonzero": int((ie_m.cpu().numpy() != 0).sum()),
                                 "cycle_ms": 1e3 * el_m,
                                 "integration_ms": tm_m["total_ms"] + tm_m["rates_ms"],
                                 "selection_mixing_renormalisation_ms":
                                     1e3 * el_m - tm_m["total_ms"] - tm_m["rates_ms"],
                                 "model": "synthetic star: T9 3 -> 0.01 centre to surface, "
                                          "convective core, shell and envelope; zoneselect "
                                          "T9 > 0.05; matrixdiffus (GNV / EVL form)"},
                 "rate_physics": {"zone_timesteps_per_s": nz / el_p,
                                  "rates_ms": tm_p["rates_ms"],
                                  "rates_ms_bare_reaclib": tm["rates_ms"],
                                  "ierr_nonzero": int((ie_p != 0).sum()),
                                  "what": "the same 2,048-zone cycle with partition functions, "
                                          "detailed balance and Chugunov screening on (the "
                                          "reference's defaults); the screened rates make the "
                                          "zones stiffer"}}
        log(f"mppnp cycle: {el_m * 1e3:.1f} ms; rate physics cycle {el_p * 1e3:.1f} ms")
        # the same 2,048-zone cycle under the dynamic network (cyminformal 1e-30, the
        # reference's default): the mask and reduced entries recomputed every iteration
        ctx_d = hip.Context(net, hip.default_knobs(cyminformal=1.0e-30, min_rate_ov_dt=1.0e-99))
        yd = x0.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx_d.integrate_batch_mixed_dev(nz, yd.data_ptr(), *(t.data_ptr() for t in (t9_0, t9_1,
                                        r0, r1)), ye.data_ptr(), dt.data_ptr(), ierr.data_ptr(),
                                        stats.data_ptr())
        torch.cuda.synchronize()
        el_d = time.perf_counter() - t0
        std = stats.cpu().numpy().view(hip.STATS_DTYPE).reshape(-1)
        extra["dynamic_network"] = {
            "zone_timesteps_per_s": nz / el_d, "ierr_nonzero": int((ierr.cpu().numpy() != 0).sum()),
            "nvar1_min_median_max": [int(std["nvar1"].min()), int(np.median(std["nvar1"])),
                                     int(std["nvar1"].max())],
            "what": "the 2,048-zone cycle with cyminformal = 1e-30, min_rate_ov_dt = 1e-99"}
        ctx_d.close()
        log(f"dynamic network cycle {el_d * 1e3:.1f} ms")
        extra["large_network"] = large_network(pkg, hip, plain=True)

    if a.save:
        np.savez(a.save, y_mixed=y_mixed, y_fp64=y64, nsubt_mixed=st_mixed["nsubt"],
                 nsubt_fp64=st64["nsubt"], newton_mixed=st_mixed["n_newton"],
                 newton_fp64=st64["n_newton"], fp64_mixed=st_mixed["n_factor_fp64"])
    sums = np.abs(y_mixed.sum(axis=1) - 1.0)
    out = {
        "metric": "zone-timesteps/s @1,100 species, mixed vs FP64 GPU, at <=1e-8 rel err",
        "value": value,
        "unit": "zone-timesteps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * elapsed / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 factors + f64 state/residual/refinement",
        "data": "synthetic (seeded generator of the 2,048-zone x 1,100-species workload)",
        "config": {"workload": "mppnp-style batch: 2,048 zones x 1,100 species (7,128 reactions), "
                               "one cycle dt=1e4 s, backward Euler, batched block factorisation "
                               "FP32 + FP64 refinement",
                   "zones": nz, "species": net.nsp, "reactions": net.nre, "path": "mixed"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "traffic_source": "profiles/" + os.path.basename(TRAFFIC_PROFILE),
                     "kernel": "the Newton worker, one launch per pass: "
                               "zone_newton_team_kernel<32, 4> (four wavefronts per zone; two "
                               "such workgroups per CU while the list holds more zones than "
                               "the device has CUs)",
                     "launches": launches, "avg_launch_ms": 1e3 * avg_launch_s,
                     # time-weighted: the same algorithmic bytes over the whole timed step
                     "pipeline_frac": bytes_step / (elapsed / a.steps) / 1e9 / HBM_PEAK_GBS},
        "cpu_baseline": cpu,
        "detail": {
            "fp64_factor_path_zone_timesteps_per_s": nz / el64,
            "max_rel_err_mixed_vs_fp64_factors": err_mixed_vs_fp64,
            "max_rel_err_mixed_vs_fp64_same_decisions": err_same,
            "zones_same_decisions_as_fp64": int(same.sum()),
            "zones_rel_err_above_1e-8": int((rel.max(axis=1) > 1e-8).sum()),
            "max_abs_sumX_minus_1": float(sums.max()),
            "max_abs_sumX_minus_1_fp64": float(np.abs(y64.sum(axis=1) - 1.0).max()),
            "ierr_nonzero": int((ie_mixed != 0).sum()),
            "newton_iterations": int(st_mixed["n_newton"].sum()),
            "max_newton_per_zone": int(st_mixed["n_newton"].max()),
            "factorisations": int(st_mixed["n_factor"].sum()),
            "fp64_factorisations": int(st_mixed["n_factor_fp64"].sum()),
            "zones_on_fp64": int((st_mixed["n_factor_fp64"] > 0).sum()),
            "solves": int(st_mixed["n_solve"].sum()),
            "refinement_residuals": int(st_mixed["n_refine"].sum()),
            "rates_ms": tm["rates_ms"], "pipeline_ms": tm["total_ms"],
            "newton_ms": tm["newton_ms"], "control_ms": tm["control_ms"],
            "passes": tm["passes"], "algorithmic_bytes_per_step": bytes_step,
            "plain_sample": plain,
            "uneven_8192_zones_one_cycle": uneven,
            "graph_launches": tm.get("graph_launches"),
            "extra": extra,
            "cpu_one_core": None if cpu_one is None else {
                "value": cpu_one["value"], "unit": cpu_one["unit"], "sample": cpu_one["sample"]},
        },
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
