"""Benchmark: zone-timesteps/s of one mppnp-style cycle on one MI355X.

Workload: 2,048 independent zones sharing the synthetic 1,100-species network (7,128 reactions),
T9 log-uniform 0.05..3, rho 1e0..1e8, one cycle of dt = 1e4 s (backward-Euler Newton with the
device GEPP, FP64).  A step is one cycle over all zones: the rate kernel, then the sub-step /
Newton pipeline until every zone is done.  Inputs are resident in HBM when the timed region
starts; each step restores the initial abundances with a device-to-device copy.

Prints one JSON line (contract of the driver).  `roofline` is for the dominant kernel (the
per-zone Newton iteration: Jacobian, rhs, GEPP, solve) against HBM bandwidth using algorithmic
bytes; `cpu_baseline` times the CPU restatement of the same path (oracle driver + the reference
SuperLU 5.0 compiled from source) on a bounded sample of the same zones on the host cores.
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
    p.add_argument("--steps", type=int, default=1)
    p.add_argument("--warmup", type=int, default=0)
    p.add_argument("--zones", type=int, default=2048)
    p.add_argument("--cpu-sample", type=int, default=32, help="zones in the CPU-baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    return p.parse_args()


def algorithmic_bytes(net, n, nnzA, stats):
    """Bytes per Newton iteration summed over all iterations of the step (definitions of the
    measurement plan): Jacobian assembly 8R + 8N + 8 nnz(A); rhs 8R + 16N; LU 12 nnz(L+U)
    (value + index); solve 8 nnz(L+U) + 16N."""
    R = net.nre
    iters = int(stats["n_newton"].sum())
    lu = int(stats["lu_nnz"].sum())
    per_iter = (8 * R + 8 * n + 8 * nnzA) + (8 * R + 16 * n) + 16 * n
    return iters * per_iter + 20 * lu, iters, lu


def cpu_baseline(w, sample, cores):
    """Oracle driver + reference SuperLU on `sample` zones spread over the batch, one process per
    core (the reference's one-zone-per-worker model)."""
    from concurrent.futures import ProcessPoolExecutor
    idx = np.linspace(0, w.nz - 1, sample).round().astype(int)
    t = time.time()
    with ProcessPoolExecutor(max_workers=cores, initializer=_cpu_init, initargs=(w.name,)) as ex:
        results = list(ex.map(_cpu_zone, idx.tolist()))
    dt = time.time() - t
    ok = sum(1 for r in results if r == 0)
    return {"value": len(idx) / dt, "unit": "zone-timesteps/s", "cores": cores, "kind": "port",
            "sample": f"{len(idx)} of the {w.nz} zones (every ~{w.nz // len(idx)}th), one cycle "
                      f"each; oracle integrate_network + reference SuperLU 5.0 built from source; "
                      f"{ok} ok; {dt:.1f} s wall"}


_CPU = {}


def _cpu_init(name):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    from __graft_entry__ import _load_package
    pkg = _load_package()
    w = pkg.network.config(1)
    ls = oracle_lib.LS_REF if oracle_lib.ref_available() else oracle_lib.LS_GEPP
    _CPU["w"] = w
    _CPU["o"] = oracle_lib.Oracle(w.net, oracle_lib.default_knobs(cyminformal=-1.0), ls)


def _cpu_zone(z):
    w, o = _CPU["w"], _CPU["o"]
    ierr, y, ye, st = o.integrate_zone(w.x0[z], w.t9[z, 0], w.t9[z, 1], w.rho[z, 0], w.rho[z, 1],
                                       w.dt[z])
    return ierr


def main():
    a = _args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (the path shards "
              f"zones across GPUs, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch
    from __graft_entry__ import _load_package
    pkg = _load_package()
    hip = pkg.hip
    w = pkg.network.config(1, nz=a.zones)
    net = w.net
    ctx = hip.Context(net)
    dev = torch.device("cuda:0")
    nz = w.nz
    x0 = torch.tensor(w.x0, dtype=torch.float64, device=dev)
    yps = x0.clone()
    t9_0 = torch.tensor(w.t9[:, 0], dtype=torch.float64, device=dev)
    t9_1 = torch.tensor(w.t9[:, 1], dtype=torch.float64, device=dev)
    r0 = torch.tensor(w.rho[:, 0], dtype=torch.float64, device=dev)
    r1 = torch.tensor(w.rho[:, 1], dtype=torch.float64, device=dev)
    dt = torch.tensor(w.dt, dtype=torch.float64, device=dev)
    ye = torch.full((nz,), 0.5, dtype=torch.float64, device=dev)
    ierr = torch.zeros(nz, dtype=torch.int32, device=dev)
    stats = torch.zeros((nz, hip.STATS_DTYPE.itemsize // 8), dtype=torch.int64, device=dev)

    def step():
        yps.copy_(x0)
        torch.cuda.synchronize()
        ctx.integrate_batch_dev(nz, yps.data_ptr(), t9_0.data_ptr(), t9_1.data_ptr(),
                                r0.data_ptr(), r1.data_ptr(), ye.data_ptr(), dt.data_ptr(),
                                ierr.data_ptr(), stats.data_ptr(), 0)

    for i in range(a.warmup):
        step()
        torch.cuda.synchronize()
        print(f"warmup step {i + 1}/{a.warmup} done", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    timings = []
    t0 = time.perf_counter()
    for i in range(a.steps):
        step()
        timings.append(ctx.last_timing())
        print(f"step {i + 1}/{a.steps}: {timings[-1]['total_ms']:.1f} ms", file=sys.stderr,
              flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    st = np.frombuffer(stats.cpu().numpy().tobytes(), dtype=hip.STATS_DTYPE)
    ie = ierr.cpu().numpy()
    y = yps.cpu().numpy()
    sums = np.abs(y.sum(axis=1) - 1.0)

    # dominant kernel: the Newton-iteration kernel, HIP events around each launch
    tm = timings[-1]
    bytes_step, iters, lu = algorithmic_bytes(net, ctx.n, ctx.nnzA, st)
    newton_s = tm["newton_ms"] / 1e3
    launches = max(1, int(tm["newton_launches"]))
    avg_launch_s = newton_s / launches
    achieved = (bytes_step / launches) / avg_launch_s / 1e9 if newton_s > 0 else 0.0

    cpu = None
    if not a.no_cpu:
        cores = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(w, min(a.cpu_sample, nz), cores)

    value = nz * a.steps / elapsed
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
        "dtype": "f64",
        "data": "synthetic (seeded generator of the 2,048-zone x 1,100-species workload)",
        "config": {"workload": "mppnp-style batch: 2,048 zones x 1,100 species (7,128 reactions), "
                               "one cycle dt=1e4 s, backward Euler + device GEPP, plain FP64",
                   "zones": nz, "species": net.nsp, "reactions": net.nre, "path": "plain-fp64"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "zone_newton_kernel", "launches": launches,
                     "avg_launch_ms": 1e3 * avg_launch_s},
        "cpu_baseline": cpu,
        "detail": {"newton_iterations": iters, "lu_nnz_total": lu,
                   "max_newton_per_zone": int(st["n_newton"].max()),
                   "max_substeps_per_zone": int(st["nsubt"].max()),
                   "ierr_nonzero": int((ie != 0).sum()), "max_abs_sumX_minus_1": float(sums.max()),
                   "rates_ms": tm["rates_ms"], "pipeline_ms": tm["total_ms"],
                   "newton_ms": tm["newton_ms"], "control_ms": tm["control_ms"],
                   "passes": tm["passes"]},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
