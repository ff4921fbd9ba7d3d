"""LARE3D Lagrangian-remap step on one MI355X: cell-updates/s.

Headline workload (`value`): the 256^3 resistive Brio-Wu problem
(x-directed shock tube, rho/p/By = 1/1/1 left and 0.125/0.1/-1 right,
Bx = 0.75, gamma 2, 1% seeded density noise, eta = 1e-3, zero-gradient x and
periodic y/z boundaries), FP64, inputs resident in HBM.  A step is one full
LARE3D timestep (startOfTimestep + halfTimestep + endOfTimestep).

Also measured, in the same JSON line:
  "batched": 256 Orszag-Tang grids, sizes log-uniform in [32^3, 128^3],
             plasma beta U[0.1, 10], perturbation amplitude U[0, 0.05], FP64,
             stepped as one batch (one launch per kernel, graph-captured
             step) against the same grids solved one after another, each
             with its own captured step graph and on the plain path (every
             kernel launched from the host);
  "serving": seeded Poisson streams of solves (sizes log-uniform in
             [32^3, 192^3], 50-500 steps, Orszag-Tang or Brio-Wu, FP64 or
             FP32 storage) served with continuous batching under a 200 GB
             HBM budget, swept over arrival rates up to saturation:
             throughput and p50/p99 completion latency at each rate;
  "fp32":    FP32 storage against FP64 on Orszag-Tang with 1e-3 Gaussian
             velocity perturbations: time per step and per-field error.
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s FP64 256^3 + %HBM roofline; batched-serve p50/p99 latency"
HBM_PEAK_GBS = 8000.0
# algorithmic HBM bytes per cell-update of the whole step: state read and
# written once per phase (viscosity+CFL 72, Lagrange 144, x 168, y 160,
# z 152) plus two resistive substeps of 72
STEP_BYTES_IDEAL = 696.0
STEP_BYTES_RESISTIVE = 840.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(n, steps, threads):
    """The reference's own OpenMP path on the host cores, on the workload
    itself: oracle/_ref/ref_driver_res_fast, the reference's LARE sources
    built with its optimisation flags, with the element-wise copies at
    lagran.cpp:482-484 that let its resistive path run (it throws there
    otherwise)."""
    from oracle import oracle as O
    exe = O.REF_DRIVER_RES_FAST
    if os.path.exists(exe):
        with tempfile.TemporaryDirectory() as td:
            meta = O.run_reference(td, problem="bw", nx=n, steps=steps, noise=0.01, seed=1, eta=1e-3,
                                   threads=threads, fast=True)
        el = meta["elapsed_s"]
        return {"value": n ** 3 * steps / el, "unit": "cell-updates/s", "cores": threads,
                "kind": "reference",
                "sample": f"reference OpenMP LARE3D (-O3 -march=x86-64-v3, resistive path made runnable), "
                          f"the workload itself (resistive Brio-Wu {n}^3, eta 1e-3, 1% density noise) for "
                          f"its first {steps} steps, step loop {el:.2f} s"}
    # fall back to the C restatement (the port), resistive as the workload
    import sams_amd.lare as L
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from lare_testlib import make_oracle
    O.set_threads(threads)
    p = L.problem("bw", n=n, seed=1, noise=0.01, eta=1e-3)
    o = make_oracle(p)
    t0 = time.perf_counter()
    for _ in range(steps):
        o.step()
    el = time.perf_counter() - t0
    return {"value": n ** 3 * steps / el, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"C restatement of the reference (oracle/), Brio-Wu {n}^3 resistive, {steps} steps"}


def bench_batched(steps, seq_steps, ngrids=256, seed=2024):
    """Config: 256 Orszag-Tang grids as one batch vs solved one after another."""
    import math
    import numpy as np
    import sams_amd.lare as L
    rng = np.random.default_rng(seed)
    sizes = [int(round(math.exp(rng.uniform(math.log(32), math.log(128))))) for _ in range(ngrids)]
    betas = rng.uniform(0.1, 10.0, ngrids)
    amps = rng.uniform(0.0, 0.05, ngrids)
    t0 = time.perf_counter()
    grids = [L.Lare.from_problem(L.problem("ot", n=sizes[i], seed=1 + i, amp=float(amps[i]), beta=float(betas[i])))
             for i in range(ngrids)]
    setup = time.perf_counter() - t0
    cells = float(sum(n ** 3 for n in sizes))
    b = L.Batch(grids)
    t0 = time.perf_counter()
    tune = b.autotune(steps=2)  # launch configuration for this batch (state restored)
    tune_s = time.perf_counter() - t0
    b.steps(2)  # capture and warm
    b.sync()
    t0 = time.perf_counter()
    b.steps(steps)
    b.sync()
    tb = time.perf_counter() - t0
    prof_ms, launches = b.profile(1)
    b.close()
    for g in grids:  # capture each grid's own step graph
        g.steps(1, record_dt=False)
    t0 = time.perf_counter()
    for g in grids:
        g.steps(seq_steps, record_dt=False)
    ts = time.perf_counter() - t0
    # the plain path: each grid alone, every kernel launched from the host
    for g in grids:
        g.set_graph(False)
    t0 = time.perf_counter()
    for g in grids:
        g.steps(seq_steps, record_dt=False)
    te = time.perf_counter() - t0
    for g in grids:
        g.close()
    rb, rs, re_ = cells * steps / tb, cells * seq_steps / ts, cells * seq_steps / te
    log(f"bench: batched {ngrids} grids ({cells / 1e6:.1f} M cells): {rb / 1e9:.3f} G cell-updates/s batched, "
        f"{rs / 1e9:.3f} sequential with a captured graph per grid (x{rb / rs:.2f}), {re_ / 1e9:.3f} sequential "
        f"plain path (x{rb / re_:.2f}) (set-up {setup:.1f} s, autotune {tune_s:.1f} s: {tune})")
    return {"workload": f"{ngrids} Orszag-Tang grids, n log-uniform in [32, 128], beta U[0.1, 10], "
                        f"amp U[0, 0.05], FP64, seed {seed}",
            "cells": cells, "value": rb, "unit": "cell-updates/s", "steps": steps,
            "ms_per_step": 1e3 * tb / steps, "sequential_value": rs, "sequential_steps_per_grid": seq_steps,
            "speedup_vs_sequential": rb / rs, "sequential_plain_value": re_,
            "speedup_vs_sequential_plain": rb / re_, "launches_per_step": launches,
            "kernel_ms_per_step_eager": prof_ms, "tune": tune, "autotune_s": tune_s}


def bench_serving(qps_list, duration, seed, budget, k):
    """Serving at each arrival rate of qps_list (the same seeded stream shape,
    `duration` seconds of arrivals, every request run to completion):
    throughput, p50/p99 completion latency, the fraction of the wall time
    the server spent in batch ticks, peak residency."""
    from sams_amd import serve as S
    points = []
    for qps in qps_list:
        reqs = S.poisson_requests(qps, duration, seed)
        srv = S.Server(budget_bytes=budget, steps_per_tick=k)
        t0 = time.perf_counter()
        done = srv.run(reqs)
        wall = time.perf_counter() - t0
        out = S.summarize(done, wall)
        out.update({"qps": qps, "ticks": srv.ticks, "tick_busy_frac": srv.busy / wall if wall > 0 else None,
                    "peak_resident_grids": srv.peak_active, "peak_resident_bytes": srv.peak_used})
        points.append(out)
        log(f"bench: serving {qps} req/s: {out['requests']} requests in {wall:.1f} s, "
            f"{out['throughput'] / 1e9:.3f} G cell-updates/s, p50 {out['p50_latency_s']:.2f} s, "
            f"p99 {out['p99_latency_s']:.2f} s, busy {out['tick_busy_frac']:.0%}, peak {srv.peak_active} grids")
    best = max(points, key=lambda o: o["throughput"])
    return {"workload": f"Poisson arrivals for {duration} s at each rate (seed {seed}): n log-uniform in [32, 192] "
                        f"(n^3 grids), 50-500 steps, Orszag-Tang or Brio-Wu (resistive half the time), "
                        f"FP64 or FP32 storage 50/50",
            "hbm_budget_bytes": budget, "steps_per_tick": k, "unit": "cell-updates/s",
            "value": best["throughput"], "at_qps": best["qps"], "p50_latency_s": best["p50_latency_s"],
            "p99_latency_s": best["p99_latency_s"], "sweep": points}


def bench_fp32(n, steps):
    """FP32 storage vs FP64 on Orszag-Tang n^3 with 1e-3 velocity noise."""
    import numpy as np
    import sams_amd.lare as L
    res = {}
    runs = {}
    for name, prec in (("fp64", L.FP64), ("fp32", L.FP32_STORAGE)):
        g = L.Lare.from_problem(L.problem("ot", n=n, seed=7, amp=1e-3, precision=prec))
        g.steps(1, record_dt=False)
        g.sync()
        t0 = time.perf_counter()
        g.steps(steps - 1, record_dt=False)
        g.sync()
        el = time.perf_counter() - t0
        res[name + "_ms_per_step"] = 1e3 * el / (steps - 1)
        res[name + "_energy"] = list(g.energy())
        runs[name] = g
    err = {}
    for f in L.STATE_FIELDS:
        x = runs["fp64"].download(f)[2:-2, 2:-2, 2:-2]
        y = runs["fp32"].download(f)[2:-2, 2:-2, 2:-2]
        d = x - y
        nx2, nxi = float(np.sqrt(np.sum(x * x))), float(np.max(np.abs(x)))
        e2, ei = float(np.sqrt(np.sum(d * d))), float(np.max(np.abs(d)))
        err[f] = {"rel_l2": e2 / nx2 if nx2 > 0 else e2, "rel_linf": ei / nxi if nxi > 0 else ei}
        del x, y, d
    for g in runs.values():
        g.close()
    res.update({"workload": f"Orszag-Tang {n}^3 + N(0, 1e-3^2) velocity perturbation, {steps} steps",
                "error_vs_fp64": err, "speedup": res["fp64_ms_per_step"] / res["fp32_ms_per_step"]})
    log(f"bench: fp32 storage {n}^3: {res['fp32_ms_per_step']:.2f} ms/step vs fp64 {res['fp64_ms_per_step']:.2f}; "
        f"max rel L2 error {max(e['rel_l2'] for e in err.values()):.2e}")
    return res


def measured_traffic(kernel, n):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, scaled by a
    plain copy's reading) from the rocprofv3 counter runs summarised in
    profiles/ for this workload; None where no measurement matches."""
    path = os.path.join(ROOT, "profiles", f"bw{n}_resistive_traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and not line.startswith("#") and parts[0] == kernel:
                return int(parts[1]) + int(parts[2])
    return None


def kernel_roofline(dev, nsteps, cells, ms_step):
    """Per-kernel HIP-event timing on the grid's stream (after the timed
    region); the dominant kernel's algorithmic GB/s against the HBM peak."""
    prof = dev.profile(nsteps)
    dom = max(prof.items(), key=lambda kv: kv[1][1])
    name, (calls, tot_ms, tot_bytes) = dom
    avg_ms = tot_ms / calls
    achieved = (tot_bytes / calls) / (avg_ms * 1e-3) / 1e9
    for k, (c, t, b) in sorted(prof.items(), key=lambda kv: -kv[1][1])[:60]:
        log(f"  {k:20s} calls {c:4d} avg {t / c * 1e3:9.1f} us  share {t / sum(v[1] for v in prof.values()):6.1%}"
            f"  {b / c / (t / c * 1e-3) / 1e9 if t > 0 else 0:8.1f} GB/s")
    step_bytes = STEP_BYTES_RESISTIVE * cells
    traffic = measured_traffic(name, round(cells ** (1 / 3)))
    roofline = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "algorithmic_bytes_per_launch": round(tot_bytes / calls),
                "avg_launch_us": round(avg_ms * 1e3, 2),
                "step_algorithmic_GBs": round(step_bytes / (ms_step * 1e-3) / 1e9, 1),
                "step_frac": round(step_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    # every kernel of the step: launches per step, average launch, its
    # algorithmic GB/s and the measured HBM bytes per launch where profiled
    n = round(cells ** (1 / 3))
    roofline["kernels"] = [
        {"kernel": k, "launches_per_step": round(c / nsteps, 2), "avg_us": round(t / c * 1e3, 1),
         "algorithmic_GBs": round(b / c / (t / c * 1e-3) / 1e9, 1) if t > 0 and b > 0 else None,
         "traffic_bytes": measured_traffic(k, n)}
        for k, (c, t, b) in sorted(prof.items(), key=lambda kv: -kv[1][1])]
    return roofline


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--profile-steps", type=int, default=2)
    ap.add_argument("--cpu-steps", type=int, default=12)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=100)
    ap.add_argument("--batched-seq-steps", type=int, default=10)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--serve-qps", default="1,2,4,6",
                    help="comma-separated arrival rates of the serving sweep (requests/s)")
    ap.add_argument("--serve-duration", type=float, default=15.0)
    ap.add_argument("--serve-steps-per-tick", type=int, default=8)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--fp32-n", type=int, default=512)
    ap.add_argument("--fp32-steps", type=int, default=100)
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch configuration instead of tuning it before the timed steps")
    ap.add_argument("--tune", default=None,
                    help="a fixed launch configuration instead of tuning, as key=value pairs "
                         "(walk=1,march_k=16,zx_k=16,zline=1,order=1,ct_fuse=1)")
    ap.add_argument("--eager", action="store_true",
                    help="launch every kernel from the host instead of replaying the captured step graph")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="first run one plain device copy of this many MiB (a known-traffic kernel for "
                         "scaling hardware byte counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {a.gpus}); "
              "the LARE3D step is not sharded across GPUs here", file=sys.stderr)
        sys.exit(2)

    import sams_amd.lare as L
    if a.copy_probe_mb:
        import torch
        src = torch.ones(a.copy_probe_mb << 17, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"bench: copy probe {a.copy_probe_mb} MiB read + {a.copy_probe_mb} MiB written")
        del src, dst
    n = a.n
    p = L.problem("bw", n=n, seed=1, noise=0.01, eta=1e-3)
    log(f"bench: Brio-Wu {n}^3 resistive FP64, warmup {a.warmup}, steps {a.steps}")
    dev = L.Lare.from_problem(p)
    tune = None
    if a.tune:
        dev.set_tune(**{k: int(v) for k, v in (kv.split("=") for kv in a.tune.split(","))})
        tune = dev.get_tune()
        log(f"bench: launch configuration {tune}")
    elif not a.no_autotune:
        tune = dev.autotune(steps=3)  # state saved and restored around the candidates
        log(f"bench: autotuned launch configuration {tune}")
    if a.eager:
        dev.set_graph(False)
    if a.warmup:
        dev.steps(a.warmup, record_dt=False)
    dev.sync()
    t0 = time.perf_counter()
    dev.steps(a.steps, record_dt=False)
    dev.sync()
    el = time.perf_counter() - t0
    cells = n ** 3
    value = cells * a.steps / el
    ms_step = 1e3 * el / a.steps
    log(f"bench: {ms_step:.3f} ms/step, {value / 1e9:.3f} G cell-updates/s")

    roofline = None
    if a.profile_steps > 0:
        roofline = kernel_roofline(dev, a.profile_steps, cells, ms_step)

    cpu = None
    if not a.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
        try:
            cpu = cpu_baseline(n, a.cpu_steps, threads)
            log(f"bench: cpu baseline {cpu['value'] / 1e6:.2f} M cell-updates/s ({cpu['kind']}, {threads} threads)")
        except Exception as e:  # report, never fake
            log(f"bench: cpu baseline failed: {e}")

    extra = {}
    if not a.no_batched:
        extra["batched"] = bench_batched(a.batched_steps, a.batched_seq_steps)
    if not a.no_serving:
        extra["serving"] = bench_serving([float(q) for q in a.serve_qps.split(",")], a.serve_duration, 7, 200e9,
                                         a.serve_steps_per_tick)
    if not a.no_fp32:
        extra["fp32"] = bench_fp32(a.fp32_n, a.fp32_steps)

    out = {"metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"brio-wu {n}^3 resistive (eta=1e-3, 1% density noise), FP64",
                      "nx": n, "ny": n, "nz": n, "parallelism": "single-gpu"},
           "roofline": roofline, "cpu_baseline": cpu, "tune": tune}
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
