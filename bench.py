"""LARE3D Lagrangian-remap step on one MI355X: cell-updates/s.

Workload: the 256^3 resistive Brio-Wu problem (x-directed shock tube,
rho/p/By = 1/1/1 left and 0.125/0.1/-1 right, Bx = 0.75, gamma 2, 1% seeded
density noise, eta = 1e-3, zero-gradient x and periodic y/z boundaries), FP64,
inputs resident in HBM.  A step is one full LARE3D timestep
(startOfTimestep + halfTimestep + endOfTimestep).  Prints one JSON line.
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
    """The reference's own OpenMP path (oracle/_ref/ref_driver_fast, built from
    the reference sources) on the host cores.  The reference's resistive path
    throws (lagran.cpp:482), so its ideal-MHD twin of the workload is timed."""
    from oracle import oracle as O
    exe = O.REF_DRIVER_FAST
    if os.path.exists(exe):
        with tempfile.TemporaryDirectory() as td:
            meta = O.run_reference(td, problem="bw", nx=n, steps=steps, noise=0.01, seed=1,
                                   threads=threads, fast=True)
        el = meta["elapsed_s"]
        return {"value": n ** 3 * steps / el, "unit": "cell-updates/s", "cores": threads,
                "kind": "reference",
                "sample": f"reference OpenMP LARE3D (-O3 -march=x86-64-v3), Brio-Wu {n}^3 ideal-MHD twin "
                          f"(eta=0: the reference throws when resistive), {steps} steps, "
                          f"step loop {el:.2f} s"}
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--profile-steps", type=int, default=2)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {a.gpus}); "
              "the LARE3D step is not sharded across GPUs here", file=sys.stderr)
        sys.exit(2)

    import sams_amd.lare as L
    n = a.n
    p = L.problem("bw", n=n, seed=1, noise=0.01, eta=1e-3)
    log(f"bench: Brio-Wu {n}^3 resistive FP64, warmup {a.warmup}, steps {a.steps}")
    dev = L.Lare.from_problem(p)
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

    # per-kernel HIP-event timing on the grid's stream (after the timed region)
    prof = dev.profile(a.profile_steps)
    dom = max(prof.items(), key=lambda kv: kv[1][1])
    name, (calls, tot_ms, tot_bytes) = dom
    avg_ms = tot_ms / calls
    achieved = (tot_bytes / calls) / (avg_ms * 1e-3) / 1e9
    for k, (c, t, b) in sorted(prof.items(), key=lambda kv: -kv[1][1])[:12]:
        log(f"  {k:20s} calls {c:4d} avg {t / c * 1e3:9.1f} us  share {t / sum(v[1] for v in prof.values()):6.1%}"
            f"  {b / c / (t / c * 1e-3) / 1e9 if t > 0 else 0:8.1f} GB/s")
    step_bytes = STEP_BYTES_RESISTIVE * cells
    roofline = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "avg_launch_us": round(avg_ms * 1e3, 2),
                "step_algorithmic_GBs": round(step_bytes / (ms_step * 1e-3) / 1e9, 1),
                "step_frac": round(step_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}

    cpu = None
    if not a.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
        try:
            cpu = cpu_baseline(n, a.cpu_steps, threads)
            log(f"bench: cpu baseline {cpu['value'] / 1e6:.2f} M cell-updates/s ({cpu['kind']}, {threads} threads)")
        except Exception as e:  # report, never fake
            log(f"bench: cpu baseline failed: {e}")

    out = {"metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"brio-wu {n}^3 resistive (eta=1e-3, 1% density noise), FP64",
                      "nx": n, "ny": n, "nz": n, "parallelism": "single-gpu"},
           "roofline": roofline, "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
