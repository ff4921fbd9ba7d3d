"""Continuous-batching service for LARE3D solves on one GPU.

A request is one MHD solve: a problem (Orszag-Tang or Brio-Wu with its
parameters, grid size, storage precision) and a step count.  The server keeps
every admitted grid resident in HBM and advances all active grids together:
each tick runs a few steps of one batch per storage precision (every kernel
one launch over the batch, lare_batch_*), then retires the grids that reached
their step count and admits waiting requests, first come first served, while
their device memory fits the HBM budget.  Requests arriving while others run
join at the next tick.  A grid's result does not depend on which grids share
its batch (per-grid arithmetic, per-grid dt), so a solve served in a mix
equals the same solve run alone, bit for bit.

The reference has no serving layer; this is the request loop around the
step, in the shape of the runner's time loop (runner.h:564-586) applied to
many independent simulations.
"""
import concurrent.futures
import dataclasses
import math
import time

import numpy as np

from . import lare as L


@dataclasses.dataclass(eq=False)
class Request:
    rid: int
    problem: object          # L.LareProblem
    steps: int
    arrival: float           # seconds from the start of the run
    # filled by the server
    admitted: float = math.nan
    finished: float = math.nan
    grid: object = None
    done_steps: int = 0
    dt_last: float = math.nan
    prefill: object = None   # future of the host initial conditions

    @property
    def cells(self):
        p = self.problem
        return int(p.nx) * int(p.ny) * int(p.nz)

    @property
    def latency(self):
        return self.finished - self.arrival


def device_bytes(p):
    """HBM held by one grid of problem p: 43 state/work arrays in the storage
    type and 6 FP64 flux arrays over the padded box (7 with the kinetic-energy
    correction), each rounded up to 256 bytes, plus axes and scalars (the
    pool of lare_ctx_create with its default layout), the field-transfer
    staging buffer ((nx+5)(ny+5)(nz+5) doubles, held from the first upload
    until the grid is destroyed) and 64 KiB for the launch tables."""
    pz = ((int(p.nz) + 5) + 7) // 8 * 8
    box = (int(p.nx) + 5) * (int(p.ny) + 5) * pz
    t = 4 if p.precision == L.FP32_STORAGE else 8
    lmax = max(int(p.nx), int(p.ny), int(p.nz)) + 5

    def arr(nbytes):
        return (nbytes + 255) // 256 * 256

    stage = arr((int(p.nx) + 5) * (int(p.ny) + 5) * (int(p.nz) + 5) * 8)
    return 43 * arr(box * t) + (7 if p.rke else 6) * arr(box * 8) + 6 * lmax * 8 + 256 + stage + 65536


def poisson_requests(qps, duration, seed, nmin=32, nmax=192, smin=50, smax=500):
    """Seeded Poisson stream: sizes log-uniform in [nmin, nmax]^3, step
    counts uniform in [smin, smax], Orszag-Tang or Brio-Wu and FP64 or FP32
    storage 50/50, seeded problem parameters."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / qps)
        if t >= duration:
            break
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        steps = int(rng.integers(smin, smax + 1))
        prec = L.FP64 if rng.random() < 0.5 else L.FP32_STORAGE
        if rng.random() < 0.5:
            p = L.problem("ot", n=n, seed=int(rng.integers(1, 2**31)), amp=float(rng.uniform(0.0, 0.05)),
                          beta=float(rng.uniform(0.1, 10.0)), precision=prec)
        else:
            # x-directed shock tube on the n^3 grid, resistive with
            # probability 1/2
            eta = 1e-3 if rng.random() < 0.5 else 0.0
            p = L.problem("bw", n=n, seed=int(rng.integers(1, 2**31)),
                          noise=float(rng.uniform(0.0, 0.02)), eta=eta, precision=prec)
        out.append(Request(len(out), p, steps, t))
    return out


class Server:
    """Admission by HBM budget, continuous batching, per-request latency."""

    def __init__(self, budget_bytes=200e9, steps_per_tick=8, clock=time.perf_counter, keep_results=False,
                 prefill_ahead=4):
        self.budget = float(budget_bytes)
        self.k = int(steps_per_tick)
        self.clock = clock
        self.keep = keep_results
        self.used = 0.0
        self.active = []
        self.waiting = []
        self.done = []
        self.ticks = 0
        self.cell_updates = 0.0
        self.peak_active = 0
        self.peak_used = 0.0
        self.busy = 0.0  # seconds from a tick's first enqueue to its last sync
        self._batches = {}
        self._members = {}
        # host initial conditions of the next waiting requests are computed
        # on worker threads while the GPU steps (the fill releases the GIL)
        self.ahead = int(prefill_ahead)
        self._pool = concurrent.futures.ThreadPoolExecutor(max_workers=2)

    @staticmethod
    def _fill(p):
        return {f: L.problem_fill(p, f) for f in L.STATE_FIELDS}

    def _prefetch(self):
        for r in self.waiting[:self.ahead]:
            if r.prefill is None:
                r.prefill = self._pool.submit(self._fill, r.problem)

    def _admit(self, now):
        while self.waiting:
            r = self.waiting[0]
            need = device_bytes(r.problem)
            if need > self.budget:
                raise ValueError(f"request {r.rid} needs {need / 1e9:.1f} GB, above the whole budget")
            if self.used + need > self.budget:
                break
            self.waiting.pop(0)
            if r.prefill is None:
                r.prefill = self._pool.submit(self._fill, r.problem)
            host = r.prefill.result()
            r.prefill = None
            g = L.Lare(L.problem_desc(r.problem))
            g.set_problem_bcs(r.problem)
            for f, a in host.items():
                g.upload(f, a)
            r.grid = g
            r.admitted = now
            self.used += need
            self.active.append(r)
            self.peak_active = max(self.peak_active, len(self.active))
            self.peak_used = max(self.peak_used, self.used)

    def _batch(self, prec, members):
        key = tuple(id(r) for r in members)
        if self._members.get(prec) != key:
            old = self._batches.pop(prec, None)
            if old is not None:
                old.close()
            self._batches[prec] = L.Batch([r.grid for r in members])
            self._members[prec] = key
        return self._batches[prec]

    def run(self, requests, time_scale=1.0):
        """Serve `requests`; arrivals are released against the wall clock
        (time_scale < 1 compresses the schedule).  While a tick's steps run
        on the GPU, the host admits waiting requests (initial conditions and
        upload, on each new grid's own stream); they join the next tick.
        Returns the finished requests."""
        pending = sorted(requests, key=lambda r: r.arrival)
        t0 = self.clock()
        i = 0

        def release():
            nonlocal i
            now = self.clock() - t0
            while i < len(pending) and pending[i].arrival * time_scale <= now:
                self.waiting.append(pending[i])
                i += 1
            self._prefetch()
            return now

        while i < len(pending) or self.waiting or self.active:
            self._admit(release())
            if not self.active:
                if i < len(pending):
                    time.sleep(max(0.0, pending[i].arrival * time_scale - (self.clock() - t0)))
                continue
            # steps this tick: at most k, and no grid past its step count
            running = list(self.active)
            k = min(self.k, min(r.steps - r.done_steps for r in running))
            groups = {}
            for r in running:
                groups.setdefault(int(r.problem.precision), []).append(r)
            tick0 = self.clock()
            for prec, mem in groups.items():
                self._batch(prec, mem).steps(k)
            # overlap: admissions while the GPU works
            self._admit(release())
            for prec in groups:
                self._batches[prec].sync()
            self.ticks += 1
            now = self.clock() - t0
            self.busy += self.clock() - tick0
            finished = []
            for r in running:
                r.done_steps += k
                self.cell_updates += r.cells * k
                if r.done_steps >= r.steps:
                    finished.append(r)
            # a batch must not outlive its members: drop the batches that
            # hold a finished grid before releasing the grid
            for prec in {int(r.problem.precision) for r in finished}:
                self._batches.pop(prec).close()
                self._members.pop(prec)
            for r in finished:
                r.finished = now
                r.dt_last = r.grid.get_dt()
                if not self.keep:
                    r.grid.close()
                    r.grid = None
                self.used -= device_bytes(r.problem)
                self.done.append(r)
            gone = {id(r) for r in finished}
            self.active = [r for r in self.active if id(r) not in gone]
        for b in self._batches.values():
            b.close()
        self._batches.clear()
        self._members.clear()
        self._pool.shutdown(wait=True)
        return self.done


def summarize(done, wall):
    lat = np.array([r.latency for r in done])
    cu = float(sum(r.cells * r.steps for r in done))
    return {"requests": len(done), "wall_s": wall, "cell_updates": cu,
            "throughput": cu / wall if wall > 0 else math.nan,
            "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else math.nan,
            "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else math.nan,
            "mean_latency_s": float(lat.mean()) if len(lat) else math.nan}
