"""The continuous-batching server: every request served in a mix (admitted
while others run, sharing batches with grids of other sizes, problems and
boundary conditions) ends bit-identical to the same solve run alone; the
HBM budget bounds what is resident at once."""
import pytest

import sams_amd.lare as L
from sams_amd import serve as S
from lare_testlib import bit_diff, make_oracle

pytestmark = pytest.mark.gpu


def _requests():
    specs = [
        (L.problem("ot", n=12, seed=3, amp=0.02), 17, 0.00),
        (L.problem("bw", n=24, ny=4, nz=4, seed=5, noise=0.01, eta=1e-3), 9, 0.00),
        (L.problem("ot", n=10, ny=14, nz=8, seed=8, amp=0.04, precision=L.FP32_STORAGE), 13, 0.00),
        (L.problem("bw", n=20, ny=5, nz=4, seed=2, noise=0.02, bcx=L.BC_MIRROR), 21, 0.05),
        (L.problem("bw", n=16, ny=4, nz=4, seed=6, eta=1e-3, precision=L.FP32_STORAGE), 7, 0.05),
        (L.problem("ot", n=14, seed=4, beta=0.5), 11, 0.10),
    ]
    return [S.Request(i, p, st, t) for i, (p, st, t) in enumerate(specs)]


def _check_alone(done):
    for r in done:
        alone = L.Lare.from_problem(r.problem)
        dts = alone.steps(r.steps)
        assert dts[-1].hex() == r.dt_last.hex(), f"request {r.rid} dt"
        for f in L.STATE_FIELDS:
            assert bit_diff(r.grid.download(f), alone.download(f)) is None, (r.rid, f)


def test_served_results_equal_solo_runs():
    srv = S.Server(budget_bytes=50e9, steps_per_tick=4, keep_results=True)
    done = srv.run(_requests())
    assert len(done) == 6
    assert srv.peak_active >= 3  # the first three arrive together and share the device
    _check_alone(done)
    summ = S.summarize(done, max(r.finished for r in done))
    assert summ["p99_latency_s"] >= summ["p50_latency_s"] > 0


def test_budget_limits_residency():
    reqs = _requests()
    need = sorted(S.device_bytes(r.problem) for r in reqs)
    budget = need[-1] + need[-2]  # room for the two largest at most
    srv = S.Server(budget_bytes=budget, steps_per_tick=3, keep_results=True)
    done = srv.run(reqs)
    assert len(done) == 6
    assert srv.peak_used <= budget
    _check_alone(done)


def test_served_results_at_lane_walking_sizes_equal_oracle():
    """Requests of 32^3 to 56^3 (the serving workload's lower sizes): the
    first three FP64 requests arrive together, so their batch holds 349 568
    cells and runs the lane-walking group kernels (300 000 cells and up);
    later ticks mix other sizes and precisions.  Every FP64 result is
    bit-identical to the oracle (dt and every state field), and every
    result to the same solve run alone."""
    from oracle import oracle as O
    specs = [
        (L.problem("ot", n=48, seed=31, amp=0.03), 12, 0.00),
        (L.problem("bw", n=56, seed=32, noise=0.01, eta=1e-3), 10, 0.00),
        (L.problem("ot", n=44, ny=40, nz=36, seed=33, amp=0.02, beta=0.4, eta=0.01, eta0=0.05, j_max=2.0), 9, 0.00),
        (L.problem("bw", n=40, seed=34, noise=0.02, bcx=L.BC_MIRROR, rke=True), 14, 0.02),
        (L.problem("ot", n=36, seed=35, amp=0.05, precision=L.FP32_STORAGE), 8, 0.02),
        (L.problem("bw", n=32, seed=36, noise=0.01, eta=1e-3, precision=L.FP32_STORAGE), 11, 0.04),
    ]
    assert sum(int(p.nx * p.ny * p.nz) for p, _, _ in specs[:3]) >= 300000
    reqs = [S.Request(i, p, st, t) for i, (p, st, t) in enumerate(specs)]
    srv = S.Server(budget_bytes=50e9, steps_per_tick=3, keep_results=True)
    done = srv.run(reqs)
    assert len(done) == len(specs)
    _check_alone(done)
    O.set_threads(16)
    for r in done:
        if r.problem.precision != L.FP64:
            continue
        o = make_oracle(r.problem)
        for _ in range(r.steps):
            o.step()
        assert o.get_dt().hex() == r.dt_last.hex(), f"request {r.rid} dt against the oracle"
        for f in L.STATE_FIELDS:
            assert bit_diff(r.grid.download(f), o.download(f)) is None, (r.rid, f)
        print(f"request {r.rid}: equal to the oracle", flush=True)
