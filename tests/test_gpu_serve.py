"""The continuous-batching server: every request served in a mix (admitted
while others run, sharing batches with grids of other sizes, problems and
boundary conditions) ends bit-identical to the same solve run alone; the
HBM budget bounds what is resident at once."""
import pytest

import sams_amd.lare as L
from sams_amd import serve as S
from lare_testlib import bit_diff

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
