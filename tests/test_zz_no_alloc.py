"""lare_energy_noalloc, the energy diagnostic that allocates no device
memory, against lare_energy through the C-ABI: the same four sums, bit for
bit, on two calls in a row (the second on a smaller grid, so it reads less
of the shared scratch array than the first left behind), on FP64 and FP32
storage, and a clear error for a grid beyond the scratch array's size."""
import numpy as np
import pytest

import sams_amd.lare as L

pytestmark = pytest.mark.gpu


def _same(a, b):
    return np.array(a, dtype=np.float64).view(np.uint64).tolist() == np.array(b, dtype=np.float64).view(np.uint64).tolist()


@pytest.mark.parametrize("precision", [L.FP64, L.FP32_STORAGE])
def test_energy_noalloc_equals_energy_on_two_calls(precision):
    big = L.Lare.from_problem(L.problem("ot", n=96, ny=80, nz=72, seed=3, amp=0.03, precision=precision))
    small = L.Lare.from_problem(L.problem("bw", n=24, ny=6, nz=5, seed=4, noise=0.01, eta=1e-3,
                                          precision=precision))
    big.steps(3)
    small.steps(2)
    for g, name in ((big, "96x80x72"), (small, "24x6x5")):
        want = g.energy()
        got = g.energy_noalloc()
        print(name, got, flush=True)
        assert _same(got, want), (name, got, want)
        assert _same(g.energy_noalloc(), want), name  # a repeat reads the same
    # the first grid again after the smaller one
    assert _same(big.energy_noalloc(), big.energy())


def test_energy_noalloc_rejects_a_grid_beyond_its_scratch():
    g = L.Lare(L.problem_desc(L.problem("ot", n=580, ny=580, nz=584, precision=L.FP32_STORAGE)))
    with pytest.raises(RuntimeError, match="bytes of scratch"):
        g.energy_noalloc()
    g.close()
