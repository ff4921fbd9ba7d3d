"""Batches: grids of uneven sizes, problems, boundary conditions and physics
(ideal and resistive, with and without the kinetic-energy correction)
stepped together by one launch per kernel give, per grid, the same bits as
the grid stepped alone and as the oracle."""
import numpy as np
import pytest

import sams_amd.lare as L
from lare_testlib import bit_diff, make_oracle

pytestmark = pytest.mark.gpu

MIX = [
    L.problem("ot", n=12, ny=9, nz=14, seed=3, amp=0.04),
    L.problem("bw", n=20, ny=4, nz=5, seed=4, noise=0.01, eta=1e-3),
    L.problem("ot", n=17, seed=9, amp=0.02, beta=2.0),
    L.problem("bw", n=16, ny=5, nz=3, seed=11, noise=0.02, bcx=L.BC_CLAMP),
    L.problem("bw", n=16, ny=4, nz=4, bcx=L.BC_MIRROR),
    L.problem("ot", n=8, ny=10, nz=6, seed=1, amp=0.05),
    L.problem("ot", n=11, ny=13, nz=9, seed=6, amp=0.08, rke=True),
    L.problem("bw", n=18, ny=4, nz=6, seed=8, noise=0.02, eta=1e-3, rke=True),
]


def _diffs(a, b):
    out = {}
    for f in L.STATE_FIELDS:
        d = bit_diff(a.download(f), b.download(f))
        if d:
            out[f] = d
    return out


@pytest.mark.parametrize("graph", [True, False])
def test_batch_matches_each_grid_alone(graph):
    steps = 7
    mem = [L.Lare.from_problem(p) for p in MIX]
    alone = [L.Lare.from_problem(p) for p in MIX]
    b = L.Batch(mem)
    b.set_graph(graph)
    b.steps(steps)
    b.sync()
    for i, a in enumerate(alone):
        dts = a.steps(steps)
        assert mem[i].get_dt().hex() == dts[-1].hex(), f"member {i} dt"
        d = _diffs(mem[i], a)
        assert not d, f"member {i}: {d}"


def test_batch_matches_oracle_and_continues():
    """Two batch calls in a row, then the members continue on their own."""
    sel = [MIX[0], MIX[1], MIX[4], MIX[6], MIX[7]]
    mem = [L.Lare.from_problem(p) for p in sel]
    orc = [make_oracle(p) for p in sel]
    b = L.Batch(mem)
    b.steps(3)
    b.steps(2)
    b.sync()
    for o in orc:
        for _ in range(5):
            o.step()
    for i in range(len(sel)):
        assert mem[i].get_dt().hex() == orc[i].get_dt().hex(), f"member {i} dt"
        assert not _diffs(mem[i], orc[i]), f"member {i}"
    for i in range(len(sel)):
        assert mem[i].step().hex() == orc[i].step().hex()
        assert not _diffs(mem[i], orc[i]), f"member {i} after a solo step"


def test_energy_diagnostic_matches_host_sum():
    """lare_energy against the same sums on the host in numpy (FP64, a
    different summation order: relative 1e-12)."""
    p = L.problem("bw", n=18, ny=6, nz=5, seed=2, noise=0.02)
    g = L.Lare.from_problem(p)
    g.steps(3)
    e = g.energy()
    s = g.state()
    d = L.problem_desc(p)
    nx, ny, nz = 18, 6, 5
    dx = (d.x_max - d.x_min) / nx
    dy = (d.y_max - d.y_min) / ny
    dz = (d.z_max - d.z_min) / nz
    cv = dx * dy * dz

    def c(a, lo):  # cells 1..n of an array whose first index is lo
        return a[1 - lo:1 - lo + nx, 1 - lo:1 - lo + ny, 1 - lo:1 - lo + nz]

    rho = s["rho"]
    internal = np.sum(c(rho, -1) * (c(s["energy_electron"], -1) + c(s["energy_ion"], -1))) * cv
    # vertex density at vertices 1..n: mean of the 8 cells (uniform cv)
    r8 = sum(rho[1 + i + 1:1 + i + 1 + nx, 1 + j + 1:1 + j + 1 + ny, 1 + k + 1:1 + k + 1 + nz]
             for i in (0, 1) for j in (0, 1) for k in (0, 1)) / 8.0
    v2 = c(s["vx"], -2) ** 2 + c(s["vy"], -2) ** 2 + c(s["vz"], -2) ** 2
    kinetic = np.sum(0.5 * r8 * v2) * cv
    bx = 0.5 * (s["bx"][3:3 + nx, 2:2 + ny, 2:2 + nz] + s["bx"][2:2 + nx, 2:2 + ny, 2:2 + nz])
    by = 0.5 * (s["by"][2:2 + nx, 3:3 + ny, 2:2 + nz] + s["by"][2:2 + nx, 2:2 + ny, 2:2 + nz])
    bz = 0.5 * (s["bz"][2:2 + nx, 2:2 + ny, 3:3 + nz] + s["bz"][2:2 + nx, 2:2 + ny, 2:2 + nz])
    magnetic = np.sum(bx * bx + by * by + bz * bz) / (2.0 * d.mu0) * cv
    want = (internal, kinetic, magnetic, internal + kinetic + magnetic)
    for k in range(4):
        assert e[k] == pytest.approx(want[k], rel=1e-12), k
