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


# A batch above the lane-walking threshold (300000 cells per launch): with the
# default launch configuration every group launch of this batch, including
# the resistive and the kinetic-energy-correction subsets (about 340000 and
# 306000 cells), takes the lane-walking forms, while each grid alone (at most
# 118000 cells) runs the one-cell forms.
BIG = [
    L.problem("ot", n=40, seed=21, amp=0.03),
    L.problem("bw", n=72, ny=40, nz=36, seed=22, noise=0.01, eta=1e-3),
    L.problem("ot", n=48, ny=56, nz=44, seed=23, amp=0.02, beta=0.5, eta=0.01, eta0=0.05, j_max=2.0),
    L.problem("bw", n=64, ny=40, nz=40, seed=24, noise=0.02, bcx=L.BC_MIRROR, rke=True),
    L.problem("ot", n=44, seed=25, amp=0.05, rke=True),
    L.problem("bw", n=56, ny=44, nz=48, seed=26, noise=0.01, eta=1e-3, rke=True),
]


def _check_batch_against_alone_and_oracle(probs, steps, tune=None, oracle_members=None):
    from oracle import oracle as O
    mem = [L.Lare.from_problem(p) for p in probs]
    b = L.Batch(mem)
    if tune:
        b.set_tune(**tune)
    b.steps(steps)
    b.sync()
    print("batch tune:", b.get_tune(), flush=True)
    for i, p in enumerate(probs):
        alone = L.Lare.from_problem(p)
        dts = alone.steps(steps)
        assert mem[i].get_dt().hex() == dts[-1].hex(), f"member {i} dt against its solo run"
        d = _diffs(mem[i], alone)
        assert not d, f"member {i} against its solo run: {d}"
        alone.close()
    O.set_threads(16)
    for i in (oracle_members if oracle_members is not None else range(len(probs))):
        o = make_oracle(probs[i])
        for _ in range(steps):
            o.step()
        assert mem[i].get_dt().hex() == o.get_dt().hex(), f"member {i} dt against the oracle"
        d = _diffs(mem[i], o)
        assert not d, f"member {i} against the oracle: {d}"
        print(f"member {i}: equal to its solo run and to the oracle", flush=True)


def test_large_batch_lane_walking_group_kernels_match_alone_and_oracle():
    """Default configuration above the threshold: the lane-walking group
    kernels over six uneven grids (ideal, resistive with and without the
    eta0/j_max branch, rke, mirror edges) against each grid's one-cell solo
    run and the oracle."""
    _check_batch_against_alone_and_oracle(BIG, 4)


def test_small_batch_forced_lane_walking_matches_alone_and_oracle():
    """walk=1 forces the lane-walking forms on the small mixed batch."""
    _check_batch_against_alone_and_oracle(MIX, 7, tune={"walk": 1}, oracle_members=[0, 1, 3, 4, 6, 7])


def test_large_batch_one_cell_forms_match_alone():
    """walk=2 keeps the one-cell forms above the threshold too."""
    _check_batch_against_alone_and_oracle(BIG[:3], 3, tune={"walk": 2}, oracle_members=[])


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
