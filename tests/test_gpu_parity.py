"""Bit-exact parity of the MI355X kernels with the oracle and the reference.

Every test drives the device through the C-ABI (sams_amd/liblare_hip.so) and
compares with the CPU restatement (oracle/) on the same seeded inputs, or
with the reference's own dt sequences and state digests
(tests/golden/goldens.json).  The bar is bit-identity: FP64 fields with equal
bits, dt sequences with equal bits."""
import numpy as np
import pytest

import sams_amd.lare as L
from lare_testlib import bit_diff, case_bcs, case_problem, goldens, make_oracle, sha
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = goldens()


def _state_diffs(dev, orc, fields=L.STATE_FIELDS):
    out = {}
    for f in fields:
        d = bit_diff(dev.download(f), orc.download(f))
        if d:
            out[f] = d
    return out


def _skip_set0_rke(c, kernel_set):
    if kernel_set == 0 and (c.get("rke") or c.get("lare_bc")):
        pytest.skip("the kinetic-energy correction and map conditions run on kernel sets 1-3 only")
    if kernel_set != 0 and c.get("geometry", "cartesian") != "cartesian":
        pytest.skip("cylindrical and spherical grids run kernel set 0 only")


@pytest.mark.parametrize("kernel_set", [0, 1, 2])
@pytest.mark.parametrize("name", ["ot_pert_10x8x12", "bw_zerograd_24x4x6", "bw_walls_16x5x3",
                                  "bw_mirror_16x4x4", "ot_rke_12", "bw_rke_mirror_16x5x4", "ot_larebc_rke_12",
                                  "bw_larebc_16x5x4"])
def test_events_of_first_two_steps_match_oracle(name, kernel_set):
    """Event by event (startOfTimestep, halfTimestep, endOfTimestep) for two
    steps: dt, the half-step velocities and every state field; for the
    one-kernel-per-reference-kernel set (0), the fused set (1: one cell per
    lane at these sizes) and the fused set with lane-walking kernels (2);
    the rke cases add the kinetic-energy correction to endOfTimestep."""
    _skip_set0_rke(CASES[name], kernel_set)
    p = case_problem(CASES[name])
    dev = L.Lare.from_problem(p)
    dev.set_kernel_set(kernel_set)
    case_bcs(CASES[name], dev)
    orc = case_bcs(CASES[name], make_oracle(p))
    for step in (1, 2):
        dt_d = dev.start_of_timestep()
        dt_o = orc.start_of_timestep()
        assert dt_d.hex() == dt_o.hex(), f"step {step} dt"
        diffs = _state_diffs(dev, orc, L.STATE_FIELDS + ["vx1", "vy1", "vz1"])
        assert not diffs, f"step {step} after startOfTimestep: {diffs}"
        dev.half_timestep()
        orc.half_timestep()
        diffs = _state_diffs(dev, orc)
        assert not diffs, f"step {step} after halfTimestep: {diffs}"
        dev.end_of_timestep()
        orc.end_of_timestep()
        diffs = _state_diffs(dev, orc)
        assert not diffs, f"step {step} after endOfTimestep: {diffs}"


@pytest.mark.parametrize("kernel_set", [0, 1, 2])
@pytest.mark.parametrize("name", sorted(CASES))
def test_reference_goldens(name, kernel_set):
    """Whole runs against the reference: the dt sequence and state digests."""
    c = CASES[name]
    _skip_set0_rke(c, kernel_set)
    dev = L.Lare.from_problem(case_problem(c))
    dev.set_kernel_set(kernel_set)
    case_bcs(c, dev)
    snaps = c["snapshots"]
    done = 0
    for s in sorted(int(k) for k in snaps):
        if s > done:
            dts = dev.steps(s - done)
            for i, dt in enumerate(dts):
                assert dt.hex() == float.fromhex(c["dt"][done + i]).hex(), f"dt of step {done + i + 1}"
            done = s
        for f in L.STATE_FIELDS:
            assert sha(dev.download(f)) == snaps[str(s)][f]["sha256"], f"{f} after step {s}"


def test_oracle_parity_random_perturbed_orszag_tang():
    """A strongly perturbed state (velocity noise 0.2) on a 20x18x16 grid for
    15 steps: every state field and dt bit-identical with the oracle."""
    p = L.problem("ot", n=20, ny=18, nz=16, seed=99, amp=0.2)
    dev = L.Lare.from_problem(p)
    orc = make_oracle(p)
    for s in range(15):
        assert dev.step().hex() == orc.step().hex(), f"dt of step {s + 1}"
    assert not _state_diffs(dev, orc)


@pytest.mark.parametrize("kernel_set", [0, 1, 2])
def test_resistive_brio_wu_matches_oracle(kernel_set):
    """Resistive MHD (eta = 1e-3): two half-dt substeps per step."""
    p = L.problem("bw", n=24, ny=5, nz=4, seed=4, noise=0.01, eta=1e-3)
    dev = L.Lare.from_problem(p)
    dev.set_kernel_set(kernel_set)
    orc = make_oracle(p)
    for s in range(6):
        assert dev.step().hex() == orc.step().hex(), f"dt of step {s + 1}"
    assert not _state_diffs(dev, orc)


def test_kinetic_energy_correction_changes_only_the_energies():
    """rke on and off from the same state for one step: rho, v and B are
    bit-identical (the correction only heats), e_e and e_i differ and never
    decrease, and the device matches the oracle in both runs."""
    runs = {}
    for rke in (False, True):
        p = L.problem("ot", n=14, ny=12, nz=10, seed=5, amp=0.1, rke=rke)
        dev = L.Lare.from_problem(p)
        orc = make_oracle(p)
        for s in range(3):
            assert dev.step().hex() == orc.step().hex(), f"rke={rke} dt of step {s + 1}"
        assert not _state_diffs(dev, orc), f"rke={rke}"
        runs[rke] = dev
    # same first steps: the correction enters the state only through the energies
    p0 = L.problem("ot", n=14, ny=12, nz=10, seed=5, amp=0.1)
    a, b = L.Lare.from_problem(p0), L.Lare.from_problem(L.problem("ot", n=14, ny=12, nz=10, seed=5, amp=0.1,
                                                                    rke=True))
    a.step()
    b.step()
    for f in ("rho", "vx", "vy", "vz", "bx", "by", "bz"):
        assert bit_diff(a.download(f), b.download(f)) is None, f
    for f in ("energy_electron", "energy_ion"):
        ea, eb = a.download(f)[2:-2, 2:-2, 2:-2], b.download(f)[2:-2, 2:-2, 2:-2]
        assert np.all(eb >= ea) and np.any(eb > ea), f


def test_uniform_state_first_dt_closed_form():
    """A uniform state at rest: the first dt is 0.8*dx/sqrt(gamma p/rho +
    B^2/(mu0 rho)) (lagran.cpp:410 with p_visc = 0; dt2..dt4 are cv/1e-10cv)."""
    p = L.problem("ot", n=16)
    d = L.problem_desc(p)
    dev = L.Lare(d)
    rho, pr, b = 1.3, 0.7, 0.4
    g = d.gas_gamma
    for f in L.STATE_FIELDS:
        shape = dev.shape(f)
        v = {"rho": rho, "energy_electron": pr / ((g - 1.0) * rho) / 2.0,
             "energy_ion": pr / ((g - 1.0) * rho) / 2.0, "bx": b}.get(f, 0.0)
        dev.upload(f, np.full(shape, v))
    dt = dev.start_of_timestep()
    dx = 1.0 / 16
    p_e = (g - 1.0) * rho * (pr / ((g - 1.0) * rho) / 2.0)
    press = p_e + p_e
    expect = 0.8 * (dx / (np.sqrt(0.0) + np.sqrt(g * press / rho + (b * b) / 1.0 / rho + 0.0)))
    assert dt == pytest.approx(expect, rel=1e-15)


@pytest.mark.parametrize("kernel_set", [1, 2])
def test_graph_replays_match_eager_steps(kernel_set):
    """The captured-step HIP graph, replayed with host synchronisation and
    downloads between replays, gives after every replay the same bits as the
    eagerly launched step and as the oracle."""
    p = L.problem("bw", n=20, ny=6, nz=5, seed=8, noise=0.02, eta=1e-3)
    g = L.Lare.from_problem(p)
    e = L.Lare.from_problem(p)
    for d in (g, e):
        d.set_kernel_set(kernel_set)
    e.set_graph(False)
    orc = make_oracle(p)
    for s in range(5):
        dg = g.steps(1)[0]
        de = e.steps(1)[0]
        do = orc.step()
        assert dg.hex() == de.hex() == do.hex(), f"dt of replay {s + 1}"
        assert not _state_diffs(g, e), f"replay {s + 1} vs eager"
        assert not _state_diffs(g, orc), f"replay {s + 1} vs oracle"


@pytest.mark.parametrize("kernel_set", [0, 1, 2])
def test_split_start_of_timestep_matches_whole_event(kernel_set):
    """lare_viscosity_dt followed by lare_lagrange with the dt it returned
    (the runner's timestep protocol, runner.h:158-169) gives the same bits
    as lare_start_of_timestep, on the resistive path where the substep count
    int(dt/dtr)+1 is computed on the host."""
    p = L.problem("bw", n=20, ny=4, nz=6, seed=5, noise=0.02, eta=1e-3)
    split = L.Lare.from_problem(p)
    whole = L.Lare.from_problem(p)
    for d in (split, whole):
        d.set_kernel_set(kernel_set)
    for s in range(4):
        dt, dtr = split.viscosity_dt()
        assert dtr == dt
        split.lagrange(dt)
        dt_w = whole.start_of_timestep()
        assert dt.hex() == dt_w.hex(), f"dt of step {s + 1}"
        assert not _state_diffs(split, whole, L.STATE_FIELDS + ["vx1", "vy1", "vz1"])
        for d in (split, whole):
            d.half_timestep()
            d.end_of_timestep()
        assert not _state_diffs(split, whole), f"step {s + 1}"


@pytest.mark.parametrize("kernel_set", [2, 3])
@pytest.mark.parametrize("ratio", [0.5, 2.0])
def test_odd_substep_counts_match_kernel_set0(ratio, kernel_set):
    """A runner dt other than the package's own (the minimum over packages,
    runner.h:158-169) changes the resistive substep count int(dt/dtr)+1:
    1 substep for dt = dtr/2, 3 for dt = 2 dtr.  The fused sets (2: B
    alternating with bx1..bz1 in one kernel per substep; 3: one cell per
    lane) must equal the one-kernel-per-reference-kernel set, whose substeps
    follow lagran.cpp:479-500 literally (its events are pinned to the oracle
    above)."""
    p = L.problem("bw", n=20, ny=4, nz=6, seed=9, noise=0.02, eta=1e-3)
    devs = [L.Lare.from_problem(p) for _ in range(2)]
    for k, d in zip((0, kernel_set), devs):
        d.set_kernel_set(k)
    for s in range(2):
        dts = [d.viscosity_dt() for d in devs]
        assert dts[0][0].hex() == dts[1][0].hex(), f"dt of step {s + 1}"
        for d in devs:
            d.lagrange(dts[0][0] * ratio)
        assert not _state_diffs(devs[0], devs[1], L.STATE_FIELDS + ["vx1", "vy1", "vz1"]), f"step {s + 1}"
        for d in devs:
            d.half_timestep()
            d.end_of_timestep()
        assert not _state_diffs(devs[0], devs[1]), f"step {s + 1}"


@pytest.mark.parametrize("kernel_set", [0, 1, 2])
def test_edges_without_conditions_match_oracle(kernel_set):
    """Edges with no boundary condition keep their ghost values (the
    reference fills nothing there): all z edges, and the upper x edge of rho
    and energy_ion only."""
    p = L.problem("ot", n=10, ny=8, nz=9, seed=6, amp=0.03)
    dev = L.Lare.from_problem(p)
    dev.set_kernel_set(kernel_set)
    orc = make_oracle(p)
    for g in (dev, orc):
        for f in L.FIELDS:
            for e in (0, 1):
                g.set_bc(f, 2, e, L.BC_NONE)
        for f in ("rho", "energy_ion"):
            g.set_bc(f, 0, 1, L.BC_NONE)
    for s in range(3):
        assert dev.step().hex() == orc.step().hex(), f"dt of step {s + 1}"
        assert not _state_diffs(dev, orc), f"step {s + 1}"


Z_EDGE_KINDS = [
    ((L.BC_PERIODIC, 0.0), (L.BC_PERIODIC, 0.0)),
    ((L.BC_ZERO_GRADIENT, 0.0), (L.BC_MIRROR, 0.0)),
    ((L.BC_MIRROR, 0.0), (L.BC_CLAMP, 0.3)),
    ((L.BC_CLAMP, 0.1), (L.BC_ZERO_GRADIENT, 0.0)),
    ((L.BC_PERIODIC, 0.0), (L.BC_CLAMP, 0.2)),
    ((L.BC_ZERO_GRADIENT, 0.0), (L.BC_PERIODIC, 0.0)),
    ((L.BC_MIRROR, 0.0), (L.BC_PERIODIC, 0.0)),
]


@pytest.mark.parametrize("kernel_set", [1, 2])
def test_z_edges_of_every_kind_match_oracle(kernel_set):
    """The z fills (one z-row end per lane, all sources loaded before any
    ghost is written) against the oracle's fills in the reference order, for
    every kind on either edge and mixed pairs, including the staggered
    periodic upper edge whose source at index 0 the lower fill has just
    written: each field's z edges get one pair of the list in turn."""
    p = L.problem("bw", n=16, ny=6, nz=7, seed=12, noise=0.02, eta=1e-3)
    dev = L.Lare.from_problem(p)
    dev.set_kernel_set(kernel_set)
    orc = make_oracle(p)
    for g in (dev, orc):
        for i, f in enumerate(L.FIELDS):
            (k0, v0), (k1, v1) = Z_EDGE_KINDS[i % len(Z_EDGE_KINDS)]
            g.set_bc(f, 2, 0, k0, v0)
            g.set_bc(f, 2, 1, k1, v1)
    for s in range(2):
        assert dev.step().hex() == orc.step().hex(), f"dt of step {s + 1}"
        assert not _state_diffs(dev, orc), f"step {s + 1}"


# every pair of Z_EDGE_KINDS but the staggered periodic-lower / clamp-upper
# one, whose fill the engine keeps in per-dimension launches
ALL_DIM_KINDS = [k for k in Z_EDGE_KINDS if k[0][0] != L.BC_PERIODIC or k[1][0] != L.BC_CLAMP]


@pytest.mark.parametrize("kinds", ["no_periodic_clamp", "all"])
@pytest.mark.parametrize("kernel_set", [1, 2])
def test_edges_of_every_kind_on_all_dims_match_oracle(kernel_set, kinds):
    """Fills of all three dimensions in one launch (each ghost cell computed
    from the cells it derives from through the x, then y, then z fills)
    against the oracle's fills in the reference order: every field gets a
    different pair of edge kinds on each dimension, so corners and edges of
    the box compose periodic, zero-gradient, mirror and clamp fills of
    centred and staggered lines.  With "all", some variables have a
    staggered periodic-lower / clamp-upper line, and their groups take the
    per-dimension fills."""
    lst = ALL_DIM_KINDS if kinds == "no_periodic_clamp" else Z_EDGE_KINDS
    p = L.problem("bw", n=14, ny=7, nz=6, seed=21, noise=0.02, eta=1e-3)
    dev = L.Lare.from_problem(p)
    dev.set_kernel_set(kernel_set)
    orc = make_oracle(p)
    for g in (dev, orc):
        for i, f in enumerate(L.FIELDS):
            for d in range(3):
                (k0, v0), (k1, v1) = lst[(i + 3 * d) % len(lst)]
                g.set_bc(f, d, 0, k0, v0)
                g.set_bc(f, d, 1, k1, v1)
    for s in range(2):
        assert dev.step().hex() == orc.step().hex(), f"dt of step {s + 1}"
        assert not _state_diffs(dev, orc), f"step {s + 1}"


def test_curvilinear_grids_run_kernel_set_0_only():
    """Cylindrical and spherical grids run the one-kernel-per-reference-kernel
    set (its stored metric carries the metric factors); the fused sets,
    batches and the Cartesian energy diagnostic refuse them with an error."""
    p = L.problem("bw", n=16, ny=6, nz=5, seed=3, noise=0.01, geometry="spherical", x0=2.0, y0=0.7)
    g = L.Lare.from_problem(p)
    with pytest.raises(RuntimeError, match="kernel set 0 only"):
        g.set_kernel_set(1)
    with pytest.raises(RuntimeError, match="Cartesian"):
        L.Batch([g])
    with pytest.raises(RuntimeError, match="Cartesian"):
        g.energy()
    orc = make_oracle(p)
    for s in range(3):
        assert g.step().hex() == orc.step().hex(), f"dt of step {s + 1}"
    assert not _state_diffs(g, orc)
