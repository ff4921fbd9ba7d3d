"""The step at the benchmark's full size against the reference itself.

The benchmark workload (resistive Brio-Wu 256^3, 1% density noise, seed 1,
eta = 1e-3) for its 200 steps against the reference's own run of it
(oracle/_ref/ref_driver_res: the reference with the three element-wise
copies at lagran.cpp:482-484 that let its resistive path run): the dt of
every step and SHA-256 digests of every state field, ghosts included,
after steps 1, 50, 100 and 200 (tests/golden/goldens.json "full_cases",
written by tests/golden/make_goldens.py --fullsize).

The ideal-MHD twin (the same grid with eta = 0) runs the reference itself
on the host cores (oracle/_ref/ref_driver, the -ffp-contract=off build: the
-march=x86-64-v3 timing build contracts multiply-adds, already in the noisy
initial density) for 8 steps, every state field compared after each step.
At this size the lane-walking kernels run (the small golden cases use the
one-cell forms), so these tests pin them against the reference's output.
"""
import os
import tempfile

import numpy as np
import pytest

import sams_amd.lare as L
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _run_against_digests(c, dev, upto=None):
    import sys
    sys.path.insert(0, os.path.dirname(__file__))
    from lare_testlib import sha
    done = 0
    for s in sorted(int(k) for k in c["snapshots"]):
        if upto is not None and s > upto:
            break
        dts = dev.steps(s - done)
        for i, dt in enumerate(dts):
            assert dt.hex() == float.fromhex(c["dt"][done + i]).hex(), f"dt of step {done + i + 1}"
        done = s
        for f in L.STATE_FIELDS:
            assert sha(dev.download(f)) == c["snapshots"][str(s)][f]["sha256"], f"{f} after step {s}"
        print(f"step {s}: dt and every state field equal to the reference", flush=True)


def _full_case(name):
    import sys
    sys.path.insert(0, os.path.dirname(__file__))
    from lare_testlib import case_problem, full_goldens
    c = full_goldens().get(name)
    if c is None:
        pytest.skip(f"no full-size golden {name}")
    return c, case_problem(c)


def test_benchmark_workload_200_steps_matches_reference():
    """The benchmark workload, every one of its 200 steps, against the
    reference: the default launch configuration (lane-walking kernels at
    this size)."""
    c, p = _full_case("bw_res_256_200")
    dev = L.Lare.from_problem(p)
    _run_against_digests(c, dev)


@pytest.mark.parametrize("tune", [
    # 8 cells per lane in the marches, plane-major workgroup order, the z
    # sweep without the lane exchange, CT flux and update as separate kernels
    dict(walk=1, march_k=8, zx_k=8, zline=0, order=0, ct_fuse=0),
    # the configuration the autotuner picks for this workload (16 cells per
    # lane everywhere, the CT of all three sweeps fused)
    dict(walk=1, march_k=16, zx_k=16, zline=1, order=1, ct_fuse=2),
])
def test_benchmark_workload_other_launch_configurations_match_reference(tune):
    """The other launch configurations the autotuner can choose give the
    same bits for the first 50 steps."""
    c, p = _full_case("bw_res_256_200")
    dev = L.Lare.from_problem(p)
    dev.set_tune(**tune)
    _run_against_digests(c, dev, upto=50)


@pytest.mark.skipif(not os.path.exists(O.REF_DRIVER), reason="oracle/_ref/ref_driver not built")
def test_brio_wu_256_cubed_matches_reference():
    n, steps = 256, 8
    with tempfile.TemporaryDirectory() as td:
        meta = O.run_reference(td, problem="bw", nx=n, steps=steps, dumps=tuple(range(1, steps + 1)), noise=0.01,
                               seed=1, threads=min(16, os.cpu_count() or 1))
        dev = L.Lare.from_problem(L.problem("bw", n=n, seed=1, noise=0.01))
        for s in range(1, steps + 1):
            dt = dev.steps(1)[0]
            assert dt.hex() == float.fromhex(meta["dt"][s - 1]).hex(), f"dt of step {s}"
            for f in L.STATE_FIELDS:
                ref = np.fromfile(os.path.join(td, f"s{s:05d}_{f}.f64"), dtype="<f8")
                got = np.ascontiguousarray(dev.download(f), dtype="<f8").ravel()
                assert ref.size == got.size, f
                bad = np.flatnonzero(ref.view(np.uint64) != got.view(np.uint64))
                assert bad.size == 0, f"{f} after step {s}: {bad.size} values differ"


def test_benchmark_workload_matches_oracle():
    """The benchmark workload itself (Brio-Wu 256^3, 1% density noise, seed 1,
    eta = 1e-3: resistive, two half-dt substeps per step), which the reference
    cannot run, against the oracle's restatement: dt and every state field
    bit-identical after each of six steps (steps 4-6 send a few quotients of
    the resistive substep, numerators below 2^-300, through the full
    division)."""
    import sys
    sys.path.insert(0, os.path.dirname(__file__))
    from lare_testlib import make_oracle
    p = L.problem("bw", n=256, seed=1, noise=0.01, eta=1e-3)
    dev = L.Lare.from_problem(p)
    orc = make_oracle(p)
    for s in range(1, 7):
        assert dev.step().hex() == orc.step().hex(), f"dt of step {s}"
        for f in L.STATE_FIELDS:
            a = np.ascontiguousarray(dev.download(f), dtype="<f8").view(np.uint64)
            b = np.ascontiguousarray(orc.download(f), dtype="<f8").view(np.uint64)
            assert np.array_equal(a, b), f"{f} after step {s}"
