"""The oracle against the reference's own outputs (tests/golden/goldens.json,
written by tests/golden/make_goldens.py from the reference sources compiled
in place).  Bit-exact: same dt sequence and same SHA-256 of every state field,
ghosts included, at every recorded step."""
import pytest

from lare_testlib import case_bcs, case_problem, goldens, make_oracle, sha
from oracle import oracle as O

CASES = goldens()


@pytest.mark.parametrize("name", [n for n in sorted(CASES) if CASES[n]["nx"] <= 32])
def test_oracle_matches_reference_goldens(name):
    c = CASES[name]
    o = case_bcs(c, make_oracle(case_problem(c)))
    snaps = c["snapshots"]
    if "0" in snaps:
        for f in O.STATE_FIELDS:
            assert sha(o.download(f)) == snaps["0"][f]["sha256"], f"initial {f}"
    for s in range(1, c["steps"] + 1):
        dt = o.step()
        assert dt.hex() == float.fromhex(c["dt"][s - 1]).hex(), f"dt of step {s}"
        if str(s) in snaps:
            for f in O.STATE_FIELDS:
                assert sha(o.download(f)) == snaps[str(s)][f]["sha256"], f"{f} after step {s}"


@pytest.mark.slow
def test_oracle_config0_64cube_100_steps():
    """Orszag-Tang 64^3, 100 steps: the whole dt sequence and the final state."""
    c = CASES["ot_64_config0"]
    O.set_threads(8)
    o = case_bcs(c, make_oracle(case_problem(c)))
    for s in range(1, c["steps"] + 1):
        dt = o.step()
        assert dt.hex() == float.fromhex(c["dt"][s - 1]).hex(), f"dt of step {s}"
        if str(s) in c["snapshots"]:
            for f in O.STATE_FIELDS:
                assert sha(o.download(f)) == c["snapshots"][str(s)][f]["sha256"], f"{f} after step {s}"


@pytest.mark.parametrize("name", [n for n in sorted(CASES) if CASES[n].get("geometry", "cartesian") != "cartesian"])
def test_geometry_goldens_depend_on_the_metric_factors(name):
    """The curvilinear goldens pin the metric factors: the same problem on a
    Cartesian grid (every factor 1) misses the reference's first dt or its
    first recorded state."""
    import copy
    c = copy.deepcopy(CASES[name])
    c["geometry"] = "cartesian"
    o = case_bcs(c, make_oracle(case_problem(c)))
    first = min(int(s) for s in c["snapshots"] if int(s) > 0)
    same = True
    for s in range(1, first + 1):
        same = same and o.step().hex() == float.fromhex(c["dt"][s - 1]).hex()
    same = same and all(sha(o.download(f)) == c["snapshots"][str(first)][f]["sha256"] for f in O.STATE_FIELDS)
    assert not same
