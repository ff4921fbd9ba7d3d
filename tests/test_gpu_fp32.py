"""FP32 storage: every state and work array stored in FP32, the arithmetic,
the CFL reduction, the remap flux arrays (dm and the five sweep fluxes) and
the energy diagnostic in FP64.  Its oracle is the FP64 run of the same
problem: the per-field error after the stated number of steps must stay
within the bounds written here, and the FP32 run must be deterministic and
identical alone and in a batch."""
import numpy as np
import pytest

import sams_amd.lare as L
from lare_testlib import bit_diff

pytestmark = pytest.mark.gpu


def _pair(p, steps):
    p64 = L.problem(**p, precision=L.FP64)
    p32 = L.problem(**p, precision=L.FP32_STORAGE)
    a, b = L.Lare.from_problem(p64), L.Lare.from_problem(p32)
    da, db = a.steps(steps), b.steps(steps)
    return a, b, np.array(da), np.array(db)


VECTOR = {"vx": ("vx", "vy", "vz"), "vy": ("vx", "vy", "vz"), "vz": ("vx", "vy", "vz"),
          "bx": ("bx", "by", "bz"), "by": ("bx", "by", "bz"), "bz": ("bx", "by", "bz")}


def field_errors(a, b):
    """Error of b against a per state field over the domain (ghosts
    excluded): relative L2 and L-inf, each normalised by the FP64 field's own
    L2 norm / max, and for velocity and B components also by the norm / max
    of the whole vector (a component that is small in a problem, vz of
    Orszag-Tang, carries the absolute error of the large ones)."""
    xs = {f: a.download(f)[2:-2, 2:-2, 2:-2] for f in L.STATE_FIELDS}
    ys = {f: b.download(f)[2:-2, 2:-2, 2:-2] for f in L.STATE_FIELDS}
    out = {}
    for f in L.STATE_FIELDS:
        d = xs[f] - ys[f]
        comps = VECTOR.get(f, (f,))
        own2, owni = np.sqrt(np.sum(xs[f] ** 2)), np.max(np.abs(xs[f]))
        vec2 = np.sqrt(sum(np.sum(xs[c] ** 2) for c in comps) / len(comps))
        veci = max(np.max(np.abs(xs[c])) for c in comps)
        e2, ei = np.sqrt(np.sum(d * d)), np.max(np.abs(d))
        out[f] = dict(l2=e2 / own2 if own2 > 0 else e2, linf=ei / owni if owni > 0 else ei,
                      l2v=e2 / vec2 if vec2 > 0 else e2, linfv=ei / veci if veci > 0 else ei)
    return out


# Bounds on the vector-normalised relative L2 and L-inf error per field
# after the stated steps.  FP32 rounding of a stored value is 6e-8 relative;
# the bounds leave room for its growth through ~200 stores per cell per step
# and, for Brio-Wu, for the shock's amplification of perturbations.
CASES = [
    ("ot", dict(kind="ot", n=24, seed=5, amp=0.01), 20, 2e-6, 1e-5),
    ("bw", dict(kind="bw", n=48, ny=4, nz=4, seed=3, noise=0.01, eta=1e-3), 20, 2e-5, 2e-4),
]


@pytest.mark.parametrize("name,p,steps,l2max,linfmax", CASES)
def test_fp32_storage_error_per_field(name, p, steps, l2max, linfmax):
    a, b, da, db = _pair(p, steps)
    err = field_errors(a, b)
    for f, e in err.items():
        print(f"{name} {f}: rel L2 {e['l2']:.3e} Linf {e['linf']:.3e}; "
              f"vector-normalised L2 {e['l2v']:.3e} Linf {e['linfv']:.3e}")
    rel_dt = np.max(np.abs(db - da) / da)
    print(f"{name} dt: max rel {rel_dt:.3e}")
    assert rel_dt < 1e-5
    for f, e in err.items():
        assert e["l2v"] < l2max, f"{f} rel L2 {e['l2v']:.3e}"
        assert e["linfv"] < linfmax, f"{f} rel Linf {e['linfv']:.3e}"
    ea, eb = a.energy(), b.energy()
    assert abs(eb[3] - ea[3]) / abs(ea[3]) < 1e-5


def test_fp32_storage_batch_equals_alone():
    ps = [L.problem("ot", n=10, ny=12, nz=9, seed=2, amp=0.03, precision=L.FP32_STORAGE),
          L.problem("bw", n=20, ny=4, nz=5, seed=1, noise=0.01, eta=1e-3, precision=L.FP32_STORAGE)]
    mem = [L.Lare.from_problem(p) for p in ps]
    alone = [L.Lare.from_problem(p) for p in ps]
    b = L.Batch(mem)
    b.steps(6)
    b.sync()
    for i, g in enumerate(alone):
        g.steps(6)
        for f in L.STATE_FIELDS:
            assert bit_diff(mem[i].download(f), g.download(f)) is None, (i, f)


def test_fp32_storage_kernel_forms_agree():
    """The lane-walking kernels (set 2) and the one-cell-per-lane kernels
    (set 3) round to FP32 at the same stores: bit-identical FP32 runs."""
    ps = [L.problem("ot", n=12, ny=10, nz=9, seed=3, amp=0.05, precision=L.FP32_STORAGE),
          L.problem("bw", n=24, ny=5, nz=4, seed=2, noise=0.01, eta=1e-3, precision=L.FP32_STORAGE)]
    for i, p in enumerate(ps):
        runs = [L.Lare.from_problem(p) for _ in range(2)]
        for k, g in zip((2, 3), runs):
            g.set_kernel_set(k)
        da, db = runs[0].steps(6), runs[1].steps(6)
        assert [d.hex() for d in da] == [d.hex() for d in db], i
        for f in L.STATE_FIELDS:
            assert bit_diff(runs[0].download(f), runs[1].download(f)) is None, (i, f)


def test_fp32_and_fp64_do_not_share_a_batch():
    a = L.Lare.from_problem(L.problem("ot", n=8))
    b = L.Lare.from_problem(L.problem("ot", n=8, precision=L.FP32_STORAGE))
    with pytest.raises(RuntimeError, match="one storage precision"):
        L.Batch([a, b])
