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


def test_config2_fp32_storage_512_cubed_100_steps():
    """The FP32-storage workload at its size: Orszag-Tang 512^3 with seeded
    N(0, 1e-3^2) velocity perturbations, 100 steps, FP32 storage against
    FP64.  Per field (domain cells) the relative L2 and L-inf error,
    normalised by the FP64 vector's norm for velocity and B components, must
    stay below the bounds below; the dt sequences within 1e-4 and the total
    energy (FP64 diagnostic) within 1e-6."""
    n, steps = 512, 100
    a = L.Lare.from_problem(L.problem("ot", n=n, seed=7, amp=1e-3, precision=L.FP64))
    da = np.array(a.steps(steps))
    b = L.Lare.from_problem(L.problem("ot", n=n, seed=7, amp=1e-3, precision=L.FP32_STORAGE))
    db = np.array(b.steps(steps))
    rel_dt = float(np.max(np.abs(db - da) / da))
    print(f"dt: max rel {rel_dt:.3e}", flush=True)
    ea, eb = a.energy(), b.energy()
    rel_e = abs(eb[3] - ea[3]) / abs(ea[3])
    print(f"total energy: rel {rel_e:.3e}", flush=True)
    # field by field, so that at most two full fields are on the host at once
    norms = {}
    errs = {}
    for f in L.STATE_FIELDS:
        x = a.download(f)[2:-2, 2:-2, 2:-2]
        y = b.download(f)[2:-2, 2:-2, 2:-2]
        d = x - y
        norms[f] = (float(np.sum(x * x)), float(np.max(np.abs(x))))
        errs[f] = (float(np.sqrt(np.sum(d * d))), float(np.max(np.abs(d))))
        del x, y, d
    # measured on MI355X (vector-normalised L2 / Linf): rho 3.0e-7 / 7.0e-6,
    # e_e 2.4e-7 / 7.8e-6, e_i 3.6e-7 / 1.7e-5, vx 4.1e-7 / 1.9e-5,
    # vy 5.7e-7 / 5.8e-5, vz 1.9e-6 / 7.3e-5, bx 8.7e-7 / 1.7e-5,
    # by 9.7e-7 / 2.1e-5, bz 2.9e-6 / 6.3e-5; dt 1.5e-5, total energy 4.7e-8
    bounds = {"rho": (2e-6, 1e-4), "energy_electron": (2e-6, 1e-4), "energy_ion": (2e-6, 1e-4),
              "vx": (1e-5, 3e-4), "vy": (1e-5, 3e-4), "vz": (1e-5, 3e-4),
              "bx": (1e-5, 3e-4), "by": (1e-5, 3e-4), "bz": (1e-5, 3e-4)}
    bad = []
    for f in L.STATE_FIELDS:
        comps = VECTOR.get(f, (f,))
        n2 = np.sqrt(sum(norms[c][0] for c in comps) / len(comps))
        ni = max(norms[c][1] for c in comps)
        e2, ei = errs[f][0] / n2, errs[f][1] / ni
        own2 = errs[f][0] / np.sqrt(norms[f][0]) if norms[f][0] > 0 else errs[f][0]
        owni = errs[f][1] / norms[f][1] if norms[f][1] > 0 else errs[f][1]
        print(f"{f}: rel L2 {own2:.3e} Linf {owni:.3e}; vector-normalised L2 {e2:.3e} Linf {ei:.3e}", flush=True)
        if e2 >= bounds[f][0] or ei >= bounds[f][1]:
            bad.append(f"{f}: rel L2 {e2:.3e}, Linf {ei:.3e}")
    assert rel_dt < 1e-4, f"dt rel {rel_dt:.3e}"
    assert rel_e < 1e-6, f"total energy rel {rel_e:.3e}"
    assert not bad, bad


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
