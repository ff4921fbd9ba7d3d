"""Shared helpers of the LARE3D tests: build oracle and device grids for the
same seeded problem and compare them bit for bit."""
import hashlib
import json
import os

import numpy as np

from oracle import oracle as O
import sams_amd.lare as L

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "goldens.json")
BCX = {"": 0, "zerograd": L.BC_ZERO_GRADIENT, "mirror": L.BC_MIRROR, "clamp": L.BC_CLAMP,
       "periodic": L.BC_PERIODIC}


def goldens():
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


def case_problem(c):
    """LareProblem for a golden case."""
    return L.problem(c["problem"], n=c["nx"], ny=c["ny"], nz=c["nz"], seed=c["seed"],
                     amp=c["amp"], noise=c["noise"], bcx=BCX[c["bcx"]], rke=c.get("rke", False))


def make_oracle(p):
    """Oracle grid holding problem p at t = 0 (product initial conditions)."""
    o = O.Oracle(L.problem_desc(p))
    O.problem_bcs(o, p.kind, p.bcx)
    for n in O.STATE_FIELDS:
        o.upload(n, L.problem_fill(p, n))
    return o


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def bit_diff(a, b):
    """None if a and b are bit-identical, else a short description."""
    ua = np.ascontiguousarray(a).view(np.uint64)
    ub = np.ascontiguousarray(b).view(np.uint64)
    if ua.shape == ub.shape and np.array_equal(ua, ub):
        return None
    bad = np.argwhere(ua != ub)
    i = tuple(bad[0])
    rel = np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))
    return f"{len(bad)} of {a.size} differ, first at {i}: {a[i]!r} vs {b[i]!r}, max rel {rel:.3e}"


def rel_linf(a, b):
    scale = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / scale) if scale > 0 else float(np.max(np.abs(a - b)))
