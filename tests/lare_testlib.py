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


def full_goldens():
    """Reference runs at a benchmark's full size and step count."""
    with open(GOLDEN) as f:
        return json.load(f).get("full_cases", {})


def ic_goldens():
    """Runs of the reference's own initial-condition packages."""
    with open(GOLDEN) as f:
        return json.load(f)["ic_cases"]


def case_problem(c):
    """LareProblem for a golden case."""
    return L.problem(c["problem"], n=c["nx"], ny=c["ny"], nz=c["nz"], seed=c["seed"],
                     amp=c["amp"], noise=c["noise"], bcx=BCX[c["bcx"]], rke=c.get("rke", False),
                     eta=c.get("eta", 0.0), eta0=c.get("eta0", 0.0), j_max=c.get("j_max") or 0.0,
                     geometry=c.get("geometry", "cartesian"), x0=c.get("x0", 0.0), y0=c.get("y0", 0.0))


def case_bcs(c, grid):
    """The golden case's boundary conditions beyond the problem's own (the
    LARE initial-condition package's conditions on every edge) on a device
    or oracle grid."""
    if c.get("lare_bc"):
        apply_lare_bc(grid, c["nx"], c["ny"], c["nz"])
    return grid


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


KEEP, VALUE = O.MAP_KEEP, O.MAP_VALUE


def lare_bc_rules(nx, ny, nz):
    """The BC_OTHER conditions of the LARE initial-condition package
    (src/InitialConditions/LARE/boundary.cpp:31-1120, attached to every edge
    by lareic.h:124-146) as map edges: {(field, dim, edge): (src[3], value)}
    with the edge's ghost layers in order (lower -2..0 or -1..0, upper n..n+2
    or n+1..n+2).  Restated from the source; pinned by the reference's own
    runs (goldens *_larebc_*)."""
    n3 = (nx, ny, nz)
    out = {}
    for f in O.FIELDS:
        st = L.STAGGER[f]
        for d in range(3):
            n, s = n3[d], st[d]
            if f == "dm":  # dm_x_bcs (boundary.cpp:1063-1081), x only
                if d == 0:
                    out[(f, d, 0)] = ((0, KEEP, KEEP), 0.0)     # dm(-1) = dm(0)
                    out[(f, d, 1)] = ((n, KEEP, KEEP), 0.0)     # dm(n+1) = dm(n)
                continue
            if f in ("vx", "vy", "vz", "vx1", "vy1", "vz1") and d == 0:
                # velocities: x walls set to 0 over -2..0 and n..n+2
                out[(f, d, 0)] = ((VALUE, VALUE, VALUE), 0.0)
                out[(f, d, 1)] = ((VALUE, VALUE, VALUE), 0.0)
                continue
            if s:  # staggered along d: -2 <- 2, -1 <- 1; 0 and n kept
                lo = (2, 1, KEEP)
                hi = (KEEP, n, n - 1) if f in ("by", "bz") or (f == "bx" and d != 0) else (KEEP, n - 1, n - 2)
            else:  # centred along d: -1 <- 2, 0 <- 1
                lo = (2, 1, KEEP)
                hi = (n, n - 1, KEEP) if f in ("bx", "by", "bz") else (n - 1, n - 2, KEEP)
            out[(f, d, 0)] = (lo, 0.0)
            out[(f, d, 1)] = (hi, 0.0)
    return out


def apply_lare_bc(grid, nx, ny, nz):
    """Attach lare_bc_rules to a device grid or an oracle grid (any object
    with set_bc / set_bc_map); the dm edges along y and z get none."""
    rules = lare_bc_rules(nx, ny, nz)
    for f in O.FIELDS:
        for d in range(3):
            for e in (0, 1):
                if (f, d, e) in rules:
                    src, val = rules[(f, d, e)]
                    grid.set_bc_map(f, d, e, src, val)
                else:
                    grid.set_bc(f, d, e, 0)
