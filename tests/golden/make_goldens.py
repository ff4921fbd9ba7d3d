"""Write tests/golden/goldens.json from the reference itself.

Runs oracle/_ref/ref_driver -- the reference's own LARE sources compiled in
place with -ffp-contract=off by oracle/Makefile, driven through the
reference runner -- on small seeded problems, and records for each case the
dt sequence (hex floats) and, at the dumped steps, the SHA-256 of every state
field's raw little-endian float64 bytes over its full reference extent
(ghosts included), with L-inf/L2 norms for reading.  Requires the reference
tree; the JSON it writes is the fixture the tests use.

Cases with the remap kinetic-energy correction (rke) run the reference on one
OpenMP thread: its atomic adds into delta_ke (xremap.cpp:552-555 and twins)
sum in an arbitrary order when threads share a cell, and one thread's order
is the one the oracle and the kernels reproduce.  Cases with lare_bc attach
the LARE initial-condition package's BC_OTHER conditions to every edge
(src/InitialConditions/LARE/lareic.h:124-146, boundary.cpp) instead of the
problem's built-in ones.

Resistive cases run oracle/_ref/ref_driver_res, the reference with the three
element-wise copies at lagran.cpp:482-484 that let its resistive path run
(oracle/Makefile writes that copy of lagran.cpp at build time).

    make -C oracle ref ic && python tests/golden/make_goldens.py
    python tests/golden/make_goldens.py --fullsize   (the 256^3 benchmark workload)
    python tests/golden/make_goldens.py --ic-long    (the initial-condition packages over many steps)
"""
import hashlib
import json
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

FIELDS = O.STATE_FIELDS

# name, problem, nx, ny, nz, steps, dumps, extra
CASES = [
    ("ot_12", "ot", 12, 12, 12, 10, (0, 1, 2, 10), {}),
    ("ot_pert_10x8x12", "ot", 10, 8, 12, 12, (0, 1, 12), {"amp": 0.05, "seed": 7}),
    ("ot_pert_16", "ot", 16, 16, 16, 40, (0, 40), {"amp": 0.01, "seed": 5}),
    ("bw_zerograd_24x4x6", "bw", 24, 4, 6, 15, (0, 1, 15), {"noise": 0.01, "seed": 3}),
    ("bw_mirror_16x4x4", "bw", 16, 4, 4, 10, (0, 10), {"bcx": "mirror"}),
    ("bw_walls_16x5x3", "bw", 16, 5, 3, 8, (0, 8), {"bcx": "clamp", "noise": 0.02, "seed": 11}),
    ("bw_plain_32x3x3", "bw", 32, 3, 3, 20, (0, 20), {}),
    ("ot_64_config0", "ot", 64, 64, 64, 100, (1, 100), {}),
    ("ot_rke_12", "ot", 12, 12, 12, 10, (0, 1, 10), {"amp": 0.05, "seed": 7, "rke": True}),
    ("ot_rke_10x8x12", "ot", 10, 8, 12, 12, (0, 12), {"amp": 0.02, "seed": 2, "rke": True}),
    ("bw_rke_zerograd_24x4x6", "bw", 24, 4, 6, 15, (0, 15), {"noise": 0.01, "seed": 3, "rke": True}),
    ("bw_rke_mirror_16x5x4", "bw", 16, 5, 4, 10, (0, 10), {"bcx": "mirror", "noise": 0.02, "seed": 4,
                                                           "rke": True}),
    ("ot_larebc_rke_12", "ot", 12, 12, 12, 10, (0, 1, 10), {"amp": 0.05, "seed": 7, "rke": True, "lare_bc": True}),
    ("ot_larebc_10x8x12", "ot", 10, 8, 12, 12, (0, 12), {"amp": 0.02, "seed": 2, "lare_bc": True}),
    ("bw_larebc_16x5x4", "bw", 16, 5, 4, 10, (0, 10), {"noise": 0.01, "seed": 3, "lare_bc": True}),
    # two-cell dimensions, as the reference's quasi-2D initial-condition packages use (nz = 2, or ny = nz = 2)
    ("ot_2d_14x12x2", "ot", 14, 12, 2, 12, (0, 1, 12), {"amp": 0.05, "seed": 3}),
    ("bw_mirror_32x2x2", "bw", 32, 2, 2, 20, (0, 20), {"bcx": "mirror", "noise": 0.01, "seed": 5}),
    ("bw_rke_2d_20x6x2", "bw", 20, 6, 2, 12, (0, 12), {"noise": 0.02, "seed": 6, "rke": True}),
    ("ot_larebc_12x10x2", "ot", 12, 10, 2, 10, (0, 10), {"amp": 0.05, "seed": 8, "lare_bc": True}),
    # resistive MHD (oracle/_ref/ref_driver_res: the reference with the element-wise copies at
    # lagran.cpp:482-484 that let its resistive path run); eta0 with a small j_max makes
    # eta_calc's j > j_max branch (lagran.cpp:473) fire in part of the domain (about 20-50%
    # of the cells for the Orszag-Tang case, the shock front for Brio-Wu)
    ("bw_res_24x5x4", "bw", 24, 5, 4, 12, (0, 1, 12), {"noise": 0.01, "seed": 3, "eta": 1e-3}),
    ("bw_res_strong_20x6x5", "bw", 20, 6, 5, 10, (0, 10), {"noise": 0.02, "seed": 5, "eta": 0.05}),
    ("ot_res_16x12x10", "ot", 16, 12, 10, 8, (0, 1, 8), {"amp": 0.05, "seed": 7, "eta": 0.05}),
    ("ot_res_jmax_16x12x10", "ot", 16, 12, 10, 10, (1, 10), {"amp": 0.05, "seed": 7, "eta": 0.01, "eta0": 0.05,
                                                             "j_max": 2.0}),
    ("bw_res_jmax_24x5x4", "bw", 24, 5, 4, 12, (1, 12), {"noise": 0.01, "seed": 9, "eta": 1e-3, "eta0": 0.02,
                                                         "j_max": 5.0}),
    ("ot_res_rke_12x10x8", "ot", 12, 10, 8, 10, (0, 10), {"amp": 0.05, "seed": 2, "eta": 0.02, "rke": True}),
    ("ot_res_larebc_12x10x8", "ot", 12, 10, 8, 10, (0, 10), {"amp": 0.05, "seed": 4, "eta": 0.02,
                                                             "lare_bc": True}),
    ("bw_res_mirror_2d_20x6x2", "bw", 20, 6, 2, 12, (0, 12), {"bcx": "mirror", "noise": 0.02, "seed": 6,
                                                              "eta": 0.01, "eta0": 0.02, "j_max": 3.0}),
    # cylindrical (x, y, z = r, phi, z) and spherical (r, theta, phi) grids: the domain shifted to
    # r > 0 (and theta inside (0, pi)); metric factors of setup.cpp:276-335, geometric forces of
    # lagran.cpp:711-721
    ("bw_cyl_16x6x5", "bw", 16, 6, 5, 10, (0, 1, 10), {"noise": 0.01, "seed": 3, "geometry": "cylindrical",
                                                       "x0": 2.0}),
    ("ot_cyl_12x10x8", "ot", 12, 10, 8, 10, (1, 10), {"amp": 0.05, "seed": 3, "geometry": "cylindrical",
                                                      "x0": 1.0}),
    ("bw_walls_cyl_16x5x4", "bw", 16, 5, 4, 8, (8,), {"bcx": "clamp", "noise": 0.02, "seed": 4,
                                                      "geometry": "cylindrical", "x0": 1.5}),
    ("bw_sph_16x6x5", "bw", 16, 6, 5, 10, (0, 1, 10), {"noise": 0.01, "seed": 3, "geometry": "spherical",
                                                       "x0": 2.0, "y0": 0.7}),
    ("ot_sph_12x10x8", "ot", 12, 10, 8, 8, (1, 8), {"amp": 0.05, "seed": 3, "geometry": "spherical",
                                                    "x0": 1.0, "y0": 0.3}),
    ("bw_res_sph_16x6x5", "bw", 16, 6, 5, 8, (8,), {"noise": 0.01, "seed": 5, "eta": 0.01,
                                                    "geometry": "spherical", "x0": 2.0, "y0": 0.7}),
]

# The benchmark workload at its full size and step count: resistive Brio-Wu
# 256^3, 1% density noise (seed 1), eta = 1e-3, 200 steps (--fullsize).
FULL_CASES = [
    ("bw_res_256_200", "bw", 256, 256, 256, 200, (1, 50, 100, 200), {"noise": 0.01, "seed": 1, "eta": 1e-3}),
]


# The reference's own initial-condition packages (src/InitialConditions/*),
# run by oracle/_ref/ic_driver as the reference's src/main.cpp runs them:
# package, steps, OpenMP threads (EmeryWindTunnel turns rke on: one thread)
IC_CASES = [
    ("SodShockTube", 10, 8),
    ("BrioAndWu", 10, 8),
    ("MHDRotor", 10, 8),
    ("OrszagTang", 10, 8),
    ("OrszagTang3D", 3, 8),
    ("EmeryWindTunnel", 4, 1),
    ("KarmanVortex", 4, 8),
]
IC_DRIVER = os.path.join(ROOT, "oracle", "_ref", "ic_driver")

# The same packages over many steps (or to the package's own end time, when
# it comes first): package, steps, OpenMP threads (--ic-long)
IC_LONG_CASES = [
    ("SodShockTube", 500, 8),
    ("BrioAndWu", 500, 8),
    ("MHDRotor", 300, 8),
    ("OrszagTang", 300, 8),
    ("OrszagTang3D", 60, 8),
    ("KarmanVortex", 200, 8),
    ("EmeryWindTunnel", 120, 1),
]


def run_ic(exe, out_dir, ic, steps, threads):
    """Run an initial-condition driver build; return its meta dict."""
    import subprocess
    os.makedirs(out_dir, exist_ok=True)
    subprocess.run([exe, "--ic", ic, "--steps", str(steps), "--dump", str(steps), "--out", out_dir], cwd=out_dir,
                   env=dict(os.environ, OMP_NUM_THREADS=str(threads)), check=True, capture_output=True)
    with open(os.path.join(out_dir, "meta.json")) as f:
        return json.load(f)


def digest(path):
    a = np.fromfile(path, dtype="<f8")
    return {"sha256": hashlib.sha256(a.tobytes()).hexdigest(),
            "linf": float(np.max(np.abs(a))), "l2": float(np.sqrt(np.sum(a * a)))}


def run_case(name, prob, nx, ny, nz, steps, dumps, extra, threads=8):
    """One reference run: the case's golden record."""
    with tempfile.TemporaryDirectory() as td:
        meta = O.run_reference(td, problem=prob, nx=nx, ny=ny, nz=nz, steps=steps, dumps=dumps,
                               seed=extra.get("seed", 1), amp=extra.get("amp", 0.0),
                               noise=extra.get("noise", 0.0), bcx=extra.get("bcx"),
                               threads=1 if extra.get("rke") else threads, rke=extra.get("rke", False),
                               lare_bc=extra.get("lare_bc", False), eta=extra.get("eta", 0.0),
                               eta0=extra.get("eta0", 0.0), j_max=extra.get("j_max"),
                               geometry=extra.get("geometry"), x0=extra.get("x0", 0.0), y0=extra.get("y0", 0.0))
        snaps = {}
        for s in dumps:
            snaps[str(s)] = {f: digest(os.path.join(td, f"s{s:05d}_{f}.f64")) for f in FIELDS}
    print(name, "done", flush=True)
    return {"problem": prob, "nx": nx, "ny": ny, "nz": nz, "steps": steps,
            "seed": extra.get("seed", 1), "amp": extra.get("amp", 0.0),
            "noise": extra.get("noise", 0.0), "bcx": extra.get("bcx", ""),
            "rke": bool(extra.get("rke", False)),
            "lare_bc": bool(extra.get("lare_bc", False)),
            "eta": extra.get("eta", 0.0), "eta0": extra.get("eta0", 0.0), "j_max": extra.get("j_max", 0.0),
            "geometry": extra.get("geometry", "cartesian"), "x0": extra.get("x0", 0.0), "y0": extra.get("y0", 0.0),
            "dt": meta["dt"], "snapshots": snaps}


def main(argv):
    """No arguments: every small case and initial-condition case.
    --only a,b: those cases, merged into the existing file.
    --fullsize: the full-size cases (long: the 256^3 case takes about half an
    hour on 8 cores), merged into the existing file."""
    path = os.path.join(ROOT, "tests", "golden", "goldens.json")
    if argv and argv[0] == "--ic-long":
        with open(path) as f:
            out = json.load(f)
        out["ic_long_cases"] = {}
        for ic, steps, threads in IC_LONG_CASES:
            with tempfile.TemporaryDirectory() as td:
                meta = run_ic(IC_DRIVER, td, ic, steps, threads)
                n = len(meta["dt"])
                if n < steps:  # the package's end time came first
                    meta = run_ic(IC_DRIVER, td, ic, n, threads)
                snaps = {str(n): {f: digest(os.path.join(td, f"s{n:05d}_{f}.f64")) for f in FIELDS}}
            out["ic_long_cases"][ic] = {"steps": n, "threads": threads, "dt": meta["dt"], "snapshots": snaps}
            print(ic, n, "done", flush=True)
        with open(path, "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)
        return
    if argv and argv[0] in ("--only", "--fullsize"):
        with open(path) as f:
            out = json.load(f)
        if argv[0] == "--only":
            want = set(argv[1].split(","))
            for case in CASES:
                if case[0] in want:
                    out["cases"][case[0]] = run_case(*case)
        else:
            out.setdefault("full_cases", {})
            for case in FULL_CASES:
                out["full_cases"][case[0]] = run_case(*case)
    else:
        out = {"about": "Reference LARE3D (oracle/_ref/ref_driver, ref_driver_res) dt sequences and "
                        "state-field SHA-256 digests; see tests/golden/make_goldens.py", "cases": {}}
        if os.path.exists(path):
            with open(path) as f:
                out["full_cases"] = json.load(f).get("full_cases", {})
        for case in CASES:
            out["cases"][case[0]] = run_case(*case)
        out["ic_cases"] = {}
        for ic, steps, threads in IC_CASES:
            with tempfile.TemporaryDirectory() as td:
                meta = run_ic(IC_DRIVER, td, ic, steps, threads)
                snaps = {str(steps): {f: digest(os.path.join(td, f"s{steps:05d}_{f}.f64")) for f in FIELDS}}
            out["ic_cases"][ic] = {"steps": steps, "threads": threads, "dt": meta["dt"], "snapshots": snaps}
            print(ic, "done", flush=True)
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:])
