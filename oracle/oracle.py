"""Python wrapper of the C oracle (oracle/liblare_oracle.so).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker -- never by the product path.
The oracle restates the reference LARE3D step on the CPU; see
oracle/lare_oracle.c for the reference file:line of every kernel.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblare_oracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")
REF_DRIVER_FAST = os.path.join(HERE, "_ref", "ref_driver_fast")
# The reference with its resistive path made runnable (the three assigns at
# lagran.cpp:482-484 copy element by element; oracle/Makefile patches a copy
# of lagran.cpp at build time): resistive runs only.
REF_DRIVER_RES = os.path.join(HERE, "_ref", "ref_driver_res")
REF_DRIVER_RES_FAST = os.path.join(HERE, "_ref", "ref_driver_res_fast")
# The same driver and reference runner with the MI355X LARE3D package in place
# of the reference's step sources (needs a GPU to run).
DROPIN_DRIVER = os.path.join(HERE, "_ref", "lare3d_dropin")

FIELDS = ["rho", "energy_electron", "energy_ion", "vx", "vy", "vz", "bx", "by", "bz",
          "vx1", "vy1", "vz1", "dm"]
STATE_FIELDS = FIELDS[:9]
MAP_KEEP, MAP_VALUE = -1000, -1001  # map-edge table entries besides a source layer


class LoDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_long), ("ny", ctypes.c_long), ("nz", ctypes.c_long),
                ("x_min", ctypes.c_double), ("x_max", ctypes.c_double),
                ("y_min", ctypes.c_double), ("y_max", ctypes.c_double),
                ("z_min", ctypes.c_double), ("z_max", ctypes.c_double),
                ("gas_gamma", ctypes.c_double), ("visc1", ctypes.c_double),
                ("visc2", ctypes.c_double), ("mu0", ctypes.c_double),
                ("dt_multiplier", ctypes.c_double), ("resistive", ctypes.c_int),
                ("eta_background", ctypes.c_double), ("eta0", ctypes.c_double),
                ("j_max", ctypes.c_double), ("rke", ctypes.c_int), ("geometry", ctypes.c_int)]


_lib = None


def build():
    """Compile liblare_oracle.so with the oracle Makefile."""
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        vp = ctypes.c_void_p
        L.lo_create.restype = vp
        L.lo_create.argtypes = [P(LoDesc)]
        L.lo_destroy.argtypes = [vp]
        L.lo_field.restype = P(ctypes.c_double)
        L.lo_field.argtypes = [vp, ctypes.c_int, P(ctypes.c_long), P(ctypes.c_long)]
        L.lo_array.restype = P(ctypes.c_double)
        L.lo_array.argtypes = [vp, ctypes.c_char_p, P(ctypes.c_long), P(ctypes.c_long)]
        L.lo_axis.restype = P(ctypes.c_double)
        L.lo_axis.argtypes = [vp, ctypes.c_char_p, P(ctypes.c_long), P(ctypes.c_long)]
        L.lo_set_bc.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.lo_set_bc_map.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, P(ctypes.c_int), ctypes.c_double]
        L.lo_get_dt.restype = ctypes.c_double
        L.lo_get_dt.argtypes = [vp]
        L.lo_set_dt.argtypes = [vp, ctypes.c_double]
        for f in ("lo_step", "lo_start_of_timestep", "lo_set_dt_only"):
            getattr(L, f).restype = ctypes.c_double
            getattr(L, f).argtypes = [vp]
        L.lo_half_timestep.argtypes = [vp]
        L.lo_end_of_timestep.argtypes = [vp]
        L.lo_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def set_threads(n):
    lib().lo_set_threads(int(n))


class Oracle:
    """CPU restatement of one LARE3D grid; same field names as sams_amd.lare."""

    def __init__(self, desc):
        """desc: any object with the lare_desc scalar attributes."""
        d = LoDesc(desc.nx, desc.ny, desc.nz, desc.x_min, desc.x_max, desc.y_min, desc.y_max,
                   desc.z_min, desc.z_max, desc.gas_gamma, desc.visc1, desc.visc2, desc.mu0,
                   desc.dt_multiplier, int(desc.resistive), desc.eta_background, desc.eta0,
                   desc.j_max, int(getattr(desc, "rke", 0)), int(getattr(desc, "geometry", 0)))
        self.h = lib().lo_create(ctypes.byref(d))
        self.nx, self.ny, self.nz = desc.nx, desc.ny, desc.nz

    def close(self):
        if getattr(self, "h", None):
            lib().lo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def view(self, name):
        """Writable numpy view of a registered variable (reference extent)."""
        lb = (ctypes.c_long * 3)()
        ub = (ctypes.c_long * 3)()
        p = lib().lo_field(self.h, FIELDS.index(name), lb, ub)
        shape = tuple(ub[k] - lb[k] + 1 for k in range(3))
        return np.ctypeslib.as_array(p, shape=shape)

    def array(self, name):
        lb = (ctypes.c_long * 3)()
        ub = (ctypes.c_long * 3)()
        p = lib().lo_array(self.h, name.encode(), lb, ub)
        if not p:
            raise KeyError(name)
        shape = tuple(ub[k] - lb[k] + 1 for k in range(3))
        return np.ctypeslib.as_array(p, shape=shape), tuple(lb)

    def axis(self, name):
        lb = ctypes.c_long()
        ub = ctypes.c_long()
        p = lib().lo_axis(self.h, name.encode(), ctypes.byref(lb), ctypes.byref(ub))
        return np.ctypeslib.as_array(p, shape=(ub.value - lb.value + 1,)).copy(), lb.value

    def upload(self, name, arr):
        v = self.view(name)
        v[...] = arr

    def download(self, name):
        return self.view(name).copy()

    def state(self):
        return {n: self.download(n) for n in STATE_FIELDS}

    def set_bc(self, name, dim, edge, kind, value=0.0):
        lib().lo_set_bc(self.h, FIELDS.index(name), dim, edge, kind, value)

    def set_bc_map(self, name, dim, edge, src, value=0.0):
        """Map edge: src = 3 layer entries (index, MAP_KEEP or MAP_VALUE)."""
        arr = (ctypes.c_int * 3)(*[int(x) for x in (list(src) + [MAP_KEEP] * 3)[:3]])
        lib().lo_set_bc_map(self.h, FIELDS.index(name), dim, edge, arr, value)

    def get_dt(self):
        return lib().lo_get_dt(self.h)

    def set_dt(self, dt):
        lib().lo_set_dt(self.h, float(dt))

    def step(self):
        return lib().lo_step(self.h)

    def start_of_timestep(self):
        return lib().lo_start_of_timestep(self.h)

    def half_timestep(self):
        lib().lo_half_timestep(self.h)

    def end_of_timestep(self):
        lib().lo_end_of_timestep(self.h)


def problem_bcs(o, kind, bcx=0):
    """The boundary conditions of the benchmark problems (mirrors
    lare_problem_bcs): Orszag-Tang all periodic; Brio-Wu x = bcx
    (0 -> zero-gradient; 4 -> walls: velocities clamped, rest zero-gradient),
    y/z periodic."""
    if kind == 0:
        bx = 1
    else:
        bx = 2 if bcx == 0 else bcx
    for name in FIELDS:
        k = bx
        if bx == 4 and name not in ("vx", "vy", "vz", "vx1", "vy1", "vz1"):
            k = 2
        for e in (0, 1):
            o.set_bc(name, 0, e, k)
            o.set_bc(name, 1, e, 1)
            o.set_bc(name, 2, e, 1)


def run_reference(out_dir, problem="ot", nx=16, ny=None, nz=None, steps=10, dumps=(), seed=1,
                  amp=0.0, noise=0.0, bcx=None, threads=4, fast=False, beta=-1.0, exe=None, rke=False,
                  lare_bc=False, eta=0.0, eta0=0.0, j_max=None, geometry=None, x0=0.0, y0=0.0):
    """Run the reference driver (oracle/_ref, built from the reference's own
    sources; exe selects another build of it, such as DROPIN_DRIVER) and
    return its meta dict.  Resistive runs (eta or eta0 > 0) use the build
    whose resistive path runs (REF_DRIVER_RES).  Raises if it has not been
    built."""
    import json
    resistive = eta > 0.0 or eta0 > 0.0
    if resistive:
        exe = exe or (REF_DRIVER_RES_FAST if fast else REF_DRIVER_RES)
    exe = exe or (REF_DRIVER_FAST if fast else REF_DRIVER)
    if not os.path.exists(exe):
        raise FileNotFoundError(exe)
    os.makedirs(out_dir, exist_ok=True)
    args = [exe, "--problem", problem, "--nx", str(nx), "--ny", str(ny or nx), "--nz", str(nz or nx),
            "--steps", str(steps), "--out", out_dir, "--seed", str(seed), "--amp", repr(amp),
            "--noise", repr(noise), "--beta", repr(beta)]
    if dumps:
        args += ["--dump", ",".join(str(s) for s in dumps)]
    if bcx:
        args += ["--bcx", bcx]
    if rke:
        args += ["--rke"]
    if lare_bc:
        args += ["--lare-bc"]
    if geometry and geometry != "cartesian":
        args += ["--geometry", geometry]
    if x0 or y0:
        args += ["--x0", repr(float(x0)), "--y0", repr(float(y0))]
    if resistive:
        args += ["--eta", repr(float(eta)), "--eta0", repr(float(eta0))]
        if j_max is not None:
            args += ["--j-max", repr(float(j_max))]
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    subprocess.run(args, cwd=out_dir, env=env, check=True, capture_output=True)
    with open(os.path.join(out_dir, "meta.json")) as f:
        return json.load(f)
