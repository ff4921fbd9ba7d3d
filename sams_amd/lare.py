"""Python binding of the MI355X LARE3D C-ABI (include/lare_hip.h).

Loads sams_amd/liblare_hip.so (built in-tree by sams_amd/build.py) and
exposes a context class whose methods mirror the reference package's
events: start_of_timestep / half_timestep / end_of_timestep
(src/PhysicsPackages/LARE/shared_data.h:397-419), plus field transfer in the
reference's array layout.  There is no fallback: if the library or a GPU is
missing the calls raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblare_hip.so")

FIELDS = ["rho", "energy_electron", "energy_ion", "vx", "vy", "vz", "bx", "by", "bz",
          "vx1", "vy1", "vz1", "dm"]
STATE_FIELDS = FIELDS[:9]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}
# stagger (1 = HALF_CELL) per dimension, src/PhysicsPackages/LARE/setup.cpp:45-69
STAGGER = {"rho": (0, 0, 0), "energy_electron": (0, 0, 0), "energy_ion": (0, 0, 0),
           "vx": (1, 1, 1), "vy": (1, 1, 1), "vz": (1, 1, 1), "bx": (1, 0, 0), "by": (0, 1, 0),
           "bz": (0, 0, 1), "vx1": (1, 1, 1), "vy1": (1, 1, 1), "vz1": (1, 1, 1), "dm": (0, 0, 0)}

BC_NONE, BC_PERIODIC, BC_ZERO_GRADIENT, BC_MIRROR, BC_CLAMP, BC_MAP = 0, 1, 2, 3, 4, 5
MAP_KEEP, MAP_VALUE = -1000, -1001  # map-edge layer entries besides a source layer
FP64, FP32_STORAGE = 0, 1
ORSZAG_TANG, BRIO_WU = 0, 1
CARTESIAN, CYLINDRICAL, SPHERICAL = 0, 1, 2


class LareDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int64), ("ny", ctypes.c_int64), ("nz", ctypes.c_int64),
                ("x_min", ctypes.c_double), ("x_max", ctypes.c_double),
                ("y_min", ctypes.c_double), ("y_max", ctypes.c_double),
                ("z_min", ctypes.c_double), ("z_max", ctypes.c_double),
                ("gas_gamma", ctypes.c_double), ("visc1", ctypes.c_double),
                ("visc2", ctypes.c_double), ("mu0", ctypes.c_double),
                ("dt_multiplier", ctypes.c_double), ("resistive", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("eta_background", ctypes.c_double),
                ("eta0", ctypes.c_double), ("j_max", ctypes.c_double),
                ("rke", ctypes.c_int32), ("geometry", ctypes.c_int32)]


class LareTune(ctypes.Structure):
    """lare_tune: launch configuration of the fused kernel set."""
    _fields_ = [("walk", ctypes.c_int32), ("march_k", ctypes.c_int32), ("zx_k", ctypes.c_int32),
                ("zline", ctypes.c_int32), ("order", ctypes.c_int32), ("ct_fuse", ctypes.c_int32),
                ("ms_per_step", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class LareProblem(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("bcx", ctypes.c_int32),
                ("nx", ctypes.c_int64), ("ny", ctypes.c_int64), ("nz", ctypes.c_int64),
                ("seed", ctypes.c_uint64), ("amp", ctypes.c_double), ("beta", ctypes.c_double),
                ("noise", ctypes.c_double), ("eta", ctypes.c_double),
                ("precision", ctypes.c_int32), ("rke", ctypes.c_int32),
                ("eta0", ctypes.c_double), ("j_max", ctypes.c_double),
                ("geometry", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("x0", ctypes.c_double), ("y0", ctypes.c_double)]


GEOMETRY = {"cartesian": CARTESIAN, "cylindrical": CYLINDRICAL, "spherical": SPHERICAL}


def problem(kind="ot", n=16, ny=None, nz=None, seed=1, amp=0.0, beta=-1.0, noise=0.0, eta=0.0,
            bcx=0, precision=FP64, rke=False, eta0=0.0, j_max=0.0, geometry=CARTESIAN, x0=0.0, y0=0.0):
    """A LareProblem: kind 'ot' (Orszag-Tang) or 'bw' (Brio-Wu); rke turns the
    remap kinetic-energy correction on; eta (background) or eta0 (added where
    |j| > j_max; j_max <= 0 means never) turn resistive MHD on; geometry
    (0/'cartesian', 1/'cylindrical', 2/'spherical') with the domain's x and y
    extents shifted by x0, y0."""
    k = {"ot": ORSZAG_TANG, "bw": BRIO_WU}[kind] if isinstance(kind, str) else int(kind)
    gm = GEOMETRY[geometry] if isinstance(geometry, str) else int(geometry)
    return LareProblem(k, bcx, n, ny if ny is not None else n, nz if nz is not None else n,
                       seed, amp, beta, noise, eta, precision, int(bool(rke)), eta0, j_max, gm, 0, x0, y0)


_lib = None


def lib():
    """The loaded C-ABI library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("liblare_hip.so is not built: run python -m sams_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        vp = ctypes.c_void_p
        sig = {
            "lare_last_error": (ctypes.c_char_p, []),
            "lare_ctx_create": (ctypes.c_int, [P(LareDesc), P(vp)]),
            "lare_ctx_destroy": (ctypes.c_int, [vp]),
            "lare_field_bounds": (ctypes.c_int, [vp, ctypes.c_int, P(ctypes.c_int64), P(ctypes.c_int64)]),
            "lare_upload_field": (ctypes.c_int, [vp, ctypes.c_int, P(ctypes.c_double)]),
            "lare_download_field": (ctypes.c_int, [vp, ctypes.c_int, P(ctypes.c_double)]),
            "lare_set_bc_map": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, P(ctypes.c_int32),
                                               ctypes.c_double]),
            "lare_set_bc": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_double]),
            "lare_get_dt": (ctypes.c_int, [vp, P(ctypes.c_double)]),
            "lare_set_dt": (ctypes.c_int, [vp, ctypes.c_double]),
            "lare_start_of_timestep": (ctypes.c_int, [vp, P(ctypes.c_double)]),
            "lare_viscosity_dt": (ctypes.c_int, [vp, P(ctypes.c_double), P(ctypes.c_double)]),
            "lare_lagrange": (ctypes.c_int, [vp, ctypes.c_double]),
            "lare_half_timestep": (ctypes.c_int, [vp]),
            "lare_end_of_timestep": (ctypes.c_int, [vp]),
            "lare_step": (ctypes.c_int, [vp, P(ctypes.c_double)]),
            "lare_steps": (ctypes.c_int, [vp, ctypes.c_int32, P(ctypes.c_double)]),
            "lare_sync": (ctypes.c_int, [vp]),
            "lare_problem_desc": (ctypes.c_int, [P(LareProblem), P(LareDesc)]),
            "lare_problem_fill": (ctypes.c_int, [P(LareProblem), ctypes.c_int, P(ctypes.c_double)]),
            "lare_problem_bcs": (ctypes.c_int, [vp, P(LareProblem)]),
            "lare_problem_init": (ctypes.c_int, [vp, P(LareProblem)]),
            "lare_set_kernel_set": (ctypes.c_int, [vp, ctypes.c_int32]),
            "lare_set_graph": (ctypes.c_int, [vp, ctypes.c_int32]),
            "lare_profile_begin": (ctypes.c_int, [vp]),
            "lare_profile_end": (ctypes.c_int, [vp]),
            "lare_profile_count": (ctypes.c_int, [vp, P(ctypes.c_int32)]),
            "lare_profile_get": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32,
                                                P(ctypes.c_int32), P(ctypes.c_double), P(ctypes.c_double)]),
            "lare_energy": (ctypes.c_int, [vp, P(ctypes.c_double)]),
            "lare_energy_noalloc": (ctypes.c_int, [vp, P(ctypes.c_double)]),
            "lare_div_check": (ctypes.c_int, [P(ctypes.c_double), P(ctypes.c_double), P(ctypes.c_double),
                                              ctypes.c_int64, ctypes.c_int32]),
            "lare_batch_create": (ctypes.c_int, [P(vp), ctypes.c_int32, P(vp)]),
            "lare_batch_destroy": (ctypes.c_int, [vp]),
            "lare_batch_steps": (ctypes.c_int, [vp, ctypes.c_int32]),
            "lare_batch_sync": (ctypes.c_int, [vp]),
            "lare_batch_set_graph": (ctypes.c_int, [vp, ctypes.c_int32]),
            "lare_batch_energy": (ctypes.c_int, [vp, P(ctypes.c_double)]),
            "lare_batch_profile_begin": (ctypes.c_int, [vp]),
            "lare_batch_profile_end": (ctypes.c_int, [vp, P(ctypes.c_double), P(ctypes.c_int32)]),
            "lare_autotune": (ctypes.c_int, [vp, ctypes.c_int32, P(LareTune)]),
            "lare_get_tune": (ctypes.c_int, [vp, P(LareTune)]),
            "lare_set_tune": (ctypes.c_int, [vp, P(LareTune)]),
            "lare_batch_autotune": (ctypes.c_int, [vp, ctypes.c_int32, P(LareTune)]),
            "lare_batch_get_tune": (ctypes.c_int, [vp, P(LareTune)]),
            "lare_batch_set_tune": (ctypes.c_int, [vp, P(LareTune)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        msg = lib().lare_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")


def field_shape(name, nx, ny, nz):
    """Reference extent of a field: (-2 or -1)..n+2 per dimension."""
    st = STAGGER[name]
    n3 = (nx, ny, nz)
    return tuple(n3[k] + 2 - (-2 if st[k] else -1) + 1 for k in range(3))


def problem_desc(p):
    d = LareDesc()
    _check(lib().lare_problem_desc(ctypes.byref(p), ctypes.byref(d)), "lare_problem_desc")
    return d


def problem_fill(p, name):
    """Initial value of one field over its reference extent (host numpy)."""
    a = np.empty(field_shape(name, p.nx, p.ny, p.nz), dtype=np.float64)
    _check(lib().lare_problem_fill(ctypes.byref(p), FIELD_ID[name],
                                   a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
           "lare_problem_fill")
    return a


def div_check(a, d, mode=0):
    """a / d computed on the device as the kernels compute quotients: by the
    division every kernel uses (mode 0) or the shared-divisor form (mode 1);
    mode 2 gives sqrt(a) as the kernels compute it (d unused)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    d = np.ascontiguousarray(d, dtype=np.float64)
    if a.shape != d.shape:
        raise ValueError("div_check: a and d differ in shape")
    q = np.empty_like(a)
    P = ctypes.POINTER(ctypes.c_double)
    _check(lib().lare_div_check(a.ctypes.data_as(P), d.ctypes.data_as(P), q.ctypes.data_as(P), a.size, mode),
           "lare_div_check")
    return q


class Lare:
    """One device-resident LARE3D grid."""

    def __init__(self, desc):
        self.desc = desc
        h = ctypes.c_void_p()
        _check(lib().lare_ctx_create(ctypes.byref(desc), ctypes.byref(h)), "lare_ctx_create")
        self.h = h
        self.nx, self.ny, self.nz = desc.nx, desc.ny, desc.nz

    @classmethod
    def from_problem(cls, p):
        self = cls(problem_desc(p))
        _check(lib().lare_problem_init(self.h, ctypes.byref(p)), "lare_problem_init")
        return self

    def close(self):
        if getattr(self, "h", None):
            lib().lare_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, name):
        return field_shape(name, self.nx, self.ny, self.nz)

    def upload(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != self.shape(name):
            raise ValueError(f"{name}: expected shape {self.shape(name)}, got {a.shape}")
        _check(lib().lare_upload_field(self.h, FIELD_ID[name], a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
               "lare_upload_field")

    def download(self, name):
        a = np.empty(self.shape(name), dtype=np.float64)
        _check(lib().lare_download_field(self.h, FIELD_ID[name], a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
               "lare_download_field")
        return a

    def state(self):
        return {n: self.download(n) for n in STATE_FIELDS}

    def set_bc(self, name, dim, edge, kind, value=0.0):
        _check(lib().lare_set_bc(self.h, FIELD_ID[name], dim, edge, kind, value), "lare_set_bc")

    def set_bc_map(self, name, dim, edge, src, value=0.0):
        """A map edge (LARE-style condition): src lists the edge's ghost
        layers' sources (layer index, MAP_VALUE or MAP_KEEP)."""
        arr = (ctypes.c_int32 * 3)(*[int(x) for x in (list(src) + [MAP_KEEP] * 3)[:3]])
        _check(lib().lare_set_bc_map(self.h, FIELD_ID[name], dim, edge, arr, value), "lare_set_bc_map")

    def set_problem_bcs(self, p):
        _check(lib().lare_problem_bcs(self.h, ctypes.byref(p)), "lare_problem_bcs")

    def set_kernel_set(self, k):
        """1: fused kernels (default; lane-walking forms by launch size or as
        tuned); 2 / 3: fused with the lane-walking forms always / never;
        0: one kernel per reference kernel."""
        _check(lib().lare_set_kernel_set(self.h, int(k)), "lare_set_kernel_set")

    def autotune(self, steps=3):
        """Time every candidate launch configuration (state saved and
        restored), keep the fastest; returns it as a dict."""
        t = LareTune()
        _check(lib().lare_autotune(self.h, int(steps), ctypes.byref(t)), "lare_autotune")
        return t.as_dict()

    def get_tune(self):
        t = LareTune()
        _check(lib().lare_get_tune(self.h, ctypes.byref(t)), "lare_get_tune")
        return t.as_dict()

    def set_tune(self, **kw):
        t = LareTune()
        _check(lib().lare_get_tune(self.h, ctypes.byref(t)), "lare_get_tune")
        for k, v in kw.items():
            setattr(t, k, v)
        _check(lib().lare_set_tune(self.h, ctypes.byref(t)), "lare_set_tune")

    def set_graph(self, on):
        """Replay one captured HIP graph per step in steps() (default on)."""
        _check(lib().lare_set_graph(self.h, int(bool(on))), "lare_set_graph")

    def get_dt(self):
        d = ctypes.c_double()
        _check(lib().lare_get_dt(self.h, ctypes.byref(d)), "lare_get_dt")
        return d.value

    def set_dt(self, dt):
        _check(lib().lare_set_dt(self.h, float(dt)), "lare_set_dt")

    def start_of_timestep(self):
        d = ctypes.c_double()
        _check(lib().lare_start_of_timestep(self.h, ctypes.byref(d)), "lare_start_of_timestep")
        return d.value

    def viscosity_dt(self):
        """Prologue, shock viscosity and CFL reduction of startOfTimestep;
        returns (dt, dtr) as LARE3D::set_dt hands them to the runner."""
        d, r = ctypes.c_double(), ctypes.c_double()
        _check(lib().lare_viscosity_dt(self.h, ctypes.byref(d), ctypes.byref(r)), "lare_viscosity_dt")
        return d.value, r.value

    def lagrange(self, dt):
        """Rest of startOfTimestep with the runner's dt: resistive substeps
        and predictor."""
        _check(lib().lare_lagrange(self.h, float(dt)), "lare_lagrange")

    def half_timestep(self):
        _check(lib().lare_half_timestep(self.h), "lare_half_timestep")

    def end_of_timestep(self):
        _check(lib().lare_end_of_timestep(self.h), "lare_end_of_timestep")

    def step(self):
        d = ctypes.c_double()
        _check(lib().lare_step(self.h, ctypes.byref(d)), "lare_step")
        return d.value

    def steps(self, n, record_dt=True):
        if record_dt:
            out = np.empty(n, dtype=np.float64)
            _check(lib().lare_steps(self.h, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "lare_steps")
            return out
        _check(lib().lare_steps(self.h, n, None), "lare_steps")
        return None

    def sync(self):
        _check(lib().lare_sync(self.h), "lare_sync")

    def profile(self, nsteps):
        """Run nsteps with every kernel bracketed by HIP events; returns
        {kernel: (calls, total_ms, algorithmic_bytes)}."""
        _check(lib().lare_profile_begin(self.h), "lare_profile_begin")
        try:
            self.steps(nsteps, record_dt=False)
        finally:
            _check(lib().lare_profile_end(self.h), "lare_profile_end")
        n = ctypes.c_int32()
        _check(lib().lare_profile_count(self.h, ctypes.byref(n)), "lare_profile_count")
        out = {}
        buf = ctypes.create_string_buffer(128)
        for i in range(n.value):
            calls = ctypes.c_int32()
            ms = ctypes.c_double()
            by = ctypes.c_double()
            _check(lib().lare_profile_get(self.h, i, buf, 128, ctypes.byref(calls), ctypes.byref(ms),
                                          ctypes.byref(by)), "lare_profile_get")
            out[buf.value.decode()] = (calls.value, ms.value, by.value)
        return out


    def energy(self):
        """FP64 energy diagnostic: (internal, kinetic, magnetic, total)."""
        out = (ctypes.c_double * 4)()
        _check(lib().lare_energy(self.h, out), "lare_energy")
        return tuple(out)

    def energy_noalloc(self):
        """The same diagnostic through lare_energy_noalloc, which allocates no
        device memory (a fixed per-GPU scratch array; grids up to about
        560^3 cells)."""
        out = (ctypes.c_double * 4)()
        _check(lib().lare_energy_noalloc(self.h, out), "lare_energy_noalloc")
        return tuple(out)


class Batch:
    """Grids of one storage precision that advance together: every kernel of
    a step is one launch over all members.  Members keep their own state,
    dt and boundary conditions."""

    def __init__(self, members):
        self.members = list(members)
        arr = (ctypes.c_void_p * len(self.members))(*[m.h.value for m in self.members])
        h = ctypes.c_void_p()
        _check(lib().lare_batch_create(arr, len(self.members), ctypes.byref(h)), "lare_batch_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().lare_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def steps(self, n):
        """Enqueue n steps of every member (returns at once)."""
        _check(lib().lare_batch_steps(self.h, int(n)), "lare_batch_steps")

    def sync(self):
        _check(lib().lare_batch_sync(self.h), "lare_batch_sync")

    def set_graph(self, on):
        _check(lib().lare_batch_set_graph(self.h, int(bool(on))), "lare_batch_set_graph")

    def autotune(self, steps=3):
        """lare_batch_autotune: the fastest launch configuration for this
        batch (members' state saved and restored); returns it as a dict."""
        t = LareTune()
        _check(lib().lare_batch_autotune(self.h, int(steps), ctypes.byref(t)), "lare_batch_autotune")
        return t.as_dict()

    def get_tune(self):
        t = LareTune()
        _check(lib().lare_batch_get_tune(self.h, ctypes.byref(t)), "lare_batch_get_tune")
        return t.as_dict()

    def set_tune(self, **kw):
        t = LareTune()
        _check(lib().lare_batch_get_tune(self.h, ctypes.byref(t)), "lare_batch_get_tune")
        for k, v in kw.items():
            setattr(t, k, v)
        _check(lib().lare_batch_set_tune(self.h, ctypes.byref(t)), "lare_batch_set_tune")

    def energy(self):
        out = (ctypes.c_double * (4 * len(self.members)))()
        _check(lib().lare_batch_energy(self.h, out), "lare_batch_energy")
        return np.array(out).reshape(len(self.members), 4)

    def profile(self, nsteps):
        """Eager steps with HIP events around every launch: (ms of kernel time
        per step, launches per step)."""
        _check(lib().lare_batch_profile_begin(self.h), "lare_batch_profile_begin")
        self.steps(nsteps)
        ms, n = ctypes.c_double(), ctypes.c_int32()
        _check(lib().lare_batch_profile_end(self.h, ctypes.byref(ms), ctypes.byref(n)), "lare_batch_profile_end")
        return ms.value / nsteps, n.value / nsteps
