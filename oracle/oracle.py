"""ctypes wrapper of the CPU oracle (oracle/efs_oracle.cpp) — test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this module, as the
checker; the product never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libefs_oracle.so")

F_VEL, F_RHO, F_GAMMA, F_BETA, F_DIV, F_PRESSURE, F_FORCES, F_BOUNDARY, F_TMPSCAL, F_TMPVEC = range(10)
FIELDS = {"vel": F_VEL, "rho": F_RHO, "gamma": F_GAMMA, "beta": F_BETA, "divergence": F_DIV,
          "pressure": F_PRESSURE, "forces": F_FORCES, "boundary": F_BOUNDARY, "tmpscal": F_TMPSCAL, "tmpvec": F_TMPVEC}


class OraParams(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_float), ("gravity_force", ctypes.c_float), ("lambda_minus", ctypes.c_float),
        ("lambda_plus", ctypes.c_float), ("eta", ctypes.c_float), ("liquid_threshold", ctypes.c_float),
        ("rho_d", ctypes.c_float), ("vel_decay", ctypes.c_float), ("gravity_enabled", ctypes.c_int32),
        ("solver_type", ctypes.c_int32), ("max_steps", ctypes.c_int32), ("solver_threshold", ctypes.c_float),
        ("disable_mask", ctypes.c_uint32), ("key64", ctypes.c_int32), ("sharpen_mode", ctypes.c_int32),
        ("re_calc_div", ctypes.c_int32),
    ]


class OraStats(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("iterations", ctypes.c_int32), ("err", ctypes.c_double)]


class OraPrim(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("solid", ctypes.c_int32), ("a", ctypes.c_double * 3), ("b", ctypes.c_double * 3)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    vp, i32, f64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    P = ctypes.POINTER
    sigs = {
        "ora3_create": (vp, [i32, i32, i32, i32]), "ora3_destroy": (None, [vp, i32]),
        "ora3_set_params": (None, [vp, i32, P(OraParams)]), "ora3_field": (vp, [vp, i32, i32]),
        "ora3_step": (i32, [vp, i32, P(OraStats)]), "ora3_stage": (i32, [vp, i32, i32, P(OraStats)]),
        "ora3_solve": (i32, [vp, i32, P(OraStats)]), "ora3_divergence": (None, [vp, i32]),
        "ora3_enforce_walls": (None, [vp, i32]), "ora3_matmul": (None, [vp, i32, vp, vp]),
        "ora3_trace": (None, [vp, i32, P(f64), f64, i32, P(f64)]), "ora3_vinterp": (None, [vp, i32, P(f64), P(f64)]),
        "ora3_sinterp": (f64, [vp, i32, i32, P(f64)]), "ora3_num_entries": (ctypes.c_size_t, [vp, i32]),
        "ora2_create": (vp, [i32, i32, i32]), "ora2_destroy": (None, [vp, i32]),
        "ora2_set_params": (None, [vp, i32, P(OraParams)]), "ora2_field": (vp, [vp, i32, i32]),
        "ora2_step": (i32, [vp, i32, P(OraStats)]), "ora2_stage": (i32, [vp, i32, i32, P(OraStats)]),
        "ora2_solve": (i32, [vp, i32, P(OraStats)]), "ora2_matmul": (None, [vp, i32, vp, vp]),
        "ora3_cg64": (i32, [vp, vp, f64, i32, i32, P(f64)]),
        "ora3_gradient": (None, [vp, i32]), "ora3_sharpen_delta": (None, [vp, i32]),
        "ora3_trace_field": (None, [vp, i32, P(f64), P(f64)]), "ora3_scatter": (None, [vp, i32, P(f64), f64]),
        "ora2_gradient": (None, [vp, i32]), "ora2_enforce_walls": (None, [vp, i32]),
        "ora2_divergence": (None, [vp, i32]), "ora2_sharpen_delta": (None, [vp, i32]),
        "ora2_sinterp": (f64, [vp, i32, i32, P(f64)]), "ora2_trace": (None, [vp, i32, P(f64), f64, i32, P(f64)]),
        "ora_uniform": (f64, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]),
        "ora_normal": (f64, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]),
        "ora_rasterise": (None, [P(i32), P(OraPrim), i32, vp, vp]),
        "ora_face_noise": (None, [P(i32), vp, ctypes.c_uint64, f64, vp]),
    }
    for k, (r, a) in sigs.items():
        fn = getattr(lib, k)
        fn.restype = r
        fn.argtypes = a
    _lib = lib
    return lib


def params_from_scene(scene, **over):
    p = OraParams()
    p.dt, p.gravity_force = scene.dt, scene.gravity_force
    p.lambda_minus, p.lambda_plus, p.eta, p.liquid_threshold, p.rho_d, p.vel_decay = 0.5, 0.1, 1.0, 0.5, 3.0, 0.8
    p.gravity_enabled, p.solver_type, p.max_steps = 1, scene.solver_type, scene.max_steps
    p.solver_threshold, p.disable_mask, p.key64, p.sharpen_mode, p.re_calc_div = scene.solver_threshold, 0, 0, 0, 0
    for k, v in over.items():
        setattr(p, k, v)
    return p


def default_params(**over):
    p = OraParams()
    p.dt, p.gravity_force = 0.0167, 10.0
    p.lambda_minus, p.lambda_plus, p.eta, p.liquid_threshold, p.rho_d, p.vel_decay = 0.5, 0.1, 1.0, 0.5, 3.0, 0.8
    p.gravity_enabled, p.solver_type, p.max_steps, p.solver_threshold = 1, 2, 2000, 1e-5
    p.disable_mask, p.key64, p.sharpen_mode, p.re_calc_div = 0, 0, 0, 0
    for k, v in over.items():
        setattr(p, k, v)
    return p


def prims_array(prims):
    arr = (OraPrim * max(1, len(prims)))()
    for i, p in enumerate(prims):
        arr[i].kind, arr[i].solid = p.kind, p.solid
        for k in range(3):
            arr[i].a[k], arr[i].b[k] = p.a[k], p.b[k]
    return arr


class Oracle3D:
    """CPUSimulator3D restated; dbl=True gives the FP64 instance (ref64)."""

    def __init__(self, W, H, D, params=None, dbl=False):
        self.lib = load()
        self.dims = (W, H, D)
        self.dbl = int(bool(dbl))
        self.dtype = np.float64 if dbl else np.float32
        self.h = self.lib.ora3_create(W, H, D, self.dbl)
        self.set_params(params if params is not None else default_params())

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ora3_destroy(self.h, self.dbl)
            self.h = None

    def set_params(self, p):
        self.params = p
        self.lib.ora3_set_params(self.h, self.dbl, ctypes.byref(p))

    def field(self, name):
        """numpy view (no copy) of a field: (D, H, W) or (D, H, W, 3)."""
        W, H, D = self.dims
        f = FIELDS[name]
        ptr = self.lib.ora3_field(self.h, self.dbl, f)
        shape = (D, H, W, 3) if f in (F_VEL, F_FORCES, F_TMPVEC) else (D, H, W)
        n = int(np.prod(shape))
        ct = ctypes.c_double if self.dbl else ctypes.c_float
        return np.ctypeslib.as_array((ct * n).from_address(ptr)).reshape(shape)

    def step(self):
        st = OraStats()
        self.lib.ora3_step(self.h, self.dbl, ctypes.byref(st))
        return st

    def stage(self, s):
        st = OraStats()
        self.lib.ora3_stage(self.h, self.dbl, s, ctypes.byref(st))
        return st

    def solve(self):
        st = OraStats()
        self.lib.ora3_solve(self.h, self.dbl, ctypes.byref(st))
        return st

    def divergence(self):
        self.lib.ora3_divergence(self.h, self.dbl)

    def matmul(self, x):
        x = np.ascontiguousarray(x, dtype=self.dtype)
        y = np.empty_like(x)
        self.lib.ora3_matmul(self.h, self.dbl, x.ctypes.data, y.ctypes.data)
        return y

    def trace(self, start, dt, substeps=4):
        s = (ctypes.c_double * 3)(*start)
        o = (ctypes.c_double * 3)()
        self.lib.ora3_trace(self.h, self.dbl, s, dt, substeps, o)
        return np.array(list(o))

    def vinterp(self, pos):
        s = (ctypes.c_double * 3)(*pos)
        o = (ctypes.c_double * 3)()
        self.lib.ora3_vinterp(self.h, self.dbl, s, o)
        return np.array(list(o))

    def sinterp(self, name, pos):
        s = (ctypes.c_double * 3)(*pos)
        return self.lib.ora3_sinterp(self.h, self.dbl, FIELDS[name], s)

    def num_entries(self):
        return self.lib.ora3_num_entries(self.h, self.dbl)

    def solve64(self, x0=None, rel_tol=1e-11, max_iter=20000, threads=None):
        """FP64 CG on the current divergence (FP64 instance only), OpenMP-parallel, to
        ||b - A p|| / ||b|| <= rel_tol; returns (iterations, final relative residual)."""
        assert self.dbl, "solve64 runs on the FP64 instance"
        if threads is None:
            threads = max(1, min(16, os.cpu_count() or 1))
        xp = None
        if x0 is not None:
            x0 = np.ascontiguousarray(x0, dtype=np.float64)
            xp = x0.ctypes.data
        rel = ctypes.c_double()
        it = self.lib.ora3_cg64(self.h, xp, rel_tol, max_iter, threads, ctypes.byref(rel))
        return it, rel.value

    def gradient(self):
        """project()'s pressure-gradient subtraction alone, on the current pressure."""
        self.lib.ora3_gradient(self.h, self.dbl)

    def enforce_walls(self):
        self.lib.ora3_enforce_walls(self.h, self.dbl)

    def sharpen_delta(self):
        """rho_sharpen's per-cell delta rho (before any is applied); rho is unchanged."""
        self.lib.ora3_sharpen_delta(self.h, self.dbl)
        return self.field("tmpscal").copy()

    def trace_field(self, start):
        s = (ctypes.c_double * 3)(*start)
        o = (ctypes.c_double * 3)()
        self.lib.ora3_trace_field(self.h, self.dbl, s, o)
        return np.array(list(o))

    def scatter(self, pos, value):
        s = (ctypes.c_double * 3)(*pos)
        self.lib.ora3_scatter(self.h, self.dbl, s, value)

    def load_scene(self, scene):
        """Rasterise the scene's primitives and add the seeded face noise (float inputs)."""
        W, H, D = self.dims
        dims = (ctypes.c_int * 3)(W, H, D)
        rho = np.zeros((D, H, W), np.float32)
        bnd = np.zeros((D, H, W), np.float32)
        pa = prims_array(scene.prims)
        self.lib.ora_rasterise(dims, pa, len(scene.prims), rho.ctypes.data, bnd.ctypes.data)
        vel = np.zeros((D, H, W, 3), np.float32)
        if scene.noise_sigma:
            self.lib.ora_face_noise(dims, bnd.ctypes.data, scene.seed, scene.noise_sigma, vel.ctypes.data)
        self.field("rho")[...] = rho
        self.field("boundary")[...] = bnd
        self.field("vel")[...] = vel
        return rho, bnd, vel


class Oracle2D:
    """CPUSimulator2D restated."""

    def __init__(self, W, H, params=None, dbl=False):
        self.lib = load()
        self.dims = (W, H)
        self.dbl = int(bool(dbl))
        self.dtype = np.float64 if dbl else np.float32
        self.h = self.lib.ora2_create(W, H, self.dbl)
        self.set_params(params if params is not None else default_params())

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ora2_destroy(self.h, self.dbl)
            self.h = None

    def set_params(self, p):
        self.params = p
        self.lib.ora2_set_params(self.h, self.dbl, ctypes.byref(p))

    def field(self, name):
        W, H = self.dims
        f = FIELDS[name]
        ptr = self.lib.ora2_field(self.h, self.dbl, f)
        shape = (H, W, 2) if f in (F_VEL, F_FORCES, F_TMPVEC) else (H, W)
        n = int(np.prod(shape))
        ct = ctypes.c_double if self.dbl else ctypes.c_float
        return np.ctypeslib.as_array((ct * n).from_address(ptr)).reshape(shape)

    def step(self):
        st = OraStats()
        self.lib.ora2_step(self.h, self.dbl, ctypes.byref(st))
        return st

    def stage(self, s):
        st = OraStats()
        self.lib.ora2_stage(self.h, self.dbl, s, ctypes.byref(st))
        return st

    def solve(self):
        st = OraStats()
        self.lib.ora2_solve(self.h, self.dbl, ctypes.byref(st))
        return st

    def matmul(self, x):
        x = np.ascontiguousarray(x, dtype=self.dtype)
        y = np.empty_like(x)
        self.lib.ora2_matmul(self.h, self.dbl, x.ctypes.data, y.ctypes.data)
        return y

    def gradient(self):
        self.lib.ora2_gradient(self.h, self.dbl)

    def enforce_walls(self):
        self.lib.ora2_enforce_walls(self.h, self.dbl)

    def divergence(self):
        self.lib.ora2_divergence(self.h, self.dbl)

    def sharpen_delta(self):
        self.lib.ora2_sharpen_delta(self.h, self.dbl)
        return self.field("tmpscal").copy()

    def sinterp(self, name, pos):
        s = (ctypes.c_double * 2)(*pos)
        return self.lib.ora2_sinterp(self.h, self.dbl, FIELDS[name], s)

    def trace(self, start, dt, substeps=4):
        s = (ctypes.c_double * 2)(*start)
        o = (ctypes.c_double * 2)()
        self.lib.ora2_trace(self.h, self.dbl, s, dt, substeps, o)
        return np.array(list(o))

    def load_scene(self, scene):
        W, H = self.dims
        dims = (ctypes.c_int * 3)(W, H, 0)
        rho = np.zeros((H, W), np.float32)
        bnd = np.zeros((H, W), np.float32)
        pa = prims_array(scene.prims)
        self.lib.ora_rasterise(dims, pa, len(scene.prims), rho.ctypes.data, bnd.ctypes.data)
        vel = np.zeros((H, W, 2), np.float32)
        if scene.noise_sigma:
            self.lib.ora_face_noise(dims, bnd.ctypes.data, scene.seed, scene.noise_sigma, vel.ctypes.data)
        self.field("rho")[...] = rho
        self.field("boundary")[...] = bnd
        self.field("vel")[...] = vel
        return rho, bnd, vel
