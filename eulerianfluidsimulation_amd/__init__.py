"""MI355X (gfx950) HIP backend for the Eulerian liquid step of MarQMark/EulerianFluidSimulation.

`HipSimulator3D` / `HipSimulator2D` mirror the reference's ISimulator3D / ISimulator2D backend
interface (include/simulation/simulator/ISimulator.h:8-53): `update()` runs the seven stages in
order with the DISABLE_* gating, the protected stage methods are exposed under the same names,
and `grid()` gives the IGrid3D / IGrid2D field accessors (include/simulation/data/IGrid.h) as
host mirrors in the reference layout.  Every call goes through the C ABI of include/efs.h; the
library must be built (there is no CPU fallback).
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import (FIELD_BETA, FIELD_BOUNDARY, FIELD_DIVERGENCE, FIELD_FORCES, FIELD_GAMMA, FIELD_PRESSURE, FIELD_RHO,
                   FIELD_VEL, Params, Prim, Stats, EfsError, check)

__all__ = ["HipSimulator3D", "HipGrid3D", "HipSimulator2D", "HipGrid2D", "Params", "Stats", "EfsError", "default_params",
           "load_library"]

_FIELD_NAMES = {
    "vel": FIELD_VEL, "rho": FIELD_RHO, "gamma": FIELD_GAMMA, "beta": FIELD_BETA,
    "divergence": FIELD_DIVERGENCE, "pressure": FIELD_PRESSURE, "forces": FIELD_FORCES, "boundary": FIELD_BOUNDARY,
}


def load_library():
    return _abi.load()


def default_params() -> Params:
    p = Params()
    _abi.load().efs_default_params(ctypes.byref(p))
    return p


class HipGrid3D:
    """IGrid3D over the device fields: each accessor downloads the field (reference layout,
    float32, shape (D, H, W) or (D, H, W, 3)); `set` uploads one."""

    def __init__(self, sim):
        self._sim = sim

    def _shape(self, f):
        W, H, D = self._sim.dims
        return (D, H, W, 3) if f in (FIELD_VEL, FIELD_FORCES) else (D, H, W)

    def _get(self, f):
        shape = self._shape(f)
        a = np.empty(shape, dtype=np.float32)
        check(self._sim.lib.efs_download(self._sim.h, f, a.ctypes.data, a.nbytes))
        return a

    def set(self, name, array):
        f = _FIELD_NAMES[name]
        a = np.ascontiguousarray(array, dtype=np.float32)
        check(self._sim.lib.efs_upload(self._sim.h, f, a.ctypes.data, a.nbytes))

    def vel(self):
        return self._get(FIELD_VEL)

    def rho(self):
        return self._get(FIELD_RHO)

    def gamma(self):
        return self._get(FIELD_GAMMA)

    def beta(self):
        return self._get(FIELD_BETA)

    def divergence(self):
        return self._get(FIELD_DIVERGENCE)

    def pressure(self):
        return self._get(FIELD_PRESSURE)

    def forces(self):
        return self._get(FIELD_FORCES)

    def boundary(self):
        return self._get(FIELD_BOUNDARY)


class HipGrid2D(HipGrid3D):
    """IGrid2D over the device fields: shape (H, W) or (H, W, 2) (Grid2D's y*W + x order)."""

    def _shape(self, f):
        W, H = self._sim.dims
        return (H, W, 2) if f in (FIELD_VEL, FIELD_FORCES) else (H, W)


class HipSimulator3D:
    """ISimulator3D(width, height, depth) on one GPU."""

    def __init__(self, width, height, depth, params: Params = None):
        self.lib = _abi.load()
        self.dims = (int(width), int(height), int(depth))
        self.params = params if params is not None else default_params()
        h = ctypes.c_void_p()
        check(self.lib.efs_create_3d(width, height, depth, ctypes.byref(self.params), ctypes.byref(h)))
        self.h = h
        self._init_state()

    def _init_state(self):
        self.last_stats = Stats()
        self.paused = False
        self.frames = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.efs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- parameters (PROPERTY_* reads) --
    def set_params(self, params: Params):
        self.params = params
        check(self.lib.efs_set_params(self.h, ctypes.byref(params)))

    # -- ISimulator::update --
    def update(self) -> int:
        st = Stats()
        rc = check(self.lib.efs_step(self.h, ctypes.byref(st)), allowed=(0, 1, 2))
        self.last_stats = st
        if rc == 1:  # CPUSimulator3D.cpp:542-546: non-convergence pauses the simulation
            self.paused = True
        self.frames += 1  # SIMULATION_FRAMES
        return rc

    def _stage(self, s):
        st = Stats()
        rc = check(self.lib.efs_stage(self.h, s, ctypes.byref(st)), allowed=(0, 1, 2))
        self.last_stats = st
        return rc

    def vel_decay(self):
        return self._stage(_abi.STAGE_VEL_DECAY)

    def vel_extrapolate(self):
        return self._stage(_abi.STAGE_VEL_EXTRAPOLATE)

    def rho_sharpen(self):
        return self._stage(_abi.STAGE_RHO_SHARPEN)

    def rho_advect(self):
        return self._stage(_abi.STAGE_RHO_ADVECT)

    def vel_advect(self):
        return self._stage(_abi.STAGE_VEL_ADVECT)

    def add_force(self):
        return self._stage(_abi.STAGE_ADD_FORCE)

    def project(self):
        return self._stage(_abi.STAGE_PROJECT)

    def project_divergence(self):
        """Wall enforcement and divergence only (the first half of project)."""
        check(self.lib.efs_project_divergence(self.h))

    # -- ISolver3D::solve on the current divergence --
    def solve_pressure(self) -> Stats:
        st = Stats()
        check(self.lib.efs_solve_pressure(self.h, ctypes.byref(st)), allowed=(0, 1, 2))
        self.last_stats = st
        return st

    # -- IGrid3D / getRenderData --
    def grid(self) -> HipGrid3D:
        return HipGrid3D(self)

    def get_render_data(self):
        return {"dim": 3, "grid": self.grid()}

    def render_feed(self, vel=True):
        """The marching-cubes feed (MarchingCubes.cpp:75-105): the density, shaped like the grid,
        and with vel the velocity as vec4 per cell (w = 0), packed on the device."""
        import numpy as np
        n = int(np.prod(self.dims))
        rho = np.empty(n, np.float32)
        vel4 = np.empty((n, 4), np.float32) if vel else None
        check(self.lib.efs_render_feed(self.h, rho.ctypes.data, vel4.ctypes.data if vel else None))
        shape = tuple(reversed(self.dims)) if len(self.dims) == 3 else (self.dims[1], self.dims[0])
        return rho.reshape(shape), (vel4.reshape(shape + (4,)) if vel else None)

    # -- scenes --
    def load_scene(self, scene):
        prims = (Prim * max(1, len(scene.prims)))()
        for i, p in enumerate(scene.prims):
            prims[i].kind = p.kind
            prims[i].solid = p.solid
            for k in range(3):
                prims[i].a[k] = p.a[k]
                prims[i].b[k] = p.b[k]
        check(self.lib.efs_scene_rasterise(self.h, prims, len(scene.prims)))
        if scene.noise_sigma:
            check(self.lib.efs_scene_face_noise(self.h, scene.seed, scene.noise_sigma))

    def synchronize(self):
        check(self.lib.efs_synchronize(self.h))


class HipSimulator2D(HipSimulator3D):
    """ISimulator2D(width, height) on one GPU with CPUSimulator2D's semantics
    (src/simulation/simulator/CPUSimulator2D.cpp)."""

    def __init__(self, width, height, params: Params = None):
        self.lib = _abi.load()
        self.dims = (int(width), int(height))
        self.params = params if params is not None else default_params()
        h = ctypes.c_void_p()
        check(self.lib.efs_create_2d(width, height, ctypes.byref(self.params), ctypes.byref(h)))
        self.h = h
        self._init_state()

    def grid(self) -> HipGrid2D:
        return HipGrid2D(self)

    def get_render_data(self):
        return {"dim": 2, "grid": self.grid()}
