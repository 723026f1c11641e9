"""Seeded scene inputs for the workloads (the reference ships none; its editor rasterises box
spawns in grid-index space, Editor.cpp:176-182, 246-257, and has no RNG).

A scene is a list of primitives evaluated at interior cell centres ((i - 0.5) h, h = 1/n for an
n-cell interior) plus the step parameters.  Coordinates are in the reference's axes: gravity
acts along -y (CPUSimulator3D.cpp:474), so a workload's "height" / vertical axis is y.

The counter-based generator (splitmix64 finaliser, Box-Muller in double) is restated in the
library's host code and in the oracle; tests check all three agree bit for bit.
"""
from dataclasses import dataclass, field
import math
from typing import List, Tuple

MASK = (1 << 64) - 1
BOX, SPHERE = 0, 1


def mix64(z: int) -> int:
    z = (z + 0x9E3779B97F4A7C15) & MASK
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def uniform01(seed: int, stream: int, ctr: int) -> float:
    key = mix64(seed & MASK) ^ ((stream * 0xD1B54A32D192ED03) & MASK)
    return (mix64((key + ctr * 0x9E3779B97F4A7C15) & MASK) >> 11) * 2.0 ** -53


def uniform(seed, stream, ctr, lo, hi):
    return lo + (hi - lo) * uniform01(seed, stream, ctr)


@dataclass
class Prim:
    kind: int
    solid: int
    a: Tuple[float, float, float]
    b: Tuple[float, float, float]


@dataclass
class Scene:
    name: str
    dims: Tuple[int, int, int]          # grid incl. the one-cell border; dims[2] == 0 for 2D
    prims: List[Prim]
    dt: float
    gravity_force: float                # cells/s^2 (GRAVITY_FORCE_PHYSICAL / CELL_SIZE)
    noise_sigma: float = 1e-3
    seed: int = 0
    max_steps: int = 2000
    solver_threshold: float = 1e-5
    solver_type: int = 2                # SELF_CG
    steps: int = 100
    extra: dict = field(default_factory=dict)

    def params(self, lib_params_cls):
        p = lib_params_cls()
        p.dt = self.dt
        p.gravity_force = self.gravity_force
        p.lambda_minus = 0.5
        p.lambda_plus = 0.1
        p.eta = 1.0
        p.liquid_threshold = 0.5
        p.rho_d = 3.0
        p.vel_decay = 0.8
        p.gravity_enabled = 1
        p.solver_type = self.solver_type
        p.max_steps = self.max_steps
        p.solver_threshold = self.solver_threshold
        p.disable_mask = 0
        return p


BIG = 4.0


def _box(lo, hi, solid=0):
    return Prim(BOX, solid, tuple(lo), tuple(hi))


def dam2d(n=128, solver="cg"):
    """Config 1: 2D dam break, fluid x < 0.4, y < 0.6, h = 1/n, dt = 1/120, 100 steps."""
    return Scene(name=f"dam2d_{n}", dims=(n + 2, n + 2, 0),
                 prims=[_box((-BIG, -BIG, -BIG), (0.4, 0.6, BIG))],
                 dt=1.0 / 120.0, gravity_force=9.81 * n, steps=100,
                 solver_type=3 if solver == "jacobi" else 2,
                 max_steps=400 if solver == "jacobi" else 2000)


def pool2d(n=64, seed=0, nboxes=4, solver="cg"):
    """2D counterpart of the sphere-over-pool scene: pool of depth 0.3, a fluid disc r = 0.15 at
    height 0.7 and `nboxes` seeded solid boxes (side U(0.05, 0.2), centre U(0, 1)^2, stream 5)."""
    prims = [_box((-BIG, -BIG, -BIG), (BIG, 0.3, BIG)), Prim(SPHERE, 0, (0.5, 0.7, 0.0), (0.15, 0.0, 0.0))]
    for k in range(nboxes):
        side = [uniform(seed, 5, 4 * k + a, 0.05, 0.2) for a in range(2)]
        cen = [uniform01(seed, 5, 4 * k + 2 + a) for a in range(2)]
        prims.append(_box([cen[0] - side[0] / 2, cen[1] - side[1] / 2, -BIG],
                          [cen[0] + side[0] / 2, cen[1] + side[1] / 2, BIG], solid=1))
    return Scene(name=f"pool2d_{n}", dims=(n + 2, n + 2, 0), prims=prims, dt=1.0 / 120.0, gravity_force=9.81 * n,
                 steps=100, seed=seed, solver_type=3 if solver == "jacobi" else 2,
                 max_steps=400 if solver == "jacobi" else 2000)


def dam3d(n=128):
    """Config 2: 3D dam break, fluid x < 0.4 and (vertical) y < 0.5 over the full depth,
    dt = 1/60, CG tolerance 1e-5, at most 500 iterations, 200 steps."""
    return Scene(name=f"dam3d_{n}", dims=(n + 2, n + 2, n + 2),
                 prims=[_box((-BIG, -BIG, -BIG), (0.4, 0.5, BIG))],
                 dt=1.0 / 60.0, gravity_force=9.81 * n, steps=200, max_steps=500)


def sphere_pool(n=256, seed=0, nboxes=8):
    """Config 3: pool of depth 0.3, a fluid sphere r = 0.15 centred at height 0.7, and 8 seeded
    solid boxes (side U(0.05, 0.2), centre U(0, 1)^3, stream 2)."""
    prims = [_box((-BIG, -BIG, -BIG), (BIG, 0.3, BIG)),
             Prim(SPHERE, 0, (0.5, 0.7, 0.5), (0.15, 0.0, 0.0))]
    for k in range(nboxes):
        side = [uniform(seed, 2, 6 * k + a, 0.05, 0.2) for a in range(3)]
        cen = [uniform01(seed, 2, 6 * k + 3 + a) for a in range(3)]
        prims.append(_box([c - s / 2 for c, s in zip(cen, side)], [c + s / 2 for c, s in zip(cen, side)], solid=1))
    return Scene(name=f"sphere_pool_{n}", dims=(n + 2, n + 2, n + 2), prims=prims,
                 dt=1.0 / 60.0, gravity_force=9.81 * n, steps=100, seed=seed)


def dam_spheres(n=256, seed=0, nspheres=12):
    """Config 5: dam break (x < 0.4, y < 0.5) around a static solid = union of 12 seeded
    spheres (radius U(0.03, 0.1), centre U(0, 1)^3, stream 3)."""
    prims = [_box((-BIG, -BIG, -BIG), (0.4, 0.5, BIG))]
    for k in range(nspheres):
        r = uniform(seed, 3, 4 * k, 0.03, 0.1)
        cen = tuple(uniform01(seed, 3, 4 * k + 1 + a) for a in range(3))
        prims.append(Prim(SPHERE, 1, cen, (r, 0.0, 0.0)))
    return Scene(name=f"dam_spheres_{n}", dims=(n + 2, n + 2, n + 2), prims=prims,
                 dt=1.0 / 60.0, gravity_force=9.81 * n, steps=500, seed=seed)


def training_scene(n=192, seed=0):
    """Config 4 sample: 1-4 fluid boxes, 0-10 solid boxes (stream 4), face velocities
    N(0, 0.5)."""
    nf = 1 + int(uniform01(seed, 4, 0) * 4)
    ns = int(uniform01(seed, 4, 1) * 11)
    prims = []
    ctr = 2
    for k in range(nf):
        lo = [uniform(seed, 4, ctr + a, 0.0, 0.7) for a in range(3)]
        ext = [uniform(seed, 4, ctr + 3 + a, 0.1, 0.5) for a in range(3)]
        ctr += 6
        prims.append(_box(lo, [l + e for l, e in zip(lo, ext)]))
    for k in range(ns):
        side = [uniform(seed, 4, ctr + a, 0.05, 0.2) for a in range(3)]
        cen = [uniform01(seed, 4, ctr + 3 + a) for a in range(3)]
        ctr += 6
        prims.append(_box([c - s / 2 for c, s in zip(cen, side)], [c + s / 2 for c, s in zip(cen, side)], solid=1))
    return Scene(name=f"train_{n}_{seed}", dims=(n + 2, n + 2, n + 2), prims=prims,
                 dt=1.0 / 60.0, gravity_force=9.81 * n, noise_sigma=0.5, seed=seed, steps=1)


CONFIGS = {
    "dam2d": dam2d,
    "dam3d": dam3d,
    "sphere_pool": sphere_pool,
    "dam_spheres": dam_spheres,
    "training": training_scene,
}
