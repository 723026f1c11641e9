"""Loader of the reference's 3D scene files (assets/scenes/*.json, written by Scene::save,
src/simulation/scene/Scene.cpp:140-165) for the HIP backend.

It rebuilds what Editor::init does on a reset (src/core/Editor.cpp:21-40, 90-258) with the
backend's entry points:
  * the grid size is the SimulationDomainNode's "resolution" (SimulationDomainNode.cpp:21-29);
    its "position" / "scale" are the world box the meshes are voxelised in;
  * the "properties" block gives the step parameters (include/simulation/Properties.h:136-200
    names; GRAVITY_FORCE = Gravity / Cell Size as src/Application.cpp:89 sets it);
  * fluid: ModelFluidSpawnNode meshes are voxelised into rho (1 inside, 0 outside), then every
    BoundingBoxFluidSpawnNode sets rho = 1 on the cells [position, position + scale) in grid
    units (Editor.cpp:176-182; BoundingBoxFluidSpawnNode::getFluidSpawn);
  * solids: the ModelNode meshes with "boundary": true are voxelised into the boundary field,
    the domain border is made solid and rho is cleared under every solid (Editor.cpp:191-257).
Meshes are Wavefront OBJ files ("v" and "f" records, polygons fanned into triangles), placed by
the node's translate * rotate * scale transform (Transform.cpp:70-76; "rotation" is stored as
x, y, z, w, ModelNode.cpp:29-30).

ForceFieldNodes (ForceFieldNode.cpp:24-45) become the forces field as
Editor::updateForceFields sets it (Editor.cpp:260-330): each field's cells are the cell centres
inside its transformed unit cube [-0.5, 0.5]^3 (the voxeliser's inside test,
assets/shaders/forcefield_cells.comp), a field over "Max ForceField cells" (default 1000) is
skipped, and the active fields write direction * strength on their cells in scene order, the
direction being the normalised first column of the node's transform.  QueryNodes are returned
as query.QueryNode objects; other node types are kept as plain records.
"""
import ctypes
import json
import math
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import _abi
from ._abi import check

# "properties" names -> efs_params fields
_FLOAT_PROPS = {"DT": "dt", "Lambda-": "lambda_minus", "Lambda+": "lambda_plus", "Eta": "eta",
                "Liquid Threshold": "liquid_threshold", "D": "rho_d", "Vel Decay": "vel_decay",
                "Threshold": "solver_threshold"}
_INT_PROPS = {"Max Steps": "max_steps", "Gravity Enabled": "gravity_enabled", "Re-Calc Div": "re_calc_div",
              "Solver": "solver_type"}
_DISABLE = ["Disable Velocity Decay", "Disable Velocity Extrapolation", "Disable Density Sharpen",
            "Disable Density Advection", "Disable Velocity Advection", "Disable Add Force", "Disable Projection"]


@dataclass
class MeshNode:
    name: str
    path: str
    position: Tuple[float, float, float]
    rotation_xyzw: Tuple[float, float, float, float]
    scale: Tuple[float, float, float]
    boundary: bool
    fluid_spawn: bool


@dataclass
class ForceFieldSpec:
    name: str
    position: Tuple[float, float, float]
    rotation_xyzw: Tuple[float, float, float, float]
    scale: Tuple[float, float, float]
    active: bool = True
    strength: float = 5.0

    def matrix(self) -> np.ndarray:
        """The 3x3 linear part rotate * scale of the node's transform."""
        return (quat_matrix(*self.rotation_xyzw) * np.asarray(self.scale, np.float32)[None, :]).astype(np.float32)

    def direction(self) -> np.ndarray:
        """normalize(vec3(transform[0])) (ForceFieldNode.cpp:89-93)."""
        c = self.matrix()[:, 0].astype(np.float32)
        return (c / np.float32(np.sqrt(np.float32(c @ c)))).astype(np.float32)

    def triangles(self) -> np.ndarray:
        """The unit cube's 12 triangles (include/assets/Primitives.h BoxGeometry) in world space,
        (12, 9) float32."""
        from .meshes import box
        cube = np.asarray(box((-0.5, -0.5, -0.5), (0.5, 0.5, 0.5)), np.float32)
        return np.ascontiguousarray((cube @ self.matrix().T + np.asarray(self.position, np.float32)).reshape(-1, 9), np.float32)


@dataclass
class SceneSpec:
    resolution: Tuple[int, int, int]
    domain_center: Tuple[float, float, float]
    domain_size: Tuple[float, float, float]
    properties: Dict[str, object]
    box_spawns: List[Tuple[Tuple[float, float, float], Tuple[float, float, float]]] = field(default_factory=list)
    meshes: List[MeshNode] = field(default_factory=list)
    force_fields: List[ForceFieldSpec] = field(default_factory=list)
    queries: list = field(default_factory=list)
    other_nodes: List[dict] = field(default_factory=list)
    base_dir: str = "."

    def params(self, lib_params_cls=None):
        """efs_params from the reference defaults overridden by the scene's properties."""
        from . import default_params
        p = default_params() if lib_params_cls is None else lib_params_cls()
        if lib_params_cls is not None:
            _abi.load().efs_default_params(ctypes.byref(p))
        props = self.properties
        for k, f in _FLOAT_PROPS.items():
            if k in props:
                setattr(p, f, float(props[k]))
        for k, f in _INT_PROPS.items():
            if k in props:
                setattr(p, f, int(props[k]))
        if "Gravity" in props and "Cell Size" in props and float(props["Cell Size"]) > 0:
            p.gravity_force = float(np.float32(props["Gravity"]) / np.float32(props["Cell Size"]))
        elif "Gravity Force" in props:
            p.gravity_force = float(props["Gravity Force"])
        mask = 0
        for bit, k in enumerate(_DISABLE):
            if props.get(k, False):
                mask |= 1 << bit
        p.disable_mask = mask
        return p


def _vec(node, key, n, default):
    v = node.get(key, default)
    return tuple(float(x) for x in v[:n])


def parse(path_or_dict, base_dir=None) -> SceneSpec:
    """Parses a scene file (or its already-loaded JSON)."""
    if isinstance(path_or_dict, dict):
        d = path_or_dict
        base = base_dir or "."
    else:
        with open(path_or_dict) as f:
            d = json.load(f)
        base = base_dir or os.path.dirname(os.path.abspath(path_or_dict))
    res, center, size = None, (0.0, 0.0, 0.0), (2.0, 2.0, 2.0)
    spec = SceneSpec(resolution=(0, 0, 0), domain_center=center, domain_size=size, properties=dict(d.get("properties", {})),
                     base_dir=base)
    for n in d.get("nodes", []):
        t = n.get("type")
        if t == "SimulationDomainNode":
            res = tuple(int(x) for x in n.get("resolution", [10, 10, 10])[:3])
            spec.domain_center = _vec(n, "position", 3, [0, 0, 0])
            spec.domain_size = _vec(n, "scale", 3, [2, 2, 2])
        elif t == "BoundingBoxFluidSpawnNode":
            spec.box_spawns.append((_vec(n, "position", 3, [0, 0, 0]), _vec(n, "scale", 3, [1, 1, 1])))
        elif t in ("ModelNode", "ModelFluidSpawnNode"):
            spec.meshes.append(MeshNode(name=n.get("name", ""), path=n.get("path", ""),
                                        position=_vec(n, "position", 3, [0, 0, 0]),
                                        rotation_xyzw=_vec(n, "rotation", 4, [0, 0, 0, 0]),
                                        scale=_vec(n, "scale", 3, [1, 1, 1]),
                                        boundary=bool(n.get("boundary", False)) and t == "ModelNode",
                                        fluid_spawn=(t == "ModelFluidSpawnNode")))
        elif t == "ForceFieldNode":
            spec.force_fields.append(ForceFieldSpec(name=n.get("name", ""), position=_vec(n, "position", 3, [0, 0, 0]),
                                                    rotation_xyzw=_vec(n, "rotation", 4, [0, 0, 0, 1]),
                                                    scale=_vec(n, "scale", 3, [0.25, 0.25, 0.25]),
                                                    active=bool(n.get("active", True)), strength=float(n.get("strength", 5.0))))
        elif t == "QueryNode":
            from .query import QueryNode
            spec.queries.append(QueryNode.from_json(n))
        else:
            spec.other_nodes.append(n)
    if res is None:
        raise ValueError("scene has no SimulationDomainNode")
    spec.resolution = res
    return spec


def load_obj(path) -> np.ndarray:
    """Triangles (n, 3, 3) float32 of a Wavefront OBJ file: "v" records, "f" records with
    1-based (or negative, relative) vertex indices, polygons fanned from their first vertex."""
    verts, tris = [], []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            if parts[0] == "v":
                verts.append([float(parts[1]), float(parts[2]), float(parts[3])])
            elif parts[0] == "f":
                idx = []
                for tok in parts[1:]:
                    k = int(tok.split("/")[0])
                    idx.append(k - 1 if k > 0 else len(verts) + k)
                for j in range(1, len(idx) - 1):
                    tris.append([idx[0], idx[j], idx[j + 1]])
    v = np.asarray(verts, np.float32)
    return v[np.asarray(tris, np.int64)] if tris else np.zeros((0, 3, 3), np.float32)


def quat_matrix(x, y, z, w) -> np.ndarray:
    """glm::mat4_cast of the quaternion (w, x, y, z) (no normalisation: the zero quaternion the
    scene files store for "no rotation" gives the identity)."""
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]], np.float32)


def world_triangles(tris: np.ndarray, node: MeshNode) -> np.ndarray:
    """translate * rotate * scale applied to every vertex (Transform::updateTransform)."""
    R = quat_matrix(*node.rotation_xyzw)
    S = np.asarray(node.scale, np.float32)
    T = np.asarray(node.position, np.float32)
    return ((tris * S) @ R.T + T).astype(np.float32)


def resolve(spec: SceneSpec, path: str) -> str:
    """A mesh path of the scene: absolute, or relative to the scene's directory or one of its
    parents (the reference stores paths relative to the application directory, e.g.
    "assets/models/Bucket.obj" beside "assets/scenes/")."""
    if os.path.isabs(path):
        return path
    d = os.path.abspath(spec.base_dir)
    while True:
        cand = os.path.normpath(os.path.join(d, path))
        if os.path.exists(cand):
            return cand
        parent = os.path.dirname(d)
        if parent == d:
            return os.path.normpath(os.path.join(spec.base_dir, path))
        d = parent


def mesh_triangles(spec: SceneSpec, loader=load_obj):
    """(boundary triangles, fluid-spawn triangles), world space, (n, 9) float32 each."""
    bnd, flu = [], []
    for m in spec.meshes:
        if not (m.boundary or m.fluid_spawn):
            continue
        tri = world_triangles(loader(resolve(spec, m.path)), m).reshape(-1, 9)
        (bnd if m.boundary else flu).append(tri)
    cat = lambda L: np.ascontiguousarray(np.concatenate(L) if L else np.zeros((0, 9), np.float32), np.float32)
    return cat(bnd), cat(flu)


def _voxelise(sim, tris, spec, f, vin, vout):
    c = (ctypes.c_float * 3)(*spec.domain_center)
    s = (ctypes.c_float * 3)(*spec.domain_size)
    check(sim.lib.efs_scene_voxelise(sim.h, tris.ctypes.data if len(tris) else None, len(tris), c, s, f, vin, vout))


def apply(sim, spec: SceneSpec, loader=load_obj):
    """Sets rho and the boundary of a HipSimulator3D (whose dims must be the scene's resolution)
    the way Editor::init does on a reset."""
    W, H, D = spec.resolution
    if tuple(sim.dims) != (W, H, D):
        raise ValueError(f"simulator dims {sim.dims} != scene resolution {spec.resolution}")
    bnd_tris, fluid_tris = mesh_triangles(spec, loader)
    g = sim.grid()
    # fluid (Editor::initFluid): meshes first, then the boxes
    if len(fluid_tris):
        _voxelise(sim, fluid_tris, spec, _abi.FIELD_RHO, 1.0, 0.0)
        rho = g.rho()
    else:
        rho = np.zeros((D, H, W), np.float32)
    for (px, py, pz), (sx, sy, sz) in spec.box_spawns:
        x0, x1 = max(0, int(px)), min(W, int(px + sx))
        y0, y1 = max(0, int(py)), min(H, int(py + sy))
        z0, z1 = max(0, int(pz)), min(D, int(pz + sz))
        if x1 > x0 and y1 > y0 and z1 > z0:
            rho[z0:z1, y0:y1, x0:x1] = 1.0
    # solids (Editor::initBoundaries): meshes, the domain border, no fluid under a solid
    if len(bnd_tris):
        _voxelise(sim, bnd_tris, spec, _abi.FIELD_BOUNDARY, 1.0, 0.0)
        bnd = g.boundary()
    else:
        bnd = np.zeros((D, H, W), np.float32)
    bnd[0, :, :] = bnd[-1, :, :] = 1.0
    bnd[:, 0, :] = bnd[:, -1, :] = 1.0
    bnd[:, :, 0] = bnd[:, :, -1] = 1.0
    rho[bnd == 1.0] = 0.0
    g.set("boundary", bnd)
    g.set("rho", rho)
    return rho, bnd


def apply_force_fields(sim, spec: SceneSpec, max_cells=None):
    """Sets the forces field of a 3D simulator from the scene's force fields (cleared first).
    Returns the cell count of each field; a field over max_cells (default: the scene's
    "Max ForceField cells", else 1000) is skipped, as the reference skips it."""
    if max_cells is None:
        max_cells = int(spec.properties.get("Max ForceField cells", 1000))
    W, H, D = spec.resolution
    sim.grid().set("forces", np.zeros((D, H, W, 3), np.float32))
    c = (ctypes.c_float * 3)(*spec.domain_center)
    s = (ctypes.c_float * 3)(*spec.domain_size)
    counts = []
    for ff in spec.force_fields:
        tris = ff.triangles()
        f = ff.direction() * np.float32(ff.strength) if ff.active else np.zeros(3, np.float32)
        n = ctypes.c_int32(0)
        if ff.active:
            check(sim.lib.efs_scene_force_region(sim.h, tris.ctypes.data, len(tris), c, s, (ctypes.c_float * 3)(*f.tolist()),
                                                 int(max_cells), ctypes.byref(n)))
        counts.append(int(n.value))
    return counts


def simulator_from_scene(path, loader=load_obj, **param_over):
    """A HipSimulator3D set up from a scene file, with the scene's parameters (overridable)."""
    from . import HipSimulator3D, Params
    spec = parse(path)
    p = spec.params(Params)
    for k, v in param_over.items():
        setattr(p, k, v)
    sim = HipSimulator3D(*spec.resolution, params=p)
    apply(sim, spec, loader)
    if spec.force_fields:
        apply_force_fields(sim, spec)
    return sim, spec
