"""Scene setup on the GPU: the HIP mesh voxelisation against the FP64 winding-number oracle, and
a scene file in the reference's schema (box spawn + mesh obstacle) loaded and stepped."""
import ctypes
import json

import numpy as np
import pytest

from eulerianfluidsimulation_amd import HipSimulator3D, _abi, meshes, scene_json
from oracle import voxel_ref

pytestmark = pytest.mark.gpu


def test_voxelise_matches_winding_oracle(hiplib):
    dims, center, size = (34, 30, 26), (0.0, 0.0, 0.0), (1.0625, 0.9375, 0.8125)
    tris = np.ascontiguousarray(np.concatenate([meshes.cup((-0.4, -0.45, -0.3), (0.1, 0.1, 0.3), 0.1),
                                                meshes.box((0.2, -0.2, -0.2), (0.45, 0.05, 0.25))]).reshape(-1, 9))
    sim = HipSimulator3D(*dims)
    c = (ctypes.c_float * 3)(*center)
    s = (ctypes.c_float * 3)(*size)
    assert sim.lib.efs_scene_voxelise(sim.h, tris.ctypes.data, len(tris), c, s, _abi.FIELD_BOUNDARY, 1.0, 0.0) == 0
    got = sim.grid().boundary() == 1.0
    ref, w = voxel_ref.voxelise(dims, tris.reshape(-1, 3, 3), center, size)
    near = np.abs(np.abs(w) - np.pi) < 1e-3  # decisions within rounding of the threshold
    print(f"[voxelise] inside gpu {int(got.sum())} oracle {int(ref.sum())} mismatches {int((got != ref).sum())} near {int(near.sum())}")
    assert ref.sum() > 500
    assert np.array_equal(got[~near], ref[~near])
    sim.close()


def test_scene_file_loads_and_steps(hiplib, tmp_path):
    tris = meshes.cup((-0.4, -0.45, -0.3), (0.1, 0.1, 0.3), 0.1)
    meshes.to_obj(tris, tmp_path / "cup.obj")
    scene = {
        "nodes": [
            {"name": "Simulation Domain", "type": "SimulationDomainNode", "position": [0.0, 0.0, 0.0],
             "resolution": [34, 30, 26], "scale": [1.0625, 0.9375, 0.8125]},
            {"name": "Cup", "type": "ModelNode", "path": "cup.obj", "position": [0.0, 0.0, 0.0],
             "rotation": [0.0, 0.0, 0.0, 0.0], "scale": [1.0, 1.0, 1.0], "boundary": True},
            {"name": "Fluid Spawn", "type": "BoundingBoxFluidSpawnNode", "position": [4.0, 12.0, 5.0],
             "scale": [12.0, 14.0, 16.0]},
        ],
        "properties": {"DT": 0.01, "Gravity": 9.81, "Cell Size": 0.03125, "Lambda-": 50.0, "Lambda+": 10.0, "Eta": 100.0,
                       "D": 32.0, "Max Steps": 2000, "Threshold": 1e-5, "Disable Velocity Decay": True},
    }
    path = tmp_path / "scene.json"
    path.write_text(json.dumps(scene))
    sim, spec = scene_json.simulator_from_scene(str(path))
    g = sim.grid()
    bnd, rho = g.boundary(), g.rho()
    ref, w = voxel_ref.voxelise(spec.resolution, tris, spec.domain_center, spec.domain_size)
    near = np.abs(np.abs(w) - np.pi) < 1e-3
    ref = ref.copy()
    ref[0, :, :] = ref[-1, :, :] = ref[:, 0, :] = ref[:, -1, :] = ref[:, :, 0] = ref[:, :, -1] = True
    assert np.array_equal((bnd == 1.0)[~near], ref[~near])
    assert np.all(rho[bnd == 1.0] == 0.0)
    box = np.zeros_like(rho, bool)
    box[5:21, 12:26, 4:16] = True
    assert np.array_equal(rho == 1.0, box & (bnd != 1.0))
    m0 = float(rho.astype(np.float64).sum())
    for k in range(4):
        assert sim.update() == 0
    m1 = float(sim.grid().rho().astype(np.float64).sum())
    print(f"[scene] solids {int((bnd == 1).sum())} fluid cells {int(m0)} mass {m0:.3f} -> {m1:.3f} its {sim.last_stats.iterations}")
    assert abs(m1 - m0) <= 1e-3 * m0
    sim.close()
