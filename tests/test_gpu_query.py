"""Query boxes and force fields on the GPU.

efs_query_box (QueryNode::update's statistics) against the numpy query oracle on the fields of
a stepped dam break: count, min and max bit-exact; means within 1e-6 relative (FP64 sums on
both sides, rounded to FP32 once).  The reference's own FP32 loop-order sums are compared too,
within 1e-4 of the field scale.  efs_scene_force_region (Editor::updateForceFields) against the
FP64 winding-number oracle on rotated cubes: the forced cells, the overwrite order of
overlapping fields, the MAX_FORCEFIELD_CELLS skip, and add_force consuming the result.
"""
import ctypes
import math

import numpy as np
import pytest

from eulerianfluidsimulation_amd import HipSimulator2D, HipSimulator3D, query, scene_json, scenes
from oracle import query_ref, voxel_ref
from tests.helpers import gpu_params

pytestmark = pytest.mark.gpu

BOXES = [((0, 0, 0), (40, 40, 40)), ((3, 1, 5), (20, 12, 30)), ((-4, 2, 2), (6, 33, 7)), ((10, 10, 10), (11, 11, 11)),
         ((12, 12, 12), (12, 20, 20)), ((30, 30, 30), (50, 50, 50))]


@pytest.fixture(scope="module")
def dam(hiplib):
    sc = scenes.dam3d(32)
    sim = HipSimulator3D(*sc.dims, params=gpu_params(sc))
    sim.load_scene(sc)
    for _ in range(4):
        assert sim.update() == 0
    g = sim.grid()
    vel = g.vel()
    f = dict(rho=g.rho(), u=vel[..., 0].copy(), v=vel[..., 1].copy(), w=vel[..., 2].copy(), p=g.pressure())
    print(f"[query] dam3d dims={sc.dims} fluid cells {(f['rho'] >= 0.5).sum()}", flush=True)
    yield sim, f
    sim.close()


def _gpu_query(sim, lo, hi, only_fluid):
    from eulerianfluidsimulation_amd import _abi
    r = _abi.QueryResult()
    assert sim.lib.efs_query_box(sim.h, (ctypes.c_int32 * 3)(*lo), (ctypes.c_int32 * 3)(*hi), int(only_fluid),
                                 ctypes.byref(r)) == 0
    return r


@pytest.mark.parametrize("box", BOXES)
@pytest.mark.parametrize("only_fluid", [True, False])
def test_query_box_matches_oracle(dam, box, only_fluid):
    sim, f = dam
    lo, hi = box
    r = _gpu_query(sim, lo, hi, only_fluid)
    ref = query_ref.query_box(f["rho"], f["u"], f["v"], f["w"], f["p"], lo, hi, only_fluid)
    print(f"[query] box {box} only_fluid={only_fluid} count gpu {r.count} oracle {ref['count']}")
    assert r.count == ref["count"]
    assert list(r.max_pos) == list(ref["max_pos"]) and list(r.min_pos) == list(ref["min_pos"])
    np.testing.assert_allclose(np.array(r.avg_pos), ref["avg_pos"], rtol=1e-6, atol=0)
    scale_v = max(1.0, float(np.abs(np.stack([f["u"], f["v"], f["w"]])).max()))
    scale_p = max(1.0, float(np.abs(f["p"]).max()))
    np.testing.assert_allclose(np.array(r.vel), ref["vel"], rtol=1e-6, atol=1e-7 * scale_v)
    assert abs(r.pressure - ref["pressure"]) <= 1e-6 * abs(ref["pressure"]) + 1e-7 * scale_p
    # the reference's FP32 loop-order sums
    r32 = query_ref.query_box(f["rho"], f["u"], f["v"], f["w"], f["p"], lo, hi, only_fluid, np.float32)
    np.testing.assert_allclose(np.array(r.vel), r32["vel"], atol=1e-4 * scale_v)
    assert abs(r.pressure - r32["pressure"]) <= 1e-4 * scale_p


def test_query_node_records_frames(dam):
    sim, f = dam
    q = query.QueryNode(position=(2, 1, 2), scale=(20, 12, 28), mode=query.SINGLE_SIMULATION)
    st = query.RunState(gravity=9.81)
    q.update(sim, st)
    q.update(sim, st)
    assert len(q.history) == 2
    ref = query_ref.query_box(f["rho"], f["u"], f["v"], f["w"], f["p"], (2, 1, 2), (22, 13, 30), True)
    np.testing.assert_allclose(q.history[0].avg_position, ref["avg_pos"], rtol=1e-6)
    assert q.history[0].gravity == pytest.approx(9.81)


def test_query_box_2d(hiplib):
    sc = scenes.dam2d(40)
    sim = HipSimulator2D(sc.dims[0], sc.dims[1], params=gpu_params(sc))
    sim.load_scene(sc)
    for _ in range(3):
        assert sim.update() == 0
    g = sim.grid()
    vel = g.vel()
    rho, p = g.rho()[None], g.pressure()[None]
    u, v = vel[None, ..., 0].copy(), vel[None, ..., 1].copy()
    for lo, hi in [((0, 0, 0), (50, 50, 1)), ((4, 3, 0), (20, 30, 5))]:
        r = _gpu_query(sim, lo, hi, True)
        ref = query_ref.query_box(rho, u, v, None, p, lo, hi, True)
        assert r.count == ref["count"] > 0
        assert list(r.max_pos) == list(ref["max_pos"]) and list(r.min_pos) == list(ref["min_pos"])
        np.testing.assert_allclose(np.array(r.avg_pos), ref["avg_pos"], rtol=1e-6)
        np.testing.assert_allclose(np.array(r.vel)[:2], ref["vel"][:2], rtol=1e-6, atol=1e-6)
        assert r.vel[2] == 0.0
    sim.close()


DIMS, CENTER, SIZE = (34, 30, 26), (0.0, 0.0, 0.0), (1.0625, 0.9375, 0.8125)


def _spec(fields, max_cells=100000):
    spec = scene_json.SceneSpec(resolution=DIMS, domain_center=CENTER, domain_size=SIZE,
                                properties={"Max ForceField cells": max_cells})
    spec.force_fields = fields
    return spec


def _oracle_forces(fields):
    W, H, D = DIMS
    out = np.zeros((D, H, W, 3), np.float32)
    near = np.zeros((D, H, W), bool)
    for ff in fields:
        inside, w = voxel_ref.voxelise(DIMS, ff.triangles().reshape(-1, 3, 3), CENTER, SIZE)
        near |= np.abs(np.abs(w) - np.pi) < 1e-3
        if ff.active:
            out[inside] = ff.direction() * np.float32(ff.strength)
    return out, near


def test_force_fields_match_winding_oracle(hiplib):
    s = math.sqrt(0.5)
    c15, s15 = math.cos(math.radians(15)), math.sin(math.radians(15))
    fields = [scene_json.ForceFieldSpec("a", (-0.1, 0.0, 0.05), (0.0, 0.0, s15, c15), (0.5, 0.3, 0.4), True, 5.0),
              scene_json.ForceFieldSpec("b", (0.1, 0.05, -0.05), (s, 0.0, 0.0, s), (0.35, 0.25, 0.3), True, 2.5),
              scene_json.ForceFieldSpec("c", (0.3, -0.2, 0.2), (0.0, 0.0, 0.0, 1.0), (0.2, 0.2, 0.2), False, 9.0)]
    sim = HipSimulator3D(*DIMS)
    counts = scene_json.apply_force_fields(sim, _spec(fields))
    got = sim.grid().forces()
    ref, near = _oracle_forces(fields)
    print(f"[forces] counts {counts} forced gpu {(np.abs(got).sum(-1) > 0).sum()} oracle {(np.abs(ref).sum(-1) > 0).sum()} "
          f"near {near.sum()}")
    assert counts[0] > 500 and counts[1] > 300 and counts[2] == 0
    assert np.array_equal(got[~near], ref[~near])  # overlap: the later field wins
    sim.close()


def test_force_field_over_limit_is_skipped(hiplib):
    fields = [scene_json.ForceFieldSpec("big", (0.0, 0.0, 0.0), (0.0, 0.0, 0.0, 1.0), (0.6, 0.6, 0.6), True, 5.0),
              scene_json.ForceFieldSpec("small", (0.3, 0.3, 0.2), (0.0, 0.0, 0.0, 1.0), (0.12, 0.12, 0.12), True, 3.0)]
    sim = HipSimulator3D(*DIMS)
    counts = scene_json.apply_force_fields(sim, _spec(fields, max_cells=1000))
    got = sim.grid().forces()
    ref, near = _oracle_forces(fields[1:])
    assert counts[0] > 1000 and 0 < counts[1] <= 1000
    assert np.array_equal(got[~near], ref[~near])
    sim.close()


def test_add_force_uses_force_field(hiplib):
    sc = scenes.dam3d(32)
    sim = HipSimulator3D(*sc.dims, params=gpu_params(sc, gravity_enabled=0))
    sim.load_scene(sc)
    spec = scene_json.SceneSpec(resolution=sc.dims, domain_center=(0.5, 0.5, 0.5), domain_size=(1.0, 1.0, 1.0),
                                properties={})
    spec.force_fields = [scene_json.ForceFieldSpec("f", (0.2, 0.25, 0.5), (0.0, 0.0, 0.0, 1.0), (0.1, 0.1, 0.1), True, 4.0)]
    (n,) = scene_json.apply_force_fields(sim, spec)
    g = sim.grid()
    v0, fo, bnd = g.vel(), g.forces(), g.boundary()
    sim.add_force()
    v1 = g.vel()
    dt = np.float32(sim.params.dt)
    D, H, W = bnd.shape
    interior = np.zeros_like(bnd, bool)
    interior[1:-1, 1:-1, 1:-1] = True
    upd = interior & (bnd != 1.0)
    assert 0 < n <= 1000 and (np.abs(fo).sum(-1) > 0).sum() == n
    np.testing.assert_array_equal(v1[upd], v0[upd] + fo[upd] * dt)
    np.testing.assert_array_equal(v1[~upd], v0[~upd])
    sim.close()
