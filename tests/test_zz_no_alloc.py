"""efs_query_box_noalloc (no counterpart in the reference; for callers that cannot allocate)
against efs_query_box on the same state: every statistic identical, on two calls in a row, the
second over a smaller box, and with a box large enough that the kernel's grid is clamped to the
full scratch of block partials."""
import ctypes

import numpy as np
import pytest

from eulerianfluidsimulation_amd import HipSimulator3D, scenes
from eulerianfluidsimulation_amd._abi import QueryResult
from tests.helpers import gpu_params

pytestmark = pytest.mark.gpu


def _query(sim, fn, lo, hi, only_fluid):
    r = QueryResult()
    rc = fn(sim.h, (ctypes.c_int32 * 3)(*lo), (ctypes.c_int32 * 3)(*hi), int(only_fluid), ctypes.byref(r))
    assert rc == 0, sim.lib.efs_last_error()
    return r


def _fields(r):
    return (r.count, list(r.max_pos), list(r.min_pos), list(r.avg_pos), list(r.vel), r.pressure)


def test_query_box_noalloc_matches_query_box(hiplib):
    lib = hiplib
    fn = lib.efs_query_box_noalloc
    sc = scenes.dam3d(64)
    print(f"[no-alloc] dam3d dims={sc.dims}, 3 steps", flush=True)
    sim = HipSimulator3D(*sc.dims, params=gpu_params(sc))
    sim.load_scene(sc)
    for _ in range(3):
        assert sim.update() == 0
    # whole grid (1123 blocks' worth of cells: the grid is clamped to the 512 partials), then a
    # smaller box, then a one-cell box and an empty one
    for lo, hi, fluid in [((0, 0, 0), (70, 70, 70), True), ((0, 0, 0), (70, 70, 70), False),
                          ((3, 1, 5), (20, 12, 30), True), ((10, 10, 10), (11, 11, 11), False),
                          ((5, 5, 5), (5, 9, 9), True)]:
        ref = _query(sim, lib.efs_query_box, lo, hi, fluid)
        got = _query(sim, fn, lo, hi, fluid)
        print(f"[no-alloc] box {lo}-{hi} only_fluid={fluid}: count {got.count} / {ref.count}, "
              f"pressure {got.pressure!r} / {ref.pressure!r}", flush=True)
        assert _fields(got) == _fields(ref)
    # an argument error is reported, not a fault
    assert fn(sim.h, None, None, 0, None) == -1
    sim.close()
