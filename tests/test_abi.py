"""The C-ABI library loads without a GPU and exports every entry point include/efs.h declares
(no compute calls here: there is no GPU in the build container)."""
import ctypes
import os
import re

import pytest

from eulerianfluidsimulation_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "efs.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(efs_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_abi.LIB_PATH), "build the library first (python __graft_entry__.py)"
    lib = ctypes.CDLL(_abi.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_symbols()) == set(_abi.SIGNATURES)


def test_struct_layouts_match_header():
    # efs_params: 8 floats, 3 int32, float, uint32, 5 int32 = 72 bytes; efs_stats padded to 8
    assert ctypes.sizeof(_abi.Params) == 72
    assert ctypes.sizeof(_abi.Stats) == 4 + 4 + 8 * 4 + 4 + 4 + 28 + 4
    assert ctypes.sizeof(_abi.Prim) == 56


def test_missing_library_fails_loudly(tmp_path):
    saved = _abi._lib
    try:
        _abi._lib = None
        with pytest.raises(RuntimeError, match="no CPU fallback"):
            _abi.load(str(tmp_path / "nope.so"))
    finally:
        _abi._lib = saved


def test_version_and_defaults_without_gpu():
    lib = _abi.load()
    assert b"gfx950" in lib.efs_version()
    p = _abi.Params()
    lib.efs_default_params(ctypes.byref(p))  # host-only
    assert abs(p.dt - 0.0167) < 1e-7 and p.max_steps == 2000 and p.solver_type == 2
    assert abs(p.solver_threshold - 1e-5) < 1e-12 and abs(p.lambda_minus - 0.5) < 1e-7


def test_host_adapter_builds():
    """The C++ host adapter compiles against include/efs.h and links the library."""
    from eulerianfluidsimulation_amd import build as b
    path = b.build_host()
    assert os.path.exists(path)


def test_available_solvers_without_gpu():
    """ISimulator::availableSolvers (ISimulator.cpp:5-30): SELF_CG first (the fallback), then
    JACOBI, and in 3D the reference's MULTIGRID id (Properties.h:481)."""
    lib = _abi.load()
    buf = (ctypes.c_int32 * 4)()
    assert lib.efs_available_solvers(3, buf, 4) == 3 and list(buf[:3]) == [2, 3, 4]
    assert lib.efs_available_solvers(2, buf, 4) == 2 and list(buf[:2]) == [2, 3]
    assert lib.efs_available_solvers(1, buf, 4) == -1


def test_create_rejects_grids_beyond_the_index_range_without_gpu():
    """The advection arrays hold 8 entries per cell with 32-bit indices: a grid of 2^28 cells or
    more is refused before any device call (646^3 = 269.6 M cells), and so are grids thinner than
    one interior cell."""
    lib = _abi.load()
    h = ctypes.c_void_p()
    assert lib.efs_create_3d(646, 646, 646, None, ctypes.byref(h)) == -1 and not h.value
    assert b"too large" in lib.efs_last_error()
    assert lib.efs_create_3d(2, 10, 10, None, ctypes.byref(h)) == -1
    assert b"at least 3 cells" in lib.efs_last_error()
    assert lib.efs_create_2d(16384, 16385, None, ctypes.byref(h)) == -1
