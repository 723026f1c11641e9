"""Builds the HIP backend library (libefs_hip.so) for gfx950 in-tree with hipcc.

Every source is compiled without FMA contraction (-ffp-contract=off) so the per-cell stages
match the CPU path bit for bit, and without packed-FP32 instructions.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libefs_hip.so")
SOURCES = ["efs_api.hip", "efs_step.hip", "efs_solver.hip", "efs_learned.hip", "efs_unet.hip", "efs_2d.hip", "efs_mg.hip", "efs_scene.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("EFS_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-value",
    "-Wno-unused-result",
]


HOST_DIR = os.path.join(HERE, "host")
HOST_BIN = os.path.join(HOST_DIR, "efs_run")
MIRROR_BIN = os.path.join(HOST_DIR, "efs_mirror_check")


def build_host(verbose: bool = False) -> str:
    """C++ host programs over the C++ adapter (HipSimulator3D.h), linked against the library: the
    driver efs_run and the adapter's mirror self-check efs_mirror_check."""
    for src, out in (("efs_run.cpp", HOST_BIN), ("efs_mirror_check.cpp", MIRROR_BIN)):
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", os.path.join(HOST_DIR, src), "-o", out,
               "-L", HERE, "-lefs_hip", "-Wl,-rpath,$ORIGIN/.."]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"g++ failed building {src}")
    return HOST_BIN


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libefs_hip.so")
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
