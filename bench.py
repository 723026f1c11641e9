"""Benchmark: 3D liquid steps/s at 256^3 (258^3 with the border) with the pressure CG run to
relative residual ||r|| / ||b|| <= 1e-5 (FP64 dots / alpha / beta, FP32 stencils), verified by
the FP64 true residual of every solve.  The default solver is CG preconditioned by one geometric
multigrid V-cycle per iteration, started from the previous step's pressure; the same simulation
is then continued with plain CG from x0 = b (the reference CPU backend's solve) for a few steps,
timed, and reported beside it.

Workload: the mixed-precision sphere-over-pool scene (fluid pool of depth 0.3, a fluid sphere
r = 0.15 at height 0.7, 8 seeded solid boxes), synthetic seeded inputs built on the host and
uploaded once; the timed region starts with every field resident in HBM.  A step is one full
ISimulator::update (decay, extrapolation, sharpening, mass-conserving advection, velocity
advection, forces, projection).

Prints one JSON line with steps/s, CG iterations/step, the roofline of the dominant CG kernel
(HIP events on the library's stream) and the CPU baseline (the oracle, single thread, on a
bounded sample — see `cpu_baseline.sample`).
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per interior cell of the two CG kernels (FP32 vectors, u16 row code)
K1_BYTES = 2 + 4 + 4 + 4      # row code, r, d_{i-1} read; d_i written
K2_BYTES = 2 + 4 + 4 + 4 + 4 + 4  # row code, d_i, x_i, r_i read; x_{i+1}, r_{i+1} written
K2_BYTES_PC = 2 + 4 + 8 + 4 + 8 + 4  # preconditioned CG: the FP64 iterate is read and written
# the fine level of the multigrid V(2,2)-cycle, event-timed as two brackets (bytes per cell):
# up: first sweep (row code 2 + b 4 + stored red 4 + parents' correction 4/8 read, x written 4)
# and second sweep (row code 2 + b 4 + stored red 4 read, x written 4)
MG_UP_BYTES = (2 + 4 + 4 + 4 / 8 + 4) + (2 + 4 + 4 + 4)
# down: first sweep from zero (row code 2 + b 4 read, x written 4), second sweep (row code 2 +
# b 4 + stored black 4 read, x written 4), residual restriction (row code 2 + b 4 + x 4 read,
# coarse b written 4/8)
MG_DOWN_BYTES = (2 + 4 + 4) + (2 + 4 + 4 + 4) + (2 + 4 + 4 + 4 / 8)

TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_sphere_pool_256_mg.md")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary (calibrated against a copy of known size), for the default workload only."""
    if not (ARGS.n == 256 and ARGS.config == "sphere_pool" and ARGS.precond == "mg"):
        return None
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_json:"):
                    t = json.loads(line.split(":", 1)[1])
                    key = {"cg_update": "cg_update_pcg"}.get(kernel, kernel)
                    return t.get(key)
    except OSError:
        return None
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256, help="interior resolution per axis")
    ap.add_argument("--config", default="sphere_pool", choices=["sphere_pool", "dam_spheres", "dam3d"])
    ap.add_argument("--precond", default="mg", choices=["none", "mg", "learned"])
    ap.add_argument("--criterion", default="relative", choices=["relative", "reference"])
    ap.add_argument("--guess", default="previous", choices=["previous", "divergence"],
                    help="CG initial guess: the previous step's pressure (the reference CUDA backend) or b "
                         "(the reference CPU backend)")
    ap.add_argument("--ref-steps", type=int, default=3,
                    help="extra timed steps with plain CG from x0 = b (the reference CPU backend's solve), reported beside")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=128, help="interior resolution of the CPU sample")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="before the run, one device-to-device copy of MIB MiB (a counter calibration kernel of "
                         "known traffic for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)")
    return ap.parse_args()


def cpu_baseline(scene_fn, n_gpu, gpu_iters_per_step):
    """Times the FP32 oracle (single thread) on one step of the same scene at a smaller
    resolution, split into the non-solve part and the per-iteration CG cost, and scales both
    by the cell count to the benchmark grid with the GPU's measured CG iterations per step."""
    from oracle.oracle import Oracle3D, params_from_scene
    n_cpu = ARGS.cpu_n
    sc = scene_fn(n_cpu)
    o = Oracle3D(*sc.dims, params=params_from_scene(sc))
    o.load_scene(sc)
    t0 = time.perf_counter()
    for s in range(6):
        o.stage(s)
    o.divergence()
    t_nosolve = time.perf_counter() - t0
    t0 = time.perf_counter()
    st = o.solve()
    t_solve = time.perf_counter() - t0
    iters = max(1, st.iterations)
    t_iter = t_solve / iters
    scale = ((n_gpu + 2) ** 3) / ((n_cpu + 2) ** 3)
    t_step = scale * t_nosolve + gpu_iters_per_step * scale * t_iter
    return {
        "value": 1.0 / t_step,
        "unit": "steps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"one oracle step of the same scene at {n_cpu}^3 interior: stages+divergence "
                   f"{t_nosolve:.2f}s, CG {iters} its in {t_solve:.2f}s ({1e3 * t_iter:.1f} ms/it); both scaled by the "
                   f"cell ratio {scale:.2f} to {n_gpu}^3 with the GPU's {gpu_iters_per_step:.1f} CG its/step"),
    }


def main():
    global ARGS
    ARGS = parse()
    if ARGS.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {ARGS.gpus} is not supported (the path would shard "
              "by z-slabs across GPUs, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    from eulerianfluidsimulation_amd import HipSimulator3D, Params, Stats, _abi, scenes

    scene_fn = {"sphere_pool": scenes.sphere_pool, "dam_spheres": scenes.dam_spheres, "dam3d": scenes.dam3d}[ARGS.config]
    sc = scene_fn(ARGS.n)
    p = sc.params(Params)
    p.criterion = _abi.CRIT_RELATIVE if ARGS.criterion == "relative" else _abi.CRIT_REFERENCE
    p.precond = {"none": _abi.PRECOND_NONE, "mg": _abi.PRECOND_MG, "learned": _abi.PRECOND_LEARNED}[ARGS.precond]
    p.true_residual = 1
    p.initial_guess = _abi.GUESS_PREVIOUS if ARGS.guess == "previous" else _abi.GUESS_DIVERGENCE
    if ARGS.copy_probe > 0:
        import torch
        n = ARGS.copy_probe * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] copy probe: {ARGS.copy_probe} MiB read + {ARGS.copy_probe} MiB written", file=sys.stderr, flush=True)
        del src, dst
    sim = HipSimulator3D(*sc.dims, params=p)
    t0 = time.perf_counter()
    sim.load_scene(sc)
    print(f"[bench] scene {sc.name} dims {sc.dims} built in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    lib = sim.lib
    for i in range(ARGS.warmup):
        rc = sim.update()
        print(f"[bench] warmup {i}: rc={rc} its={sim.last_stats.iterations}", file=sys.stderr, flush=True)
    lib.efs_set_kernel_timing(sim.h, 1)
    sim.synchronize()
    iters = []
    true_res = []
    t0 = time.perf_counter()
    for i in range(ARGS.steps):
        rc = sim.update()
        iters.append(sim.last_stats.iterations)
        true_res.append(sim.last_stats.true_rel_residual)
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    kt = {}
    for k, name in ((_abi.KT_CG_DIR_SPMV, "cg_dir_spmv"), (_abi.KT_CG_UPDATE, "cg_update"), (_abi.KT_OTHER, "mg_up_fine"),
                    (_abi.KT_MG_DOWN, "mg_down_fine")):
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        lib.efs_kernel_time(sim.h, k, ctypes.byref(n), ctypes.byref(ms))
        kt[name] = (n.value, ms.value)
    lib.efs_set_kernel_timing(sim.h, 0)
    # the same simulation continued with plain CG from x0 = b (the CPU backend's solve), for reference
    ref = None
    if ARGS.ref_steps > 0:
        pc = sim.params
        pc.initial_guess = _abi.GUESS_DIVERGENCE
        pc.precond = _abi.PRECOND_NONE
        sim.set_params(pc)
        sim.synchronize()
        rits, rtrue = [], []
        t1 = time.perf_counter()
        for i in range(ARGS.ref_steps):
            sim.update()
            rits.append(sim.last_stats.iterations)
            rtrue.append(sim.last_stats.true_rel_residual)
        sim.synchronize()
        rsps = ARGS.ref_steps / (time.perf_counter() - t1)
        ref = {"solver": "plain mixed-precision CG, x0 = b", "steps_per_s": rsps, "cg_iters_per_step": float(np.mean(rits)),
               "true_rel_residual_max": float(max(rtrue)), "steps": ARGS.ref_steps}
    interior = ARGS.n ** 3
    # dominant kernel = larger total time among the event-timed ones
    dom = max(kt, key=lambda k: kt[k][1])
    launches, total_ms = kt[dom]
    real = max(1, sum(iters))
    avg_ms = total_ms / max(1, launches)
    bytes_per_launch = {"cg_dir_spmv": K1_BYTES, "cg_update": K2_BYTES_PC if ARGS.precond != "none" else K2_BYTES,
                        "mg_up_fine": MG_UP_BYTES, "mg_down_fine": MG_DOWN_BYTES}[dom] * interior
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    steps_per_s = ARGS.steps / elapsed
    its = float(np.mean(iters)) if iters else 0.0
    out = {
        "metric": "3D sim steps/s at 256^3 to 1e-5 rel. residual; CG iters/step; stencil HBM GB/s",
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": ARGS.steps,
        "warmup": ARGS.warmup,
        "ms_per_step": 1e3 * elapsed / ARGS.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (stencils/advection) + f64 (CG dots, alpha/beta, true residual)",
        "data": "synthetic seeded scene",
        "config": {"workload": f"{sc.name}: {ARGS.n}^3 interior ({sc.dims[0]}^3 with border), sphere r=0.15 over pool "
                               f"depth 0.3, 8 seeded solid boxes, dt=1/60, CG to rel. residual 1e-5 (FP64-verified)",
                   "criterion": ARGS.criterion,
                   "precond": {"none": "none", "mg": "geometric multigrid V-cycle (red-black GS)",
                               "learned": "learned U-Net"}[ARGS.precond],
                   "initial_guess": "previous step pressure (CUDA backend, CudaSolver.cu:427)" if ARGS.guess == "previous"
                   else "b (CPU backend, CGSolver.cpp:429)"},
        "cg_iters_per_step": its,
        "cg_iters_each_step": iters,
        "plain_cg_reference": ref,
        "speedup_vs_plain_cg": (steps_per_s / ref["steps_per_s"]) if ref else None,
        "true_rel_residual_max": float(max(true_res)) if true_res else None,
        "kernel_events": {k: {"launches": v[0], "avg_us": 1e3 * v[1] / max(1, v[0])} for k, v in kt.items()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom), "kernel": dom,
                     "algorithmic_bytes_per_launch": bytes_per_launch, "real_iterations": real},
    }
    sim.close()
    if not ARGS.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(scene_fn, ARGS.n, its)
        except Exception as e:  # the oracle is a reported baseline; its absence does not void the run
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
