"""Benchmark: 3D liquid steps/s at 256^3 (258^3 with the border) with the pressure CG run to
relative residual ||r|| / ||b|| <= 1e-5 (FP64 dots / alpha / beta, FP32 stencils), verified by
the FP64 true residual of every solve.  The default solver is CG preconditioned by one geometric
multigrid V-cycle per iteration, started from the previous step's pressure.

Workload: the mixed-precision sphere-over-pool scene (fluid pool of depth 0.3, a fluid sphere
r = 0.15 at height 0.7, 8 seeded solid boxes), synthetic seeded inputs built on the host and
uploaded once; the timed region starts with every field resident in HBM.  A step is one full
ISimulator::update (decay, extrapolation, sharpening, mass-conserving advection, velocity
advection, forces, projection).

Plain CG is measured in runs of its own: each starts from the scene, takes the same warm-up steps
with the default solver (deterministic, so every run reaches the same state) and then times the
same number of steps with plain mixed-precision CG, once from the previous pressure (the default's
guess) and once from x0 = b (the reference CPU backend's solve).

Prints one JSON line with steps/s, CG iterations/step, the roofline of the dominant event-timed
kernel (HIP events recorded on the launches themselves, on the library's stream, in every
--kernel-timing-every-th timed step: a timed launch is followed by a gap of several microseconds,
which on every launch would slow the measured step rate by about 3 %), the rooflines of every
event-timed kernel, and the CPU baseline (the oracle, one thread, on a bounded sample; see
`cpu_baseline.sample`).
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
# Algorithmic HBM bytes per interior cell of each event-timed kernel (FP32 fields, u8 boundary
# class, u16 row code; neighbour and gather re-reads are assumed to hit the caches).
K1_BYTES = 2 + 4 + 4 + 4      # CG direction + SpMV: row code, r, d_{i-1} read; d_i written
K2_BYTES = 2 + 4 + 4 + 4 + 4 + 4  # CG update: row code, d_i, x_i, r_i read; x_{i+1}, r_{i+1} written
K2_BYTES_PC = 2 + 4 + 8 + 4 + 8 + 4  # preconditioned CG: the FP64 iterate is read and written
# the fine level of the multigrid V(2,2)-cycle, event-timed as two brackets:
# up: first sweep (row code 2 + b 4 + stored red 4 + parents' correction 4/8 read, x written 4)
# and second sweep (row code 2 + b 4 + stored red 4 read, x written 4)
MG_UP_BYTES = (2 + 4 + 4 + 4 / 8 + 4) + (2 + 4 + 4 + 4)
# down: both sweeps in one kernel (row code 2 + b 4 read, x written 4), residual restriction (row
# code 2 + b 4 + x 4 read, coarse b written 4/8)
MG_DOWN_BYTES = (2 + 4 + 4) + (2 + 4 + 4 + 4 / 8)
# rho_advect backward trace: u, v, w 12 + boundary class 1 + gamma 4 read; velocity sample 12,
# advected gamma 4, corner base 4, 8 row entries 32 written; column-sum accumulator 8 read + 8
# written (fixed-point atomics, summed per block in LDS first)
ADV_BWD_BYTES = (12 + 1 + 4) + (12 + 4 + 4 + 32) + 16
# project's walls + divergence: u, v, w 12 + rho 4 + boundary class 1 read, divergence 4 written
WALLS_DIV_BYTES = 12 + 4 + 1 + 4
# project's gradient + walls: FP64 pressure 8 + boundary class 1 read, u, v, w read and written 24
GRADIENT_BYTES = 8 + 1 + 24
JACOBI_BYTES = 2 + 4 + 4 + 4  # row code, b, p_in read; p_out written

TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_sphere_pool_256_mg.md")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary (calibrated against a copy of known size), for the default workload only."""
    if not (ARGS.n == 256 and ARGS.config == "sphere_pool" and ARGS.precond == "mg" and ARGS.solver == "cg"):
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
    ap.add_argument("--solver", default="cg", choices=["cg", "jacobi"],
                    help="SELF_CG (with --precond) or JACOBI (max_steps / 2 passes, the reference's Jacobi solve)")
    ap.add_argument("--plain-steps", type=int, default=None,
                    help="timed steps of each plain-CG comparison run (default: --steps; 0 skips them)")
    ap.add_argument("--learned-weights", default=None, metavar="NPZ",
                    help="with --precond learned: the trained network (eulerianfluidsimulation_amd.learned --out)")
    ap.add_argument("--learned-every", type=int, default=1, help="with --precond learned: apply the network every k-th "
                    "flexible CG iteration")
    ap.add_argument("--fp64-check", action="store_true",
                    help="after the timed steps, one more step whose projection is compared with the FP64 oracle "
                         "(divergence, and pressure against an FP64 CG solve to relative residual 1e-11)")
    ap.add_argument("--unet", type=int, default=0, metavar="N",
                    help="instead of the step: time the learned preconditioner's U-Net (widths 16-64) at N^3 "
                         "(inference forward and one training step) against the BF16 MFMA peak")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=80, help="interior resolution of the CPU sample's stages and solve")
    ap.add_argument("--kernel-timing-every", type=int, default=4, metavar="N",
                    help="record the kernel events in every N-th timed step (1: every step)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="before the run, one device-to-device copy of MIB MiB (a counter calibration kernel of "
                         "known traffic for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)")
    return ap.parse_args()


def host_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def cpu_baseline(scene_fn, n_gpu, state):
    """The reference CPU path's cost per step at the benchmark size, from the FP32 oracle on one
    core: the six stages and the divergence of one step of the same scene at --cpu-n (scaled by
    the cell ratio), the reference's own CG from x0 = b with its own criterion (r.r < 1e-5) and
    cap (MAX_STEPS) at --cpu-n, and the cost of one CG iteration measured at the benchmark size
    on the GPU run's final state.  If the sample's CG reaches MAX_STEPS, the reference runs
    MAX_STEPS iterations per step at the benchmark size too (its residual floor only grows with
    the grid) and then pauses (CPUSimulator3D.cpp:541-547)."""
    import numpy as np
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
    its_sample = max(1, st.iterations)
    max_steps = sc.max_steps
    capped = st.err >= sc.solver_threshold  # the CG stopped at MAX_STEPS without meeting its criterion
    del o
    # one CG iteration at the benchmark size: the reference's CG from x0 = b, 8 iterations, on
    # the GPU run's state (its divergence)
    big = scene_fn(n_gpu)
    ob = Oracle3D(*big.dims, params=params_from_scene(big, max_steps=8))
    for k in ("rho", "boundary", "divergence"):
        ob.field(k)[...] = state[k]
    t0 = time.perf_counter()
    stb = ob.solve()
    t_iter_big = (time.perf_counter() - t0) / max(1, stb.iterations)
    del ob
    scale = ((n_gpu + 2) ** 3) / ((n_cpu + 2) ** 3)
    if capped:
        its_big = max_steps
        how = (f"the sample's CG reached MAX_STEPS={max_steps} without r.r < {sc.solver_threshold:g} (r.r = {st.err:.2e}), "
               f"so the reference runs {max_steps} iterations per step at {n_gpu}^3 and pauses")
    else:
        its_big = min(max_steps, int(round(its_sample * (n_gpu + 2) / (n_cpu + 2))))
        how = (f"the sample's CG converged in {its_sample} iterations; CG iterations grow with the grid size, so "
               f"{its_big} at {n_gpu}^3 (capped at MAX_STEPS={max_steps})")
    t_step = scale * t_nosolve + its_big * t_iter_big
    return {
        "value": 1.0 / t_step,
        "unit": "steps/s",
        "cores": 1,
        "kind": "port",
        "reference_pauses": bool(capped),
        "cg_iters_per_step": its_big,
        "host": host_info(),
        "sample": (f"FP32 oracle, one core. Stages + divergence of one step at {n_cpu}^3 interior: {t_nosolve:.2f} s, "
                   f"scaled by the cell ratio {scale:.2f}. Reference CG at {n_cpu}^3 from x0 = b: {st.iterations} "
                   f"iterations in {t_solve:.2f} s; {how}. One CG iteration at {n_gpu}^3 measured on the GPU run's "
                   f"state: {1e3 * t_iter_big:.0f} ms. Step = {scale * t_nosolve:.1f} s + {its_big} x "
                   f"{t_iter_big:.3f} s = {t_step:.1f} s"),
    }


# the library's event-timed kernels, in efs_kernel_time's order
KERNELS = ("cg_dir_spmv", "cg_update", "jacobi", "mg_up_fine", "mg_down_fine", "adv_bwd", "adv_fwd", "walls_div", "gradient")


def kernel_bytes(name):
    return {"cg_dir_spmv": K1_BYTES, "cg_update": K2_BYTES_PC if ARGS.precond != "none" else K2_BYTES,
            "jacobi": JACOBI_BYTES, "mg_up_fine": MG_UP_BYTES, "mg_down_fine": MG_DOWN_BYTES, "adv_bwd": ADV_BWD_BYTES,
            "adv_fwd": None, "walls_div": WALLS_DIV_BYTES, "gradient": GRADIENT_BYTES}[name]


def run(sc, params, warmup, steps, timed_params=None, kernel_timing=False, keep_state=False, tag=""):
    """One simulator from the scene: `warmup` steps with `params`, then `steps` timed steps with
    `timed_params` (default: the same)."""
    import numpy as np
    from eulerianfluidsimulation_amd import HipSimulator3D
    sim = HipSimulator3D(*sc.dims, params=params)
    if _NET is not None and params.precond == _abi_precond_learned():
        from eulerianfluidsimulation_amd import learned
        learned.attach(sim, _NET[0], learned.default_opts(_NET[1], every=ARGS.learned_every))
    t0 = time.perf_counter()
    sim.load_scene(sc)
    print(f"[bench{tag}] scene {sc.name} dims {sc.dims} built in {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    for i in range(warmup):
        rc = sim.update()
        print(f"[bench{tag}] warmup {i}: rc={rc} its={sim.last_stats.iterations}", file=sys.stderr, flush=True)
    if timed_params is not None:
        sim.set_params(timed_params)
    lib = sim.lib
    if kernel_timing:
        lib.efs_set_kernel_timing(sim.h, max(1, ARGS.kernel_timing_every))
    sim.synchronize()
    iters, true_res, rcs, fallbacks = [], [], [], []
    t0 = time.perf_counter()
    for i in range(steps):
        rcs.append(sim.update())
        iters.append(sim.last_stats.iterations)
        true_res.append(sim.last_stats.true_rel_residual)
        fallbacks.append(sim.last_stats.fallback)
        if steps >= 100 and i % 50 == 49:
            print(f"[bench{tag}] step {i + 1}: its {iters[-1]} ({time.perf_counter() - t0:.1f}s)", file=sys.stderr, flush=True)
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    print(f"[bench{tag}] {steps} steps in {elapsed:.3f}s, its/step {np.mean(iters):.2f}", file=sys.stderr, flush=True)
    kt = {}
    if kernel_timing:
        for k, name in enumerate(KERNELS):
            n = ctypes.c_int64()
            ms = ctypes.c_double()
            lib.efs_kernel_time(sim.h, k, ctypes.byref(n), ctypes.byref(ms))
            if n.value:
                kt[name] = (n.value, ms.value)
        lib.efs_set_kernel_timing(sim.h, 0)
    state = None
    if keep_state:
        g = sim.grid()
        state = {k: getattr(g, k)() for k in ("rho", "boundary", "divergence")}
    fp64 = fp64_check(sc, sim) if (ARGS.fp64_check and not tag) else None
    sim.close()
    return {"elapsed": elapsed, "iters": iters, "true_res": true_res, "rcs": rcs, "kt": kt, "state": state,
            "fallbacks": fallbacks, "fp64": fp64}


def fp64_check(sc, sim):
    """One more step: the six stages, then the projection with the run's solver; its divergence and
    pressure against the FP64 oracle on the same pre-projection state."""
    import numpy as np
    from oracle.oracle import Oracle3D, params_from_scene
    for s in range(6):
        sim._stage(s)
    g = sim.grid()
    state = {k: getattr(g, k)() for k in ("vel", "rho", "boundary", "gamma", "beta", "forces")}
    rc = sim.project()
    div_g, p_g = g.divergence().astype(np.float64), g.pressure().astype(np.float64)
    o = Oracle3D(*sc.dims, params=params_from_scene(sc), dbl=True)
    for k, v in state.items():
        o.field(k)[...] = v
    o.divergence()
    div64 = o.field("divergence")
    e_div = float(np.linalg.norm(div_g - div64) / max(1e-300, np.linalg.norm(div64)))
    t0 = time.perf_counter()
    it64, rel64 = o.solve64(x0=p_g, rel_tol=1e-11)
    p64 = o.field("pressure")
    e_p = float(np.linalg.norm(p_g - p64) / max(1e-300, np.linalg.norm(p64)))
    print(f"[bench] FP64 check: |div-div64|/|div64| {e_div:.2e} |p-p64|/|p64| {e_p:.2e} ({it64} FP64 iterations, "
          f"{time.perf_counter() - t0:.0f}s)", file=sys.stderr, flush=True)
    return {"project_rc": rc, "div_rel_err": e_div, "p_rel_err": e_p, "fp64_cg_iters": it64, "fp64_rel_residual": rel64}


_NET = None  # (UNet, out_scale) with --precond learned


def _abi_precond_learned():
    from eulerianfluidsimulation_amd import _abi
    return _abi.PRECOND_LEARNED


def roofline(name, launches, total_ms):
    per_cell = kernel_bytes(name)
    if per_cell is None:
        return None
    avg_ms = total_ms / max(1, launches)
    b = per_cell * ARGS.n ** 3
    achieved = b / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(name), "kernel": name, "launches": launches, "avg_us": 1e3 * avg_ms,
            "algorithmic_bytes_per_cell": per_cell, "algorithmic_bytes_per_launch": b}


MFMA_BF16_PEAK_TFLOPS = 2500.0


def unet_flops(n, widths=(16, 32, 48, 64), inference=True):
    """FLOPs of one U-Net forward at n^3: 2 * 27 * cin * cout per voxel for each 3x3x3 layer; the
    output layer counts its one used channel at inference (the library computes only that one)
    and all 16 in training."""
    c = widths
    layers = [(16, c[0], 0), (c[0], c[0], 0), (c[0], c[1], 1), (c[1], c[1], 1), (c[1], c[2], 2), (c[2], c[2], 2),
              (c[2], c[3], 3), (c[3], c[3], 3), (c[3] + c[2], c[2], 2), (c[2] + c[1], c[1], 1), (c[1] + c[0], c[0], 0),
              (c[0], 1 if inference else 16, 0)]
    return sum(2 * 27 * ci * co * (n ** 3) / 8 ** lvl for ci, co, lvl in layers)


def unet_bench():
    import torch
    from eulerianfluidsimulation_amd import learned
    n = ARGS.unet
    net = learned.UNet(seed=1)
    x = torch.randn((n, n, n, 16), dtype=torch.float32, device="cuda")
    t = torch.randn((n, n, n), dtype=torch.float32, device="cuda")
    for _ in range(max(1, ARGS.warmup)):
        net.forward(x)
        net.loss_grad(x, t)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(ARGS.steps):
        net.forward(x)
    torch.cuda.synchronize()
    fwd_ms = 1e3 * (time.perf_counter() - t0) / ARGS.steps
    t0 = time.perf_counter()
    for _ in range(ARGS.steps):
        net.loss_grad(x, t)
        net.adam_step(1e-4)
    torch.cuda.synchronize()
    train_ms = 1e3 * (time.perf_counter() - t0) / ARGS.steps
    f_inf = unet_flops(n)
    f_train = 3 * unet_flops(n, inference=False)  # forward, backward data, backward weights
    tf = f_inf / (fwd_ms * 1e-3) / 1e12
    out = {"metric": "U-Net inference and training step at N^3 (learned preconditioner network)", "n": n,
           "steps": ARGS.steps, "forward_ms": fwd_ms, "train_step_ms": train_ms,
           "forward_gflop": f_inf / 1e9, "train_gflop": f_train / 1e9,
           "roofline": {"bound": "mfma", "achieved": tf, "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": tf / MFMA_BF16_PEAK_TFLOPS, "traffic": None, "kernel": "U-Net forward (all layers)"},
           "train_tflops": f_train / (train_ms * 1e-3) / 1e12, "dtype": "bf16 operands, fp32 accumulation"}
    print(json.dumps(out))


def main():
    global ARGS
    ARGS = parse()
    if ARGS.unet:
        unet_bench()
        return
    if ARGS.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {ARGS.gpus} is not supported (the step would shard by "
              "z-slabs, one per GPU, with halo-plane exchanges and all-reduced CG dot products)", file=sys.stderr)
        sys.exit(2)
    import copy

    import numpy as np
    from eulerianfluidsimulation_amd import Params, _abi, scenes

    scene_fn = {"sphere_pool": scenes.sphere_pool, "dam_spheres": scenes.dam_spheres, "dam3d": scenes.dam3d}[ARGS.config]
    sc = scene_fn(ARGS.n)
    p = sc.params(Params)
    p.criterion = _abi.CRIT_RELATIVE if ARGS.criterion == "relative" else _abi.CRIT_REFERENCE
    p.precond = {"none": _abi.PRECOND_NONE, "mg": _abi.PRECOND_MG, "learned": _abi.PRECOND_LEARNED}[ARGS.precond]
    p.solver_type = _abi.SOLVER_JACOBI if ARGS.solver == "jacobi" else _abi.SOLVER_SELF_CG
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
    if ARGS.precond == "learned":
        global _NET
        if not ARGS.learned_weights:
            print("--precond learned needs --learned-weights", file=sys.stderr)
            sys.exit(2)
        from eulerianfluidsimulation_amd import learned
        net = learned.UNet(seed=0)
        out_scale = net.load(ARGS.learned_weights)
        _NET = (net, out_scale)
    main_run = run(sc, p, ARGS.warmup, ARGS.steps, kernel_timing=True, keep_state=not ARGS.no_cpu_baseline)
    steps_per_s = ARGS.steps / main_run["elapsed"]
    iters, true_res, kt = main_run["iters"], main_run["true_res"], main_run["kt"]
    its = float(np.mean(iters)) if iters else 0.0
    # plain CG from the same warmed-up state over the same steps
    plain = None
    plain_steps = ARGS.steps if ARGS.plain_steps is None else ARGS.plain_steps
    if plain_steps > 0 and ARGS.solver == "cg" and ARGS.precond != "none":
        plain = {}
        for key, guess, label in (("x0_previous", _abi.GUESS_PREVIOUS, "plain CG, x0 = previous pressure (same guess)"),
                                  ("x0_b", _abi.GUESS_DIVERGENCE, "plain CG, x0 = b (the reference CPU backend)")):
            pt = copy.copy(p)
            pt.precond = _abi.PRECOND_NONE
            pt.initial_guess = guess
            r = run(sc, p, ARGS.warmup, plain_steps, timed_params=pt, tag=f" {key}")
            sps = plain_steps / r["elapsed"]
            plain[key] = {"solver": label, "steps": plain_steps, "steps_per_s": sps,
                          "cg_iters_per_step": float(np.mean(r["iters"])), "cg_iters_each_step": r["iters"],
                          "true_rel_residual_max": float(max(r["true_res"])),
                          "speedup_of_default": (steps_per_s / sps) if plain_steps == ARGS.steps else None}
    # dominant event-timed kernel: the largest total time
    rooflines = {name: roofline(name, *v) for name, v in kt.items()}
    timed = {k: v for k, v in rooflines.items() if v is not None}
    dom = max(timed, key=lambda k: kt[k][1])
    out = {
        "metric": "3D sim steps/s at 256^3 to 1e-5 rel. residual; CG iters/step; stencil HBM GB/s",
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": ARGS.steps,
        "warmup": ARGS.warmup,
        "ms_per_step": 1e3 * main_run["elapsed"] / ARGS.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (stencils/advection) + f64 (CG dots, alpha/beta, true residual)",
        "data": "synthetic seeded scene",
        "config": {"workload": f"{sc.name}: {ARGS.n}^3 interior ({sc.dims[0]}^3 with border), "
                               + ("sphere r=0.15 over pool depth 0.3, 8 seeded solid boxes" if ARGS.config == "sphere_pool"
                                  else ARGS.config)
                               + ", dt=1/60, CG to rel. residual 1e-5 (FP64-verified)",
                   "criterion": ARGS.criterion,
                   "solver": ARGS.solver,
                   "precond": {"none": "none", "mg": "geometric multigrid V-cycle (red-black GS)",
                               "learned": "learned U-Net"}[ARGS.precond],
                   "initial_guess": "previous step pressure (CUDA backend, CudaSolver.cu:427)" if ARGS.guess == "previous"
                   else "b (CPU backend, CGSolver.cpp:429)"},
        "cg_iters_per_step": its,
        "cg_iters_each_step": iters,
        "all_steps_converged": all(rc == 0 for rc in main_run["rcs"]),
        "learned_fallback_steps": int(sum(main_run["fallbacks"])) if ARGS.precond == "learned" else None,
        "true_rel_residual_max": float(max(true_res)) if true_res else None,
        "plain_cg": plain,
        "fp64_check": main_run["fp64"],
        "kernel_events": {k: {"launches": v[0], "avg_us": 1e3 * v[1] / max(1, v[0]), "total_ms": v[1]} for k, v in kt.items()},
        "kernel_events_every": max(1, ARGS.kernel_timing_every),
        "roofline": rooflines[dom],
        "rooflines": timed,
    }
    if not ARGS.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(scene_fn, ARGS.n, main_run["state"])
        except Exception as e:  # the oracle is a reported baseline; its absence does not void the run
            out["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
