"""Training and use of the learned pressure preconditioner (a 4-level 3D U-Net) through the C ABI.

Data: (divergence, solid mask, pressure) triples that the plain mixed-precision CG generates on
seeded random scenes (1-4 fluid boxes, 0-10 solid boxes, face velocities N(0, 0.5)), restricted
to half resolution (`efs_learned_coarse_pair`).  Training runs the library's own BF16 MFMA
forward and backward kernels with Adam (`efs_unet_loss_grad_device`, `efs_unet_adam_step`).
torch is used only to hold device buffers.
"""
import ctypes
import time

import numpy as np

from . import _abi, scenes
from ._abi import LearnedOpts, check

DEFAULT_WIDTHS = (16, 32, 48, 64)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class UNet:
    """Owner of an efs_unet handle."""

    def __init__(self, widths=DEFAULT_WIDTHS, seed=0):
        self.lib = _abi.load()
        w = (ctypes.c_int32 * 4)(*widths)
        h = ctypes.c_void_p()
        check(self.lib.efs_unet_create(w, seed, ctypes.byref(h)))
        self.h = h
        self.widths = tuple(widths)
        self.n = self.lib.efs_unet_num_params(h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.efs_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        a = np.empty(self.n, np.float32)
        check(self.lib.efs_unet_get_params(self.h, a.ctypes.data))
        return a

    def set_params(self, a):
        a = np.ascontiguousarray(a, np.float32)
        assert a.size == self.n
        check(self.lib.efs_unet_set_params(self.h, a.ctypes.data))

    def get_grads(self):
        a = np.empty(self.n, np.float32)
        check(self.lib.efs_unet_get_grads(self.h, a.ctypes.data))
        return a

    def forward(self, x):
        """x: torch CUDA float32 (D, H, W, 16) -> (D, H, W)."""
        import torch
        D, H, W, _ = x.shape
        out = torch.empty((D, H, W), dtype=torch.float32, device=x.device)
        torch.cuda.synchronize()
        check(self.lib.efs_unet_forward_device(self.h, _ptr(x), D, H, W, _ptr(out)))
        return out

    def loss_grad(self, x, target, mask=None):
        import torch
        D, H, W, _ = x.shape
        loss = ctypes.c_double()
        torch.cuda.synchronize()
        check(self.lib.efs_unet_loss_grad_device(self.h, _ptr(x), _ptr(target), _ptr(mask) if mask is not None else None,
                                                 D, H, W, ctypes.byref(loss)))
        return loss.value

    def adam_step(self, lr):
        check(self.lib.efs_unet_adam_step(self.h, ctypes.c_float(lr)))

    def save(self, path, out_scale=None):
        """Parameters (and the output unit the network was trained for) as an .npz file."""
        np.savez(path, params=self.get_params(), widths=np.array(self.widths, np.int32),
                 out_scale=np.array(np.nan if out_scale is None else out_scale, np.float64))

    def load(self, path):
        """Loads save()'s file; returns its out_scale (None if it holds none)."""
        with np.load(path, allow_pickle=False) as f:
            self.set_params(f["params"])
            s = float(f["out_scale"]) if "out_scale" in f else float("nan")
        return None if np.isnan(s) else s


def default_opts(out_scale, warm_start=1, jacobi_weight=0.6, max_iters=0, stall_window=20, stall_ratio=0.9, every=1):
    o = LearnedOpts()
    o.every = every
    o.out_scale = out_scale
    o.jacobi_weight = jacobi_weight
    o.warm_start = warm_start
    o.max_iters = max_iters
    o.stall_window = stall_window
    o.stall_ratio = stall_ratio
    return o


def attach(sim, net, opts):
    """Use `net` as the preconditioner of `sim` (an eulerianfluidsimulation_amd.HipSimulator3D)."""
    check(sim.lib.efs_set_learned_preconditioner(sim.h, net.h, ctypes.byref(opts)))
    p = sim.params
    p.precond = _abi.PRECOND_LEARNED
    sim.set_params(p)


def out_scale_for(n_coarse, s0=0.03):
    """Unit of the network output: the pressure/divergence ratio of the Poisson solve grows
    with the square of the resolution."""
    return float(s0 * n_coarse * n_coarse)


def generate_pairs(n, count, seed0=0, out_scale=None, criterion=_abi.CRIT_RELATIVE, tol=1e-5, log=None):
    """`count` training triples at interior resolution n (a multiple of 16): each is a random
    scene (scenes.training_scene), one projection with the plain CG, restricted to n/2.
    Returns torch CUDA tensors (inputs (k, n/2, n/2, n/2, 16), targets, masks) and the
    out_scale used."""
    import torch
    from . import HipSimulator3D, Params
    nc = n // 2
    if out_scale is None:
        out_scale = out_scale_for(nc)
    xs = torch.empty((count, nc, nc, nc, 16), dtype=torch.float32, device="cuda")
    ts = torch.empty((count, nc, nc, nc), dtype=torch.float32, device="cuda")
    ms = torch.empty((count, nc, nc, nc), dtype=torch.float32, device="cuda")
    sim = None
    t0 = time.time()
    for k in range(count):
        sc = scenes.training_scene(n, seed=seed0 + k)
        p = sc.params(Params)
        p.criterion = criterion
        p.solver_threshold = tol
        p.max_steps = 4000
        if sim is None:
            sim = HipSimulator3D(*sc.dims, params=p)
        else:
            sim.set_params(p)
            _reset_fields(sim)
        sim.load_scene(sc)
        # divergence of the seeded velocities and the plain solve (the project stage)
        sim.project()
        sig = ctypes.c_double()
        torch.cuda.synchronize()
        check(sim.lib.efs_learned_coarse_pair(sim.h, ctypes.c_float(out_scale), _ptr(xs[k]), _ptr(ts[k]), _ptr(ms[k]),
                                              ctypes.byref(sig)))
        if log and (k % 10 == 0 or k == count - 1):
            log(f"pair {k}: its={sim.last_stats.iterations} sigma={sig.value:.3e} ({time.time() - t0:.1f}s)")
    if sim is not None:
        sim.close()
    return xs, ts, ms, out_scale


NET_CHANNELS = 3  # live channels of the 16-channel network input: r / sigma, fluid and solid fractions


def generate_pairs_compact(n, count, seed0=0, out_scale=None, log=None):
    """Like generate_pairs, but keeps only the live input channels (inputs (k, n/2, n/2, n/2, 3)),
    so that thousands of triples at n = 192 stay on the device (17.7 MB each).  Also returns the
    plain CG iterations of each solve."""
    import torch
    from . import HipSimulator3D, Params
    nc = n // 2
    if out_scale is None:
        out_scale = out_scale_for(nc)
    x3 = torch.empty((count, nc, nc, nc, NET_CHANNELS), dtype=torch.float32, device="cuda")
    ts = torch.empty((count, nc, nc, nc), dtype=torch.float32, device="cuda")
    ms = torch.empty((count, nc, nc, nc), dtype=torch.float32, device="cuda")
    x16 = torch.empty((nc, nc, nc, 16), dtype=torch.float32, device="cuda")
    its = []
    sim = None
    t0 = time.time()
    for k in range(count):
        sc = scenes.training_scene(n, seed=seed0 + k)
        p = sc.params(Params)
        p.criterion = _abi.CRIT_RELATIVE
        p.solver_threshold = 1e-5
        p.max_steps = 4000
        if sim is None:
            sim = HipSimulator3D(*sc.dims, params=p)
        else:
            sim.set_params(p)
            _reset_fields(sim)
        sim.load_scene(sc)
        sim.project()
        its.append(sim.last_stats.iterations)
        sig = ctypes.c_double()
        torch.cuda.synchronize()
        check(sim.lib.efs_learned_coarse_pair(sim.h, ctypes.c_float(out_scale), _ptr(x16), _ptr(ts[k]), _ptr(ms[k]),
                                              ctypes.byref(sig)))
        x3[k].copy_(x16[..., :NET_CHANNELS])
        if log and (k % 50 == 0 or k == count - 1):
            log(f"pair {k}: its={its[-1]} sigma={sig.value:.3e} ({time.time() - t0:.1f}s)")
    if sim is not None:
        sim.close()
    return x3, ts, ms, out_scale, its


def train_compact(net, x3, ts, ms, steps, lr=2e-3, lr_end=2e-4, seed=0, log=None):
    """train() on compact inputs: each step expands one sample to the 16-channel input."""
    import torch
    rng = np.random.default_rng(seed)
    losses = []
    k = x3.shape[0]
    x16 = torch.zeros(tuple(x3.shape[1:4]) + (16,), dtype=torch.float32, device=x3.device)
    t0 = time.time()
    for s in range(steps):
        i = int(rng.integers(0, k))
        x16[..., :NET_CHANNELS].copy_(x3[i])
        loss = net.loss_grad(x16, ts[i], ms[i])
        a = lr_end + 0.5 * (lr - lr_end) * (1 + np.cos(np.pi * s / max(1, steps)))
        net.adam_step(a)
        losses.append(loss)
        if log and (s % 200 == 0 or s == steps - 1):
            log(f"step {s}: loss={np.mean(losses[-200:]):.4e} lr={a:.2e} ({time.time() - t0:.1f}s)")
    return losses


def evaluate(net, out_scale, n, seeds, opts_kw=None, log=None):
    """Plain CG and the learned flexible CG on held-out training-distribution scenes at n^3:
    iterations, wall time of the solve, FP64 true residual and fallback, per scene."""
    from . import HipSimulator3D, Params
    rows = []
    # the first scene is solved once more, untimed, before the timed runs: the first launch of each
    # kernel in the process loads its code, which would otherwise be charged to that scene
    for i, seed in enumerate([seeds[0]] + list(seeds) if seeds else []):
        warmup = i == 0
        sc = scenes.training_scene(n, seed=seed)
        row = {"seed": seed}
        for mode in ("plain", "learned"):
            p = sc.params(Params)
            p.criterion = _abi.CRIT_RELATIVE
            p.true_residual = 1
            p.max_steps = 4000
            sim = HipSimulator3D(*sc.dims, params=p)
            sim.load_scene(sc)
            if mode == "learned":
                attach(sim, net, default_opts(out_scale, **(opts_kw or {})))
            sim.project_divergence()
            sim.synchronize()
            t0 = time.perf_counter()
            st = sim.solve_pressure()
            sim.synchronize()
            row[mode] = {"its": st.iterations, "ms": 1e3 * (time.perf_counter() - t0), "true": st.true_rel_residual,
                         "status": st.status, "fallback": st.fallback}
            sim.close()
        if warmup:
            continue
        rows.append(row)
        if log:
            log(f"eval seed {seed}: plain {row['plain']['its']} its {row['plain']['ms']:.1f} ms | learned "
                f"{row['learned']['its']} its {row['learned']['ms']:.1f} ms fallback={row['learned']['fallback']}")
    return rows


_RESET_CACHE = {}


def _reset_fields(sim):
    """Zero velocity, density and boundary (efs_scene_rasterise then writes the border and the
    primitives) and gamma = 1, by device-to-device copies from buffers kept per grid size."""
    import torch
    W, H, D = sim.dims
    n = W * H * D
    bufs = _RESET_CACHE.get(n)
    if bufs is None:
        bufs = (torch.zeros(3 * n, dtype=torch.float32, device="cuda"), torch.ones(n, dtype=torch.float32, device="cuda"))
        _RESET_CACHE.clear()
        _RESET_CACHE[n] = bufs
    zeros, ones = bufs
    torch.cuda.synchronize()
    for f, src, nbytes in ((_abi.FIELD_VEL, zeros, 12 * n), (_abi.FIELD_RHO, zeros, 4 * n), (_abi.FIELD_BOUNDARY, zeros, 4 * n),
                           (_abi.FIELD_GAMMA, ones, 4 * n)):
        check(sim.lib.efs_upload_device(sim.h, f, _ptr(src), nbytes))


def normalize_targets(ts, ms, out_scale):
    """Rescales the targets in place to unit RMS over the fluid cells and returns the matching
    out_scale (network output unit)."""
    import torch
    w = ms.sum()
    rms = float(torch.sqrt((ts * ts * ms).sum() / torch.clamp(w, min=1e-12)))
    if rms > 0:
        ts /= rms
    return out_scale * rms


def train(net, xs, ts, ms, steps, lr=2e-3, lr_end=2e-4, seed=0, log=None):
    """Adam on single samples drawn in a seeded order; cosine learning-rate decay."""
    rng = np.random.default_rng(seed)
    losses = []
    k = xs.shape[0]
    t0 = time.time()
    for s in range(steps):
        i = int(rng.integers(0, k))
        loss = net.loss_grad(xs[i], ts[i], ms[i])
        a = lr_end + 0.5 * (lr - lr_end) * (1 + np.cos(np.pi * s / max(1, steps)))
        net.adam_step(a)
        losses.append(loss)
        if log and (s % 50 == 0 or s == steps - 1):
            log(f"step {s}: loss={np.mean(losses[-50:]):.4e} lr={a:.2e} ({time.time() - t0:.1f}s)")
    return losses


def main(argv=None):
    """Trains the preconditioner on plain-CG triples at n^3 and evaluates it on held-out scenes:
        python -m eulerianfluidsimulation_amd.learned --n 192 --pairs 2000 --steps 4000 --out net.npz
    or, with --weights, evaluates a saved network only:
        python -m eulerianfluidsimulation_amd.learned --n 192 --weights net.npz --eval 6
    Prints progress lines and one JSON summary line."""
    import argparse
    import json
    import torch
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=192)
    ap.add_argument("--pairs", type=int, default=2000)
    ap.add_argument("--steps", type=int, default=16000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval", type=int, default=6, help="held-out scenes evaluated after training")
    ap.add_argument("--out", default=None, help="where to save the trained network (.npz)")
    ap.add_argument("--weights", default=None, help="evaluate this saved network (.npz) instead of training")
    ap.add_argument("--every", type=int, default=1, help="apply the network every k-th iteration")
    a = ap.parse_args(argv)
    log = lambda m: print(f"[learned] {m}", flush=True)
    if a.weights:
        net = UNet(seed=0)
        out_scale = net.load(a.weights)
        rows = evaluate(net, out_scale, a.n, range(100000, 100000 + a.eval), opts_kw={"every": a.every}, log=log)
        summary = {"n": a.n, "weights": a.weights, "every": a.every, "eval": rows,
                   "eval_plain_iters_mean": float(np.mean([r["plain"]["its"] for r in rows])),
                   "eval_learned_iters_mean": float(np.mean([r["learned"]["its"] for r in rows])),
                   "eval_plain_ms_mean": float(np.mean([r["plain"]["ms"] for r in rows])),
                   "eval_learned_ms_mean": float(np.mean([r["learned"]["ms"] for r in rows]))}
        print(json.dumps(summary), flush=True)
        net.close()
        return
    t0 = time.time()
    x3, ts, ms, out_scale0, its = generate_pairs_compact(a.n, a.pairs, seed0=0, log=log)
    t_gen = time.time() - t0
    out_scale = normalize_targets(ts, ms, out_scale0)
    net = UNet(seed=7)
    t0 = time.time()
    losses = train_compact(net, x3, ts, ms, a.steps, lr=a.lr, log=log)
    t_train = time.time() - t0
    del x3, ts, ms
    torch.cuda.empty_cache()
    if a.out:
        net.save(a.out, out_scale)
    rows = evaluate(net, out_scale, a.n, range(100000, 100000 + a.eval), opts_kw={"every": a.every}, log=log)
    pl = [r["plain"]["its"] for r in rows]
    le = [r["learned"]["its"] for r in rows]
    summary = {"n": a.n, "pairs": a.pairs, "train_steps": a.steps, "generate_s": t_gen, "train_s": t_train,
               "plain_cg_iters_of_pairs_mean": float(np.mean(its)), "loss_first": float(np.mean(losses[:200])),
               "loss_last": float(np.mean(losses[-200:])), "out_scale": out_scale, "eval": rows,
               "eval_plain_iters_mean": float(np.mean(pl)), "eval_learned_iters_mean": float(np.mean(le)),
               "eval_plain_ms_mean": float(np.mean([r["plain"]["ms"] for r in rows])),
               "eval_learned_ms_mean": float(np.mean([r["learned"]["ms"] for r in rows]))}
    print(json.dumps(summary), flush=True)
    net.close()


if __name__ == "__main__":
    main()
