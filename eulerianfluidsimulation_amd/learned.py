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

    def save(self, path):
        np.save(path, self.get_params())

    def load(self, path):
        self.set_params(np.load(path, allow_pickle=False))


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


def _reset_fields(sim):
    W, H, D = sim.dims
    g = sim.grid()
    z = np.zeros((D, H, W), np.float32)
    g.set("rho", z)
    g.set("vel", np.zeros((D, H, W, 3), np.float32))
    g.set("gamma", np.ones((D, H, W), np.float32))
    b = np.zeros((D, H, W), np.float32)
    b[0, :, :] = b[-1, :, :] = b[:, 0, :] = b[:, -1, :] = b[:, :, 0] = b[:, :, -1] = 1.0
    g.set("boundary", b)


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
