"""Pins of the U-Net oracle (oracle/unet_ref.py): its convolution against PyTorch's CPU conv3d
and its analytic gradient against central finite differences in FP64."""
import numpy as np
import pytest

from oracle import unet_ref as U


def test_conv_matches_torch_conv3d():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = rng.standard_normal((6, 5, 7, 16))
    W = rng.standard_normal((27, 32, 16))
    b = rng.standard_normal(32)
    y = U.conv(x, W, b)
    wt = np.zeros((32, 16, 3, 3, 3))
    for t, dz, dy, dx in U.taps():
        wt[:, :, dz + 1, dy + 1, dx + 1] = W[t]
    yt = torch.nn.functional.conv3d(torch.from_numpy(x.transpose(3, 0, 1, 2))[None], torch.from_numpy(wt),
                                    torch.from_numpy(b), padding=1)[0].numpy().transpose(1, 2, 3, 0)
    np.testing.assert_allclose(y, yt, rtol=1e-12, atol=1e-10)


def test_layer_table_matches_library_layout():
    L, n = U.layer_table()
    assert [(l["cin"], l["cout"]) for l in L] == [(16, 16), (16, 16), (16, 32), (32, 32), (32, 48), (48, 48), (48, 64),
                                                   (64, 64), (112, 48), (80, 32), (48, 16), (16, 16)]
    assert n == sum(27 * l["cin"] * l["cout"] + l["cout"] for l in L)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    widths = (16, 16, 16, 16)
    L, n = U.layer_table(widths)
    p = rng.standard_normal(n) * 0.05
    x = rng.standard_normal((8, 8, 8, 16))
    t = rng.standard_normal((8, 8, 8))
    m = (rng.random((8, 8, 8)) > 0.3).astype(np.float64)
    loss, g = U.loss_grad(p, x, t, m, widths)
    # one weight and the bias of every layer, plus random entries
    idx = [l["woff"] + 7 for l in L] + [l["boff"] for l in L[:-1]] + list(rng.integers(0, n, 20))
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (U.loss_grad(pp, x, t, m, widths)[0] - U.loss_grad(pm, x, t, m, widths)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
