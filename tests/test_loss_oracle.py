"""The loss oracle (oracle/loss_oracle.py) pinned on the CPU: SSIM against an independent FP64
double loop, SSIM(x, x) = 1, the Gaussian window against its closed form, the supersampling mean
against explicit averaging, and the autograd gradient against central finite differences."""

import math
import os
import sys

import numpy as np
import torch

sys.path.insert(0, os.path.join(os.path.dirname(__file__), ".."))

from oracle import loss_oracle as LO  # noqa: E402


def _imgs(h, w, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0, 1, (3, h, w)), rng.uniform(0, 1, (3, h, w))


def test_window_closed_form():
    g = LO.window_1d().numpy()
    assert abs(g.sum() - 1.0) < 1e-15 and g.argmax() == 5
    np.testing.assert_allclose(g[5 + 2] / g[5], math.exp(-4 / 4.5), rtol=1e-14)


def test_ssim_against_double_loop():
    for h, w, seed in [(7, 9, 0), (12, 5, 1), (13, 13, 2)]:
        x, y = _imgs(h, w, seed)
        got = float(LO.ssim(torch.from_numpy(x), torch.from_numpy(y)))
        assert abs(got - LO.ssim_bruteforce(x, y)) < 1e-12


def test_ssim_identity_and_l1_zero():
    x, _ = _imgs(10, 14, 3)
    rays = x.transpose(1, 2, 0).reshape(-1, 3)
    L, g, l1, s = LO.loss(rays, x, 14, 10)
    assert abs(s - 1.0) < 1e-12 and l1 == 0.0 and abs(L) < 1e-12


def test_supersample_mean():
    rng = np.random.default_rng(4)
    W, H, fx, fy = 5, 3, 2, 3
    rays = rng.uniform(0, 1, (H * fy * W * fx, 3))
    img = LO.supersample_mean(torch.from_numpy(rays), W, H, fx, fy).numpy()
    grid = rays.reshape(H * fy, W * fx, 3)
    for i in range(H):
        for j in range(W):
            np.testing.assert_allclose(img[:, i, j], grid[i * fy:(i + 1) * fy, j * fx:(j + 1) * fx].mean(axis=(0, 1)),
                                       rtol=1e-14)


def test_gradient_against_finite_differences():
    W, H, fx, fy = 9, 7, 2, 1
    rng = np.random.default_rng(5)
    rays = rng.uniform(0, 1, (H * fy * W * fx, 3))
    y = rng.uniform(0, 1, (3, H, W))
    L, g, _, _ = LO.loss(rays, y, W, H, fx, fy)
    for k in rng.choice(rays.size, 12, replace=False):
        d = np.zeros(rays.size)
        d[k] = 1e-6
        d = d.reshape(rays.shape)
        fd = (LO.loss(rays + d, y, W, H, fx, fy)[0] - LO.loss(rays - d, y, W, H, fx, fy)[0]) / 2e-6
        assert abs(fd - g.reshape(-1)[k]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (k, fd, g.reshape(-1)[k])
