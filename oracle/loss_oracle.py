"""TEST INFRASTRUCTURE ONLY.  CPU restatement of the training loss (torch, float64 by default).

Only tests/ may import this module, as a checker -- never as a product code path.

The reference's per-iteration loss (scripts/train.py:110-112) is
    (1 - 0.2) * l1_loss(x, y) + 0.2 * (1 - ssim(x, y))
with l1_loss = mean |x - y| (utils/loss_utils.py:17-18) and ssim the SSIM index of Wang et al.
(2004) averaged over channels and pixels, computed with an 11x11 Gaussian window of standard
deviation 1.5 normalised to unit sum, zero padding, C1 = 0.01^2, C2 = 0.03^2 and
sigma^2 = E[x^2] - mu^2 (utils/loss_utils.py:24-64).  x is the rendered image after the
supersampling mean of utils/general_utils.py:489-502 (pixel (i, j) = mean of the fx*fy rays
(i*fx + a, j*fy + b) of a row-major (H*fy) x (W*fx) ray grid).  The gradient comes from torch
autograd of this restatement; `ssim_bruteforce` is an independent double loop used to pin it.
"""

from __future__ import annotations

import math

import numpy as np
import torch


def window_1d(size=11, sigma=1.5, dtype=torch.float64):
    g = torch.tensor([math.exp(-((i - size // 2) ** 2) / (2.0 * sigma * sigma)) for i in range(size)], dtype=dtype)
    return g / g.sum()


def supersample_mean(rays, width, height, fx=1, fy=1):
    """(H*fy*W*fx, 3) rays -> (3, H, W) image."""
    r = rays.reshape(height, fy, width, fx, 3)
    return r.mean(dim=(1, 3)).permute(2, 0, 1)


def ssim(x, y, size=11, sigma=1.5):
    """Mean SSIM of two (3, H, W) images (separable Gaussian window, zero padding)."""
    g = window_1d(size, sigma, x.dtype)
    w2 = (g[:, None] * g[None, :])[None, None].repeat(3, 1, 1, 1)
    conv = lambda t: torch.nn.functional.conv2d(t[None], w2, padding=size // 2, groups=3)[0]  # noqa: E731
    mu1, mu2 = conv(x), conv(y)
    s11 = conv(x * x) - mu1 * mu1
    s22 = conv(y * y) - mu2 * mu2
    s12 = conv(x * y) - mu1 * mu2
    c1, c2 = 0.01 ** 2, 0.03 ** 2
    m = ((2 * mu1 * mu2 + c1) * (2 * s12 + c2)) / ((mu1 * mu1 + mu2 * mu2 + c1) * (s11 + s22 + c2))
    return m.mean()


def loss(rays, target, width, height, fx=1, fy=1, lam=0.2, dtype=torch.float64):
    """(loss, dL/drays, l1, ssim) with rays (R,3) and target (3,H,W) numpy or torch arrays."""
    r = torch.as_tensor(np.asarray(rays), dtype=dtype).clone().requires_grad_(True)
    y = torch.as_tensor(np.asarray(target), dtype=dtype)
    x = supersample_mean(r, width, height, fx, fy)
    l1 = (x - y).abs().mean()
    s = ssim(x, y)
    L = (1 - lam) * l1 + lam * (1 - s)
    L.backward()
    return float(L), r.grad.numpy(), float(l1), float(s)


def ssim_bruteforce(x, y, size=11, sigma=1.5):
    """Independent FP64 double loop over pixels and window taps (tiny images only)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    C, H, W = x.shape
    h = size // 2
    g = np.array([math.exp(-((i - h) ** 2) / (2 * sigma * sigma)) for i in range(size)])
    g /= g.sum()
    total = 0.0
    for c in range(C):
        for i in range(H):
            for j in range(W):
                m1 = m2 = e11 = e22 = e12 = 0.0
                for a in range(-h, h + 1):
                    for b in range(-h, h + 1):
                        ii, jj = i + a, j + b
                        if 0 <= ii < H and 0 <= jj < W:
                            w = g[a + h] * g[b + h]
                            u, v = x[c, ii, jj], y[c, ii, jj]
                            m1 += w * u
                            m2 += w * v
                            e11 += w * u * u
                            e22 += w * v * v
                            e12 += w * u * v
                s11, s22, s12 = e11 - m1 * m1, e22 - m2 * m2, e12 - m1 * m2
                c1, c2 = 1e-4, 9e-4
                total += ((2 * m1 * m2 + c1) * (2 * s12 + c2)) / ((m1 * m1 + m2 * m2 + c1) * (s11 + s22 + c2))
    return total / (C * H * W)
