"""Training loss on the GPU (rg_loss_l1_ssim) against the CPU oracle of scripts/train.py:108-112
(torch float64 autograd of the restated L1 + SSIM, pinned by tests/test_loss_oracle.py)."""

import os
import sys

import numpy as np
import pytest
import torch

sys.path.insert(0, os.path.join(os.path.dirname(__file__), ".."))

from oracle import loss_oracle as LO  # noqa: E402
from ray_gauss_amd.loss import l1_ssim_forward, l1_ssim_loss  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("W,H,fx,fy", [(64, 48, 1, 1), (37, 23, 2, 2), (5, 130, 1, 3), (257, 3, 2, 1)])
def test_loss_and_gradient_match_oracle(gpu_device, W, H, fx, fy):
    rng = np.random.default_rng(W * 1000 + H)
    rays = rng.uniform(0, 1, (H * fy * W * fx, 3)).astype(np.float32)
    target = rng.uniform(0, 1, (3, H, W)).astype(np.float32)
    L, g, l1, s = LO.loss(rays.astype(np.float64), target.astype(np.float64), W, H, fx, fy)
    loss, parts, grad = l1_ssim_forward(torch.from_numpy(rays).to(gpu_device), torch.from_numpy(target).to(gpu_device),
                                        W, H, (fx, fy))
    torch.cuda.synchronize()
    assert abs(float(loss) - L) <= 2e-6 * max(1.0, abs(L)), (float(loss), L)
    assert abs(float(parts[0]) - l1) <= 2e-6 and abs(float(parts[1]) - s) <= 2e-6
    gg = grad.cpu().numpy()
    scale = np.abs(g).max()
    assert np.abs(gg - g).max() <= 1e-4 * scale, (np.abs(gg - g).max(), scale)


def test_autograd_node(gpu_device):
    W, H = 40, 30
    rng = np.random.default_rng(7)
    rays = torch.from_numpy(rng.uniform(0, 1, (H * W, 3)).astype(np.float32)).to(gpu_device).requires_grad_(True)
    target = torch.from_numpy(rng.uniform(0, 1, (3, H, W)).astype(np.float32)).to(gpu_device)
    L = 3.0 * l1_ssim_loss(rays, target, W, H)
    L.backward()
    _, g, _, _ = LO.loss(rays.detach().cpu().numpy().astype(np.float64), target.cpu().numpy().astype(np.float64), W, H)
    np.testing.assert_allclose(rays.grad.cpu().numpy(), 3.0 * g, rtol=0, atol=1e-4 * np.abs(3 * g).max())


def test_identical_images_give_zero_loss(gpu_device):
    W, H = 33, 17
    img = torch.rand((3, H, W), device=gpu_device)
    rays = img.permute(1, 2, 0).reshape(-1, 3).contiguous()
    loss, parts, grad = l1_ssim_forward(rays, img, W, H)
    torch.cuda.synchronize()
    assert abs(float(parts[1]) - 1.0) < 1e-6 and float(parts[0]) == 0.0 and abs(float(loss)) < 1e-6
    assert float(grad.abs().max()) < 1e-6
