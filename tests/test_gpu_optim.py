"""Optimiser step on the GPU: the one-launch Adam (rg_adam_step via FusedAdam / step_all) against
torch.optim.Adam on the CPU -- the implementation the reference trains with -- and the post-step
projections against the CPU oracle of scripts/train.py:168-183."""

import os
import sys

import numpy as np
import pytest
import torch

sys.path.insert(0, os.path.join(os.path.dirname(__file__), ".."))

from oracle import optim_oracle as OO  # noqa: E402
from ray_gauss_amd.optim import FusedAdam, project_params, step_all  # noqa: E402

pytestmark = pytest.mark.gpu

# ulp-level: torch's CPU Adam itself mixes a fused (vectorised) and an unfused (tail) lerp
RTOL, ATOL = 2e-6, 1e-7


def _shapes():
    return [(1_000_003,), (4097, 3), (1001, 3, 16), (5,), (64, 4, 3)]


def _grads(shape, k):
    return torch.from_numpy(np.random.default_rng(100 + k).normal(0, 1, shape).astype(np.float32))


def test_fused_adam_matches_torch_adam_with_exponential_lr(gpu_device):
    ref_params = [torch.from_numpy(np.random.default_rng(i).normal(0, 1, s).astype(np.float32))
                  for i, s in enumerate(_shapes())]
    gpu_params = [p.clone().to(gpu_device) for p in ref_params]
    ref_opts = [torch.optim.Adam([p], lr=1e-3 * (i + 1), eps=1e-15) for i, p in enumerate(ref_params)]
    gpu_opt = FusedAdam([{"params": [p], "lr": 1e-3 * (i + 1)} for i, p in enumerate(gpu_params)])
    ref_sched = [torch.optim.lr_scheduler.ExponentialLR(o, gamma=0.95) for o in ref_opts]
    gpu_sched = torch.optim.lr_scheduler.ExponentialLR(gpu_opt, gamma=0.95)
    for k in range(6):
        for i, (rp, gp) in enumerate(zip(ref_params, gpu_params)):
            g = _grads(rp.shape, 10 * k + i)
            rp.grad = g.clone()
            gp.grad = g.to(gpu_device)
        for o, s in zip(ref_opts, ref_sched):
            o.step()
            s.step()
        gpu_opt.step()
        gpu_sched.step()
        torch.cuda.synchronize()
        for i, (rp, gp) in enumerate(zip(ref_params, gpu_params)):
            rs, gs = ref_opts[i].state[rp], gpu_opt.state[gp]
            assert float(gs["step"]) == float(rs["step"]) == k + 1
            np.testing.assert_allclose(gs["exp_avg"].cpu().numpy(), rs["exp_avg"].numpy(), rtol=RTOL, atol=ATOL)
            np.testing.assert_allclose(gs["exp_avg_sq"].cpu().numpy(), rs["exp_avg_sq"].numpy(), rtol=RTOL,
                                       atol=1e-9)
            np.testing.assert_allclose(gp.detach().cpu().numpy(), rp.detach().numpy(), rtol=RTOL, atol=ATOL)


def test_step_all_equals_separate_steps(gpu_device):
    shapes = [(3000, 3), (3000, 4), (3000, 3, 16), (3000,), (3000, 4), (3000, 4, 3), (3000, 3), (3000, 1),
              (3000, 3, 4), (10,), (11,), (12,), (13,), (14,), (15,), (16,), (17,), (18,)]  # > 16: two launches
    a = [torch.from_numpy(np.random.default_rng(i).normal(0, 1, s).astype(np.float32)).to(gpu_device)
         for i, s in enumerate(shapes)]
    b = [t.clone() for t in a]
    opts_a = [FusedAdam([t], lr=0.01) for t in a]
    opts_b = [FusedAdam([t], lr=0.01) for t in b]
    for k in range(3):
        for i, (x, y) in enumerate(zip(a, b)):
            g = _grads(x.shape, k * 100 + i).to(gpu_device)
            x.grad, y.grad = g, g.clone()
        step_all(opts_a)
        for o in opts_b:
            o.step()
    torch.cuda.synchronize()
    for x, y in zip(a, b):
        assert torch.equal(x, y)


def test_projection_matches_oracle(gpu_device):
    n, L = 100_003, 4
    rng = np.random.default_rng(9)
    dens = rng.normal(0, 1, (n, 1)).astype(np.float32)
    scales = rng.normal(-4, 2, (n, 3)).astype(np.float32)
    f3d = rng.normal(-5, 1, (n, 1)).astype(np.float32)
    quats = rng.normal(0, 1, (n, 4)).astype(np.float32)
    quats[11] = 0
    axis = rng.normal(0, 1, (n, L, 3)).astype(np.float32)
    axis[5, 1] = 0
    ref = OO.project(dens, scales, f3d, -1.5, quats, axis)
    t = {k: torch.from_numpy(v.copy()).to(gpu_device) for k, v in
         dict(d=dens, s=scales, f=f3d, q=quats, a=axis).items()}
    project_params(densities=t["d"], log_scales=t["s"], filter_3d=t["f"], max_scale=-1.5, quaternions=t["q"],
                   lobe_axis=t["a"])
    torch.cuda.synchronize()
    np.testing.assert_array_equal(t["d"].cpu().numpy(), ref["densities"])
    np.testing.assert_array_equal(t["s"].cpu().numpy(), ref["log_scales"])
    np.testing.assert_allclose(t["q"].cpu().numpy(), ref["quats"], rtol=2e-7, atol=1e-7)
    np.testing.assert_allclose(t["a"].cpu().numpy(), ref["lobe_axis"], rtol=2e-7, atol=1e-7)
    assert np.array_equal(t["q"][11].cpu().numpy(), [1, 0, 0, 0])
    assert np.array_equal(t["a"][5, 1].cpu().numpy(), [0, 0, 1])


def test_projection_flags_and_optional_arrays(gpu_device):
    q = torch.tensor([[0.0, 0.0, 0.0, 2.0]], device=gpu_device)
    s = torch.tensor([[-9.0, 0.0, 3.0]], device=gpu_device)
    project_params(log_scales=s, quaternions=q, normalize_quaternion=False, max_scale=1.0)
    torch.cuda.synchronize()
    assert torch.equal(q.cpu(), torch.tensor([[0.0, 0.0, 0.0, 2.0]]))  # left alone
    assert torch.equal(s.cpu(), torch.tensor([[-9.0, 0.0, 1.0]]))      # no filter_3d: only the upper clamp
