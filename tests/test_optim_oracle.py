"""The optimiser oracle (oracle/optim_oracle.py) pinned on the CPU: against torch.optim.Adam --
the very implementation the reference trains with -- over several steps with an exponential
learning-rate schedule, against the closed form of the first step, and the projections against
the reference's own torch statements (scripts/train.py:168-183, classes/point_cloud.py:703-724)."""

import os
import sys

import numpy as np
import pytest
import torch

sys.path.insert(0, os.path.join(os.path.dirname(__file__), ".."))

from oracle import optim_oracle as OO  # noqa: E402


def _rand(shape, seed):
    return np.random.default_rng(seed).normal(0, 1, shape).astype(np.float32)


@pytest.mark.parametrize("shape", [(7,), (1000, 3), (333, 3, 16)])
def test_adam_matches_torch_adam(shape):
    p0 = _rand(shape, 0)
    tp = torch.from_numpy(p0.copy())
    opt = torch.optim.Adam([tp], lr=0.01, eps=1e-15)
    sched = torch.optim.lr_scheduler.ExponentialLR(opt, gamma=0.9)
    p, m, v = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    lr = 0.01
    for k in range(1, 6):
        g = _rand(shape, 10 + k)
        tp.grad = torch.from_numpy(g.copy())
        opt.step()
        sched.step()
        OO.adam_step(p, g, m, v, k, lr)
        lr *= 0.9
        st = opt.state[tp]
        np.testing.assert_allclose(st["exp_avg"].numpy(), m, rtol=2e-6, atol=1e-7)
        np.testing.assert_allclose(st["exp_avg_sq"].numpy(), v, rtol=2e-6, atol=1e-9)
        np.testing.assert_allclose(tp.detach().numpy(), p, rtol=1e-6, atol=1e-7)


def test_adam_first_step_closed_form():
    """At step 1, m^ = g and v^ = g^2, so the update is -lr * g / (|g| + eps) = -lr sign(g)."""
    g = _rand((512,), 3)
    p, m, v = np.zeros_like(g), np.zeros_like(g), np.zeros_like(g)
    OO.adam_step(p, g, m, v, 1, 0.05)
    np.testing.assert_allclose(p, -0.05 * np.sign(g), rtol=1e-5)


def test_manager_schedule_matches_optimizer_manager_rules():
    """donotuse steps skip the update; ExponentialLR only inside [delay, delay + decay)."""
    s = OO.ManagerSchedule(1.0, final_lr=0.01, decay=4, delay=2, donotuse=1)
    seen = []
    for _ in range(9):
        s.step(lambda lr, k: seen.append((lr, k)))
    gamma = 0.01 ** 0.25
    # step 0 skipped; steps 1..8 update; the lr decays after the updates at current = 2..5
    expect_lr = [1.0, 1.0, gamma, gamma ** 2, gamma ** 3, gamma ** 4, gamma ** 4, gamma ** 4]
    assert [k for _, k in seen] == list(range(1, 9))
    np.testing.assert_allclose([lr for lr, _ in seen], expect_lr, rtol=1e-12)


def test_projection_matches_reference_statements():
    n, L = 257, 4
    rng = np.random.default_rng(5)
    dens = rng.normal(0, 1, (n, 1)).astype(np.float32)
    scales = rng.normal(-4, 2, (n, 3)).astype(np.float32)
    f3d = rng.normal(-5, 1, (n, 1)).astype(np.float32)
    quats = rng.normal(0, 1, (n, 4)).astype(np.float32)
    quats[3] = 0
    axis = rng.normal(0, 1, (n, L, 3)).astype(np.float32)
    axis[7, 2] = 0
    got = OO.project(dens, scales, f3d, -1.5, quats, axis)
    # the reference's statements, in torch
    d, s, q, a = (torch.from_numpy(x.copy()) for x in (dens, scales, quats, axis))
    f = torch.from_numpy(f3d)
    d[d < 0] = 0
    q[torch.norm(q, dim=1) == 0] = torch.tensor([1.0, 0.0, 0.0, 0.0])
    q /= torch.norm(q, dim=1)[:, None]
    a[torch.norm(a, dim=2) == 0] = torch.tensor([0.0, 0.0, 1.0])
    a /= torch.norm(a, dim=2)[:, :, None]
    mask = s < f.repeat(1, 3)
    s[mask] = f.repeat(1, 3)[mask]
    s[s > -1.5] = -1.5
    np.testing.assert_array_equal(got["densities"], d.numpy())
    np.testing.assert_array_equal(got["log_scales"], s.numpy())
    np.testing.assert_allclose(got["quats"], q.numpy(), rtol=2e-7, atol=1e-7)
    np.testing.assert_allclose(got["lobe_axis"], a.numpy(), rtol=2e-7, atol=1e-7)
    assert np.array_equal(got["quats"][3], [1, 0, 0, 0]) and np.array_equal(got["lobe_axis"][7, 2], [0, 0, 1])
