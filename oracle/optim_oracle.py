"""TEST INFRASTRUCTURE ONLY.  CPU restatement of the training loop's optimiser step.

Only tests/ may import this module, as a checker -- never as a product code path.

The reference optimises every parameter tensor with its own ``torch.optim.Adam(lr, eps=1e-15)``
(classes/optimizer.py:203-218 init_ConstantLR / init_ExponentialLR, utils/optim_utils.py:4-83),
steps them through ``OptimizerManager.step`` (classes/optimizer.py:13-20: optional donotuse
steps, ExponentialLR applied inside [delay, delay + decay)), and then projects the parameters
(scripts/train.py:168-183).  The Adam algorithm itself lives in PyTorch (installed here:
torch 2.10); this file restates its published single-tensor update (Kingma & Ba, with PyTorch's
placement of eps after the bias-corrected square root) in float32 numpy, term by term:

    step += 1
    m  = m + (1 - b1) (g - m)
    v  = v b2 + (1 - b2) g g
    p  = p - lr / (1 - b1^step) * m / (sqrt(v) / sqrt(1 - b2^step) + eps)

and the projections of scripts/train.py:168-183 with classes/point_cloud.py:703-706 (zero
quaternion -> (1,0,0,0)) and :721-724 (zero lobe axis -> (0,0,1)).
"""

from __future__ import annotations

import numpy as np


def adam_step(p, g, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-15):
    """One Adam update in float32 (arrays updated in place); `step` is the count after it."""
    f = np.float32
    w1 = f(1.0 - beta1)
    m[...] = (m + w1 * (g - m)).astype(f)
    v[...] = (v * f(beta2) + (f(1.0 - beta2) * g) * g).astype(f)
    bc1 = 1.0 - beta1 ** step
    bc2s = f((1.0 - beta2 ** step) ** 0.5)
    denom = (np.sqrt(v) / bc2s + f(eps)).astype(f)
    p[...] = (p + (f(-(lr / bc1)) * m) / denom).astype(f)


class ManagerSchedule:
    """Learning-rate and step bookkeeping of OptimizerManager (classes/optimizer.py:4-20) with
    ExponentialLR (gamma = (final/init)^(1/decay) unless given, classes/optimizer.py:207-218)."""

    def __init__(self, init_lr, gamma=None, final_lr=None, decay=None, delay=0, donotuse=0, scheduler=True):
        self.lr = float(init_lr)
        if scheduler and gamma is None:
            gamma = (final_lr / init_lr) ** (1.0 / decay)
        self.gamma = gamma if scheduler else None
        self.decay, self.delay, self.donotuse = decay, delay, donotuse
        self.current = 0
        self.adam_steps = 0

    def step(self, do_update):
        """Calls do_update(lr, step_count) when the reference would step its optimiser."""
        if self.current < self.donotuse:
            self.current += 1
            return
        self.adam_steps += 1
        do_update(self.lr, self.adam_steps)
        if self.gamma is not None and self.delay <= self.current < self.delay + self.decay:
            self.lr = self.lr * self.gamma
        self.current += 1


def project(densities=None, log_scales=None, filter_3d=None, max_scale=np.inf, quats=None, lobe_axis=None):
    """scripts/train.py:168-183 on numpy copies (returned)."""
    out = {}
    if densities is not None:
        d = densities.copy()
        d[d < 0] = 0
        out["densities"] = d
    if quats is not None:
        q = quats.astype(np.float32).copy()
        n = np.sqrt(np.sum(q.astype(np.float64) ** 2, axis=1))
        q[n == 0] = np.array([1, 0, 0, 0], np.float32)
        n = np.sqrt(np.sum(q.astype(np.float64) ** 2, axis=1)).astype(np.float32)
        out["quats"] = (q / n[:, None]).astype(np.float32)
    if lobe_axis is not None:
        a = lobe_axis.astype(np.float32).copy()
        n = np.sqrt(np.sum(a.astype(np.float64) ** 2, axis=2))
        a[n == 0] = np.array([0, 0, 1], np.float32)
        n = np.sqrt(np.sum(a.astype(np.float64) ** 2, axis=2)).astype(np.float32)
        out["lobe_axis"] = (a / n[:, :, None]).astype(np.float32)
    if log_scales is not None:
        s = log_scales.copy()
        if filter_3d is not None:
            f3 = np.repeat(filter_3d.reshape(-1, 1), 3, axis=1)
            mask = s < f3
            s[mask] = f3[mask]
        s[s > max_scale] = np.float32(max_scale)
        out["log_scales"] = s
    return out
