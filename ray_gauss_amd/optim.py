"""Optimiser step of the training loop on the GPU: Adam over all parameter tensors in one HIP
launch (`rg_adam_step`) and the post-step parameter projections (`rg_project_params`).

Reference: every parameter tensor has its own ``torch.optim.Adam([t], lr, eps=1e-15)``
(classes/optimizer.py:203-218, utils/optim_utils.py:4-83) driven by an ``OptimizerManager``
(classes/optimizer.py:4-121) with an optional ``ExponentialLR``; after the step the training loop
clamps and normalises the parameters (scripts/train.py:168-183).

``FusedAdam`` is a ``torch.optim.Optimizer`` with torch Adam's state layout
(``state[p] = {"step", "exp_avg", "exp_avg_sq"}``), so it drops in where the reference builds
``torch.optim.Adam`` -- learning-rate schedulers, ``OptimizerManager`` (including its
``_prune_optimizer`` / ``_cat_tensor`` edits of ``exp_avg`` / ``exp_avg_sq``) and state dicts
work unchanged -- while one ``step()`` updates every parameter of every group in a single launch
(up to 16 tensors per launch).  ``step_all`` does the same across several optimisers, e.g. the
nine per-parameter managers of the reference.

The step reads host-side scalars (learning rate, bias corrections) at call time, so it is meant
to run eagerly after the (graph-captured) render step, not inside a captured graph.
"""

from __future__ import annotations

import ctypes as C

import torch

from . import _lib

_MAX_PER_LAUNCH = 16


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class FusedAdam(torch.optim.Optimizer):
    """Adam (no weight decay, no amsgrad) with torch's update formula; eps defaults to the
    reference's 1e-15.  Parameters must be contiguous float32 CUDA tensors."""

    def __init__(self, params, lr=1e-3, betas=(0.9, 0.999), eps=1e-15):
        if lr < 0.0 or eps < 0.0 or not (0.0 <= betas[0] < 1.0 and 0.0 <= betas[1] < 1.0):
            raise ValueError("lr, eps >= 0 and betas in [0, 1) required")
        super().__init__(params, dict(lr=lr, betas=tuple(betas), eps=eps))

    def _entries(self):
        """(param, state, group) for every parameter that has a gradient, with lazily created
        state."""
        out = []
        for group in self.param_groups:
            for p in group["params"]:
                if p.grad is None:
                    continue
                if p.dtype != torch.float32 or not p.is_cuda or not p.is_contiguous():
                    raise TypeError("FusedAdam: parameters must be contiguous float32 CUDA tensors")
                if p.grad.is_sparse:
                    raise TypeError("FusedAdam: sparse gradients are not supported")
                st = self.state[p]
                if len(st) == 0:
                    st["step"] = torch.tensor(0.0)
                    st["exp_avg"] = torch.zeros_like(p, memory_format=torch.preserve_format)
                    st["exp_avg_sq"] = torch.zeros_like(p, memory_format=torch.preserve_format)
                out.append((p, st, group))
        return out

    @torch.no_grad()
    def step(self, closure=None):
        loss = None
        if closure is not None:
            with torch.enable_grad():
                loss = closure()
        _launch(self._entries())
        return loss


def step_all(optimizers):
    """One Adam step of several FusedAdam optimisers, batched into as few launches as possible."""
    entries = []
    for opt in optimizers:
        if not isinstance(opt, FusedAdam):
            raise TypeError("step_all expects FusedAdam optimisers")
        entries.extend(opt._entries())
    with torch.no_grad():
        _launch(entries)


def _launch(entries):
    lib = _lib.load()
    for i in range(0, len(entries), _MAX_PER_LAUNCH):
        chunk = entries[i:i + _MAX_PER_LAUNCH]
        table = (_lib.AdamTensor * len(chunk))()
        keep = []
        for k, (p, st, group) in enumerate(chunk):
            g = p.grad if p.grad.is_contiguous() else p.grad.contiguous()
            keep.append(g)
            st["step"] += 1
            b1, b2 = group["betas"]
            t = table[k]
            t.param, t.grad = p.data_ptr(), g.data_ptr()
            t.exp_avg, t.exp_avg_sq = st["exp_avg"].data_ptr(), st["exp_avg_sq"].data_ptr()
            t.n = p.numel()
            t.lr, t.beta1, t.beta2, t.eps = float(group["lr"]), float(b1), float(b2), float(group["eps"])
            t.step = int(st["step"].item())
        rc = lib.rg_adam_step(table, len(chunk), _stream())
        if rc != 0:
            raise _lib.RayGaussError(f"rg_adam_step failed ({rc}): {lib.rg_last_error(None).decode()}")


@torch.no_grad()
def project_params(densities=None, log_scales=None, filter_3d=None, max_scale=float("inf"), quaternions=None,
                   lobe_axis=None, normalize_quaternion=True, normalize_lobe_axis=True):
    """In-place post-step projections of scripts/train.py:168-183 on raw (pre-activation)
    parameters: densities (n) or (n,1) clamped at 0; log-scales (n,3) clamped to
    [filter_3d, max_scale] (filter_3d (n,1) or (n), optional); quaternions (n,4) and lobe axes
    (n,L,3) normalised when the flags are set (zero vectors become (1,0,0,0) / (0,0,1))."""
    tensors = [t for t in (densities, log_scales, quaternions, lobe_axis) if t is not None]
    if not tensors:
        return
    n = tensors[0].shape[0]
    for t in tensors + ([filter_3d] if filter_3d is not None else []):
        if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous() or t.shape[0] != n:
            raise TypeError("project_params: contiguous float32 CUDA tensors with a common first dimension")
    shapes_ok = ((densities is None or densities.numel() == n)
                 and (filter_3d is None or filter_3d.numel() == n)
                 and (log_scales is None or tuple(log_scales.shape) == (n, 3))
                 and (quaternions is None or tuple(quaternions.shape) == (n, 4))
                 and (lobe_axis is None or (lobe_axis.dim() == 3 and lobe_axis.shape[2] == 3)))
    if not shapes_ok:
        raise ValueError("project_params: expected densities (n) or (n,1), log_scales (n,3), quaternions (n,4), "
                         "lobe_axis (n,L,3), filter_3d (n) or (n,1)")
    pr = _lib.Projection()
    pr.n = n
    pr.num_sg = int(lobe_axis.shape[1]) if lobe_axis is not None else 0
    pr.densities = densities.data_ptr() if densities is not None else None
    pr.log_scales = log_scales.data_ptr() if log_scales is not None else None
    pr.filter_3d = filter_3d.data_ptr() if filter_3d is not None else None
    pr.max_scale = float(max_scale)
    pr.quats = quaternions.data_ptr() if quaternions is not None else None
    pr.normalize_quaternion = int(bool(normalize_quaternion))
    pr.lobe_axis = lobe_axis.data_ptr() if lobe_axis is not None else None
    pr.normalize_lobe_axis = int(bool(normalize_lobe_axis))
    lib = _lib.load()
    rc = lib.rg_project_params(C.byref(pr), _stream())
    if rc != 0:
        raise _lib.RayGaussError(f"rg_project_params failed ({rc}): {lib.rg_last_error(None).decode()}")
