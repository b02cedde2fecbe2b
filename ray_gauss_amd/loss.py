"""Training loss on the GPU: supersampling mean + (1 - lambda) L1 + lambda (1 - SSIM) and its
gradient w.r.t. the ray colours in one pass of HIP kernels (`rg_loss_l1_ssim`).

Reference: scripts/train.py:108-112 -- ``reduce_supersampling`` (utils/general_utils.py:489-502),
``l1_loss`` and ``ssim`` (utils/loss_utils.py:17-64) with lambda_ssim = 0.2, differentiated by
torch autograd.  ``l1_ssim_loss`` is that expression as one autograd node: its forward returns
the scalar loss, its backward the gradient computed by the same launch.
"""

from __future__ import annotations

import ctypes as C

import torch

from . import _lib

_WS = {}


def _workspace(width, height, device):
    lib = _lib.load()
    nbytes = int(lib.rg_loss_workspace_bytes(int(width), int(height)))
    key = (device, nbytes)
    ws = _WS.get(key)
    if ws is None:
        _WS.clear()
        ws = torch.empty(nbytes, dtype=torch.uint8, device=device)
        _WS[key] = ws
    return ws


def l1_ssim_forward(ray_colors, target, width, height, supersampling=(1, 1), lambda_ssim=0.2, want_grad=True):
    """(loss (1,), parts (2,) = (L1, SSIM), dL/d ray_colors or None) for ray_colors
    ((height*ss_y)*(width*ss_x), 3) row-major and target (3, height, width)."""
    fx, fy = int(supersampling[0]), int(supersampling[1])
    R = height * fy * width * fx
    if ray_colors.shape != (R, 3) or target.shape != (3, height, width):
        raise ValueError(f"expected ray_colors ({R}, 3) and target (3, {height}, {width})")
    for t in (ray_colors, target):
        if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous():
            raise TypeError("l1_ssim_loss: contiguous float32 CUDA tensors")
    dev = ray_colors.device
    loss = torch.empty(1, dtype=torch.float32, device=dev)
    parts = torch.empty(2, dtype=torch.float32, device=dev)
    grad = torch.empty_like(ray_colors) if want_grad else None
    ws = _workspace(width, height, dev)
    lib = _lib.load()
    rc = lib.rg_loss_l1_ssim(C.c_void_p(ray_colors.data_ptr()), C.c_void_p(target.data_ptr()), int(width),
                             int(height), fx, fy, float(lambda_ssim), C.c_void_p(loss.data_ptr()),
                             C.c_void_p(parts.data_ptr()), C.c_void_p(grad.data_ptr()) if grad is not None else None,
                             C.c_void_p(ws.data_ptr()), ws.numel(),
                             C.c_void_p(torch.cuda.current_stream(dev).cuda_stream))
    if rc != 0:
        raise _lib.RayGaussError(f"rg_loss_l1_ssim failed ({rc}): {lib.rg_last_error(None).decode()}")
    return loss, parts, grad


class _L1SSIM(torch.autograd.Function):
    @staticmethod
    def forward(ctx, ray_colors, target, width, height, supersampling, lambda_ssim):
        loss, _, grad = l1_ssim_forward(ray_colors.contiguous(), target.contiguous(), width, height, supersampling,
                                        lambda_ssim, want_grad=ray_colors.requires_grad)
        ctx.save_for_backward(grad if grad is not None else loss)
        ctx.has_grad = grad is not None
        return loss[0]

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return (grad * g if ctx.has_grad else None), None, None, None, None, None


def l1_ssim_loss(ray_colors, target, width, height, supersampling=(1, 1), lambda_ssim=0.2):
    """Scalar training loss of the reference (scripts/train.py:110-112) as one autograd node."""
    return _L1SSIM.apply(ray_colors, target, int(width), int(height), tuple(supersampling), float(lambda_ssim))
