"""Viewer frames on the GPU (the reference's interactive viewer, optix_raycasting/cuda_gui):

* ``gui_camera`` -- the viewer's camera: eye and pre-scaled U, V, W, rays through pixel centres,
  direction normalize(px U + py V + W) (cuda_gui/gaussians_aabb.cu:298-309), i.e. a launch
  record with unit field-of-view tangents;
* ``render_frame`` -- one frame with a white background and deterministic samples
  (cuda_gui/gaussians_aabb.cu:418-513): RGBA8 image and expected depth sum T alpha t
  (:401-415) from the caster's forward kernel;
* ``to_rgba8`` -- colours to RGBA8 as ``make_color`` (cuda_gui/helpers.h:45-57): clamp to
  [0, 1], min(floor(256 x), 255), alpha 255; the sRGB step is commented out there and is not
  applied here either.

The viewer's per-Gaussian colour (SH of degree ``degree_sh`` + SG lobes + 0.5, :364-373) is the
caster's own colour stage; a lower display degree is obtained by passing the first
(d+1)^2 SH coefficients to ``RayCaster.prepare``."""

from __future__ import annotations

import ctypes as C

import torch

from . import _lib


def to_rgba8(rgb: torch.Tensor) -> torch.Tensor:
    """(..., 3) float32 CUDA colours -> (..., 4) uint8."""
    if rgb.dtype != torch.float32 or not rgb.is_cuda or rgb.shape[-1] != 3:
        raise TypeError("to_rgba8: float32 CUDA tensor with a last dimension of 3")
    src = rgb.contiguous()
    n = src.numel() // 3
    out = torch.empty(tuple(rgb.shape[:-1]) + (4,), dtype=torch.uint8, device=rgb.device)
    lib = _lib.load()
    rc = lib.rg_colors_to_rgba8(C.c_void_p(src.data_ptr()), n, C.c_void_p(out.data_ptr()),
                                C.c_void_p(torch.cuda.current_stream(rgb.device).cuda_stream))
    if rc != 0:
        raise _lib.RayGaussError(f"rg_colors_to_rgba8 failed ({rc}): {lib.rg_last_error(None).decode()}")
    return out


def gui_camera(eye, U, V, W, width, height) -> dict:
    """Launch record of the viewer's camera (vectors as 3-sequences, U and V already scaled by
    the field of view)."""
    import numpy as np

    f = lambda v: np.asarray(v, np.float32).reshape(3).copy()  # noqa: E731
    return {"eye": f(eye), "u": f(U), "v": f(V), "w": f(W), "tanx": np.float32(1.0), "tany": np.float32(1.0),
            "width": int(width), "height": int(height)}


def render_frame(caster, view: int = 0):
    """(rgba8 (H, W, 4) uint8, depth (H, W) float32) of view `view` of a prepared caster whose
    render parameters have a white background and no jitter or random samples."""
    p = caster.params
    if not p.white_background or int(p.jitter or 0) or int(p.rnd_sample or 0):
        raise ValueError("render_frame: the viewer renders on white with deterministic samples")
    rec = caster.views[view]
    rgb, depth, _ = caster.forward(view, None, depth=True)
    h, w = rec["height"], rec["width"]
    return to_rgba8(rgb.view(h, w, 3)), depth.view(h, w)
