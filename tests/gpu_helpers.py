"""Shared helpers of the GPU parity tests: move a synthetic scene to the device, run the HIP
path, run the CPU FP64 oracle on the same float32 inputs.

Image parity is checked against the oracle in two modes (check_image):
  * decision-matched: the oracle takes every discrete decision (slab hit sets, sample thresholds
    sigma*w > thr) in FP32 from the kernels' exported per-Gaussian records, exactly as the kernels
    do, and evaluates every value in FP64.  Every ray whose decisions are not within a few ulp of
    their boundary (native vs host exp2) or whose transmittance check is not within rounding of
    T_eps must match to 1e-4; those few flagged rays (at most a few percent) must lie within their
    explicit bound; gradients are compared against this mode with only the flagged rows masked.
  * exact decisions (FP64 throughout): every ray must lie within its explicit bound -- 1e-4 plus,
    per sample decision within EPS_W (twice the measured FP32 error of sigma*w) of the threshold,
    the most a flipped sample can move the colour (oracle.fragile_bound).
"""

from __future__ import annotations

import numpy as np
import torch

from oracle import oracle as O
from ray_gauss_amd.caster import RayCaster, TuneConfig

KEYS = ("positions", "scales", "quats", "densities", "sh", "sg", "sharpness", "lobe_axis")


def to_dev(scene, dev):
    return {k: torch.from_numpy(np.ascontiguousarray(getattr(scene, k))).to(dev) for k in KEYS}


def prepared_caster(scene, params, cams, dev, tune=None, exact_sh_grad=False, max_hits=0):
    rc = RayCaster(params, tune or TuneConfig(), exact_sh_grad=exact_sh_grad, max_hits_per_slab=max_hits)
    t = to_dev(scene, dev)
    L = scene.num_sg
    rc.prepare(t["positions"], t["scales"], t["quats"], t["densities"], t["sh"], t["sg"] if L else None,
               t["sharpness"] if L else None, t["lobe_axis"] if L else None, rebuild=True)
    rc.set_views([c if isinstance(c, dict) else c.launch_record() for c in cams])
    return rc, t


def gpu_box(rc):
    b = rc.scene_box().cpu().numpy()
    return b[:3].copy(), b[3:].copy()


def oracle_forward(scene, params, cam, box, pixels=None, record=None, colors=None):
    rgb = O.sh_forward(scene, cam["eye"]) if colors is None else np.asarray(colors, np.float64)
    return O.render(params, scene, rgb, box[0], box[1], cam, pixels, record=record)


IMG_TOL = 1e-4


def check_image(rc, scene, params, cam, rgb_gpu, pixels=None, max_flagged=0.05, box=None, colors=None):
    """Image parity of rgb_gpu (num_rays, 3) with the oracle in both modes (module docstring).
    Returns (decision-matched oracle output, mask of the rays whose gradients are compared,
    record)."""
    box = gpu_box(rc) if box is None else box
    rec = rc.export_record()
    g = np.asarray(rgb_gpu.cpu().numpy() if hasattr(rgb_gpu, "cpu") else rgb_gpu, np.float64)
    dm = oracle_forward(scene, params, cam, box, pixels, record=rec, colors=colors)
    ok = dm["fragile"] == 0
    assert ok.mean() >= 1.0 - max_flagged, ok.mean()
    err = np.abs(g - dm["rgb"]).max(axis=1)
    assert err[ok].max() <= IMG_TOL, err[ok].max()
    assert (err <= O.fragile_bound(params, dm)).all()
    ex = oracle_forward(scene, params, cam, box, pixels, colors=colors)
    err_ex = np.abs(g - ex["rgb"]).max(axis=1)
    bound = O.fragile_bound(params, ex)
    assert (err_ex <= bound).all(), (err_ex - bound).max()
    assert err_ex[ex["fragile"] == 0].max(initial=0.0) <= IMG_TOL
    return dm, ok, rec


def rel_err(a, b):
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def max_rel(a, b):
    """max |a-b| / max |b|: the per-tensor relative error used for gradients."""
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))
