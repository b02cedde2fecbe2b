"""Shared helpers of the GPU parity tests: move a synthetic scene to the device, run the HIP
path, run the CPU FP64 oracle on the same float32 inputs."""

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


def oracle_forward(scene, params, cam, box, pixels=None):
    rgb = O.sh_forward(scene, cam["eye"])
    return O.render(params, scene, rgb, box[0], box[1], cam, pixels)


def rel_err(a, b):
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def max_rel(a, b):
    """max |a-b| / max |b|: the per-tensor relative error used for gradients."""
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))
