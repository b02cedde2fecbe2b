"""MI355X-native differentiable Gaussian ray caster (the RayGauss training hot path).

Public surface:
  RenderOptixSettings, RenderOptixFunction  -- drop-in autograd boundary (render_optix.py)
  optix_utils                               -- host-glue names used by the training loop
  RayCaster, TuneConfig                     -- direct handle on the HIP library
  FusedAdam, step_all, project_params       -- optimiser step + post-step projections (optim.py)
  l1_ssim_loss                              -- the training loss and its gradient (loss.py)
  viewer, formats                           -- RGBA8 + depth frames; RayGauss PLY / checkpoints
  densify                                   -- distCUDA2, 3D filter, clone/split/prune
  scenes, cameras                           -- synthetic workloads and pinhole cameras
"""

import importlib

from .scenes import RenderParams  # noqa: F401

_LAZY = {"RenderOptixSettings": "render", "RenderOptixFunction": "render", "render": "render",
         "RayCaster": "caster", "TuneConfig": "caster", "FusedAdam": "optim", "step_all": "optim",
         "project_params": "optim", "l1_ssim_loss": "loss"}


def __getattr__(name):
    # torch-dependent pieces load lazily so the numpy-only helpers import without a GPU stack
    if name in _LAZY:
        return getattr(importlib.import_module("." + _LAZY[name], __name__), name)
    raise AttributeError(name)
