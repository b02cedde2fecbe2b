"""MI355X-native differentiable Gaussian ray caster (the RayGauss training hot path).

Public surface:
  RenderOptixSettings, RenderOptixFunction  -- drop-in autograd boundary (render_optix.py)
  optix_utils                               -- host-glue names used by the training loop
  RayCaster, TuneConfig                     -- direct handle on the HIP library
  scenes, cameras                           -- synthetic workloads and pinhole cameras
"""

from .scenes import RenderParams  # noqa: F401


def __getattr__(name):
    # torch-dependent pieces load lazily so the numpy-only helpers import without a GPU stack
    if name in ("RenderOptixSettings", "RenderOptixFunction", "render"):
        from . import render as _r

        return getattr(_r, name)
    if name in ("RayCaster", "TuneConfig"):
        from . import caster as _c

        return getattr(_c, name)
    if name == "optix_utils":
        from . import optix_utils as _o

        return _o
    raise AttributeError(name)
