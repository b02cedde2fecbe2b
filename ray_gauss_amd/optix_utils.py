"""Host-glue entry points the reference's training script calls around the ray caster
(``optix_raycasting/optix_utils.py``), with the same names and argument meanings.

The reference JIT-compiles OptiX modules and builds a GAS from AABBs; here the kernels are
compiled ahead of time for gfx950 and the LBVH is built by the HIP library from the Gaussian
parameters inside the forward call.  ``create_context`` therefore returns the object that owns
the HIP handle, and the module / program-group / pipeline / GAS / SBT functions return light
tokens so that an unchanged training loop (``scripts/train.py:75-102``) runs.
"""

from __future__ import annotations

import torch

from .caster import RayCaster, TuneConfig

SIGMA_THRESHOLD = 0.1  # reference gaussians_aabb.h:5 (checked for equality across headers there)
_STAGES = ("forward", "backward", "test", "gui")


class DeviceContext:
    """Owns the ray caster handle (replaces ox.DeviceContext, optix_utils.py:77-81)."""

    def __init__(self, log=None, tune: TuneConfig | None = None):
        self.log = log
        self.tune = tune or TuneConfig()
        self._caster = None
        self._key = None

    def caster_for(self, params, max_prim_slice=0, exact_sh_grad=False) -> RayCaster:
        key = (params.dt, params.samples_per_slab, params.t_eps, params.sigma_thr, params.t_near, params.t_far,
               bool(params.white_background), int(max_prim_slice), bool(exact_sh_grad), int(params.jitter),
               int(params.rnd_sample), int(params.iteration))
        if self._caster is None:
            self._caster = RayCaster(params, self.tune, exact_sh_grad, max_prim_slice)
        elif key != self._key:
            c = self._caster
            c.params, c.max_hits_per_slab, c.exact_sh_grad = params, int(max_prim_slice), bool(exact_sh_grad)
            c.set_tune(c.tune)
        self._key = key
        return self._caster


class _Token:
    def __init__(self, kind, **kw):
        self.kind = kind
        self.__dict__.update(kw)


def create_context(log=None, tune: TuneConfig | None = None) -> DeviceContext:
    return DeviceContext(log, tune)


def create_module(ctx, pipeline_opts, stage):
    if stage not in _STAGES:
        raise ValueError("stage must be 'forward', 'backward', 'test' or 'gui'")
    return _Token("module", stage=stage)


def create_program_groups(ctx, module):
    return (_Token("raygen", module=module), _Token("miss", module=module), _Token("hit", module=module))


def create_pipeline(ctx, program_grps, pipeline_options, debug_level=None):
    return _Token("pipeline", groups=program_grps)


def create_acceleration_structure(ctx, bboxes):
    return _Token("gas", count=int(bboxes.shape[0]))


def update_acceleration_structure(gas, bboxes):
    if hasattr(gas, "count"):
        gas.count = int(bboxes.shape[0])


def create_sbt(program_grps, positions, scales, quaternions):
    return _Token("sbt")


def quaternion_to_rotation(quaternion):
    """Rows L1, L2, L3 of R(q) for q = (w, x, y, z), not normalised (optix_utils.py:34-46)."""
    w, x, y, z = quaternion[:, 0], quaternion[:, 1], quaternion[:, 2], quaternion[:, 3]
    L1 = torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], dim=1)
    L2 = torch.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], dim=1)
    L3 = torch.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], dim=1)
    return L1, L2, L3


def compute_ellipsoids_bbox(centers, scales, L1, L2, L3, densities):
    """AABBs of the bounding ellipsoids sigma*w = SIGMA_THRESHOLD (optix_utils.py:54-66)."""
    delta = torch.log((densities / SIGMA_THRESHOLD) ** 2)
    delta = torch.sqrt(torch.clamp(delta, min=0.0))
    h = torch.stack([torch.linalg.norm(delta[:, None] * scales * Li, dim=1) for Li in (L1, L2, L3)], dim=1)
    return torch.cat([centers - h, centers + h], dim=1)
