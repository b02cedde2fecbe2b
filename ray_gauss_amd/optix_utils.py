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
    return _Token("module", stage=stage, ctx=ctx)


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
    return _Token("sbt", positions=positions, scales=scales, quaternions=quaternions)


def launch_pipeline_test(pipeline, sbt, gas, bbox_min, bbox_max, camera, densities, color_features, positions,
                         scales, quaternions, max_prim_slice, rnd_sample, supersampling, white_background):
    """Forward render of one camera with caller-computed per-Gaussian colours, no gradients
    (optix_utils.py:302-373, called from scripts/test.py:112-116): returns ray_colors
    (H*W*ssx*ssy, 3) as a torch tensor.  Inputs are torch CUDA tensors (the reference passes
    CuPy views of the same data); color_features is the (n,3) colour table of compute_cupy_rgb.
    The scene box is recomputed from the Gaussians (bbox_min / bbox_max are accepted for the
    signature; the reference derives them from the same AABBs).  As in the reference's test
    kernel (cuda_test/gaussians_aabb.cu:312-333, 455-490), rays go through pixel centres and
    samples sit at slab start + i*dt whatever rnd_sample says: the argument is accepted and
    ignored there too."""
    from .render import _launch_camera
    from .scenes import RenderParams

    ctx = pipeline.groups[0].module.ctx
    params = RenderParams(white_background=bool(white_background))
    caster = ctx.caster_for(params, int(max_prim_slice or 0))
    n = int(positions.reshape(-1, 3).shape[0])
    f = lambda t, *s: t.reshape(*s).float().contiguous()  # noqa: E731
    sh0 = torch.zeros((n, 3, 1), dtype=torch.float32, device=positions.device)
    with torch.no_grad():
        caster.prepare(f(positions, n, 3), f(scales, n, 3), f(quaternions, n, 4), f(densities, n), sh0, rebuild=True)
        settings = _Token("settings", viewpoint_cam=camera, supersampling=supersampling)
        caster.set_views([_launch_camera(settings)])
        caster.set_view_colors(0, f(color_features, n, 3))
        rgb, _, _ = caster.forward(0, None)
    return rgb


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
