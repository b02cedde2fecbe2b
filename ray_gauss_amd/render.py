"""Drop-in replacement of the reference's autograd boundary
(``optix_raycasting/render_optix.py:67-190``): ``RenderOptixSettings`` and
``RenderOptixFunction`` with the same constructor / ``apply`` signatures, argument meanings,
return shapes and ``None`` conventions, backed by the gfx950 HIP library.

Differences a caller can see:

* OptiX-only settings (``context``, ``gas``, ``program_grps_*``, ``pipeline_*``,
  ``hit_prim_idx``) are accepted.  ``context`` is the object returned by
  :func:`ray_gauss_amd.optix_utils.create_context`, which owns the HIP handle; the others are
  ignored (the LBVH and hit lists live in the handle).
* The reference's compile-time constants (``DT``, ``BUFFER_SIZE``, ``TRANSMITTANCE_EPSILON``,
  ``SIGMA_THRESHOLD``; ``cuda_train/forward/gaussians_aabb.h:2-5``) are runtime parameters:
  pass ``params=RenderParams(...)`` (defaults are the reference's values).
* ``max_prim_slice`` no longer truncates hit sets: every hit is kept; slabs with more hits than
  ``max_prim_slice`` are counted and reported like the reference's overflow counter
  (``optix_utils.py:205-207``).
* ``jitter`` and ``rnd_sample`` follow the reference exactly, including its random generator
  (``tea<4>(iteration, pixel)`` seeds, LCG draws; ``forward/gaussians_aabb.h:63-94``); the
  backward of a call reuses that call's iteration.
* ``return_aux=True`` makes ``apply`` return ``(ray_colors, depth, opacity)`` with
  depth = sum T alpha t and opacity = 1 - T_final; the default return is unchanged.
"""

from __future__ import annotations

import torch

from .cameras import launch_record_from
from .scenes import RenderParams


class RenderOptixSettings:
    def __init__(self, context, update, gas, program_grps_fwd, pipeline_fwd, program_grps_bwd, pipeline_bwd,
                 viewpoint_cam, max_prim_slice, num_sph_gauss, iteration, jitter, rnd_sample, supersampling,
                 white_background, hit_prim_idx, params: RenderParams | None = None, return_aux: bool = False,
                 exact_sh_grad: bool = False):
        self.context = context
        self.update = update
        self.gas = gas
        self.program_grps_fwd = program_grps_fwd
        self.pipeline_fwd = pipeline_fwd
        self.program_grps_bwd = program_grps_bwd
        self.pipeline_bwd = pipeline_bwd
        self.viewpoint_cam = viewpoint_cam
        self.max_prim_slice = max_prim_slice
        self.num_sph_gauss = num_sph_gauss
        self.iteration = iteration
        self.jitter = jitter
        self.rnd_sample = rnd_sample
        self.supersampling = supersampling
        self.white_background = white_background
        self.hit_prim_idx = hit_prim_idx
        self.params = params
        self.return_aux = return_aux
        self.exact_sh_grad = exact_sh_grad


def _render_params(settings) -> RenderParams:
    base = settings.params if settings.params is not None else RenderParams()
    return RenderParams(dt=base.dt, samples_per_slab=base.samples_per_slab, t_eps=base.t_eps,
                        sigma_thr=base.sigma_thr, t_near=base.t_near, t_far=base.t_far,
                        white_background=bool(settings.white_background), jitter=int(settings.jitter or 0),
                        rnd_sample=int(settings.rnd_sample or 0), iteration=int(settings.iteration or 0))


def _caster(settings):
    ctx = settings.context
    if ctx is None or not hasattr(ctx, "caster_for"):
        from .optix_utils import create_context

        ctx = create_context()
        settings.context = ctx
    return ctx.caster_for(_render_params(settings), int(settings.max_prim_slice or 0), bool(settings.exact_sh_grad))


def _launch_camera(settings) -> dict:
    rec = launch_record_from(settings.viewpoint_cam)
    ssx, ssy = settings.supersampling if settings.supersampling is not None else (1, 1)
    rec["width"] = rec["width"] * int(ssx)
    rec["height"] = rec["height"] * int(ssy)
    return rec


class RenderOptixFunction(torch.autograd.Function):
    """RenderOptixFunction.apply(positions, scales, normalized_quaternions, densities,
    color_features, sph_gauss_features, bandwidth_sharpness, lobe_axis, settings) -> ray_colors
    (H*W*ssx*ssy, 3).  Forward: render_optix.py:91-132; backward: render_optix.py:135-190."""

    @staticmethod
    def forward(ctx, positions, scales, normalized_quaternions, densities, color_features, sph_gauss_features,
                bandwidth_sharpness, lobe_axis, settings):
        if int(settings.jitter or 0) not in (0, 1) or int(settings.rnd_sample or 0) not in (0, 1, 2):
            raise ValueError("jitter must be 0 or 1 and rnd_sample 0, 1 or 2")
        caster = _caster(settings)
        L = int(settings.num_sph_gauss)
        sg = sph_gauss_features if L else None
        sharp = bandwidth_sharpness if L else None
        axis = lobe_axis if L else None
        caster.prepare(positions, scales, normalized_quaternions, densities, color_features, sg, sharp, axis,
                       rebuild=not bool(settings.update))
        cam = _launch_camera(settings)
        caster.set_views([cam])
        aux = bool(settings.return_aux)
        rgb, depth, opacity = caster.forward(0, None, depth=aux, opacity=aux)
        if settings.max_prim_slice and settings.iteration is not None and settings.iteration % 1000 == 0:
            ovf = caster.stats()["overflow_slabs"]
            if ovf > 0:
                print(" ############ Number slabs with max gaussians exceeded : ", ovf)
        settings.gas = caster
        ctx.settings = settings
        ctx.context = settings.context
        ctx.max_prim_slice = int(settings.max_prim_slice or 0)
        ctx.exact_sh_grad = bool(settings.exact_sh_grad)
        ctx.token = caster.last_token
        ctx.caster = caster
        ctx.cam = cam
        ctx.generation = caster.generation
        ctx.params = caster.params
        ctx.num_sg = L
        ctx.save_for_backward(positions, scales, normalized_quaternions, densities, color_features,
                              sph_gauss_features, bandwidth_sharpness, lobe_axis, rgb.clone())
        if aux:
            ctx.mark_non_differentiable(depth, opacity)
            return rgb, depth, opacity
        return rgb

    @staticmethod
    def backward(ctx, dloss_dray_colors, *unused):
        (positions, scales, quats, densities, color_features, sg, sharp, axis, ray_colors) = ctx.saved_tensors
        L = ctx.num_sg
        # another call (e.g. a later iteration, or an inference render) may have changed the
        # handle's settings: restore this call's through the context, which keeps its own record
        # of the settings in force current
        caster = ctx.context.caster_for(ctx.params, ctx.max_prim_slice, ctx.exact_sh_grad)
        if caster.generation != ctx.generation:
            # another forward used this handle in between: restore this call's state
            caster.prepare(positions, scales, quats, densities, color_features, sg if L else None,
                           sharp if L else None, axis if L else None, rebuild=True)
            caster.set_views([ctx.cam])
        # the recorded hit lists are used only if nothing replaced them since this forward
        caster.backward(0, None, ray_colors, dloss_dray_colors.contiguous(), ctx.token)
        g = caster.finish_grads()
        if L:
            gsg, gsh, gax = g["sg"], g["sharpness"], g["lobe_axis"]
        else:
            gsg = gsh = gax = None
        return (g["positions"], g["scales"], g["quats"], g["densities"], g["sh"], gsg, gsh, gax, None)


def render(positions, scales, quats, densities, color_features, sph_gauss_features, bandwidth_sharpness, lobe_axis,
           viewpoint_cam, params: RenderParams | None = None, context=None, return_aux=False, update=False):
    """Convenience wrapper: builds the settings object and applies RenderOptixFunction."""
    L = int(sph_gauss_features.shape[2]) if sph_gauss_features is not None and sph_gauss_features.dim() == 3 else 0
    st = RenderOptixSettings(context, update, None, None, None, None, None, viewpoint_cam, 0, L, 0, 0, 0, (1, 1),
                             bool(params.white_background) if params else False, None, params=params,
                             return_aux=return_aux)
    return RenderOptixFunction.apply(positions, scales, quats, densities, color_features, sph_gauss_features,
                                     bandwidth_sharpness, lobe_axis, st)
