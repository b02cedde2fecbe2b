"""Inference and viewer outputs on the GPU against the CPU FP64 oracle: the test-stage launch
(optix_utils.py:302-373 with caller colours), the viewer camera and frame (cuda_gui), and the
RGBA8 quantisation of make_color (cuda_gui/helpers.h:45-57).

Tolerances: images max-abs <= 1e-4 over rays the oracle does not flag as fragile; depth
<= 1e-3 of the largest depth; RGBA8 bit-exact against a numpy restatement, and within one level
of the oracle's image (a colour within 1e-4 of a 1/256 step may round either way)."""

import numpy as np
import pytest
import torch

from oracle import oracle as O
from ray_gauss_amd import optix_utils as U
from ray_gauss_amd import scenes, viewer

from gpu_helpers import check_image, gpu_box, prepared_caster

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-4


def _rgba8_numpy(x):
    x = np.asarray(x, np.float32)
    c = np.fmax(np.float32(0), np.fmin(x, np.float32(1)))  # fmin(NaN, 1) = 1, as fminf
    q = np.minimum(np.floor(c * np.float32(256)), 255).astype(np.uint8)
    return np.concatenate([q, np.full(q.shape[:-1] + (1,), 255, np.uint8)], axis=-1)


def test_rgba8_quantisation(gpu_device):
    special = np.array([-1.0, -0.0, 0.0, 1 / 256 - 1e-7, 1 / 256, 0.5, 255 / 256 - 1e-7, 255 / 256, 1 - 1e-7, 1.0,
                        2.0, np.inf, -np.inf, np.nan, 1e-30, 0.99999994], np.float32)
    rnd = np.random.default_rng(0).uniform(-0.2, 1.2, 3 * 9997).astype(np.float32)
    x = np.concatenate([special, rnd, special[:2]]).reshape(-1, 3)
    got = viewer.to_rgba8(torch.from_numpy(x).to(gpu_device)).cpu().numpy()
    assert got.shape == (x.shape[0], 4)
    assert np.array_equal(got, _rgba8_numpy(x))
    img = torch.from_numpy(x[:12].reshape(2, 2, 3, 3)[..., 0, :].copy()).to(gpu_device)
    assert viewer.to_rgba8(img).shape == (2, 2, 4)
    with pytest.raises(TypeError):
        viewer.to_rgba8(torch.zeros(4, 4, device=gpu_device))


@pytest.mark.parametrize("white,ss", [(False, (1, 1)), (True, (2, 1))])
def test_launch_pipeline_test_vs_oracle(gpu_device, white, ss):
    """The test stage renders caller colours at pixel centres with deterministic samples; the
    rnd_sample argument does not change the result (the reference's test kernel never reads
    it)."""
    wl = scenes.workload("small4k")
    sc = wl.scene
    n = sc.n
    cam = wl.cameras[0]
    colors = np.random.default_rng(7).uniform(0, 1, (n, 3)).astype(np.float32)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(gpu_device)  # noqa: E731
    ctx = U.create_context()
    module = U.create_module(ctx, None, "test")
    groups = U.create_program_groups(ctx, module)
    pipe = U.create_pipeline(ctx, groups, None)
    pos, scl, q, dens = dev(sc.positions), dev(sc.scales), dev(sc.quats), dev(sc.densities)
    L1, L2, L3 = U.quaternion_to_rotation(q)
    boxes = U.compute_ellipsoids_bbox(pos, scl, L1, L2, L3, dens)
    gas = U.create_acceleration_structure(ctx, boxes)
    sbt = U.create_sbt(groups, pos, scl, q)
    args = (pipe, sbt, gas, boxes[:, :3].min(0).values, boxes[:, 3:].max(0).values, cam, dens, dev(colors), pos, scl,
            q)
    out = U.launch_pipeline_test(*args, max_prim_slice=0, rnd_sample=0, supersampling=ss, white_background=white)
    out1 = U.launch_pipeline_test(*args, max_prim_slice=0, rnd_sample=1, supersampling=ss, white_background=white)
    torch.cuda.synchronize()
    rec = cam.launch_record()
    rec["width"] *= ss[0]
    rec["height"] *= ss[1]
    assert out.shape == (rec["width"] * rec["height"], 3)
    assert torch.equal(out, out1)
    prm = scenes.RenderParams(white_background=white)
    check_image(ctx._caster, sc, prm, rec, out, colors=colors, max_flagged=0.01)


def test_viewer_frame_vs_oracle(gpu_device):
    """Viewer camera (pre-scaled U, V; unit tangents), white background, RGBA8 and depth."""
    wl = scenes.workload("small4k")
    rec = wl.cameras[0].launch_record()
    gui = viewer.gui_camera(rec["eye"], rec["u"] * rec["tanx"], rec["v"] * rec["tany"], rec["w"], rec["width"],
                            rec["height"])
    prm = scenes.RenderParams(white_background=True)
    rc, _ = prepared_caster(wl.scene, prm, [gui], gpu_device)
    rgba, depth = viewer.render_frame(rc)
    torch.cuda.synchronize()
    h, w = rec["height"], rec["width"]
    assert rgba.shape == (h, w, 4) and rgba.dtype == torch.uint8 and depth.shape == (h, w)
    rgb_dev, _, _ = rc.forward(0)
    ref, ok, _ = check_image(rc, wl.scene, prm, gui, rgb_dev, max_flagged=0.01)
    got = rgba.cpu().numpy().reshape(-1, 4).astype(np.int32)
    assert np.array_equal(got, _rgba8_numpy(rgb_dev.cpu().numpy()).astype(np.int32))
    exp = _rgba8_numpy(ref["rgb"].astype(np.float32)).astype(np.int32)
    assert np.abs(got - exp)[ok].max() <= 1
    assert (got == exp)[ok].mean() > 0.99
    d = depth.cpu().numpy().reshape(-1)
    assert np.abs(d - ref["depth"])[ok].max() <= 1e-4 * max(1.0, ref["depth"].max())
    with pytest.raises(ValueError):
        rc2, _ = prepared_caster(wl.scene, scenes.RenderParams(), [gui], gpu_device)
        viewer.render_frame(rc2)


def test_set_view_colors_shape_check(gpu_device):
    wl = scenes.workload("small4k")
    rc, _ = prepared_caster(wl.scene, wl.params, [wl.cameras[0].launch_record()], gpu_device)
    with pytest.raises((ValueError, TypeError)):
        rc.set_view_colors(0, torch.zeros((wl.scene.n, 4), device=gpu_device))
    with pytest.raises(Exception):
        rc.set_view_colors(3, torch.zeros((wl.scene.n, 3), device=gpu_device))
