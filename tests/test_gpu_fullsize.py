"""The 1M-Gaussian, 8-view workload at full size: oracle parity on seeded crops, and
size-independent properties over all 8.13M rays (batched == per-view bit for bit, opacity in
[0,1], white-background identity for rays that miss everything, gradient of a constant loss is
zero, linearity of the backward in dL/dC)."""

import numpy as np
import pytest
import torch

from oracle import oracle as O
from ray_gauss_amd import scenes
from ray_gauss_amd.batch import make_ray_batch

from gpu_helpers import check_image, gpu_box, max_rel, prepared_caster

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def big(gpu_device):
    wl = scenes.workload("unbounded1m_8views")
    cams = [c.launch_record() for c in wl.cameras]
    rc, t = prepared_caster(wl.scene, wl.params, cams, gpu_device)
    return wl, cams, rc, t


def test_bvh_and_crop_parity(big):
    wl, cams, rc, t = big
    assert rc.check_bvh() == 0
    v = 3
    px = scenes.crop_pixels(wl.cameras[v], 600, 380, 32, 32)
    pxt = torch.from_numpy(px)
    rgb, depth, opac = rc.forward(v, pxt, depth=True, opacity=True)
    tok = rc.last_token
    ref, ok, rec = check_image(rc, wl.scene, wl.params, cams[v], rgb, px)
    # gradients on the crop against the decision-matched oracle, only its flagged rays masked
    dL = np.random.default_rng(3).normal(0, 1, (len(px), 3)) * ok[:, None]
    rc.backward(v, pxt, rgb, torch.from_numpy(dL.astype(np.float32)).cuda(), tok)
    assert rc.stats()["dropped"] == 0
    g = rc.finish_grads()
    box = gpu_box(rc)
    og = O.full_backward(wl.params, wl.scene, cams[v]["eye"], box[0], box[1], cams[v],
                         rgb.cpu().numpy().astype(np.float64), dL, px, record=rec)
    for k in ("positions", "scales", "quats", "densities", "sh", "sg", "sharpness", "lobe_axis"):
        e = max_rel(g[k].cpu().numpy(), og[k])
        assert e <= 1e-3, (k, e)


def test_full_batch_properties(big):
    wl, cams, rc, t = big
    rv, rp, offs = make_ray_batch(cams, None, t["positions"].device)
    rgb, depth, opac = rc.forward_batch(rv, rp, depth=True, opacity=True)
    assert rgb.shape[0] == 8 * 1237 * 822
    assert torch.isfinite(rgb).all() and torch.isfinite(depth).all()
    assert (opac >= 0).all() and (opac <= 1).all()
    # batched == per-view, bit for bit, on two views
    for v in (0, 5):
        ref, _, _ = rc.forward(v, rp[offs[v]:offs[v + 1]])
        assert torch.equal(ref, rgb[offs[v]:offs[v + 1]])
    # rays with zero opacity show the white background exactly
    z = opac == 0
    if z.any():
        assert torch.equal(rgb[z], torch.ones_like(rgb[z]))
    st = rc.stats()
    assert st["pairs"] > 0 and st["slab_hits"] >= st["pairs"]


def test_backward_linearity(big):
    wl, cams, rc, t = big
    v = 1
    px = torch.from_numpy(scenes.crop_pixels(wl.cameras[v], 500, 300, 64, 64))
    rgb, _, _ = rc.forward(v, px)
    gen = torch.Generator(device=rgb.device).manual_seed(4)
    a = torch.randn(rgb.shape, generator=gen, device=rgb.device)
    b = torch.randn(rgb.shape, generator=gen, device=rgb.device)
    rc.backward(v, px, rgb, a)
    ga = rc.finish_grads()
    rc.backward(v, px, rgb, b)
    gb = rc.finish_grads()
    rc.backward(v, px, rgb, 2.0 * a - 3.0 * b)
    gab = rc.finish_grads()
    for k in ("positions", "densities", "sh", "quats"):
        lin = (2.0 * ga[k] - 3.0 * gb[k]).cpu().numpy()
        assert max_rel(gab[k].cpu().numpy(), lin) <= 1e-4, k
    rc.backward(v, px, rgb, torch.zeros_like(rgb))
    g0 = rc.finish_grads()
    for k in ("positions", "densities", "sh"):
        assert torch.count_nonzero(g0[k]) == 0
