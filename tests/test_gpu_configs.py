"""The two remaining workloads at full size on the GPU.

blobs300k_16views: 300k Gaussians, 16 views of uneven size (widths 400-1600, 4:3), each a random
25-100 % subset of its pixels (8.27M rays): the batched launch equals the per-view launches bit for
bit, and seeded 32x32 crops of several views match the oracle (images and all eight gradient
tensors).

unbounded3m_4views: 3M Gaussians with SH degree 3 and EIGHT spherical-Gaussian lobes, 4 views of
2048x1536 (12.6M rays): crop parity of images and gradients including the eight-lobe SG terms,
properties of the full batch, and the training-step protocol of scripts/train.py:164-183 --
perturb every mean by N(0, 1e-3) and every log-scale by N(0, 1e-2), refit the tree without a
rebuild, then hit sets bit-exact against the FP32 brute force and images against the oracle, for
both tree builders.
"""

import numpy as np
import pytest
import torch

from oracle import oracle as O
from ray_gauss_amd import scenes
from ray_gauss_amd.batch import make_ray_batch
from ray_gauss_amd.caster import TuneConfig

from gpu_helpers import check_image, gpu_box, max_rel, prepared_caster

pytestmark = pytest.mark.gpu

KEYS8 = ("positions", "scales", "quats", "densities", "sh", "sg", "sharpness", "lobe_axis")


def _crop_parity(rc, wl, v, px, seed, token=True):
    """Image and gradient parity of one crop of view v against the oracle in both modes
    (gpu_helpers.check_image); gradients against the decision-matched oracle."""
    cam = wl.cameras[v].launch_record()
    pxt = torch.from_numpy(px)
    rgb, depth, opac = rc.forward(v, pxt, depth=True, opacity=True)
    tok = rc.last_token
    box = gpu_box(rc)
    ref, ok, rec = check_image(rc, wl.scene, wl.params, cam, rgb, px)
    assert np.abs(opac.cpu().numpy() - ref["opacity"])[ok].max() <= 1e-4
    dL = np.random.default_rng(seed).normal(0, 1, (len(px), 3)) * ok[:, None]
    rc.backward(v, pxt, rgb, torch.from_numpy(dL.astype(np.float32)).cuda(), tok if token else 0)
    st = rc.stats()
    assert st["dropped"] == 0 and st["bad_rays"] == 0
    g = rc.finish_grads()
    og = O.full_backward(wl.params, wl.scene, cam["eye"], box[0], box[1], cam,
                         rgb.cpu().numpy().astype(np.float64), dL, px, record=rec)
    for k in KEYS8:
        e = max_rel(g[k].cpu().numpy(), og[k])
        assert e <= 1e-3, (v, k, e)
    return ref


# ------------------------------------------------------------------ blobs300k_16views --------

@pytest.fixture(scope="module")
def c3(gpu_device):
    wl = scenes.workload("blobs300k_16views")
    cams = [c.launch_record() for c in wl.cameras]
    rc, t = prepared_caster(wl.scene, wl.params, cams, gpu_device, TuneConfig(replay_words=0, tree_builder=1))
    rv, rp, offs = make_ray_batch(cams, wl.pixels, gpu_device)
    return wl, cams, rc, t, rv, rp, offs


def test_c3_batched_equals_per_view(c3):
    wl, cams, rc, t, rv, rp, offs = c3
    assert len(cams) == 16 and rv.shape[0] == sum(len(p) for p in wl.pixels)
    widths = [c["width"] for c in cams]
    assert min(widths) >= 400 and max(widths) <= 1600
    out, _, opac = rc.forward_batch(rv, rp, opacity=True)
    st = rc.stats()
    assert st["bad_rays"] == 0 and st["pairs"] > 0
    assert torch.isfinite(out).all() and (opac >= 0).all() and (opac <= 1).all()
    for v in range(16):
        ref, _, _ = rc.forward(v, rp[offs[v]:offs[v + 1]])
        assert torch.equal(ref, out[offs[v]:offs[v + 1]]), v


@pytest.mark.parametrize("v", [0, 7, 13])
def test_c3_crop_parity(c3, v):
    wl, cams, rc, t, rv, rp, offs = c3
    w, h = cams[v]["width"], cams[v]["height"]
    px = scenes.crop_pixels(wl.cameras[v], w // 2 - 16, h // 2 - 16, 32, 32)
    _crop_parity(rc, wl, v, px, seed=10 + v)


# ------------------------------------------------------------------ unbounded3m_4views -------

@pytest.fixture(scope="module")
def c5(gpu_device):
    wl = scenes.workload("unbounded3m_4views")
    assert wl.scene.num_sg == 8 and wl.scene.n == 3_000_000
    cams = [c.launch_record() for c in wl.cameras]
    rc, t = prepared_caster(wl.scene, wl.params, cams, gpu_device, TuneConfig(replay_words=0, tree_builder=1))
    return wl, cams, rc, t


def test_c5_crop_parity_eight_lobes(c5):
    wl, cams, rc, t = c5
    px = scenes.crop_pixels(wl.cameras[1], 1000, 700, 32, 32)
    _crop_parity(rc, wl, 1, px, seed=21)


def test_c5_full_batch_properties(c5):
    wl, cams, rc, t = c5
    rv, rp, offs = make_ray_batch(cams, None, t["positions"].device)
    assert rv.shape[0] == 4 * 2048 * 1536
    rgb, depth, opac = rc.forward_batch(rv, rp, depth=True, opacity=True)
    st = rc.stats()
    assert st["bad_rays"] == 0 and st["slab_hits"] >= st["pairs"] > 0
    assert torch.isfinite(rgb).all() and torch.isfinite(depth).all()
    assert (opac >= 0).all() and (opac <= 1).all()
    ref, _, _ = rc.forward(2, rp[offs[2]:offs[3]])
    assert torch.equal(ref, rgb[offs[2]:offs[3]])
    z = opac == 0
    if z.any():
        assert torch.equal(rgb[z], torch.ones_like(rgb[z]))


@pytest.mark.parametrize("builder", [0, 1])
def test_c5_refit_after_perturbation(c5, gpu_device, builder):
    """One step of the training protocol: perturb, refit (no rebuild), and the refitted tree
    still finds exactly the FP32 predicate's hit sets; images match the oracle on the perturbed
    Gaussians."""
    wl, cams, rc0, t0 = c5
    rng = np.random.default_rng(100 + builder)
    sc = wl.scene
    pos = (sc.positions.astype(np.float64) + rng.normal(0, 1e-3, sc.positions.shape)).astype(np.float32)
    scl = np.exp(np.log(sc.scales.astype(np.float64)) + rng.normal(0, 1e-2, sc.scales.shape)).astype(np.float32)
    moved = scenes.Scene(pos, scl, sc.quats, sc.densities, sc.sh, sc.sg, sc.sharpness, sc.lobe_axis)
    rc, t = prepared_caster(sc, wl.params, cams, gpu_device, TuneConfig(replay_words=0, tree_builder=builder))
    t["positions"].copy_(torch.from_numpy(pos))
    t["scales"].copy_(torch.from_numpy(scl))
    rc.prepare(t["positions"], t["scales"], t["quats"], t["densities"], t["sh"], t["sg"], t["sharpness"],
               t["lobe_axis"], rebuild=False)
    assert rc.check_bvh() == 0
    v = 0
    px = scenes.crop_pixels(wl.cameras[v], 1010, 760, 32, 32)
    sub = torch.from_numpy(px[::16].copy())
    max_slabs, max_pairs = 8192, 8192
    pairs, counts, ro, rd, te, tx = rc.debug_hits(v, sub, max_slabs, max_pairs)
    mu, Mp, valid = rc.export_predicate()
    assert np.array_equal(mu.cpu().numpy(), pos)
    ref_pairs, ref_counts = O.hits_fp32(mu.cpu().numpy(), Mp.cpu().numpy(), valid.cpu().numpy(), ro.cpu().numpy(),
                                        rd.cpu().numpy(), te.cpu().numpy(), tx.cpu().numpy(),
                                        np.float32(wl.params.dt), wl.params.samples_per_slab, max_slabs, max_pairs)
    counts = counts.cpu().numpy()
    assert counts.max() <= max_pairs and counts.sum() > 1000
    assert np.array_equal(counts, ref_counts)
    gp = pairs.cpu().numpy()
    for r in range(len(counts)):
        got = gp[r, :counts[r]]
        got = got[np.lexsort((got[:, 1], got[:, 0]))]
        assert np.array_equal(got, ref_pairs[r, :counts[r]]), r
    rgb, _, _ = rc.forward(v, torch.from_numpy(px))
    check_image(rc, moved, wl.params, cams[v], rgb, px)
