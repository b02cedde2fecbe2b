"""Parity of the HIP ray caster with the CPU FP64 oracle on the same seeded float32 inputs.

Tolerances: images max-abs <= 1e-4 against the decision-matched oracle on every ray not flagged
(a decision within a few ulp of its boundary or a transmittance check within rounding of T_eps),
every ray within its explicit bound against both oracle modes (gpu_helpers.check_image);
gradients <= 1e-3 relative (max-abs error over max-abs value, per tensor) against the
decision-matched oracle with only the flagged rays' dL/dC set to zero; hit sets and the FP32 ray
grid bit-exact.
"""

import numpy as np
import pytest
import torch

from oracle import oracle as O
from ray_gauss_amd import scenes
from ray_gauss_amd.caster import TuneConfig

from gpu_helpers import check_image, gpu_box, max_rel, oracle_forward, prepared_caster, rel_err

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-4
GRAD_TOL = 1e-3


@pytest.fixture(scope="module")
def small(gpu_device):
    wl = scenes.workload("small4k")
    cam = wl.cameras[0].launch_record()
    rc, t = prepared_caster(wl.scene, wl.params, [cam], gpu_device)
    return wl, cam, rc, t


def test_bvh_structure(small):
    _, _, rc, _ = small
    assert rc.check_bvh() == 0


def test_ray_grid_bit_exact(small):
    wl, cam, rc, _ = small
    px = torch.arange(0, 4096, 3, dtype=torch.int32)
    _, _, ro, rd, te, tx = rc.debug_hits(0, px, 1, 1)
    o, d = O.rays(cam, px.numpy())
    bmin, bmax = gpu_box(rc)
    te_ref, tx_ref = O.ray_interval(cam, bmin, bmax, wl.params.t_near, wl.params.t_far, px.numpy())
    assert np.array_equal(ro.cpu().numpy().view(np.uint32), o.view(np.uint32))
    assert np.array_equal(rd.cpu().numpy().view(np.uint32), d.view(np.uint32))
    assert np.array_equal(te.cpu().numpy().view(np.uint32), te_ref.view(np.uint32))
    assert np.array_equal(tx.cpu().numpy().view(np.uint32), tx_ref.view(np.uint32))


def _hitset_check(rc, wl, px, max_slabs=4096, max_pairs=4096):
    pairs, counts, ro, rd, te, tx = rc.debug_hits(0, px, max_slabs, max_pairs)
    mu, Mp, valid = rc.export_predicate()
    ref_pairs, ref_counts = O.hits_fp32(mu.cpu().numpy(), Mp.cpu().numpy(), valid.cpu().numpy(), ro.cpu().numpy(),
                                        rd.cpu().numpy(), te.cpu().numpy(), tx.cpu().numpy(), np.float32(wl.params.dt),
                                        wl.params.samples_per_slab, max_slabs, max_pairs)
    counts = counts.cpu().numpy()
    assert counts.max() <= max_pairs
    assert np.array_equal(counts, ref_counts)
    gp = pairs.cpu().numpy()
    for r in range(len(counts)):
        c = counts[r]
        got = gp[r, :c]
        got = got[np.lexsort((got[:, 1], got[:, 0]))]
        assert np.array_equal(got, ref_pairs[r, :c]), r
    return int(counts.sum())


def test_hit_sets_bit_exact_small(small):
    wl, _, rc, _ = small
    px = torch.arange(0, 4096, 5, dtype=torch.int32)
    total = _hitset_check(rc, wl, px)
    assert total > 1000


def test_forward_parity_small(small):
    wl, cam, rc, _ = small
    rgb, depth, opac = rc.forward(0, depth=True, opacity=True)
    st = rc.stats()
    ref, ok, _ = check_image(rc, wl.scene, wl.params, cam, rgb, max_flagged=0.01)
    assert np.abs(opac.cpu().numpy() - ref["opacity"])[ok].max() <= IMG_TOL
    assert np.abs(depth.cpu().numpy() - ref["depth"])[ok].max() <= 1e-4 * max(1.0, ref["depth"].max())
    # the same hit sets slab by slab: the slab-hit and pair counts agree exactly
    assert st["slab_hits"] == int(ref["K"].sum()) and st["pairs"] == int(ref["P"].sum())


def test_backward_parity_small(small):
    wl, cam, rc, _ = small
    rgb, _, _ = rc.forward(0)
    tok = rc.last_token
    ref, ok, rec = check_image(rc, wl.scene, wl.params, cam, rgb, max_flagged=0.01)
    target = np.random.default_rng(1).uniform(0, 1, (4096, 3))
    dL = 2.0 * (rgb.cpu().numpy().astype(np.float64) - target) * ok[:, None]
    rc.backward(0, None, rgb, torch.from_numpy(dL.astype(np.float32)).cuda(), tok)
    assert rc.stats()["dropped"] == 0
    g = rc.finish_grads()
    box = gpu_box(rc)
    og = O.full_backward(wl.params, wl.scene, cam["eye"], box[0], box[1], cam, rgb.cpu().numpy().astype(np.float64),
                         dL, record=rec)
    for k in ("positions", "scales", "quats", "densities", "sh"):
        e = max_rel(g[k].cpu().numpy(), og[k])
        assert e <= GRAD_TOL, (k, e, rel_err(g[k].cpu().numpy(), og[k]))


@pytest.mark.parametrize("depth,seg,packet", [(8, 1, False), (32, 64, True), (16, 8, False), (8, 4, True)])
def test_tuning_invariance(gpu_device, depth, seg, packet):
    # the reference case uses the 6-entry forward list, the others the backward's depth
    """Hit-buffer depth, segment length and packet traversal change the schedule, not the image:
    every schedule sums each slab's hits in the same (slot) order, so the images are
    bit-identical."""
    wl = scenes.workload("small4k")
    cam = wl.cameras[0].launch_record()
    rc0, _ = prepared_caster(wl.scene, wl.params, [cam], gpu_device, TuneConfig(16, 8, packet=True))
    rc1, _ = prepared_caster(wl.scene, wl.params, [cam], gpu_device,
                             TuneConfig(depth, seg, packet=packet, fwd_hit_depth=depth))
    a = rc0.forward(0)[0].cpu().numpy()
    b = rc1.forward(0)[0].cpu().numpy()
    assert np.array_equal(a, b), np.abs(a - b).max()


def test_traversal_modes_agree(gpu_device):
    """Per-lane walks and packet walks reach the same leaves in the same order, including the
    slabs whose hit sets overflow the list (their per-slab walks and the backward's wave-merged
    scatter over them): identical images, and gradients equal up to the order of the atomic
    sums."""
    wl = scenes.workload("blobs100k")
    cam = wl.cameras[0].launch_record()
    px = torch.from_numpy(scenes.crop_pixels(wl.cameras[0], 384, 384, 32, 32))
    target = torch.from_numpy(np.random.default_rng(5).uniform(0, 1, (len(px), 3)).astype(np.float32)).cuda()
    out = {}
    for mode in (False, True):
        rc, _ = prepared_caster(wl.scene, wl.params, [cam], gpu_device, TuneConfig(8, 8, packet=mode, fwd_hit_depth=6))
        rgb, _, _ = rc.forward(0, px)
        st = rc.stats()
        rc.backward(0, px, rgb, 2.0 * (rgb - target))
        assert rc.stats()["dropped"] == 0
        g = {k: v.cpu().numpy() for k, v in rc.finish_grads().items() if v is not None}
        out[mode] = (rgb.cpu().numpy(), g, st)
    assert out[False][2]["fallback_slabs"] > 0  # the direct (overflow) path is exercised
    assert np.array_equal(out[True][0], out[False][0])
    for k, v in out[False][1].items():
        assert max_rel(out[True][1][k], v) <= 1e-5, k


@pytest.mark.parametrize("builder", [0, 1])
def test_blobs_crop_parity(gpu_device, builder):
    wl = scenes.workload("blobs100k")
    cam = wl.cameras[0].launch_record()
    rc, _ = prepared_caster(wl.scene, wl.params, [cam], gpu_device, TuneConfig(tree_builder=builder))
    assert rc.check_bvh() == 0
    px = torch.from_numpy(scenes.crop_pixels(wl.cameras[0], 384, 384, 32, 32))
    rgb, _, _ = rc.forward(0, px)
    tok = rc.last_token
    ref, ok, rec = check_image(rc, wl.scene, wl.params, cam, rgb, px.numpy(), max_flagged=0.02)
    dL = np.random.default_rng(2).normal(0, 1, (px.shape[0], 3)) * ok[:, None]
    rc.backward(0, px, rgb, torch.from_numpy(dL.astype(np.float32)).cuda(), tok)
    assert rc.stats()["dropped"] == 0
    g = rc.finish_grads()
    box = gpu_box(rc)
    og = O.full_backward(wl.params, wl.scene, cam["eye"], box[0], box[1], cam, rgb.cpu().numpy().astype(np.float64),
                         dL, px.numpy(), record=rec)
    for k in ("positions", "scales", "quats", "densities", "sh", "sg", "sharpness", "lobe_axis"):
        e = max_rel(g[k].cpu().numpy(), og[k])
        assert e <= GRAD_TOL, (k, e)
    _hitset_check(rc, wl, px[::7].contiguous())


@pytest.mark.parametrize("jitter,rnd", [(1, 0), (0, 1), (1, 1), (0, 2)])
def test_random_sampling_parity(gpu_device, jitter, rnd):
    """jitter / rnd_sample (forward/gaussians_aabb.cu:405-498, backward :505-775): the jittered
    FP32 rays are bit-exact with the oracle's restatement of the reference generator, and images
    and gradients match the oracle, which draws the same per-slab sample offsets."""
    wl = scenes.workload("small4k")
    prm = scenes.RenderParams(**{**wl.params.__dict__, "jitter": jitter, "rnd_sample": rnd, "iteration": 17})
    cam = wl.cameras[0].launch_record()
    rc, _ = prepared_caster(wl.scene, prm, [cam], gpu_device)
    px = torch.arange(0, 4096, 7, dtype=torch.int32)
    _, _, ro, rd, _, _ = rc.debug_hits(0, px, 1, 1)
    o, d = O.rays_rng(cam, px.numpy(), jitter, 17)
    assert np.array_equal(rd.cpu().numpy().view(np.uint32), d.view(np.uint32))
    rgb, _, _ = rc.forward(0)
    tok = rc.last_token
    ref, ok, rec = check_image(rc, wl.scene, prm, cam, rgb, max_flagged=0.01)
    base = oracle_forward(wl.scene, wl.params, cam, gpu_box(rc))
    assert np.abs(ref["rgb"] - base["rgb"]).max() > 1e-4  # the settings do change the image
    target = np.random.default_rng(3).uniform(0, 1, (4096, 3))
    dL = 2.0 * (rgb.cpu().numpy().astype(np.float64) - target) * ok[:, None]
    rc.backward(0, None, rgb, torch.from_numpy(dL.astype(np.float32)).cuda(), tok)
    assert rc.stats()["dropped"] == 0
    g = rc.finish_grads()
    box = gpu_box(rc)
    og = O.full_backward(prm, wl.scene, cam["eye"], box[0], box[1], cam, rgb.cpu().numpy().astype(np.float64), dL,
                         record=rec)
    for k in ("positions", "scales", "quats", "densities", "sh"):
        e = max_rel(g[k].cpu().numpy(), og[k])
        assert e <= GRAD_TOL, (k, e, rel_err(g[k].cpu().numpy(), og[k]))
