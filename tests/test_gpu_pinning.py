"""Parts of the GPU path that anchor everything else, each checked against the oracle on its own:
the scene box (it fixes every ray's sample grid) and the max_prim_slice overflow counter.

Scene box: render_optix.py:105-106 takes the min / max over compute_ellipsoids_bbox
(optix_utils.py:54-66), whose boxes are mu +- ||delta s * R_i|| with delta = sqrt(max(0, ln((sigma /
thr)^2))): Gaussians at or below the threshold are point boxes at their means and still count.

Overflow: __anyhit__ah (cuda_train/forward/gaussians_aabb.cu:621-641) adds one to
max_gaussians_exceeded when a slab's hit count reaches max_prim_slice, once per such slab, and the
launch reports the total (optix_utils.py:192-207).  Here no hit is ever dropped, so the counter
equals the number of processed slabs whose hit set is larger than the limit, which the FP32
brute-force hit sets give exactly when every slab is processed (t_eps = 0).
"""

import numpy as np
import pytest
import torch

from oracle import oracle as O
from ray_gauss_amd import scenes

from gpu_helpers import prepared_caster

pytestmark = pytest.mark.gpu


def _ulp(x):
    return np.spacing(np.abs(np.asarray(x, np.float32))).astype(np.float64)


def _box_check(rc, scene, thr):
    got = rc.scene_box().cpu().numpy().astype(np.float64)
    b = O.aabb(scene, thr)  # FP64 per-Gaussian boxes from the same float32 inputs
    lo, hi = b[:, :3].min(axis=0), b[:, 3:].max(axis=0)
    tol_lo = 4 * _ulp(lo) + 1e-12
    tol_hi = 4 * _ulp(hi) + 1e-12
    assert (np.abs(got[:3] - lo) <= tol_lo).all(), (got[:3], lo)
    assert (np.abs(got[3:] - hi) <= tol_hi).all(), (got[3:], hi)
    # every Gaussian's box and mean lies inside, up to the same rounding
    assert (b[:, :3] >= got[:3] - tol_lo).all() and (b[:, 3:] <= got[3:] + tol_hi).all()
    pos = scene.positions.astype(np.float64)
    assert (pos >= got[:3] - tol_lo).all() and (pos <= got[3:] + tol_hi).all()
    return got


@pytest.mark.parametrize("name", ["small4k", "blobs100k"])
def test_scene_box_matches_oracle(gpu_device, name):
    wl = scenes.workload(name)
    rc, _ = prepared_caster(wl.scene, wl.params, [wl.cameras[0].launch_record()], gpu_device)
    _box_check(rc, wl.scene, wl.params.sigma_thr)


@pytest.mark.parametrize("builder", [0, 1])
def test_scene_box_counts_gaussians_without_ellipsoid(gpu_device, builder):
    """Gaussians with |sigma| <= thr are never hit, but their means still extend the scene box:
    two such Gaussians placed outside every other box define four of its six faces."""
    from ray_gauss_amd.caster import TuneConfig

    base = scenes.scene_uniform(2000, seed=7)
    extra = scenes.Scene(positions=np.array([[4.5, -0.2, 3.25], [-3.75, 5.5, -0.1]], np.float32),
                         scales=np.full((2, 3), 0.3, np.float32), quats=np.array([[1, 0, 0, 0]] * 2, np.float32),
                         densities=np.array([0.05, 0.1], np.float32), sh=base.sh[:2], sg=base.sg[:2],
                         sharpness=base.sharpness[:2], lobe_axis=base.lobe_axis[:2])
    sc = scenes.Scene(*(np.concatenate([getattr(base, k), getattr(extra, k)]) for k in
                        ("positions", "scales", "quats", "densities", "sh", "sg", "sharpness", "lobe_axis")))
    prm = scenes.RenderParams(dt=0.01, t_eps=1e-4)
    cam = scenes.look_at_camera(scenes.orbit_eye(6.0, 30.0, 20.0), width=32, height=32).launch_record()
    rc, _ = prepared_caster(sc, prm, [cam], gpu_device, TuneConfig(tree_builder=builder))
    got = _box_check(rc, sc, prm.sigma_thr)
    # min x, max x, max y, max z are the two point boxes (sigma = 0.05 and sigma = thr exactly)
    assert got[0] == np.float32(-3.75) and got[3] == np.float32(4.5)
    assert got[4] == np.float32(5.5) and got[5] == np.float32(3.25)
    assert rc.check_bvh() == 0


def _overflow_case(rc, wl, prm_full, cam, px, max_slabs=4096, max_pairs=4096):
    """Per-slab hit counts from the FP32 brute force over the GPU's exported predicate."""
    pairs, counts, ro, rd, te, tx = rc.debug_hits(0, px, max_slabs, max_pairs)
    mu, Mp, valid = rc.export_predicate()
    ref_pairs, ref_counts = O.hits_fp32(mu.cpu().numpy(), Mp.cpu().numpy(), valid.cpu().numpy(), ro.cpu().numpy(),
                                        rd.cpu().numpy(), te.cpu().numpy(), tx.cpu().numpy(),
                                        np.float32(prm_full.dt), prm_full.samples_per_slab, max_slabs, max_pairs)
    assert ref_counts.max() <= max_pairs
    per_slab = []
    for r in range(len(ref_counts)):
        sl = ref_pairs[r, :ref_counts[r], 0]
        if len(sl):
            per_slab.append(np.bincount(sl))
    return np.concatenate(per_slab) if per_slab else np.zeros(0, np.int64)


@pytest.mark.parametrize("name", ["small4k", "blobs100k"])
def test_overflow_counter_matches_oracle(gpu_device, name):
    wl = scenes.workload(name)
    prm = scenes.RenderParams(**{**wl.params.__dict__, "t_eps": 0.0})  # every slab is processed
    cam = wl.cameras[0].launch_record()
    if name == "small4k":
        px = torch.arange(0, 4096, 3, dtype=torch.int32)
    else:
        px = torch.from_numpy(scenes.crop_pixels(wl.cameras[0], 384, 384, 32, 32))
    rc0, _ = prepared_caster(wl.scene, prm, [cam], gpu_device)
    hits = _overflow_case(rc0, wl, prm, cam, px)
    assert hits.max() >= 4
    img0, _, _ = rc0.forward(0, px)
    assert rc0.stats()["overflow_slabs"] == 0  # no limit set: nothing counted
    for limit in (int(hits.max()) - 1, int(np.percentile(hits[hits > 0], 75)), 2):
        rc, _ = prepared_caster(wl.scene, prm, [cam], gpu_device, max_hits=limit)
        img, _, _ = rc.forward(0, px)
        st = rc.stats()
        expect = int((hits > limit).sum())
        assert expect > 0
        assert st["overflow_slabs"] == expect, (limit, st["overflow_slabs"], expect)
        # no hit is dropped: the image does not depend on the limit
        assert torch.equal(img, img0)
