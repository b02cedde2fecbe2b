"""The binned-SAH tree (tree_builder 1) against the Morton-order LBVH (tree_builder 0).

Both trees hold every Gaussian with a bounding ellipsoid once, so the hit sets are identical;
the trees order the leaves differently, so each slab's hits are summed in another order: images
agree to 1e-5 absolute, gradients to 1e-4 relative, and the slab-hit and pair counts agree up to
rays whose transmittance crosses the cut-off within that rounding (at most 1e-5 of the counts).
The oracle parity of the SAH tree itself is tested in test_gpu_parity.py (blobs crop, both
builders).  Edge cases: one, two and three Gaussians, Gaussians without an ellipsoid
(density at or below the threshold) mixed in or alone, and coincident centres (median splits).
"""

import numpy as np
import pytest
import torch

from ray_gauss_amd import scenes
from ray_gauss_amd.caster import TuneConfig

from gpu_helpers import max_rel, prepared_caster

pytestmark = pytest.mark.gpu


def _tune(builder):
    return TuneConfig(16, 8, packet=True, fwd_hit_depth=12, replay_words=0, tree_builder=builder)


def _render(sc, params, cam, dev, builder, px=None, target=None):
    rc, _ = prepared_caster(sc, params, [cam], dev, _tune(builder))
    assert rc.check_bvh() == 0
    rgb, _, _ = rc.forward(0, px)
    st = rc.stats()
    g = None
    if target is not None:
        rc.backward(0, px, rgb, 2.0 * (rgb - target))
        g = {k: v.cpu().numpy() for k, v in rc.finish_grads().items() if v is not None}
    return rgb.cpu().numpy(), st, g


def test_sah_matches_lbvh_blobs(gpu_device):
    wl = scenes.workload("blobs100k")
    cam = wl.cameras[0].launch_record()
    px = torch.from_numpy(scenes.crop_pixels(wl.cameras[0], 384, 384, 48, 48))
    target = torch.from_numpy(np.random.default_rng(3).uniform(0, 1, (len(px), 3)).astype(np.float32)).cuda()
    a, sa, ga = _render(wl.scene, wl.params, cam, gpu_device, 0, px, target)
    b, sb, gb = _render(wl.scene, wl.params, cam, gpu_device, 1, px, target)
    assert abs(sa["slab_hits"] - sb["slab_hits"]) <= 1e-5 * sa["slab_hits"]
    assert abs(sa["pairs"] - sb["pairs"]) <= 1e-5 * sa["pairs"]
    assert sa["slab_hits"] > 100000 and sb["fallback_slabs"] > 0
    assert np.abs(a - b).max() <= 1e-5
    for k, v in ga.items():
        assert max_rel(gb[k], v) <= 1e-4, k


@pytest.mark.parametrize("case", ["n1", "n2", "n3", "n64_some_empty", "all_empty", "coincident", "n5000"])
def test_sah_edge_cases(gpu_device, case):
    wl = scenes.workload("small4k")
    n = {"n1": 1, "n2": 2, "n3": 3, "n64_some_empty": 64, "all_empty": 40, "coincident": 300, "n5000": 5000}[case]
    sc = scenes.scene_uniform(n, seed=n + 7, log_scale_mean=-1.5 if n < 100 else -3.0)
    if case == "n64_some_empty":
        sc.densities[::3] = np.float32(wl.params.sigma_thr)  # at the threshold: no ellipsoid
    if case == "all_empty":
        sc.densities[:] = np.float32(0.5 * wl.params.sigma_thr)
    if case == "coincident":
        sc.positions[:200] = sc.positions[0]
    cam = wl.cameras[0].launch_record()
    a, sa, _ = _render(sc, wl.params, cam, gpu_device, 0)
    b, sb, _ = _render(sc, wl.params, cam, gpu_device, 1)
    assert sa["slab_hits"] == sb["slab_hits"] and sa["pairs"] == sb["pairs"]
    if case == "all_empty":
        assert sb["slab_hits"] == 0
    elif case != "n1":
        assert sb["slab_hits"] > 0
    assert np.abs(a - b).max() <= 1e-5
