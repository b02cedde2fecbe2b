"""The tree's surface-area cost (rg_tree_cost) and BatchedTrainer's refit-quality rebuild policy.

The cost is pinned by a closed form: when every Gaussian is the same, every node box is the root
box and the cost is the node count, 2n - 1, for both builders.  The policy: a refit over Gaussians
moved far from where the tree was built raises the cost past the threshold, the trainer rebuilds
(outside its graph), the cost returns to a fresh build's, and the gradients still equal an eager
pass over the same tensors."""

import numpy as np
import pytest
import torch

from ray_gauss_amd import scenes
from ray_gauss_amd.batch import BatchedTrainer, make_ray_batch
from ray_gauss_amd.caster import RayCaster, TuneConfig

from gpu_helpers import max_rel, prepared_caster, to_dev

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("builder", [0, 1])
def test_tree_cost_of_identical_gaussians_is_the_node_count(gpu_device, builder):
    n = 1000
    f = np.float32
    sc = scenes.Scene(np.tile(np.array([[0.1, -0.2, 0.3]], f), (n, 1)), np.full((n, 3), 0.05, f),
                      np.tile(np.array([[1, 0, 0, 0]], f), (n, 1)), np.full(n, 5.0, f), np.zeros((n, 3, 1), f),
                      np.zeros((n, 3, 0), f), np.zeros((n, 0), f), np.zeros((n, 0, 3), f))
    prm = scenes.RenderParams(dt=0.01, t_eps=1e-4, white_background=True)
    cam = scenes.look_at_camera(scenes.orbit_eye(2.0, 30.0, 10.0), width=16, height=16)
    rc, _ = prepared_caster(sc, prm, [cam], gpu_device, TuneConfig(tree_builder=builder))
    assert rc.check_bvh() == 0
    assert rc.tree_cost() == pytest.approx(2 * n - 1, rel=1e-12)


@pytest.mark.parametrize("builder", [0, 1])
def test_refit_quality_rebuild_policy(gpu_device, builder):
    sc = scenes.scene_blobs(20000, 16, seed=5, num_sg=2)
    prm = scenes.RenderParams(dt=0.01, t_eps=1e-4, white_background=True)
    cams = [scenes.look_at_camera(scenes.orbit_eye(4.0, a, 20.0), width=64, height=48).launch_record()
            for a in (20.0, 140.0)]
    rv, rp, _ = make_ray_batch(cams, None, gpu_device)
    gen = torch.Generator(device=gpu_device).manual_seed(3)
    target = torch.rand((rv.shape[0], 3), generator=gen, device=gpu_device)
    t = to_dev(sc, gpu_device)
    tune = TuneConfig(tree_builder=builder)
    tr = BatchedTrainer(RayCaster(prm, tune), t, cams, rv, rp, target, use_graph=True, rebuild_ratio=1.2,
                        check_every=1)
    base = tr.base_cost
    assert base > 1.0
    tr.step()
    assert tr.rebuilds == 0 and tr.cost == pytest.approx(base, rel=1e-6)
    # move every Gaussian onto another's position: the refitted boxes span the scene
    perm = torch.randperm(sc.n, generator=torch.Generator().manual_seed(0)).to(gpu_device)
    t["positions"].copy_(t["positions"][perm].clone())
    first = tr.graph
    tr.step()
    assert tr.rebuilds == 1
    assert tr.rc.check_bvh() == 0  # the refitted tree: every box exactly its children's union
    fresh = RayCaster(prm, tune)
    fresh.prepare(t["positions"], t["scales"], t["quats"], t["densities"], t["sh"], t["sg"], t["sharpness"],
                  t["lobe_axis"], rebuild=True)
    assert tr.base_cost == pytest.approx(fresh.tree_cost(), rel=1e-9)
    if builder == 0:
        assert tr.graph is first  # same tensors and buffers: the captured graph stays valid
    else:
        assert tr.graph is None  # the SAH tree's refit launches follow its levels: capture again
    eager = BatchedTrainer(RayCaster(prm, tune), t, cams, rv, rp, target, use_graph=False,
                           rebuild_ratio=None)
    ge = {k: (v.clone() if v is not None else None) for k, v in eager.step().items()}
    g = tr.step()
    torch.cuda.synchronize()
    assert tr.rebuilds == 1
    assert tr.graph is not None
    for k in ("positions", "scales", "quats", "densities", "sh", "sg"):
        assert max_rel(g[k].cpu().numpy(), ge[k].cpu().numpy()) <= 1e-5, k
