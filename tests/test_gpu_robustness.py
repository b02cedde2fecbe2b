"""Error paths and state handling of the HIP library on the GPU: rays with out-of-range indices,
the 32-bit merge-key limit, gradients left by an unfinished backward across a tree rebuild, and
the batched trainer's graph after the Gaussians change (densification)."""

import ctypes as C

import numpy as np
import pytest
import torch

from ray_gauss_amd import _lib, scenes
from ray_gauss_amd.batch import BatchedTrainer, make_ray_batch
from ray_gauss_amd.caster import RayCaster, TuneConfig

from gpu_helpers import max_rel, prepared_caster, to_dev

pytestmark = pytest.mark.gpu


def _two_views(dev):
    wl = scenes.workload("small4k")
    cams = [wl.cameras[0].launch_record(),
            scenes.look_at_camera(scenes.orbit_eye(3.0, 120.0, 10.0), width=48, height=40).launch_record()]
    rc, t = prepared_caster(wl.scene, wl.params, cams, dev)
    rv, rp, offs = make_ray_batch(cams, None, dev)
    return wl, cams, rc, t, rv, rp


def test_bad_rays_are_skipped_and_counted(gpu_device):
    """A ray naming a view that was not set, or a pixel outside its view, is skipped (NaN colour,
    no gradient) and counted in bad_rays; the other rays are unaffected."""
    wl, cams, rc, t, rv, rp = _two_views(gpu_device)
    good, _, _ = rc.forward_batch(rv, rp)
    assert rc.stats()["bad_rays"] == 0
    rv_b, rp_b = rv.clone(), rp.clone()
    bad = torch.tensor([5, 700, 4100, 6000], device=gpu_device)
    rv_b[bad[0]] = 2          # view index == num_views
    rv_b[bad[1]] = -1         # negative view
    rp_b[bad[2]] = 48 * 40    # one past the last pixel of view 1 (rows >= 4096 are view 1)
    rp_b[bad[3]] = -3         # negative pixel (view 1)
    out, depth, opac = rc.forward_batch(rv_b, rp_b, depth=True, opacity=True)
    st = rc.stats()
    assert st["bad_rays"] == 4
    assert torch.isnan(out[bad]).all() and torch.isnan(depth[bad]).all() and torch.isnan(opac[bad]).all()
    keep = torch.ones(rv.shape[0], dtype=torch.bool, device=gpu_device)
    keep[bad] = False
    assert torch.equal(out[keep], good[keep])
    # backward: the bad rays contribute nothing, whatever their dL/dC
    gen = torch.Generator(device=gpu_device).manual_seed(0)
    dL = torch.randn(out.shape, generator=gen, device=gpu_device)
    rc.backward_batch(rv_b, rp_b, torch.nan_to_num(out), dL)
    st = rc.stats()
    assert st["bad_rays"] == 4 and st["dropped"] == 0
    g_bad = rc.finish_grads()
    dL_masked = dL.clone()
    dL_masked[bad] = 0
    rc.backward_batch(rv, rp, good, dL_masked)
    g_ref = rc.finish_grads()
    for k in ("positions", "scales", "quats", "densities", "sh"):
        assert max_rel(g_bad[k].cpu().numpy(), g_ref[k].cpu().numpy()) <= 1e-5, k


def test_pixel_list_out_of_range_is_counted(gpu_device):
    wl, cams, rc, t, rv, rp = _two_views(gpu_device)
    px = torch.tensor([0, 17, 4096, -1, 4095], dtype=torch.int32, device=gpu_device)
    out, _, _ = rc.forward(0, px)
    assert rc.stats()["bad_rays"] == 2
    assert torch.isnan(out[2:4]).all() and torch.isfinite(out[[0, 1, 4]]).all()


def test_merge_key_limit_is_enforced(gpu_device):
    """The backward merges contributions by the 32-bit key view * n + slot: rg_set_views refuses
    num_views * n >= 2^31 - 1 with RG_ERR_ARG instead of letting the key wrap."""
    wl = scenes.workload("small4k")
    rc, _ = prepared_caster(wl.scene, wl.params, [wl.cameras[0].launch_record()], gpu_device)
    L = _lib.load()
    too_many = (2 ** 31 - 1) // wl.scene.n + 1
    cams = (_lib.CameraS * too_many)()  # contents never read: the size check comes first
    rc_code = L.rg_set_views(rc._h, too_many, cams, C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc_code == -1
    assert b"num_views * n" in L.rg_last_error(rc._h)
    # the handle still works with its previous views
    out, _, _ = rc.forward(0)
    assert torch.isfinite(out).all()


def test_tree_rebuild_clears_unfinished_gradients(gpu_device):
    """Gradient accumulators are per slot; a rebuild renumbers slots, so what an unfinished
    backward accumulated is cleared, colour gradients included."""
    wl = scenes.workload("small4k")
    cam = wl.cameras[0].launch_record()
    rc, t = prepared_caster(wl.scene, wl.params, [cam], gpu_device)
    rgb, _, _ = rc.forward(0)
    rc.backward(0, None, rgb, torch.ones_like(rgb))  # accumulated, never finished
    L = wl.scene.num_sg
    rc.prepare(t["positions"], t["scales"], t["quats"], t["densities"], t["sh"], rebuild=True)
    rc.set_views([cam])
    g = rc.finish_grads()
    for k in ("positions", "scales", "quats", "densities", "sh"):
        assert torch.count_nonzero(g[k]) == 0, k


def _densified(t, extra, seed):
    """Appends `extra` perturbed copies of existing Gaussians (new tensors, as densification makes)."""
    gen = torch.Generator(device=t["positions"].device).manual_seed(seed)
    idx = torch.randint(0, t["positions"].shape[0], (extra,), generator=gen, device=t["positions"].device)
    out = {}
    for k, v in t.items():
        if v is None:
            out[k] = None
            continue
        add = v[idx].clone()
        if k == "positions":
            add += 0.01 * torch.randn(add.shape, generator=gen, device=add.device)
        out[k] = torch.cat([v, add]).contiguous()
    return out


def test_trainer_recaptures_after_densification(gpu_device):
    """BatchedTrainer's graph bakes in the handle's buffers and the parameter tensors; after a
    rebuild over a larger set of Gaussians it captures again, and its gradients equal an eager
    pass over the same tensors (three replays, with host work in between)."""
    sc = scenes.scene_blobs(20000, 16, seed=3, num_sg=4)
    prm = scenes.RenderParams(dt=0.01, t_eps=1e-4, white_background=True)
    cams = [scenes.look_at_camera(scenes.orbit_eye(4.0, a, 20.0), width=64, height=48).launch_record()
            for a in (10.0, 130.0)]
    rv, rp, _ = make_ray_batch(cams, None, gpu_device)
    gen = torch.Generator(device=gpu_device).manual_seed(1)
    target = torch.rand((rv.shape[0], 3), generator=gen, device=gpu_device)
    t = to_dev(sc, gpu_device)
    tr = BatchedTrainer(RayCaster(prm, TuneConfig()), t, cams, rv, rp, target, use_graph=True)
    tr.step()
    first = tr.graph
    t2 = _densified(t, 3000, seed=2)
    tr.rebuild(t2)
    assert tr.graph is None  # stale: the Gaussian count and the tensors changed
    eager = BatchedTrainer(RayCaster(prm, TuneConfig()), t2, cams, rv, rp, target, use_graph=False)
    ge = {k: (v.clone() if v is not None else None) for k, v in eager.step().items()}
    for rep in range(3):
        g = tr.step()
        torch.cuda.synchronize()
        _ = float((target * 2).sum())
        assert g["positions"].shape[0] == 23000
        for k in ("positions", "scales", "quats", "densities", "sh", "sg"):
            assert max_rel(g[k].cpu().numpy(), ge[k].cpu().numpy()) <= 1e-5, (rep, k)
    assert tr.graph is not first
