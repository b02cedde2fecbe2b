"""Densification and initialisation utilities on the GPU against the CPU oracle.

Tolerances: distCUDA2 relative 1e-5 (FP32 distances against float64); 3D filter 1e-5 absolute
on the log filter over points whose threshold tests are not within 1e-5 of deciding the other
way; gradient norms relative 1e-6; densification: every copied row bit-exact, densities
bit-exact, children's positions and log-scales within 1e-5 (relative to max(1, |x|))."""

import math
import types

import numpy as np
import pytest
import torch

from oracle import densify_oracle as D
from ray_gauss_amd import densify as G
from ray_gauss_amd import scenes
from ray_gauss_amd.optim import FusedAdam

pytestmark = pytest.mark.gpu


def _knn_case(n, seed):
    rng = np.random.default_rng(seed)
    if n < 100:
        return rng.normal(size=(n, 3)).astype(np.float32)
    c = rng.normal(size=(max(1, n // 500), 3)) * 4.0
    p = c[rng.integers(0, len(c), n)] + rng.normal(size=(n, 3)) * 0.05
    p = p.astype(np.float32)
    p[: n // 50] = p[n // 50: 2 * (n // 50)]  # exact duplicates
    return p


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 63, 64, 65, 4097, 30000, 262145])
def test_distcuda2_vs_oracle(gpu_device, n):
    p = _knn_case(n, n)
    got = G.distCUDA2(torch.from_numpy(p).to(gpu_device)).cpu().numpy()
    ref = D.knn3_mean_dist(p)
    assert got.shape == (n,)
    if n < 4:
        assert np.array_equal(got, ref.astype(np.float32))
        return
    rel = np.abs(got - ref) / np.maximum(ref, 1e-30)
    assert rel.max() <= 1e-5, (rel.max(), int(rel.argmax()))


def test_distcuda2_degenerate_sets(gpu_device):
    # all points equal: every distance 0; a line: neighbours at 1, 1, 4 (ends: 1, 4, 9)
    same = torch.full((1000, 3), 2.5, device=gpu_device)
    assert torch.all(G.distCUDA2(same) == 0)
    line = torch.zeros((200, 3), device=gpu_device)
    line[:, 0] = torch.arange(200, device=gpu_device, dtype=torch.float32)
    d = G.distCUDA2(line).cpu().numpy()
    assert np.all(d[2:-2] == np.float32(2.0)) and d[0] == np.float32(14 / 3) and d[1] == np.float32(2.0)
    with pytest.raises(TypeError):
        G.distCUDA2(torch.zeros((4, 3), dtype=torch.float64, device=gpu_device))


def test_filter_3d_vs_oracle(gpu_device):
    wl = scenes.workload("blobs100k")
    pts = wl.scene.positions.astype(np.float32)
    cams = wl.cameras
    got = G.filter_3d(torch.from_numpy(pts).to(gpu_device), cams).cpu().numpy().reshape(-1)
    ref, seen, margin = D.filter_3d(pts, [{"R": c.R, "T": c.T, "FoVx": c.FoVx, "FoVy": c.FoVy,
                                           "image_width": c.image_width, "image_height": c.image_height}
                                          for c in cams])
    ok = margin > 1e-5
    assert ok.mean() > 0.999 and 0.05 < seen.mean()
    err = np.abs(got - ref)[ok].max()
    assert err <= 1e-5, err
    pc = types.SimpleNamespace(positions=torch.from_numpy(pts).to(gpu_device))
    G.compute_3D_filter(pc, cams)
    assert pc.filter_3D.shape == (len(pts), 1)
    far = torch.tensor([[0.0, 0.0, 1e5]] * 3, device=gpu_device)  # above every orbit camera, behind each
    with pytest.raises(RuntimeError):
        G.filter_3d(far, cams)


def test_accumulate_gradient(gpu_device):
    rng = np.random.default_rng(2)
    n = 10000
    g = rng.normal(size=(n, 3)).astype(np.float32)
    g[::7] = 0.0
    gd = torch.from_numpy(g).to(gpu_device)
    pc = types.SimpleNamespace(xyz_gradient_accum_norm=torch.ones(n, device=gpu_device),
                               num_accum=torch.zeros(n, device=gpu_device),
                               num_accum_gnv=torch.full((n,), 3.0, device=gpu_device))
    G.accumulate_gradient(pc, gd)
    G.accumulate_gradient(pc, gd, not_visible=False)
    nrm = np.linalg.norm(g.astype(np.float64), axis=1)
    np.testing.assert_allclose(pc.xyz_gradient_accum_norm.cpu().numpy(), 1 + 2 * nrm, rtol=1e-6)
    np.testing.assert_array_equal(pc.num_accum.cpu().numpy(), 2.0 * (nrm > 0))
    np.testing.assert_array_equal(pc.num_accum_gnv.cpu().numpy(), 3.0 + (nrm > 0))


class _Mgr:
    def __init__(self, t):
        self.optimizer = FusedAdam([t], lr=1e-3)


def _point_cloud(n, dev, seed, L=2, K=16):
    rng = np.random.default_rng(seed)
    f = np.float32
    P = {
        "positions": rng.normal(size=(n, 3)).astype(f) * f(3),
        "rgb": rng.normal(size=(n, 3)).astype(f),
        "spherical_harmonics": rng.normal(size=(n, 3, K - 1)).astype(f),
        "densities": rng.uniform(0.0, 2.0, n).astype(f),
        "scales": rng.normal(-3.5, 1.0, (n, 3)).astype(f),
        "quaternions": rng.normal(size=(n, 4)).astype(f),
        "sph_gauss_features": rng.normal(size=(n, 3, L)).astype(f),
        "bandwidth_sharpness": rng.normal(size=(n, L)).astype(f),
        "lobe_axis": rng.normal(size=(n, L, 3)).astype(f),
    }
    pc = types.SimpleNamespace(**{k: torch.from_numpy(v).to(dev).requires_grad_(True) for k, v in P.items()})
    acc = rng.exponential(1.0, n).astype(f)
    num = rng.integers(0, 5, n).astype(f)
    acc[num == 0] = 0
    pc.xyz_gradient_accum_norm = torch.from_numpy(acc).to(dev)
    pc.num_accum = torch.from_numpy(num).to(dev)
    pc.num_accum_gnv = torch.from_numpy(rng.integers(0, 3, n).astype(f)).to(dev)
    pc.filter_3D = torch.from_numpy(rng.normal(-6, 1, (n, 1)).astype(f)).to(dev)
    mgrs = {}
    for oname, attr in G.PARAMS:
        t = getattr(pc, attr)
        m = _Mgr(t)
        t.grad = torch.from_numpy(rng.normal(size=t.shape).astype(f)).to(dev)
        m.optimizer.step()  # creates the moments
        t.grad = None
        mgrs[oname] = m
    pc.optim_managers = types.SimpleNamespace(optimizer_manager_dict=mgrs)
    return pc


def _snapshot(pc):
    arrs, states = {}, []
    for oname, attr in G.PARAMS:
        t = getattr(pc, attr)
        arrs[oname] = t.detach().cpu().numpy().copy()
        st = pc.optim_managers.optimizer_manager_dict[oname].optimizer.state[t]
        for k in ("exp_avg", "exp_avg_sq"):
            arrs[f"{k}:{oname}"] = st[k].cpu().numpy().copy()
            states.append(f"{k}:{oname}")
    arrs["filter_3D"] = pc.filter_3D.cpu().numpy().copy()
    arrs["num_accum_gnv"] = pc.num_accum_gnv.cpu().numpy().copy()
    states.append("num_accum_gnv")
    return arrs, tuple(states)


@pytest.mark.parametrize("n,N", [(20000, 2), (1, 2), (5000, 3)])
def test_densify_and_prune_vs_oracle(gpu_device, n, N):
    pc = _point_cloud(n, gpu_device, seed=n + N)
    before, states = _snapshot(pc)
    acc, num = pc.xyz_gradient_accum_norm.cpu().numpy(), pc.num_accum.cpu().numpy()
    thr, min_d, extent = 0.8, 0.1, 4.0
    gen = torch.Generator(device=gpu_device).manual_seed(123)
    counts = G.densify_and_prune(pc, thr, min_d, extent, N=N, generator=gen)
    ns = counts[2]
    gen2 = torch.Generator(device=gpu_device).manual_seed(123)
    z = torch.normal(torch.zeros((N * ns, 3), device=gpu_device), torch.ones((N * ns, 3), device=gpu_device),
                     generator=gen2).cpu().numpy() if ns else np.zeros((0, 3), np.float32)
    ref, rcounts = D.densify_and_prune(before, acc, num, thr, min_d, extent, z, N=N, states=states)
    assert tuple(counts) == rcounts
    if n > 1:
        assert counts[1] > 0 and counts[2] > 0 and counts[0] < n  # clones, splits and prunes happen
    n_out = counts[0] + counts[1] + N * counts[3]
    child0 = counts[0] + counts[1]
    for oname, attr in G.PARAMS:
        t = getattr(pc, attr)
        assert t.shape[0] == n_out and t.requires_grad
        got = t.detach().cpu().numpy()
        exp = ref[oname]
        if oname in ("xyz", "scales"):
            np.testing.assert_array_equal(got[:child0], exp[:child0])
            err = np.abs(got[child0:] - exp[child0:]) / np.maximum(1.0, np.abs(exp[child0:]))
            assert err.size == 0 or err.max() <= 1e-5, (oname, err.max())
        else:
            np.testing.assert_array_equal(got, exp, err_msg=oname)
        st = pc.optim_managers.optimizer_manager_dict[oname].optimizer.state[t]
        for k in ("exp_avg", "exp_avg_sq"):
            np.testing.assert_array_equal(st[k].cpu().numpy(), ref[f"{k}:{oname}"], err_msg=f"{k}:{oname}")
    np.testing.assert_array_equal(pc.filter_3D.cpu().numpy(), ref["filter_3D"])
    np.testing.assert_array_equal(pc.num_accum_gnv.cpu().numpy(), ref["num_accum_gnv"])
    assert torch.all(pc.xyz_gradient_accum_norm == 0) and pc.num_accum.shape == (n_out,)
    # the optimisers keep working on the new tensors
    for oname, attr in G.PARAMS:
        t = getattr(pc, attr)
        t.grad = torch.ones_like(t)
        pc.optim_managers.optimizer_manager_dict[oname].optimizer.step()
