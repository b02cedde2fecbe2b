"""The allocation-free kNN entry point (rg_knn_mean_dist_noalloc) against the workspace entry
point it mirrors (rg_knn_mean_dist), both called through the C ABI: identical outputs on two
calls in a row, the second smaller (the static scratch keeps the first call's contents), and a
clear RG_ERR_ARG for a point count whose workspace does not fit the scratch array."""

import ctypes as C

import numpy as np
import pytest
import torch

from ray_gauss_amd import _lib

pytestmark = pytest.mark.gpu


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _points(n, seed, dev):
    rng = np.random.default_rng(seed)
    # clustered points (duplicates included) as an initial SfM cloud looks
    c = rng.uniform(-2, 2, (64, 3))
    p = c[rng.integers(0, 64, n)] + rng.normal(0, 0.05, (n, 3))
    p[: n // 100] = p[n // 100: 2 * (n // 100)]  # exact duplicates: zero distances
    return torch.from_numpy(p.astype(np.float32)).to(dev)


def test_noalloc_knn_matches_workspace_entry_point(gpu_device):
    L = _lib.load()
    for n, seed in ((1_200_000, 1), (3_001, 2)):  # the second call is smaller
        pts = _points(n, seed, gpu_device)
        nbytes = int(L.rg_knn_workspace_bytes(n))
        ws = torch.empty(nbytes, dtype=torch.uint8, device=gpu_device)
        ref = torch.full((n,), float("nan"), device=gpu_device)
        assert L.rg_knn_mean_dist(C.c_void_p(pts.data_ptr()), n, C.c_void_p(ref.data_ptr()),
                                  C.c_void_p(ws.data_ptr()), nbytes, _stream()) == 0
        out = torch.full((n,), float("nan"), device=gpu_device)
        rc = L.rg_knn_mean_dist_noalloc(C.c_void_p(pts.data_ptr()), n, C.c_void_p(out.data_ptr()), _stream())
        assert rc == 0, L.rg_last_error(None)
        torch.cuda.synchronize()
        assert torch.isfinite(ref).all()
        assert torch.equal(out, ref), (n, float((out - ref).abs().max()))
        assert float(out[: n // 100].min()) >= 0.0


def test_noalloc_knn_rejects_too_many_points(gpu_device):
    L = _lib.load()
    n = 4_000_000
    assert int(L.rg_knn_workspace_bytes(n)) > 64 << 20
    dummy = torch.zeros(3, device=gpu_device)  # never read: the size check comes first
    rc = L.rg_knn_mean_dist_noalloc(C.c_void_p(dummy.data_ptr()), n, C.c_void_p(dummy.data_ptr()), _stream())
    assert rc == -1
    msg = L.rg_last_error(None).decode()
    assert "scratch" in msg and "n <=" in msg
    limit = int(msg.split("n <= ")[1].split(")")[0])
    assert int(L.rg_knn_workspace_bytes(limit)) <= 64 << 20 < int(L.rg_knn_workspace_bytes(limit + 1))
    assert limit >= 1_500_000
