"""Densification and initialisation utilities of the training loop, on the GPU.

* ``distCUDA2(points)`` -- simple-knn's mean squared distance to the 3 nearest other points
  (submodules/simple-knn/simple_knn.cu:186-222), used to initialise scales and the 3D filter
  (classes/point_cloud.py:206,376,807).
* ``compute_3D_filter(pc, cameras)`` -- PointCloud.compute_3D_filter
  (classes/point_cloud.py:1152-1215): sets ``pc.filter_3D`` (n, 1).
* ``accumulate_gradient(pc, grad)`` -- PointCloud.accumulate_gradient and
  accumulate_gradient_gaussians_not_visible (classes/point_cloud.py:632-640) in one pass.
* ``densify_and_prune(pc, max_grad, min_density, extent)`` -- PointCloud.densify_and_prune
  (classes/point_cloud.py:1067-1143): clone, split, prune, with every parameter tensor and its
  Adam moments rebuilt in one streaming launch.

``pc`` is any object with the reference PointCloud's attribute names (positions, rgb,
spherical_harmonics, densities, scales, quaternions, sph_gauss_features, bandwidth_sharpness,
lobe_axis, xyz_gradient_accum_norm, num_accum, num_accum_gnv, filter_3D) and, optionally,
``optim_managers.optimizer_manager_dict[name].optimizer`` holding one Adam-like optimiser per
parameter (names xyz, rgb, sh, density, scales, quaternions, sph_gauss_features,
bandwidth_sharpness, lobe_axis) whose state is edited as the reference's ``_prune_optimizer`` /
``_cat_tensor`` do.  The split's normal draws come from ``torch.normal`` with the same shape as
the reference's call, so a seeded generator gives the reference's samples.
"""

from __future__ import annotations

import ctypes as C
import math

import torch

from . import _lib

# (optimiser name, attribute) of the per-Gaussian parameters, classes/point_cloud.py:1036-1044
PARAMS = (("xyz", "positions"), ("rgb", "rgb"), ("sh", "spherical_harmonics"), ("density", "densities"),
          ("scales", "scales"), ("quaternions", "quaternions"), ("sph_gauss_features", "sph_gauss_features"),
          ("bandwidth_sharpness", "bandwidth_sharpness"), ("lobe_axis", "lobe_axis"))


def _stream(dev=None):
    return C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _check(rc, what):
    if rc != 0:
        raise _lib.RayGaussError(f"{what} failed ({rc}): {_lib.load().rg_last_error(None).decode()}")


def _dev_f32(t, name):
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float32:
        raise TypeError(f"{name}: float32 CUDA tensor expected")
    return t.detach().contiguous()


@torch.no_grad()
def distCUDA2(points: torch.Tensor) -> torch.Tensor:
    """(n, 3) float32 CUDA points -> (n,) mean squared distance to the 3 nearest other points."""
    p = _dev_f32(points, "distCUDA2")
    if p.dim() != 2 or p.shape[1] != 3:
        raise ValueError("distCUDA2: points must be (n, 3)")
    n = p.shape[0]
    out = torch.empty(n, dtype=torch.float32, device=p.device)
    if n == 0:
        return out
    lib = _lib.load()
    wsb = int(lib.rg_knn_workspace_bytes(n))
    if wsb <= 0:
        raise ValueError(f"distCUDA2: {n} points is beyond the supported size")
    ws = torch.empty(wsb, dtype=torch.uint8, device=p.device)
    _check(lib.rg_knn_mean_dist(C.c_void_p(p.data_ptr()), n, C.c_void_p(out.data_ptr()), C.c_void_p(ws.data_ptr()),
                                wsb, _stream(p.device)), "rg_knn_mean_dist")
    return out


def filter_cameras(cameras):
    """rg_filter_camera records and the focal length of classes/point_cloud.py:1160-1195 from
    reference-style cameras (R, T, FoVx, FoVy, image_width, image_height): float32 (V, 21) table
    and float(1 / max focal)."""
    import numpy as np

    rows = []
    focal = 0.0
    for cam in cameras:
        W, H = float(cam.image_width), float(cam.image_height)
        fx = W / (2 * math.tan(cam.FoVx / 2))
        fy = H / (2 * math.tan(cam.FoVy / 2))
        R = np.asarray(cam.R, np.float64).astype(np.float32).reshape(9)
        T = np.asarray(cam.T, np.float64).astype(np.float32).reshape(3)
        rows.append(np.concatenate([R, T, np.array([fx, fy, W / 2.0, H / 2.0, -0.15 * W, W * 1.15, -0.15 * H,
                                                    1.15 * H], np.float32)]))
        focal = max(focal, fx, fy)
    table = np.stack(rows).astype(np.float32) if rows else np.zeros((0, 21), np.float32)
    return table, np.float32(1.0) / np.float32(focal) if focal > 0 else np.float32(np.inf)


@torch.no_grad()
def filter_3d(positions: torch.Tensor, cameras) -> torch.Tensor:
    """(n, 1) log filter of compute_3D_filter for (n, 3) positions."""
    p = _dev_f32(positions, "filter_3d")
    n = p.shape[0]
    table, inv_focal = filter_cameras(cameras)
    if len(table) == 0:
        raise ValueError("compute_3D_filter: no cameras")
    cams = torch.from_numpy(table).to(p.device)
    out = torch.empty(n, dtype=torch.float32, device=p.device)
    ws = torch.empty(256, dtype=torch.uint8, device=p.device)
    lib = _lib.load()
    _check(lib.rg_filter_3d(C.c_void_p(p.data_ptr()), n, C.c_void_p(cams.data_ptr()), len(table), float(inv_focal),
                            C.c_void_p(out.data_ptr()), C.c_void_p(ws.data_ptr()), 256, _stream(p.device)),
           "rg_filter_3d")
    if n > 0 and int(ws[:4].view(torch.int32).item()) == 0:
        # the reference's distance[valid_points].max() of an empty selection raises
        raise RuntimeError("compute_3D_filter: no point is seen by any camera")
    return out.view(n, 1)


def compute_3D_filter(pc, cameras):
    """PointCloud.compute_3D_filter: sets pc.filter_3D."""
    pc.filter_3D = filter_3d(pc.positions, cameras)


@torch.no_grad()
def accumulate_gradient(pc, grad: torch.Tensor, not_visible: bool = True):
    """pc.xyz_gradient_accum_norm += |grad|, pc.num_accum += (|grad| > 0) and, with not_visible,
    pc.num_accum_gnv += (|grad| > 0) (scripts/train.py:118-119)."""
    g = _dev_f32(grad, "accumulate_gradient")
    n = g.shape[0]
    for name in ("xyz_gradient_accum_norm", "num_accum") + (("num_accum_gnv",) if not_visible else ()):
        t = getattr(pc, name)
        if t.dtype != torch.float32 or not t.is_contiguous() or t.numel() != n:
            raise TypeError(f"accumulate_gradient: {name} must be a contiguous float32 tensor of {n}")
    gnv = pc.num_accum_gnv.data_ptr() if not_visible else None
    _check(_lib.load().rg_accumulate_grad_norm(C.c_void_p(g.data_ptr()), n,
                                               C.c_void_p(pc.xyz_gradient_accum_norm.data_ptr()),
                                               C.c_void_p(pc.num_accum.data_ptr()), C.c_void_p(gnv),
                                               _stream(g.device)), "rg_accumulate_grad_norm")


def _optimizer_of(pc, name):
    mgr = getattr(pc, "optim_managers", None)
    if mgr is None:
        return None
    d = getattr(mgr, "optimizer_manager_dict", None)
    if d is None or name not in d:
        return None
    return d[name].optimizer


@torch.no_grad()
def densify_and_prune(pc, max_grad, min_density, extent, N: int = 2, generator=None, quiet=True):
    """PointCloud.densify_and_prune (classes/point_cloud.py:1128-1143) on the GPU; returns the
    counts (kept originals, kept clones, split, kept split)."""
    lib = _lib.load()
    pos = _dev_f32(pc.positions, "positions")
    dev = pos.device
    n = pos.shape[0]
    dens = pc.densities
    din = _lib.DensifyIn()
    din.n = n
    keep = []

    def ptr(t, name, numel):
        t = _dev_f32(t, name)
        if t.numel() != numel:
            raise ValueError(f"densify_and_prune: {name} has {t.numel()} values, expected {numel}")
        keep.append(t)
        return t.data_ptr()

    din.grad_accum = ptr(pc.xyz_gradient_accum_norm, "xyz_gradient_accum_norm", n)
    din.num_accum = ptr(pc.num_accum, "num_accum", n)
    din.positions = ptr(pc.positions, "positions", 3 * n)
    din.log_scales = ptr(pc.scales, "scales", 3 * n)
    din.quats = ptr(pc.quaternions, "quaternions", 4 * n)
    din.densities = ptr(dens, "densities", n)
    din.grad_threshold = float(max_grad)
    din.scale_limit = float(0.01 * extent)
    din.min_density = float(min_density)
    din.split_n = int(N)
    wsb = int(lib.rg_densify_workspace_bytes(n))
    if wsb <= 0:
        raise ValueError(f"densify_and_prune: {n} Gaussians is beyond the supported size")
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    counts_dev = torch.zeros(4, dtype=torch.int64, device=dev)
    s = _stream(dev)
    _check(lib.rg_densify_plan(C.byref(din), C.c_void_p(ws.data_ptr()), wsb, C.c_void_p(counts_dev.data_ptr()), s),
           "rg_densify_plan")
    ko, kc, ns, kch = (int(v) for v in counts_dev.cpu().tolist())
    n_out = ko + kc + N * kch
    # the reference's torch.normal(mean=zeros, std=stds) of shape (N * ns, 3) draws these normals
    normals = torch.normal(torch.zeros((N * ns, 3), device=dev), torch.ones((N * ns, 3), device=dev),
                           generator=generator) if ns > 0 else None

    arrays = []
    new_params = {}
    new_states = {}
    for oname, attr in PARAMS:
        t = _dev_f32(getattr(pc, attr), attr)
        keep.append(t)
        w = t.numel() // n if n else 1
        out = torch.empty((n_out,) + tuple(t.shape[1:]), dtype=torch.float32, device=dev)
        new_params[oname] = out
        arrays.append((t, out, w, 0))
        opt = _optimizer_of(pc, oname)
        st = opt.state.get(opt.param_groups[0]["params"][0], None) if opt is not None else None
        if st is not None and "exp_avg" in st:
            m = {}
            for k in ("exp_avg", "exp_avg_sq"):
                src = st[k].detach().contiguous()
                keep.append(src)
                m[k] = torch.empty_like(out)
                arrays.append((src, m[k], w, 1))
            new_states[oname] = m
    for attr, zero in (("filter_3D", 0), ("num_accum_gnv", 1)):
        t = getattr(pc, attr, None)
        if t is not None:
            t = _dev_f32(t, attr)
            keep.append(t)
            out = torch.empty((n_out,) + tuple(t.shape[1:]), dtype=torch.float32, device=dev)
            arrays.append((t, out, t.numel() // n if n else 1, zero))
            new_params["__" + attr] = out
    table = (_lib.DensifyArray * len(arrays))()
    for k, (src, dst, w, z) in enumerate(arrays):
        table[k].src, table[k].dst, table[k].width, table[k].zero_new = src.data_ptr(), dst.data_ptr(), int(w), z
    counts_host = (C.c_int64 * 4)(ko, kc, ns, kch)
    _check(lib.rg_densify_apply(C.byref(din), C.c_void_p(ws.data_ptr()), wsb, counts_host, table, len(arrays),
                                C.c_void_p(normals.data_ptr() if normals is not None else None),
                                C.c_void_p(new_params["xyz"].data_ptr()), C.c_void_p(new_params["scales"].data_ptr()),
                                C.c_void_p(new_params["density"].data_ptr()), s), "rg_densify_apply")

    # install the new tensors, moving optimiser state as _prune_optimizer / _cat_tensor do
    for oname, attr in PARAMS:
        newp = new_params[oname].requires_grad_(True)
        opt = _optimizer_of(pc, oname)
        if opt is not None:
            group = opt.param_groups[0]
            old = group["params"][0]
            st = opt.state.pop(old, None)
            group["params"][0] = newp
            if st is not None:
                if oname in new_states:
                    st["exp_avg"] = new_states[oname]["exp_avg"]
                    st["exp_avg_sq"] = new_states[oname]["exp_avg_sq"]
                opt.state[newp] = st
        setattr(pc, attr, newp)
    if "__filter_3D" in new_params:
        pc.filter_3D = new_params["__filter_3D"]
    if "__num_accum_gnv" in new_params:
        pc.num_accum_gnv = new_params["__num_accum_gnv"]
    pc.xyz_gradient_accum_norm = torch.zeros(n_out, dtype=torch.float32, device=dev)
    pc.num_accum = torch.zeros(n_out, dtype=torch.float32, device=dev)
    if not quiet:
        print(f"densify: {kc} cloned, {ns} split, {n + kc + N * ns - n_out - ns} pruned, {n_out} remaining")
    return ko, kc, ns, kch
