"""TEST INFRASTRUCTURE ONLY.  CPU restatement of the densification and initialisation utilities.

Only tests/ may import this module, as a checker -- never as a product code path.

* ``knn3_mean_dist`` -- distCUDA2 of simple-knn (submodules/simple-knn/simple_knn.cu:186-222;
  a pinned git submodule of the reference, vendored under submodules/).  Its bucket search
  (boxMeanDist, :147-184) never skips a box that may hold one of the 3 nearest points, so the
  result is the exact mean of the 3 smallest squared distances to OTHER points (duplicates at
  distance 0 count); with fewer than 3 other points the missing ones are FLT_MAX and the float32
  sum (b0 + b1 + b2) / 3 overflows as there.  Here: brute force in float64 (a scipy k-d tree
  above 20000 points).
* ``filter_3d`` -- PointCloud.compute_3D_filter (classes/point_cloud.py:1152-1215) in float64.
* ``densify_and_prune`` -- classes/point_cloud.py:1067-1143 (densify_and_clone, then
  densify_and_split, then the density prune), with the split's normal draws given, over plain
  numpy arrays; returns the new arrays in the reference's row order.
"""

from __future__ import annotations

import math

import numpy as np

FLT_MAX = np.float32(3.4028234663852886e38)


def knn3_mean_dist(points):
    p = np.asarray(points, np.float32).astype(np.float64)
    n = p.shape[0]
    out = np.empty(n, np.float64)
    if n == 0:
        return out.astype(np.float32)
    if n > 20000:  # the same quantity from a k-d tree (scipy), pinned against brute force in tests
        from scipy.spatial import cKDTree

        d, _ = cKDTree(p).query(p, k=4)
        return (d[:, 1:4] ** 2).mean(axis=1)
    if n < 4:
        res = np.empty(n, np.float32)
        for i in range(n):
            d = sorted(float(np.sum((p[j] - p[i]) ** 2)) for j in range(n) if j != i)
            b = [np.float32(v) for v in d] + [FLT_MAX] * (3 - len(d))
            with np.errstate(over="ignore"):
                res[i] = (b[0] + b[1] + b[2]) / np.float32(3.0)
        return res
    chunk = max(1, 2_000_000 // n)
    sq = np.sum(p * p, axis=1)
    for a in range(0, n, chunk):
        q = p[a:a + chunk]
        d = sq[a:a + chunk, None] + sq[None, :] - 2.0 * q @ p.T
        d = np.maximum(d, 0.0)
        # exact recomputation of the candidates (the expansion above loses digits)
        kk = min(4, n - 1)
        idx = np.argpartition(d, kk, axis=1)[:, :kk + 1]
        for r in range(q.shape[0]):
            i = a + r
            cand = [j for j in idx[r] if j != i]
            dd = np.sort(np.sum((p[cand] - p[i]) ** 2, axis=1))
            out[i] = dd[:3].mean()
    return out


def filter_3d(points, cams):
    """cams: dicts with R (3,3), T (3), FoVx, FoVy, image_width, image_height.  Returns
    (log filter (n,), seen mask, margin) with the reference's selection tests in float64;
    margin = the smallest relative distance of any of a point's tests to its threshold (a float32
    evaluation may decide such a test either way when it is below ~1e-5)."""
    x = np.asarray(points, np.float32).astype(np.float64)
    n = x.shape[0]
    dist = np.full(n, 100000.0)
    seen = np.zeros(n, bool)
    margin = np.full(n, np.inf)
    focal = 0.0
    for c in cams:
        R = np.asarray(c["R"], np.float64)
        T = np.asarray(c["T"], np.float64)
        xc = x @ R + T[None, :]
        z = xc[:, 2]
        front = z > 0.2
        zz = np.maximum(z, 0.001)
        W, H = float(c["image_width"]), float(c["image_height"])
        fx = W / (2 * math.tan(c["FoVx"] / 2))
        fy = H / (2 * math.tan(c["FoVy"] / 2))
        sx = xc[:, 0] / zz * fx + W / 2.0
        sy = xc[:, 1] / zz * fy + H / 2.0
        ins = (sx >= -0.15 * W) & (sx <= W * 1.15) & (sy >= -0.15 * H) & (sy <= 1.15 * H)
        v = front & ins
        rel = lambda a, b: np.abs(a - b) / np.maximum(1.0, np.abs(b))  # noqa: E731
        margin = np.minimum.reduce([margin, rel(z, 0.2), rel(sx, -0.15 * W), rel(sx, W * 1.15), rel(sy, -0.15 * H),
                                    rel(sy, 1.15 * H)])
        dist[v] = np.minimum(dist[v], zz[v])
        seen |= v
        focal = max(focal, fx, fy)
    if not seen.any():
        raise RuntimeError("no point is seen")
    dist[~seen] = dist[seen].max()
    return np.log(0.5 * dist / focal), seen, margin


def build_rotation(q):
    """utils/general_utils.py:466-487: rotation of q / |q|, q = (w, x, y, z)."""
    q = np.asarray(q, np.float64)
    q = q / np.linalg.norm(q, axis=1)[:, None]
    r, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.empty((q.shape[0], 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - r * z)
    R[:, 0, 2] = 2 * (x * z + r * y)
    R[:, 1, 0] = 2 * (x * y + r * z)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - r * x)
    R[:, 2, 0] = 2 * (x * z - r * y)
    R[:, 2, 1] = 2 * (y * z + r * x)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def densify_and_prune(arrs, grad_accum, num_accum, max_grad, min_density, extent, normals, N=2, states=()):
    """arrs: dict name -> (n, ...) float32 arrays with keys xyz, scales (log), quaternions,
    density and any others (copied); states: names of arrays that are Adam moments (new rows
    zero), e.g. ("exp_avg:xyz", ...).  normals: (N * n_split, 3) standard normal draws.
    Returns (new dict, (kept originals, kept clones, n split, kept split))."""
    f32 = np.float32
    with np.errstate(invalid="ignore", divide="ignore"):
        g = (np.asarray(grad_accum, f32) / np.asarray(num_accum, f32)).astype(f32)
    g[np.isnan(g)] = 0
    scale = np.exp(arrs["scales"].astype(np.float64))
    smax = scale.max(axis=1)
    lim = f32(0.01 * extent)
    n = g.shape[0]
    # densify_and_clone (:1100-1126)
    clone = (g >= f32(max_grad)) & (smax <= lim)
    dens = arrs["density"].astype(f32).reshape(n).copy()
    dens[clone] = dens[clone] / f32(2)
    cur = {k: v.copy() for k, v in arrs.items()}
    cur["density"] = dens.reshape(arrs["density"].shape)
    for k in cur:
        add = np.zeros_like(cur[k][clone]) if k in states else cur[k][clone]
        cur[k] = np.concatenate([cur[k], add], axis=0)
    # densify_and_split (:1067-1098) on the grown arrays; clones have no gradient
    n1 = cur["xyz"].shape[0]
    g1 = np.zeros(n1, f32)
    g1[:n] = g
    smax1 = np.exp(cur["scales"].astype(np.float64)).max(axis=1)
    split = (g1 >= f32(max_grad)) & (smax1 > lim)
    ns = int(split.sum())
    s = np.exp(cur["scales"][split].astype(np.float64))
    stds = np.tile(s, (N, 1))
    samples = np.asarray(normals, np.float64).reshape(N * ns, 3) * stds
    rots = np.tile(build_rotation(cur["quaternions"][split]), (N, 1, 1))
    new = {}
    for k in cur:
        rows = cur[k][split]
        new[k] = np.zeros_like(np.tile(rows, (N,) + (1,) * (rows.ndim - 1))) if k in states else \
            np.tile(rows, (N,) + (1,) * (rows.ndim - 1))
    new["xyz"] = (np.einsum("kij,kj->ki", rots, samples) + np.tile(cur["xyz"][split].astype(np.float64),
                                                                     (N, 1))).astype(f32)
    new["scales"] = np.log(stds / (0.8 * N)).astype(f32)
    new["density"] = (np.tile(cur["density"][split].reshape(-1), N) / f32(2)).astype(f32).reshape(
        (-1,) + cur["density"].shape[1:])
    for k in cur:
        cur[k] = np.concatenate([cur[k], new[k]], axis=0)
    keep = np.concatenate([~split, np.ones(N * ns, bool)])
    for k in cur:
        cur[k] = cur[k][keep]
    # density prune (:1137)
    keep2 = ~(cur["density"].reshape(-1) <= f32(min_density))
    for k in cur:
        cur[k] = cur[k][keep2]
    n_orig_kept = int((keep2[: n - ns]).sum())
    n_clone_kept = int(keep2[n - ns: n - ns + int(clone.sum())].sum())
    n_child_kept = int(keep2[n - ns + int(clone.sum()):].sum())
    return cur, (n_orig_kept, n_clone_kept, ns, n_child_kept // N if N else 0)
