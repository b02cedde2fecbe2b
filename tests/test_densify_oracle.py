"""The densification oracle pinned by an independent k-d tree, closed forms and hand-worked
cases (classes/point_cloud.py:1067-1215, submodules/simple-knn/simple_knn.cu:186-222)."""

import math

import numpy as np
import pytest
from scipy.spatial import cKDTree

from oracle import densify_oracle as D


def _kdtree_mean3(p):
    d, _ = cKDTree(np.asarray(p, np.float64)).query(np.asarray(p, np.float64), k=4)
    return (d[:, 1:4] ** 2).mean(axis=1)


@pytest.mark.parametrize("n", [4, 5, 37, 1500, 6000])
def test_knn_brute_force_matches_kdtree(n):
    rng = np.random.default_rng(n)
    p = rng.normal(size=(n, 3)).astype(np.float32) * np.float32(3.0)
    p[: n // 10] = p[n // 10: 2 * (n // 10)]  # duplicates: distance 0 counts
    np.testing.assert_allclose(D.knn3_mean_dist(p), _kdtree_mean3(p), rtol=1e-12, atol=1e-300)


def test_knn_large_path_is_the_same_quantity():
    rng = np.random.default_rng(9)
    p = rng.uniform(-5, 5, (20001, 3)).astype(np.float32)
    small = D.knn3_mean_dist(p[:5000])
    np.testing.assert_allclose(small, _kdtree_mean3(p[:5000]), rtol=1e-12)
    big = D.knn3_mean_dist(p)
    assert big.shape == (20001,) and np.all(np.isfinite(big))


def test_knn_fewer_than_four_points():
    inf = np.float32(np.inf)
    assert D.knn3_mean_dist(np.zeros((1, 3), np.float32))[0] == inf
    two = np.array([[0, 0, 0], [1, 0, 0]], np.float32)
    assert np.all(D.knn3_mean_dist(two) == inf)
    three = np.array([[0, 0, 0], [1, 0, 0], [0, 2, 0]], np.float32)
    assert np.all(D.knn3_mean_dist(three) == np.float32(3.4028234663852886e38) / np.float32(3))
    # a regular grid: every interior point has its 3 nearest at distance 1
    g = np.stack(np.meshgrid(*[np.arange(5)] * 3, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)
    assert np.all(D.knn3_mean_dist(g) == 1.0)


def _cam(R, T, fov=math.pi / 2, w=100, h=100):
    return {"R": np.asarray(R, float), "T": np.asarray(T, float), "FoVx": fov, "FoVy": fov, "image_width": w,
            "image_height": h}


def test_filter_hand_worked():
    # identity camera, 90 degree field of view, 100x100: focal 50, centre 50
    pts = np.array([[0, 0, 2], [0, 0, 0.1], [100, 0, 1], [1, 0, 4]], np.float32)
    lf, seen, margin = D.filter_3d(pts, [_cam(np.eye(3), [0, 0, 0])])
    assert list(seen) == [True, False, False, True]
    # seen: depths 2 and 4; unseen points take the largest seen depth (4)
    np.testing.assert_allclose(np.exp(lf), [0.5 * 2 / 50, 0.5 * 4 / 50, 0.5 * 4 / 50, 0.5 * 4 / 50], rtol=1e-12)
    # a second camera 1 unit behind sees the points at depth +1 (x R + T with T = (0, 0, 1)), and
    # its wider view (focal 25) sets the focal length to max(50, 25) = 50
    lf2, seen2, _ = D.filter_3d(pts, [_cam(np.eye(3), [0, 0, 1], fov=2 * math.atan(2.0))])
    np.testing.assert_allclose(np.exp(lf2[0]), 0.5 * 3 / 25, rtol=1e-12)
    lf3, _, _ = D.filter_3d(pts, [_cam(np.eye(3), [0, 0, 0]), _cam(np.eye(3), [0, 0, 1], fov=2 * math.atan(2.0))])
    # min over cameras of the depth (2 from the first), focal max(50, 25)
    np.testing.assert_allclose(np.exp(lf3[0]), 0.5 * 2 / 50, rtol=1e-12)
    # x R: a camera turned 180 degrees about y sees -z
    Ry = np.diag([-1.0, 1.0, -1.0])
    _, seen4, _ = D.filter_3d(np.array([[0, 0, -3], [0, 0, 3]], np.float32), [_cam(Ry, [0, 0, 0])])
    assert list(seen4) == [True, False]
    with pytest.raises(RuntimeError):
        D.filter_3d(np.array([[0, 0, -3]], np.float32), [_cam(np.eye(3), [0, 0, 0])])


def test_densify_hand_worked():
    f = np.float32
    # G0 clone (large gradient, small scale), G1 split (large gradient, large scale), G2 kept,
    # G3 pruned (density 0.05), G4 never had a gradient (0/0 -> 0), G5 cloned and both copies
    # pruned (0.15 / 2 <= 0.1)
    n = 6
    xyz = np.array([[0, 0, 0], [1, 2, 3], [5, 5, 5], [6, 6, 6], [7, 7, 7], [8, 8, 8]], f)
    ls = np.log(np.array([[0.001] * 3, [0.5, 0.25, 0.125], [0.01] * 3, [0.01] * 3, [0.01] * 3, [0.001] * 3], f))
    q = np.tile(np.array([[2, 0, 0, 0]], f), (n, 1))  # identity after normalisation
    dens = np.array([1.0, 0.8, 0.5, 0.05, 0.5, 0.15], f)
    extra = np.arange(n, dtype=f)[:, None] * f(10)
    mom = np.full((n, 3), 7.0, f)
    accum = np.array([1.0, 1.0, 0.0, 0.0, 0.0, 1.0], f)
    num = np.array([1.0, 1.0, 1.0, 1.0, 0.0, 1.0], f)
    z = np.array([[1, -1, 2], [0.5, 0.5, -1]], f)  # child 0 and child 1 of the one split Gaussian
    arrs = {"xyz": xyz, "scales": ls, "quaternions": q, "density": dens, "extra": extra, "m": mom}
    out, counts = D.densify_and_prune(arrs, accum, num, max_grad=0.5, min_density=0.1, extent=10.0, normals=z, N=2,
                                      states=("m",))
    # survivors: G0 (halved), G2, G4, clone of G0, two children of G1
    assert counts == (3, 1, 1, 1)
    np.testing.assert_array_equal(out["extra"][:, 0], [0, 20, 40, 0, 10, 10])
    np.testing.assert_array_equal(out["density"], f([0.5, 0.5, 0.5, 0.5, 0.4, 0.4]))
    np.testing.assert_array_equal(out["m"][:, 0], [7, 7, 7, 0, 0, 0])
    s = np.array([0.5, 0.25, 0.125])
    np.testing.assert_allclose(out["xyz"][4], [1 + 0.5, 2 - 0.25, 3 + 0.25], rtol=1e-6)
    np.testing.assert_allclose(out["xyz"][5], [1 + 0.25, 2 + 0.125, 3 - 0.125], rtol=1e-6)
    np.testing.assert_allclose(np.exp(out["scales"][4]), s / 1.6, rtol=1e-6)
    np.testing.assert_allclose(out["xyz"][3], [0, 0, 0])
    # a rotation: 90 degrees about z maps the sample (s_x z_x, ...) to (-s_y z_y, s_x z_x, s_z z_z)
    qz = np.array([[math.cos(math.pi / 4), 0, 0, math.sin(math.pi / 4)]], f)
    arrs2 = {"xyz": xyz[1:2], "scales": ls[1:2], "quaternions": qz, "density": dens[1:2]}
    out2, _ = D.densify_and_prune(arrs2, accum[1:2], num[1:2], 0.5, 0.1, 10.0, z, N=2)
    np.testing.assert_allclose(out2["xyz"][0], [1 + 0.25, 2 + 0.5, 3 + 0.25], rtol=1e-6, atol=1e-6)
