"""Scene formats: the RayGauss PLY layout (convertpth_to_ply.py:60-116 writer,
classes/point_cloud.py:208-296 reader), checkpoints (classes/point_cloud.py:922-939) and the
activations of get_data (classes/point_cloud.py:689-739)."""

import sys

import numpy as np
import pytest
import torch

from ray_gauss_amd import formats as F


def _reference_layout_file(path, n=3, K=4, L=2):
    """A PLY laid out as the reference's converter writes it, built by hand: header text
    'format binary_<byteorder>_endian 1.0', 'property float32 <name>' lines in the converter's
    order, then packed float32 records.  Every value is distinct and encodes (field, row)."""
    names = ["x", "y", "z"]
    for c in "rgb":
        names += [f"sh{k}{c}" for k in range(K)]
    names += ["density", "sx", "sy", "sz", "qx", "qy", "qz", "qw"]
    for c in "rgb":
        names += [f"sg{l}{c}" for l in range(L)]
    names += [f"bandwidth_sharpness{l}" for l in range(L)]
    for l in range(L):
        names += [f"lobe_{l}x", f"lobe_{l}y", f"lobe_{l}z"]
    names += ["gradient_accum_norm"]
    vals = np.array([[1000.0 * (j + 1) + i for j in range(len(names))] for i in range(n)], np.float32)
    head = ["ply", f"format binary_{sys.byteorder}_endian 1.0", f"element vertex {n}"]
    head += [f"property float32 {k}" for k in names] + ["end_header"]
    with open(path, "wb") as f:
        f.write(("\n".join(head) + "\n").encode())
        f.write(vals.astype(np.float32).tobytes())
    return names, vals


def test_ply_reference_layout(tmp_path):
    p = tmp_path / "point_cloud_iter1.ply"
    names, vals = _reference_layout_file(p)
    col = {k: vals[:, j] for j, k in enumerate(names)}
    g = F.load_rg_ply(p)
    assert g.n == 3 and g.sh_degree == 1 and g.num_sg == 2
    np.testing.assert_array_equal(g.positions, np.stack([col["x"], col["y"], col["z"]], 1))
    for ci, c in enumerate("rgb"):
        np.testing.assert_array_equal(g.rgb[:, ci], col[f"sh0{c}"])
        for k in range(1, 4):
            np.testing.assert_array_equal(g.sh_rest[:, ci, k - 1], col[f"sh{k}{c}"])
        for l in range(2):
            np.testing.assert_array_equal(g.sg[:, ci, l], col[f"sg{l}{c}"])
    np.testing.assert_array_equal(g.densities, col["density"])
    np.testing.assert_array_equal(g.log_scales, np.stack([col["sx"], col["sy"], col["sz"]], 1))
    np.testing.assert_array_equal(g.quats, np.stack([col["qx"], col["qy"], col["qz"], col["qw"]], 1))
    for l in range(2):
        np.testing.assert_array_equal(g.sharpness_raw[:, l], col[f"bandwidth_sharpness{l}"])
        for ai, a in enumerate("xyz"):
            np.testing.assert_array_equal(g.lobe_axis[:, l, ai], col[f"lobe_{l}{a}"])


def test_ply_round_trip_and_writer_order(tmp_path):
    rng = np.random.default_rng(3)
    n, K, L = 50, 16, 3
    g = F.RawGaussians(positions=rng.normal(size=(n, 3)).astype(np.float32),
                       rgb=rng.normal(size=(n, 3)).astype(np.float32),
                       sh_rest=rng.normal(size=(n, 3, K - 1)).astype(np.float32),
                       densities=rng.uniform(0, 50, n).astype(np.float32),
                       log_scales=rng.normal(-3, 0.5, (n, 3)).astype(np.float32),
                       quats=rng.normal(size=(n, 4)).astype(np.float32),
                       sg=rng.normal(size=(n, 3, L)).astype(np.float32),
                       sharpness_raw=rng.normal(size=(n, L)).astype(np.float32),
                       lobe_axis=rng.normal(size=(n, L, 3)).astype(np.float32))
    acc = rng.uniform(size=n).astype(np.float32)
    p = tmp_path / "s.ply"
    F.save_rg_ply(p, g, gradient_accum_norm=acc)
    ref_names, _ = _reference_layout_file(tmp_path / "ref.ply", n=1, K=K, L=L)
    assert list(F.read_ply(p).dtype.names) == ref_names
    h = F.load_rg_ply(p)
    for k in ("positions", "rgb", "sh_rest", "densities", "log_scales", "quats", "sg", "sharpness_raw", "lobe_axis"):
        np.testing.assert_array_equal(getattr(h, k), getattr(g, k), err_msg=k)
    np.testing.assert_array_equal(F.read_ply(p)["gradient_accum_norm"], acc)


def test_ply_errors(tmp_path):
    p = tmp_path / "a.ply"
    p.write_bytes(b"ply\nformat ascii 1.0\nelement vertex 0\nproperty float x\nend_header\n")
    with pytest.raises(ValueError, match="binary"):
        F.load_rg_ply(p)
    q = tmp_path / "b.ply"
    F.write_ply(q, {"x": np.zeros(2, np.float32), "y": np.zeros(2, np.float32), "z": np.zeros(2, np.float32)})
    with pytest.raises(ValueError, match="missing"):
        F.load_rg_ply(q)
    r = tmp_path / "c.txt"
    r.write_bytes(b"not a ply\n")
    with pytest.raises(ValueError, match="not a PLY"):
        F.read_ply(r)


def test_ply_big_endian(tmp_path):
    p = tmp_path / "be.ply"
    head = b"ply\nformat binary_big_endian 1.0\nelement vertex 2\nproperty float x\nproperty double y\nend_header\n"
    body = np.array([(1.5, 2.25), (-3.0, 4.0)], dtype=[("x", ">f4"), ("y", ">f8")]).tobytes()
    p.write_bytes(head + body)
    v = F.read_ply(p)
    np.testing.assert_array_equal(v["x"], [1.5, -3.0])
    np.testing.assert_array_equal(v["y"], [2.25, 4.0])


def test_checkpoint_layout(tmp_path):
    """A checkpoint laid out as save_model writes it (15-tuple), loaded without unpickling code."""
    rng = np.random.default_rng(4)
    n, K, L = 20, 9, 2
    t = lambda *s: torch.from_numpy(rng.normal(size=s).astype(np.float32))  # noqa: E731
    data = (K, t(n, 3), t(n, 3), t(n, 3, K - 1), t(n, 1), t(n, 3), t(n, 4), t(n, 3, L), t(n, L), t(n, L, 3), L,
            t(n), t(n), [{"state": {0: {"step": torch.tensor(7.0), "exp_avg": t(n, 3)}},
                          "param_groups": [{"lr": 1e-3, "params": [0]}]}], 1234)
    p = tmp_path / "chkpnt1234.pth"
    torch.save(data, p)
    g, it, opt = F.load_checkpoint(p)
    assert it == 1234 and g.n == n and g.sh_degree == 2 and g.num_sg == L
    np.testing.assert_array_equal(g.positions, data[1].numpy())
    np.testing.assert_array_equal(g.densities, data[4].numpy().reshape(-1))
    np.testing.assert_array_equal(g.lobe_axis, data[9].numpy())
    assert opt[0]["param_groups"][0]["lr"] == 1e-3
    torch.save((1, 2), tmp_path / "bad.pth")
    with pytest.raises(ValueError, match="15-tuple"):
        F.load_checkpoint(tmp_path / "bad.pth")


def test_activate_matches_get_data():
    rng = np.random.default_rng(5)
    n, L = 10, 2
    g = F.RawGaussians(positions=rng.normal(size=(n, 3)).astype(np.float32), rgb=np.zeros((n, 3), np.float32),
                       sh_rest=np.ones((n, 3, 3), np.float32), densities=np.ones(n, np.float32),
                       log_scales=np.log(np.full((n, 3), 0.25, np.float32)),
                       quats=np.tile(np.array([[0, 0, 0, 2]], np.float32), (n, 1)),
                       sg=np.zeros((n, 3, L), np.float32), sharpness_raw=np.zeros((n, L), np.float32),
                       lobe_axis=np.tile(np.array([[[3, 0, 4]]], np.float32), (n, L, 1)))
    a = F.activate(g, device="cpu")
    np.testing.assert_allclose(a["scales"].numpy(), 0.25, rtol=1e-6)
    np.testing.assert_allclose(a["quats"].numpy(), np.tile([[0, 0, 0, 1]], (n, 1)), atol=1e-7)
    np.testing.assert_allclose(a["sharpness"].numpy(), np.log(2.0), rtol=1e-6)  # softplus(0)
    np.testing.assert_allclose(a["lobe_axis"].numpy()[0, 0], [0.6, 0.0, 0.8], rtol=1e-6)
    assert a["sh"].shape == (n, 3, 4) and float(a["sh"][0, 0, 0]) == 0.0 and float(a["sh"][0, 0, 1]) == 1.0
