"""Scene formats of the reference, to load trained scenes into the caster.

* RayGauss PLY (written by convertpth_to_ply.py:60-116, read by classes/point_cloud.py:208-296
  ``load_from_rg_ply``): one binary vertex element with float properties
  x y z, sh{k}r... sh{k}g... sh{k}b... (channel-major), density, sx sy sz (log-scales),
  qx qy qz qw (the four quaternion columns in storage order), sg{l}r/g/b (channel-major),
  bandwidth_sharpness{l} (pre-softplus), lobe_{l}x/y/z, and optionally gradient_accum_norm.
* Checkpoints ``chkpnt{iter}.pth`` (classes/point_cloud.py:922-939 ``save_model``): a 15-tuple
  (harmonic_number, positions, rgb, spherical_harmonics, densities, scales, quaternions,
  sph_gauss_features, bandwidth_sharpness, lobe_axis, num_sph_gauss, xyz_gradient_accum_norm,
  num_accum, optimiser state dicts, iteration), read with ``torch.load(weights_only=True)``.

Both give the RAW parameters; ``activate`` applies the reference's activations
(classes/point_cloud.py:689-739 ``get_data``) to produce the caster's boundary inputs.
"""

from __future__ import annotations

import sys
from dataclasses import dataclass

import numpy as np

_PLY_TYPES = {"int8": "i1", "char": "i1", "uint8": "u1", "uchar": "u1", "int16": "i2", "short": "i2",
              "uint16": "u2", "ushort": "u2", "int32": "i4", "int": "i4", "uint32": "u4", "uint": "u4",
              "float32": "f4", "float": "f4", "float64": "f8", "double": "f8"}


@dataclass
class RawGaussians:
    """Raw (optimised) parameters, shapes as in classes/point_cloud.py."""

    positions: np.ndarray      # (n, 3)
    rgb: np.ndarray            # (n, 3)   SH degree-0 coefficients
    sh_rest: np.ndarray        # (n, 3, K-1)
    densities: np.ndarray      # (n,)     identity activation
    log_scales: np.ndarray     # (n, 3)   exp activation
    quats: np.ndarray          # (n, 4)   normalised on use
    sg: np.ndarray             # (n, 3, L)
    sharpness_raw: np.ndarray  # (n, L)   softplus activation
    lobe_axis: np.ndarray      # (n, L, 3) normalised on use

    @property
    def n(self):
        return self.positions.shape[0]

    @property
    def sh_degree(self):
        return int(round(np.sqrt(self.sh_rest.shape[2] + 1))) - 1

    @property
    def num_sg(self):
        return self.sg.shape[2]


def read_ply(path):
    """Vertex records of a binary PLY as a numpy structured array."""
    with open(path, "rb") as f:
        if f.readline().strip() != b"ply":
            raise ValueError(f"{path}: not a PLY file")
        fmt, count, props, in_vertex = None, None, [], False
        while True:
            line = f.readline()
            if not line:
                raise ValueError(f"{path}: truncated header")
            tok = line.split()
            if not tok:
                continue
            if tok[0] == b"format":
                fmt = tok[1].decode()
            elif tok[0] == b"element":
                in_vertex = tok[1] == b"vertex"
                if in_vertex:
                    count = int(tok[2])
            elif tok[0] == b"property" and in_vertex:
                if tok[1] == b"list":
                    raise ValueError(f"{path}: list properties are not supported")
                props.append((tok[2].decode(), _PLY_TYPES[tok[1].decode()]))
            elif tok[0] == b"end_header":
                break
        if fmt not in ("binary_little_endian", "binary_big_endian"):
            raise ValueError(f"{path}: only binary PLY is supported (got {fmt})")
        end = "<" if fmt == "binary_little_endian" else ">"
        dtype = np.dtype([(name, end + t) for name, t in props])
        return np.fromfile(f, dtype=dtype, count=count)


def write_ply(path, columns):
    """Binary PLY (the machine's byte order) of named 1-D columns, in the given order."""
    names = list(columns)
    n = len(columns[names[0]])
    dtype = np.dtype([(k, np.asarray(columns[k]).dtype.str) for k in names])
    rec = np.empty(n, dtype=dtype)
    for k in names:
        rec[k] = columns[k]
    with open(path, "wb") as f:
        f.write(b"ply\n")
        f.write(f"format binary_{sys.byteorder}_endian 1.0\n".encode())
        f.write(f"element vertex {n}\n".encode())
        for k in names:
            f.write(f"property {rec.dtype[k].name} {k}\n".encode())
        f.write(b"end_header\n")
        rec.tofile(f)


def _count(fields, fmt):
    k = 0
    while all(fmt(k, c) in fields for c in "rgb"):
        k += 1
    return k


def load_rg_ply(path) -> RawGaussians:
    """classes/point_cloud.py:208-296 (load_from_rg_ply)."""
    v = read_ply(path)
    fields = set(v.dtype.names)
    need = {"x", "y", "z", "density", "sx", "sy", "sz", "qx", "qy", "qz", "qw"}
    if not need <= fields:
        raise ValueError(f"{path}: missing fields {sorted(need - fields)}")
    K = _count(fields, lambda k, c: f"sh{k}{c}")
    L = _count(fields, lambda k, c: f"sg{k}{c}")
    if K == 0:
        raise ValueError(f"{path}: no SH colour fields")
    f32 = np.float32
    sh = np.stack([np.stack([v[f"sh{k}{c}"] for k in range(K)], axis=1) for c in "rgb"], axis=1).astype(f32)
    sg = np.zeros((len(v), 3, L), f32)
    sharp = np.zeros((len(v), L), f32)
    axis = np.zeros((len(v), L, 3), f32)
    for l in range(L):
        for ci, c in enumerate("rgb"):
            sg[:, ci, l] = v[f"sg{l}{c}"]
        sharp[:, l] = v[f"bandwidth_sharpness{l}"]
        for ai, a in enumerate("xyz"):
            axis[:, l, ai] = v[f"lobe_{l}{a}"]
    return RawGaussians(
        positions=np.stack([v["x"], v["y"], v["z"]], axis=1).astype(f32), rgb=sh[:, :, 0].copy(),
        sh_rest=sh[:, :, 1:].copy(), densities=v["density"].astype(f32),
        log_scales=np.stack([v["sx"], v["sy"], v["sz"]], axis=1).astype(f32),
        quats=np.stack([v["qx"], v["qy"], v["qz"], v["qw"]], axis=1).astype(f32), sg=sg, sharpness_raw=sharp,
        lobe_axis=axis)


def save_rg_ply(path, g: RawGaussians, gradient_accum_norm=None):
    """The field order of convertpth_to_ply.py:60-116."""
    cols = {"x": g.positions[:, 0], "y": g.positions[:, 1], "z": g.positions[:, 2]}
    sh = np.concatenate([g.rgb[:, :, None], g.sh_rest], axis=2)
    for ci, c in enumerate("rgb"):
        for k in range(sh.shape[2]):
            cols[f"sh{k}{c}"] = sh[:, ci, k]
    cols["density"] = g.densities.reshape(-1)
    for ai, a in enumerate(("sx", "sy", "sz")):
        cols[a] = g.log_scales[:, ai]
    for ai, a in enumerate(("qx", "qy", "qz", "qw")):
        cols[a] = g.quats[:, ai]
    for ci, c in enumerate("rgb"):
        for l in range(g.num_sg):
            cols[f"sg{l}{c}"] = g.sg[:, ci, l]
    for l in range(g.num_sg):
        cols[f"bandwidth_sharpness{l}"] = g.sharpness_raw[:, l]
    for l in range(g.num_sg):
        for ai, a in enumerate("xyz"):
            cols[f"lobe_{l}{a}"] = g.lobe_axis[:, l, ai]
    if gradient_accum_norm is not None:
        cols["gradient_accum_norm"] = np.asarray(gradient_accum_norm).reshape(-1)
    write_ply(path, {k: np.ascontiguousarray(v, np.float32) for k, v in cols.items()})


def load_checkpoint(path):
    """(RawGaussians, iteration, optimiser state dicts) of a chkpnt{iter}.pth (save_model layout);
    tensors are loaded with weights_only=True, so nothing in the file is executed."""
    import torch

    d = torch.load(path, map_location="cpu", weights_only=True)
    if not isinstance(d, (tuple, list)) or len(d) != 15:
        raise ValueError(f"{path}: not a RayGauss checkpoint (expected a 15-tuple)")
    a = lambda t: t.detach().float().cpu().numpy()  # noqa: E731
    g = RawGaussians(positions=a(d[1]), rgb=a(d[2]), sh_rest=a(d[3]), densities=a(d[4]).reshape(-1),
                     log_scales=a(d[5]), quats=a(d[6]), sg=a(d[7]), sharpness_raw=a(d[8]), lobe_axis=a(d[9]))
    return g, int(d[14]), d[13]


def activate(g: RawGaussians, device="cuda"):
    """Boundary inputs of the caster from raw parameters (classes/point_cloud.py:689-739
    get_data): positions, exp(log s), q / |q|, densities, [rgb | SH], SG, softplus(lambda),
    axis / |axis| -- as float32 tensors on `device`."""
    import torch

    t = lambda x: torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(device)  # noqa: E731
    q = t(g.quats)
    ax = t(g.lobe_axis)
    return {"positions": t(g.positions), "scales": torch.exp(t(g.log_scales)),
            "quats": (q / torch.norm(q, dim=1)[:, None]).contiguous(), "densities": t(g.densities),
            "sh": torch.cat([t(g.rgb)[:, :, None], t(g.sh_rest)], dim=2).contiguous(), "sg": t(g.sg),
            "sharpness": torch.nn.functional.softplus(t(g.sharpness_raw)),
            "lobe_axis": (ax / torch.norm(ax, dim=2)[:, :, None]).contiguous() if g.num_sg else ax}
