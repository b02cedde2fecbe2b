"""Seeded synthetic Gaussian scenes and views for the ray caster's workloads.

Parameters are drawn in float64 with numpy and cast once to float32, so the identical float32
arrays can be handed to the HIP path and to any checker.  Every scene is returned in the
activated form the ray caster's boundary takes (reference ``classes/point_cloud.py:729-739``):

* ``scales`` = exp(log-scales);
* ``quats`` unit quaternions in (w, x, y, z) order;
* ``densities`` = softplus(density logits) -- the reference passes densities through unchanged
  (``point_cloud.py:708-710``) and keeps softplus as its commented-out choice (``:709``); the
  workloads describe densities as logits, so softplus is applied here, before the boundary;
* ``sh`` (N, 3, 16): DC drawn as stated, higher orders small;
* ``sg`` (N, 3, L) amplitudes, ``sharpness`` (N, L) already positive, ``lobe_axis`` (N, L, 3) unit.

Workload names follow the sizes they describe:

* ``small4k``   -- 4,096 Gaussians uniform in [-1,1]^3, one 64x64 view at radius 3 (black bg,
  dt 0.01, near/far [0.5, 6], T_eps 1e-4);
* ``blobs100k`` -- 100k Gaussians in 32 anisotropic blobs, SH3 + 4 SG, one 800x800 view at
  radius 4, white background;
* ``blobs300k_16views`` -- 300k Gaussians in 96 blobs, 16 views of uneven size (width in
  [400, 1600], 4:3), each a random 25-100 % subset of its pixels;
* ``unbounded1m_8views`` -- 1M Gaussians, 60 % in an inner blob mixture of radius 1 and 40 % on a
  shell with radius log-uniform in [2, 50] and scales proportional to radius; SH3 + 4 SG; 8 views
  of 1237x822 on a radius-3.5 orbit;
* ``unbounded3m_4views`` -- 3M Gaussians, same generator, SH3 + 8 SG, 4 views of 2048x1536.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .cameras import Camera, look_at_camera, orbit_eye


@dataclass
class Scene:
    positions: np.ndarray  # (N,3) f32
    scales: np.ndarray  # (N,3) f32 (activated)
    quats: np.ndarray  # (N,4) f32 w,x,y,z
    densities: np.ndarray  # (N,) f32 (activated)
    sh: np.ndarray  # (N,3,K) f32
    sg: np.ndarray  # (N,3,L) f32
    sharpness: np.ndarray  # (N,L) f32
    lobe_axis: np.ndarray  # (N,L,3) f32

    @property
    def n(self) -> int:
        return int(self.positions.shape[0])

    @property
    def num_sg(self) -> int:
        return int(self.sg.shape[2])

    @property
    def sh_degree(self) -> int:
        return int(round(math.sqrt(self.sh.shape[2]))) - 1

    def subset(self, idx) -> "Scene":
        return Scene(*(np.ascontiguousarray(a[idx]) for a in (self.positions, self.scales, self.quats,
                                                                 self.densities, self.sh, self.sg,
                                                                 self.sharpness, self.lobe_axis)))


@dataclass
class RenderParams:
    """Runtime replacements of the reference's compile-time constants (gaussians_aabb.h:2-5), plus
    the per-call sampling settings of RenderOptixSettings (render_optix.py:67-87): jitter,
    rnd_sample and the iteration that seeds the per-ray random generator."""

    dt: float = 0.0025
    samples_per_slab: int = 8
    t_eps: float = 0.003
    sigma_thr: float = 0.1
    t_near: float = 0.0
    t_far: float = float("inf")
    white_background: bool = False
    jitter: int = 0
    rnd_sample: int = 0
    iteration: int = 0


@dataclass
class Workload:
    name: str
    scene: Scene
    cameras: list
    params: RenderParams
    pixels: list = field(default_factory=list)  # per view: int32 pixel indices or None (all pixels)

    @property
    def num_rays(self) -> int:
        tot = 0
        for cam, px in zip(self.cameras, self.pixels or [None] * len(self.cameras)):
            tot += cam.image_width * cam.image_height if px is None else len(px)
        return tot


def _softplus(x):
    return np.where(x > 20, x, np.log1p(np.exp(np.minimum(x, 20))))


def _unit_quats(rng, n):
    q = rng.standard_normal((n, 4))
    return q / np.linalg.norm(q, axis=1, keepdims=True)


def _quat_to_R(q):
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    R = np.empty((q.shape[0], 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - w * z)
    R[:, 0, 2] = 2 * (x * z + w * y)
    R[:, 1, 0] = 2 * (x * y + w * z)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - w * x)
    R[:, 2, 0] = 2 * (x * z - w * y)
    R[:, 2, 1] = 2 * (y * z + w * x)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def _unit_vectors(rng, shape):
    v = rng.standard_normal(shape + (3,))
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _appearance(rng, n, num_sg, degree=3):
    K = (degree + 1) ** 2
    sh = np.empty((n, 3, K))
    sh[:, :, 0] = rng.normal(0.5, 0.2, (n, 3))
    sh[:, :, 1:] = rng.normal(0.0, 0.05, (n, 3, K - 1))
    sg = rng.normal(0.0, 0.1, (n, 3, num_sg))
    sharp = np.exp(rng.uniform(math.log(1.0), math.log(50.0), (n, num_sg)))
    axis = _unit_vectors(rng, (n, num_sg))
    return sh, sg, sharp, axis


def _blob_mixture(rng, n, num_blobs, centres):
    """Means from a mixture of anisotropic Gaussian blobs (per-axis std U[0.05,0.2], random
    orientation)."""
    std = rng.uniform(0.05, 0.2, (num_blobs, 3))
    rot = _quat_to_R(_unit_quats(rng, num_blobs))
    which = rng.integers(0, num_blobs, n)
    local = rng.standard_normal((n, 3)) * std[which]
    return centres[which] + np.einsum("nij,nj->ni", rot[which], local)


def _finish(pos, log_s, rng, num_sg, degree=3) -> Scene:
    n = pos.shape[0]
    quats = _unit_quats(rng, n)
    dens = _softplus(rng.standard_normal(n))
    sh, sg, sharp, axis = _appearance(rng, n, num_sg, degree)
    f = lambda a: np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
    return Scene(f(pos), f(np.exp(log_s)), f(quats), f(dens), f(sh), f(sg), f(sharp), f(axis))


def scene_uniform(n=4096, seed=0, log_scale_mean=-3.5, log_scale_std=0.3, num_sg=0) -> Scene:
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1.0, 1.0, (n, 3))
    log_s = rng.normal(log_scale_mean, log_scale_std, (n, 3))
    return _finish(pos, log_s, rng, num_sg)


def scene_blobs(n=100_000, num_blobs=32, seed=0, num_sg=4) -> Scene:
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-1.0, 1.0, (num_blobs, 3))
    pos = _blob_mixture(rng, n, num_blobs, centres)
    log_s = rng.normal(-4.5, 0.4, (n, 3))
    return _finish(pos, log_s, rng, num_sg)


def scene_unbounded(n=1_000_000, seed=0, num_sg=4, inner_frac=0.6, num_blobs=64) -> Scene:
    rng = np.random.default_rng(seed)
    n_in = int(round(n * inner_frac))
    n_out = n - n_in
    # inner blob mixture: centres uniform in the unit ball (scaled so blobs stay near radius 1)
    c = _unit_vectors(rng, (num_blobs,)) * (rng.uniform(0, 1, (num_blobs, 1)) ** (1 / 3)) * 0.8
    pin = _blob_mixture(rng, n_in, num_blobs, c)
    log_s_in = rng.normal(-4.5, 0.4, (n_in, 3))
    # shell: radius log-uniform in [2, 50], direction uniform, scale proportional to radius
    r = np.exp(rng.uniform(math.log(2.0), math.log(50.0), n_out))
    pout = _unit_vectors(rng, (n_out,)) * r[:, None]
    log_s_out = rng.normal(-4.5, 0.4, (n_out, 3)) + np.log(r)[:, None]
    pos = np.concatenate([pin, pout])
    log_s = np.concatenate([log_s_in, log_s_out])
    return _finish(pos, log_s, rng, num_sg)


def _orbit_views(rng, radius, count, width, height, fovx_deg=40.0, even=True, elev=15.0):
    cams = []
    for i in range(count):
        az = 360.0 * i / count + 7.0 if even else rng.uniform(0, 360)
        el = elev if even else rng.uniform(-30, 45)
        cams.append(look_at_camera(orbit_eye(radius, az, el), width=width, height=height, fovx_deg=fovx_deg))
    return cams


def workload(name: str, seed: int = 0) -> Workload:
    rng = np.random.default_rng(seed + 1000)
    if name == "small4k":
        sc = scene_uniform(4096, seed)
        cams = [look_at_camera(orbit_eye(3.0, 30.0, 20.0), width=64, height=64, fovx_deg=40.0)]
        prm = RenderParams(dt=0.01, t_eps=1e-4, t_near=0.5, t_far=6.0, white_background=False)
        return Workload(name, sc, cams, prm, [None])
    if name == "blobs100k":
        sc = scene_blobs(100_000, 32, seed, num_sg=4)
        cams = [look_at_camera(orbit_eye(4.0, 30.0, 20.0), width=800, height=800, fovx_deg=40.0)]
        prm = RenderParams(dt=0.01, t_eps=1e-4, white_background=True)
        return Workload(name, sc, cams, prm, [None])
    if name == "blobs300k_16views":
        sc = scene_blobs(300_000, 96, seed, num_sg=4)
        cams, pixels = [], []
        for i in range(16):
            w = int(rng.integers(400, 1601))
            h = int(round(0.75 * w))
            cam = look_at_camera(orbit_eye(4.0, rng.uniform(0, 360), rng.uniform(-30, 45)), width=w, height=h)
            frac = rng.uniform(0.25, 1.0)
            npx = int(round(frac * w * h))
            px = np.sort(rng.permutation(w * h)[:npx]).astype(np.int32)
            cams.append(cam)
            pixels.append(px)
        prm = RenderParams(dt=0.01, t_eps=1e-4, white_background=True)
        return Workload(name, sc, cams, prm, pixels)
    if name == "unbounded1m_8views":
        sc = scene_unbounded(1_000_000, seed, num_sg=4)
        cams = _orbit_views(rng, 3.5, 8, 1237, 822)
        prm = RenderParams(dt=0.01, t_eps=1e-4, white_background=True)
        return Workload(name, sc, cams, prm, [None] * 8)
    if name == "unbounded3m_4views":
        sc = scene_unbounded(3_000_000, seed, num_sg=8)
        cams = _orbit_views(rng, 3.5, 4, 2048, 1536)
        prm = RenderParams(dt=0.01, t_eps=1e-4, white_background=True)
        return Workload(name, sc, cams, prm, [None] * 4)
    raise KeyError(name)


WORKLOADS = ("small4k", "blobs100k", "blobs300k_16views", "unbounded1m_8views", "unbounded3m_4views")


def crop_pixels(cam: Camera, x0: int, y0: int, w: int, h: int) -> np.ndarray:
    ys, xs = np.meshgrid(np.arange(y0, y0 + h), np.arange(x0, x0 + w), indexing="ij")
    return (ys * cam.image_width + xs).reshape(-1).astype(np.int32)
