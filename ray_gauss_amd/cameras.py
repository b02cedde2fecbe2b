"""Pinhole cameras in the convention the ray caster consumes.

The ray caster reads, per view (reference ``optix_raycasting/optix_utils.py:168-181``):

* ``world_view_transform`` -- the *transpose* of the world-to-camera matrix, so that
  ``world_view_transform[:3, 0..2]`` are the camera's right (U), down (V) and forward (W) axes in
  world space (COLMAP axes: x right, y down, z forward; ``classes/cameras.py:57-60``,
  ``utils/graphics_utils.py:38-49``);
* ``camera_center`` -- the eye position;
* ``FoVx``, ``FoVy`` (radians), ``image_width``, ``image_height``.

Any object with those attributes works with :class:`ray_gauss_amd.RenderOptixFunction`; this module
provides one that can be built from a look-at description for synthetic workloads.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

try:  # torch is optional for the pure-numpy users (oracle tests)
    import torch
except Exception:  # pragma: no cover
    torch = None


def focal_to_fov(focal: float, pixels: int) -> float:
    return 2.0 * math.atan(pixels / (2.0 * focal))


def fov_to_focal(fov: float, pixels: int) -> float:
    return pixels / (2.0 * math.tan(fov / 2.0))


def look_at_rotation(eye, target, up=(0.0, 0.0, 1.0)) -> np.ndarray:
    """Camera-to-world rotation with columns (right, down, forward)."""
    eye = np.asarray(eye, np.float64)
    fwd = np.asarray(target, np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    up = np.asarray(up, np.float64)
    right = np.cross(fwd, up)
    if np.linalg.norm(right) < 1e-8:  # looking along the up axis
        right = np.cross(fwd, np.array([0.0, 1.0, 0.0]))
    right /= np.linalg.norm(right)
    down = np.cross(fwd, right)
    return np.stack([right, down, fwd], axis=1)


@dataclass
class Camera:
    """Minimal camera with the attributes the ray caster reads."""

    image_width: int
    image_height: int
    FoVx: float
    FoVy: float
    R_c2w: np.ndarray  # (3,3) columns right, down, forward
    eye: np.ndarray  # (3,)
    device: str = "cpu"

    @property
    def R(self) -> np.ndarray:
        """classes/cameras.py convention: the transpose of the world-to-camera rotation
        (x_cam = x_world R + T)."""
        return np.asarray(self.R_c2w, np.float64)

    @property
    def T(self) -> np.ndarray:
        """World-to-camera translation."""
        return -np.asarray(self.R_c2w, np.float64).T @ np.asarray(self.eye, np.float64)

    @property
    def world_view_transform(self):
        w2c = np.eye(4, dtype=np.float64)
        w2c[:3, :3] = self.R_c2w.T
        w2c[:3, 3] = -self.R_c2w.T @ self.eye
        m = np.ascontiguousarray(w2c.T).astype(np.float32)
        if torch is not None:
            return torch.from_numpy(m).to(self.device)
        return m

    @property
    def camera_center(self):
        c = np.asarray(self.eye, np.float32)
        if torch is not None:
            return torch.from_numpy(c.copy()).to(self.device)
        return c

    # ---- FP32 launch record (what the kernels consume), optix_utils.py:168-181 ----------------
    def launch_record(self) -> dict:
        R = np.asarray(self.R_c2w, np.float32)
        return {
            "eye": np.asarray(self.eye, np.float32),
            "u": R[:, 0].copy(),
            "v": R[:, 1].copy(),
            "w": R[:, 2].copy(),
            "tanx": np.float32(math.tan(self.FoVx / 2.0)),
            "tany": np.float32(math.tan(self.FoVy / 2.0)),
            "width": int(self.image_width),
            "height": int(self.image_height),
        }


def look_at_camera(eye, target=(0.0, 0.0, 0.0), width=64, height=64, fovx_deg=40.0, up=(0.0, 0.0, 1.0),
                   device="cpu") -> Camera:
    fovx = math.radians(fovx_deg)
    focal = fov_to_focal(fovx, width)
    fovy = focal_to_fov(focal, height)
    R = look_at_rotation(eye, target, up)
    return Camera(width, height, fovx, fovy, R, np.asarray(eye, np.float64), device)


def orbit_eye(radius: float, azimuth_deg: float, elevation_deg: float) -> np.ndarray:
    a, e = math.radians(azimuth_deg), math.radians(elevation_deg)
    return radius * np.array([math.cos(e) * math.cos(a), math.cos(e) * math.sin(a), math.sin(e)])


def launch_record_from(cam) -> dict:
    """FP32 launch record from any reference-style camera object (duck type)."""
    if hasattr(cam, "launch_record"):
        return cam.launch_record()
    wvt = cam.world_view_transform
    wvt = wvt.detach().cpu().numpy() if hasattr(wvt, "detach") else np.asarray(wvt)
    cc = cam.camera_center
    cc = cc.detach().cpu().numpy() if hasattr(cc, "detach") else np.asarray(cc)
    wvt = wvt.astype(np.float32)
    return {
        "eye": cc.astype(np.float32).reshape(3),
        "u": wvt[:3, 0].copy(),
        "v": wvt[:3, 1].copy(),
        "w": wvt[:3, 2].copy(),
        "tanx": np.float32(math.tan(cam.FoVx / 2.0)),
        "tany": np.float32(math.tan(cam.FoVy / 2.0)),
        "width": int(cam.image_width),
        "height": int(cam.image_height),
    }
