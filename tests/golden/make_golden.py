"""Writes regression vectors of the CPU oracle to tests/golden/oracle_small.npz.

The reference repository holds no tests, fixtures or golden vectors for the ray caster, and it
cannot run here (OptiX), so these vectors are NOT reference outputs: they freeze the oracle's
FP64 results on a fixed seeded input (256 Gaussians, a 24x24 view, every option of the
reference that the hot path uses: SH degree 3, 2 SG lobes, white background) after the oracle
has been pinned by tests/test_oracle.py.  tests/test_golden.py checks that the oracle still
reproduces them, so an accidental change of the oracle cannot go unnoticed.

Run:  python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from ray_gauss_amd import scenes  # noqa: E402
from ray_gauss_amd.cameras import look_at_camera, orbit_eye  # noqa: E402

OUT = os.path.join(HERE, "oracle_small.npz")


def inputs():
    sc = scenes.scene_blobs(256, 4, seed=11, num_sg=2)
    # make it dense enough to matter at this size
    sc.scales[:] = sc.scales * 6.0
    cam = look_at_camera(orbit_eye(3.0, 40.0, 25.0), width=24, height=24, fovx_deg=35.0).launch_record()
    prm = scenes.RenderParams(dt=0.01, t_eps=1e-4, white_background=True)
    return sc, cam, prm


def compute():
    sc, cam, prm = inputs()
    rgb = O.sh_forward(sc, cam["eye"])
    bmin, bmax = O.scene_box(sc)
    out = O.render(prm, sc, rgb, bmin, bmax, cam)
    dL = np.random.default_rng(12).normal(0, 1, out["rgb"].shape)
    g = O.full_backward(prm, sc, cam["eye"], bmin, bmax, cam, out["rgb"], dL)
    vec = {"rgb": out["rgb"], "depth": out["depth"], "opacity": out["opacity"], "K": out["K"], "P": out["P"],
           "dL": dL, "bmin": bmin, "bmax": bmax}
    vec.update({"g_" + k: v for k, v in g.items()})
    return vec


if __name__ == "__main__":
    np.savez_compressed(OUT, **compute())
    print("wrote", OUT)
