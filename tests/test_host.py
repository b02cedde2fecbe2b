"""CPU tests of the host side: the C-ABI library's exports, camera conventions, the host-glue
mirror of the reference's optix_utils, and the synthetic workload generators."""

import ctypes
import math
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from ray_gauss_amd import scenes
from ray_gauss_amd.cameras import launch_record_from, look_at_camera

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "raygauss_hip.h")
LIB = os.path.join(ROOT, "ray_gauss_amd", "_build", "libraygauss_hip.so")


def header_functions():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|int64_t|const char\*)\s+(rg_\w+)\(", txt, re.M)))


def test_header_declares_the_abi():
    fns = header_functions()
    for f in ("rg_create", "rg_prepare", "rg_set_views", "rg_forward", "rg_backward", "rg_train_batch",
              "rg_finish_grads", "rg_debug_hits"):
        assert f in fns


@pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    for f in header_functions():
        assert hasattr(lib, f), f


def test_binding_covers_the_header():
    from ray_gauss_amd import _lib

    assert sorted(_lib.EXPORTS) == header_functions()


def test_camera_convention():
    cam = look_at_camera([3.0, 1.0, 0.5], width=40, height=30, fovx_deg=40.0)
    wvt = np.asarray(cam.world_view_transform.cpu() if hasattr(cam.world_view_transform, "cpu")
                     else cam.world_view_transform)
    fwd = -np.array([3.0, 1.0, 0.5])
    fwd /= np.linalg.norm(fwd)
    assert np.allclose(wvt[:3, 2], fwd, atol=1e-6)
    right, down = wvt[:3, 0], wvt[:3, 1]
    assert np.allclose(np.cross(right, down), fwd, atol=1e-6)  # x right, y down, z forward
    assert down[2] < 0  # image y points down in a z-up world
    # world_view_transform is the transposed world-to-camera: eye maps to the camera origin
    w2c = wvt.T
    eye_h = np.array([3.0, 1.0, 0.5, 1.0])
    assert np.allclose((w2c @ eye_h)[:3], 0, atol=1e-5)
    rec = cam.launch_record()

    class Duck:
        image_width, image_height, FoVx, FoVy = cam.image_width, cam.image_height, cam.FoVx, cam.FoVy
        world_view_transform = cam.world_view_transform
        camera_center = cam.camera_center

    rec2 = launch_record_from(Duck())
    for k in ("eye", "u", "v", "w"):
        assert np.allclose(rec[k], rec2[k], atol=1e-6)
    assert rec2["tanx"] == rec["tanx"] and rec2["tany"] == rec["tany"]
    # fovy from the shared focal length
    assert math.isclose(math.tan(cam.FoVy / 2) / math.tan(cam.FoVx / 2), 30 / 40, rel_tol=1e-12)


def test_optix_utils_mirror():
    torch = pytest.importorskip("torch")
    from oracle import oracle as O
    from ray_gauss_amd import optix_utils as u

    sc = scenes.scene_uniform(200, seed=7)
    q = torch.from_numpy(sc.quats)
    L1, L2, L3 = u.quaternion_to_rotation(q)
    R = scenes._quat_to_R(sc.quats.astype(np.float64))
    assert np.allclose(L1.numpy(), R[:, 0, :], atol=1e-6)
    assert np.allclose(L3.numpy(), R[:, 2, :], atol=1e-6)
    box = u.compute_ellipsoids_bbox(torch.from_numpy(sc.positions), torch.from_numpy(sc.scales), L1, L2, L3,
                                    torch.from_numpy(sc.densities))
    ref = O.aabb(sc)
    assert np.abs(box.numpy() - ref).max() < 1e-5
    with pytest.raises(ValueError):
        u.create_module(None, None, "nope")
    ctx = u.create_context()
    assert u.create_module(ctx, None, "forward").stage == "forward"


def test_workloads_are_seeded_and_activated():
    a = scenes.workload("small4k")
    b = scenes.workload("small4k")
    assert np.array_equal(a.scene.positions, b.scene.positions)
    sc = a.scene
    assert sc.positions.dtype == np.float32 and sc.sh.shape == (4096, 3, 16)
    assert np.allclose(np.linalg.norm(sc.quats, axis=1), 1, atol=1e-6)
    assert (sc.densities > 0).all() and (sc.scales > 0).all()
    assert a.params.dt == 0.01 and a.params.t_near == 0.5 and a.params.t_far == 6.0
    c = scenes.scene_unbounded(20000, seed=1)
    r = np.linalg.norm(c.positions, axis=1)
    assert (r[12000:] > 1.5).all() and r[12000:].max() > 30  # the shell
    assert np.median(c.scales[12000:].mean(1) / r[12000:]) > 0.005  # scales grow with radius


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stdout


@pytest.mark.skipif(not os.path.exists(LIB), reason="HIP library not built")
def test_optimiser_entry_points_reject_bad_arguments_on_the_host():
    """Argument checks of rg_adam_step / rg_project_params run before any device work."""
    from ray_gauss_amd import _lib

    lib = _lib.load()
    table = (_lib.AdamTensor * 17)()
    assert lib.rg_adam_step(table, 17, None) != 0
    assert b"16" in lib.rg_last_error(None)
    one = (_lib.AdamTensor * 1)()
    one[0].n = 4
    one[0].step = 1
    assert lib.rg_adam_step(one, 1, None) != 0  # null arrays
    assert b"null" in lib.rg_last_error(None)
    one[0].n = 0
    one[0].step = 0
    assert lib.rg_adam_step(one, 1, None) != 0  # step must count the update
    assert lib.rg_project_params(None, None) != 0
    pr = _lib.Projection()
    pr.n = -1
    assert lib.rg_project_params(ctypes.byref(pr), None) != 0
