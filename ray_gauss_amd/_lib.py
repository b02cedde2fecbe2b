"""ctypes binding of the C ABI in include/raygauss_hip.h (libraygauss_hip.so).

The product path has no CPU fallback: if the library is missing this module raises, and every
entry point reports HIP errors as exceptions.
"""

from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libraygauss_hip.so")

_lib = None


class Config(C.Structure):
    _fields_ = [("dt", C.c_float), ("samples_per_slab", C.c_int32), ("t_eps", C.c_float), ("sigma_thr", C.c_float),
                ("t_near", C.c_float), ("t_far", C.c_float), ("white_background", C.c_int32),
                ("exact_sh_grad", C.c_int32), ("hit_depth", C.c_int32), ("seg_slabs", C.c_int32),
                ("max_hits_per_slab", C.c_int32), ("packet_traversal", C.c_int32),
                ("jitter", C.c_int32), ("rnd_sample", C.c_int32), ("iteration", C.c_uint32),
                ("fwd_hit_depth", C.c_int32), ("replay_words", C.c_int32),
                ("tree_builder", C.c_int32), ("ray_chunk", C.c_int32), ("refill_lanes", C.c_int32),
                ("bwd_ray_chunk", C.c_int32), ("bwd_window", C.c_float), ("fwd_round_per_slab", C.c_int32)]


class CameraS(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("u", C.c_float * 3), ("v", C.c_float * 3), ("w", C.c_float * 3),
                ("tanx", C.c_float), ("tany", C.c_float), ("width", C.c_int32), ("height", C.c_int32)]


class Gaussians(C.Structure):
    _fields_ = [("n", C.c_int64), ("sh_degree", C.c_int32), ("num_sg", C.c_int32)] + [
        (k, C.c_void_p) for k in ("positions", "scales", "quats", "densities", "sh", "sg", "sharpness", "lobe_axis")]


class Grads(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("positions", "scales", "quats", "densities", "sh", "sg", "sharpness",
                                          "lobe_axis")]


class RayBatch(C.Structure):
    _fields_ = [("num_rays", C.c_int64), ("ray_view", C.c_void_p), ("ray_pixel", C.c_void_p)]


class Stats(C.Structure):
    _fields_ = [("slab_hits", C.c_int64), ("pairs", C.c_int64), ("overflow_slabs", C.c_int64),
                ("fallback_slabs", C.c_int64), ("gathers", C.c_int64), ("replay_words", C.c_int64),
                ("replay_truncated", C.c_int64), ("dropped", C.c_int64), ("bad_rays", C.c_int64)]


class AdamTensor(C.Structure):
    _fields_ = [("param", C.c_void_p), ("grad", C.c_void_p), ("exp_avg", C.c_void_p), ("exp_avg_sq", C.c_void_p),
                ("n", C.c_int64), ("lr", C.c_double), ("beta1", C.c_double), ("beta2", C.c_double),
                ("eps", C.c_double), ("step", C.c_int64)]


class Projection(C.Structure):
    _fields_ = [("n", C.c_int64), ("num_sg", C.c_int32), ("densities", C.c_void_p), ("log_scales", C.c_void_p),
                ("filter_3d", C.c_void_p), ("max_scale", C.c_float), ("quats", C.c_void_p),
                ("normalize_quaternion", C.c_int32), ("lobe_axis", C.c_void_p), ("normalize_lobe_axis", C.c_int32)]


class FilterCamera(C.Structure):
    _fields_ = [("R", C.c_float * 9), ("T", C.c_float * 3), ("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float),
                ("cy", C.c_float), ("x_lo", C.c_float), ("x_hi", C.c_float), ("y_lo", C.c_float), ("y_hi", C.c_float)]


class DensifyIn(C.Structure):
    _fields_ = [("n", C.c_int64)] + [(k, C.c_void_p) for k in ("grad_accum", "num_accum", "positions", "log_scales",
                                                                "quats", "densities")] + [
        ("grad_threshold", C.c_float), ("scale_limit", C.c_float), ("min_density", C.c_float),
        ("split_n", C.c_int32)]


class DensifyArray(C.Structure):
    _fields_ = [("src", C.c_void_p), ("dst", C.c_void_p), ("width", C.c_int64), ("zero_new", C.c_int32)]


EXPORTS = {
    "rg_create": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(Config)]),
    "rg_destroy": (C.c_int, [C.c_void_p]),
    "rg_set_config": (C.c_int, [C.c_void_p, C.POINTER(Config)]),
    "rg_last_error": (C.c_char_p, [C.c_void_p]),
    "rg_prepare": (C.c_int, [C.c_void_p, C.POINTER(Gaussians), C.c_int32, C.c_void_p]),
    "rg_set_views": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(CameraS), C.c_void_p]),
    "rg_forward": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                             C.POINTER(C.c_uint64), C.c_void_p]),
    "rg_backward": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                              C.c_void_p]),
    "rg_forward_batch": (C.c_int, [C.c_void_p, C.POINTER(RayBatch), C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.POINTER(C.c_uint64), C.c_void_p]),
    "rg_backward_batch": (C.c_int, [C.c_void_p, C.POINTER(RayBatch), C.c_void_p, C.c_void_p, C.c_uint64,
                                    C.c_void_p]),
    "rg_train_batch": (C.c_int, [C.c_void_p, C.POINTER(RayBatch), C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                 C.c_void_p]),
    "rg_zero_grads": (C.c_int, [C.c_void_p, C.c_void_p]),
    "rg_finish_grads": (C.c_int, [C.c_void_p, C.POINTER(Grads), C.c_void_p]),
    "rg_scene_box": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "rg_get_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats), C.c_void_p]),
    "rg_alloc_epoch": (C.c_int64, [C.c_void_p]),
    "rg_device_bytes": (C.c_int64, [C.c_void_p, C.c_void_p]),
    "rg_tree_cost": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "rg_export_predicate": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "rg_debug_hits": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p,
                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "rg_debug_check_bvh": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.c_void_p]),
    "rg_adam_step": (C.c_int, [C.POINTER(AdamTensor), C.c_int32, C.c_void_p]),
    "rg_project_params": (C.c_int, [C.POINTER(Projection), C.c_void_p]),
    "rg_set_view_colors": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "rg_colors_to_rgba8": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "rg_loss_workspace_bytes": (C.c_int64, [C.c_int32, C.c_int32]),
    "rg_loss_l1_ssim": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_float,
                                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "rg_knn_workspace_bytes": (C.c_int64, [C.c_int64]),
    "rg_knn_mean_dist": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "rg_knn_mean_dist_noalloc": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "rg_filter_3d": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int32, C.c_float, C.c_void_p, C.c_void_p,
                               C.c_int64, C.c_void_p]),
    "rg_accumulate_grad_norm": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "rg_densify_workspace_bytes": (C.c_int64, [C.c_int64]),
    "rg_densify_plan": (C.c_int, [C.POINTER(DensifyIn), C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "rg_densify_apply": (C.c_int, [C.POINTER(DensifyIn), C.c_void_p, C.c_int64, C.POINTER(C.c_int64),
                                   C.POINTER(DensifyArray), C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p]),
}


def load():
    """Loads libraygauss_hip.so (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP library not built: {LIB_PATH} is missing (run __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class RayGaussError(RuntimeError):
    pass


def check(rc: int, handle) -> None:
    if rc != 0:
        msg = load().rg_last_error(handle)
        raise RayGaussError(f"raygauss_hip error {rc}: {msg.decode() if msg else ''}")
