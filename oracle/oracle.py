"""TEST INFRASTRUCTURE ONLY.  ctypes front end of the CPU FP64 oracle (raygauss_oracle.cpp).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may import this module,
and only as a checker / CPU baseline -- never as a product code path.

Every function mirrors one reference routine; see the header of raygauss_oracle.cpp for the
file:line each follows.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle_rg.so")
_lib = None


class Camera(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("u", C.c_float * 3), ("v", C.c_float * 3), ("w", C.c_float * 3),
                ("tanx", C.c_float), ("tany", C.c_float), ("width", C.c_int32), ("height", C.c_int32)]


class Params(C.Structure):
    _fields_ = [("dt", C.c_float), ("spb", C.c_int32), ("t_eps", C.c_float), ("sigma_thr", C.c_float),
                ("t_near", C.c_float), ("t_far", C.c_float), ("white_bg", C.c_int32),
                ("exact_sh_grad", C.c_int32), ("jitter", C.c_int32), ("rnd_sample", C.c_int32),
                ("iteration", C.c_uint32), ("eps_w", C.c_float)]

# Fragility margin on sigma*w / thr: a sample decision whose exact value lies within EPS_W of the
# threshold may be decided either way by an FP32 evaluation.  tests/test_oracle.py measures the
# FP32 error of both the reference's evaluation and the HIP path's closed form near the threshold
# (orc_sigmaw_fp32_error) and asserts EPS_W is at least twice the largest one seen.
EPS_W = 3e-4


class RecordS(C.Structure):
    _fields_ = [("mu", C.c_void_p), ("Mp", C.c_void_p), ("valid", C.c_void_p), ("h", C.c_void_p)]


class _RecordHolder:
    """The GPU's FP32 per-Gaussian records (RayCaster.export_record()) kept alive for C."""

    def __init__(self, rec):
        self.arrays = [np.ascontiguousarray(rec["mu"], np.float32), np.ascontiguousarray(rec["Mp"], np.float32),
                       np.ascontiguousarray(rec["valid"], np.int32), np.ascontiguousarray(rec["h"], np.float32)]
        r = RecordS()
        r.mu, r.Mp, r.valid, r.h = (a.ctypes.data for a in self.arrays)
        self.s = r


def _rec_ptr(holder):
    return None if holder is None else C.byref(holder.s)


class SceneS(C.Structure):
    _fields_ = [("n", C.c_int64), ("deg", C.c_int32), ("L", C.c_int32)] + [
        (k, C.c_void_p) for k in ("pos", "scale", "quat", "dens", "sh", "sg", "lam", "axis")]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
    return _lib


def nthreads(default=None):
    if default is not None:
        return int(default)
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except Exception:  # pragma: no cover
        return os.cpu_count() or 1


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def camera_struct(rec: dict) -> Camera:
    c = Camera()
    for k in ("eye", "u", "v", "w"):
        getattr(c, k)[:] = [float(x) for x in np.asarray(rec[k], np.float32)]
    c.tanx = float(np.float32(rec["tanx"]))
    c.tany = float(np.float32(rec["tany"]))
    c.width = int(rec["width"])
    c.height = int(rec["height"])
    return c


def params_struct(prm, exact_sh_grad=False) -> Params:
    p = Params()
    p.dt = float(np.float32(prm.dt))
    p.spb = int(prm.samples_per_slab)
    p.t_eps = float(np.float32(prm.t_eps))
    p.sigma_thr = float(np.float32(prm.sigma_thr))
    p.t_near = float(np.float32(prm.t_near))
    p.t_far = float(np.float32(prm.t_far))
    p.white_bg = int(bool(prm.white_background))
    p.exact_sh_grad = int(bool(exact_sh_grad))
    p.jitter = int(getattr(prm, "jitter", 0))
    p.rnd_sample = int(getattr(prm, "rnd_sample", 0))
    p.iteration = int(getattr(prm, "iteration", 0)) & 0xFFFFFFFF
    p.eps_w = float(getattr(prm, "eps_w", EPS_W))
    return p


class _SceneHolder:
    """Keeps the contiguous float32 arrays alive while C holds pointers into them."""

    def __init__(self, sc):
        self.arrays = [_f32(a) for a in (sc.positions, sc.scales, sc.quats, sc.densities, sc.sh, sc.sg,
                                         sc.sharpness, sc.lobe_axis)]
        s = SceneS()
        s.n = sc.positions.shape[0]
        s.deg = int(round(np.sqrt(sc.sh.shape[2]))) - 1
        s.L = int(sc.sg.shape[2])
        for k, a in zip(("pos", "scale", "quat", "dens", "sh", "sg", "lam", "axis"), self.arrays):
            setattr(s, k, a.ctypes.data)
        self.s = s


def rays_rng(cam_rec, pixels=None, jitter=0, iteration=0):
    """Rays with the reference's sub-pixel jitter (generator seeded per pixel)."""
    cs = camera_struct(cam_rec)
    n = cam_rec["width"] * cam_rec["height"] if pixels is None else len(pixels)
    px = None if pixels is None else np.ascontiguousarray(pixels, np.int32)
    o = np.empty((n, 3), np.float32)
    d = np.empty((n, 3), np.float32)
    lib().orc_rays_rng(C.byref(cs), C.c_int64(n), _p(px), C.c_int32(int(jitter)), C.c_uint32(int(iteration)),
                       _p(o), _p(d))
    return o, d


def tea(v0, v1, cycles, key):
    k = np.ascontiguousarray(key, np.uint32)
    f = lib().orc_tea
    f.restype = C.c_uint32
    return int(f(C.c_uint32(v0), C.c_uint32(v1), C.c_int32(cycles), _p(k)))


def tea4(v0, v1):
    f = lib().orc_tea4
    f.restype = C.c_uint32
    return int(f(C.c_uint32(v0), C.c_uint32(v1)))


def lcg(seed, n):
    ints = np.empty(n, np.uint32)
    floats = np.empty(n, np.float32)
    lib().orc_lcg(C.c_uint32(seed), C.c_int32(n), _p(ints), _p(floats))
    return ints, floats


def rays(cam_rec, pixels=None):
    cs = camera_struct(cam_rec)
    n = cam_rec["width"] * cam_rec["height"] if pixels is None else len(pixels)
    px = None if pixels is None else np.ascontiguousarray(pixels, np.int32)
    o = np.empty((n, 3), np.float32)
    d = np.empty((n, 3), np.float32)
    lib().orc_rays(C.byref(cs), C.c_int64(n), _p(px), _p(o), _p(d))
    return o, d


def ray_interval(cam_rec, bmin, bmax, t_near, t_far, pixels=None):
    cs = camera_struct(cam_rec)
    n = cam_rec["width"] * cam_rec["height"] if pixels is None else len(pixels)
    px = None if pixels is None else np.ascontiguousarray(pixels, np.int32)
    te = np.empty(n, np.float32)
    tx = np.empty(n, np.float32)
    lib().orc_ray_interval(C.byref(cs), C.c_int64(n), _p(px), _p(_f32(bmin)), _p(_f32(bmax)),
                           C.c_float(t_near), C.c_float(t_far), _p(te), _p(tx))
    return te, tx


def sh_forward(scene, eye, threads=None):
    h = _SceneHolder(scene)
    rgb = np.empty((scene.n, 3), np.float64)
    lib().orc_sh_forward(C.byref(h.s), _p(_f32(eye)), _p(rgb), C.c_int(nthreads(threads)))
    return rgb


def sh_backward(scene, eye, dLdrgb, exact=False, threads=None):
    h = _SceneHolder(scene)
    n, L = scene.n, scene.num_sg
    K = scene.sh.shape[2]
    dsh = np.zeros((n, 3, K))
    dsg = np.zeros((n, 3, L))
    dlam = np.zeros((n, L))
    dax = np.zeros((n, L, 3))
    dpos = np.zeros((n, 3))
    lib().orc_sh_backward(C.byref(h.s), _p(_f32(eye)), _p(_f64(dLdrgb)), C.c_int(int(exact)), _p(dsh), _p(dsg),
                          _p(dlam), _p(dax), _p(dpos), C.c_int(nthreads(threads)))
    return {"sh": dsh, "sg": dsg, "sharpness": dlam, "lobe_axis": dax, "positions": dpos}


def aabb(scene, sigma_thr=0.1, threads=None):
    h = _SceneHolder(scene)
    box = np.empty((scene.n, 6), np.float64)
    lib().orc_aabb(C.byref(h.s), C.c_float(sigma_thr), _p(box), C.c_int(nthreads(threads)))
    return box


def scene_box(scene, sigma_thr=0.1):
    """Scene AABB (render_optix.py:105-106) rounded outward to float32."""
    b = aabb(scene, sigma_thr)
    lo = b[:, :3].min(axis=0)
    hi = b[:, 3:].max(axis=0)
    lo32 = np.nextafter(lo.astype(np.float32), np.float32(-np.inf))
    hi32 = np.nextafter(hi.astype(np.float32), np.float32(np.inf))
    return lo32.astype(np.float32), hi32.astype(np.float32)


def render(prm, scene, rgb, bmin, bmax, cam_rec, pixels=None, threads=None, record=None):
    """Forward render of one view.  record (the GPU's export_record()) switches to the
    decision-matched mode: hit sets and sample thresholds decided in FP32 from those records as
    the HIP path decides them, every value still FP64; 'fragile' then only flags decisions within
    1e-6 of the threshold (native vs host exp2), early exits and origin-inside rays."""
    h = _SceneHolder(scene)
    rh = None if record is None else _RecordHolder(record)
    n = cam_rec["width"] * cam_rec["height"] if pixels is None else len(pixels)
    px = None if pixels is None else np.ascontiguousarray(pixels, np.int32)
    out = {
        "rgb": np.empty((n, 3)), "depth": np.empty(n), "opacity": np.empty(n),
        "fragile": np.empty(n, np.int32), "K": np.empty(n, np.int64), "P": np.empty(n, np.int64),
        "maxhits": np.empty(n, np.int64), "NC": np.empty(n, np.int64), "nfrag": np.empty(n, np.int64),
        "cmax": np.empty(n),
    }
    ps = params_struct(prm)
    cs = camera_struct(cam_rec)
    lib().orc_render(C.byref(ps), C.byref(h.s), _p(_f64(rgb)), _p(_f32(bmin)), _p(_f32(bmax)), C.byref(cs),
                     C.c_int64(n), _p(px), _p(out["rgb"]), _p(out["depth"]), _p(out["opacity"]), _p(out["fragile"]),
                     _p(out["K"]), _p(out["P"]), _p(out["maxhits"]), _p(out["NC"]), _p(out["nfrag"]),
                     _p(out["cmax"]), _rec_ptr(rh), C.c_int(nthreads(threads)))
    return out


# kinds of fragility (render()["fragile"] bits)
FRAG_SAMPLE, FRAG_EXIT, FRAG_QUIRK = 1, 2, 4


def fragile_bound(prm, out):
    """Per-ray bound on |C_fp32 - C_oracle| that the ray's near-threshold decisions allow: each
    counted sample decision can move the colour by at most 4 thr (1 + eps_w) dt max(cmax, 1) (its
    density changes by at most thr (1 + eps_w); see the oracle's header note), a transmittance check
    within rounding of T_eps by at most 2.002 T_eps max(cmax, 1) (one slab more or less); plus the
    1e-4 tolerance of every ray.  Rays whose origin-inside decision is fragile get +inf (the
    decision adds or removes a whole Gaussian)."""
    eps_w = float(getattr(prm, "eps_w", EPS_W))
    c = np.maximum(out["cmax"], 1.0)
    b = 1e-4 + out["nfrag"] * 4.0 * float(np.float32(prm.sigma_thr)) * (1.0 + eps_w) * float(np.float32(prm.dt)) * c
    b = b + ((out["fragile"] & FRAG_EXIT) != 0) * 2.002 * float(np.float32(prm.t_eps)) * c
    return np.where((out["fragile"] & FRAG_QUIRK) != 0, np.inf, b)


def sigmaw_fp32_error(prm, scene, bmin, bmax, cam_rec, pixels=None, band=1e-2, threads=None):
    """Largest relative FP32 error of sigma*w near the threshold (within `band`): the reference's
    FP32 evaluation and the HIP path's closed form, against FP64; and the number of samples."""
    h = _SceneHolder(scene)
    n = cam_rec["width"] * cam_rec["height"] if pixels is None else len(pixels)
    px = None if pixels is None else np.ascontiguousarray(pixels, np.int32)
    out = np.zeros(3)
    ps = params_struct(prm)
    cs = camera_struct(cam_rec)
    lib().orc_sigmaw_fp32_error(C.byref(ps), C.byref(h.s), _p(_f32(bmin)), _p(_f32(bmax)), C.byref(cs),
                                C.c_int64(n), _p(px), C.c_double(band), _p(out), C.c_int(nthreads(threads)))
    return {"reference": float(out[0]), "closed_form": float(out[1]), "samples": int(out[2])}


def render_literal(prm, scene, rgb, bmin, bmax, cam_rec, pixels=None, threads=None):
    h = _SceneHolder(scene)
    n = cam_rec["width"] * cam_rec["height"] if pixels is None else len(pixels)
    px = None if pixels is None else np.ascontiguousarray(pixels, np.int32)
    out = {"rgb": np.empty((n, 3)), "depth": np.empty(n), "K": np.empty(n, np.int64)}
    ps = params_struct(prm)
    cs = camera_struct(cam_rec)
    lib().orc_render_literal(C.byref(ps), C.byref(h.s), _p(_f64(rgb)), _p(_f32(bmin)), _p(_f32(bmax)),
                             C.byref(cs), C.c_int64(n), _p(px), _p(out["rgb"]), _p(out["depth"]), _p(out["K"]),
                             C.c_int(nthreads(threads)))
    return out


def backward(prm, scene, rgb, bmin, bmax, cam_rec, ray_colors, dLdC, pixels=None, threads=None, record=None):
    """Ray-caster backward for one view; returns gradients w.r.t. positions, scales, quats,
    densities and the per-Gaussian colour (rgb) -- before the SH/SG chain.  record: as in render()."""
    h = _SceneHolder(scene)
    rh = None if record is None else _RecordHolder(record)
    n = scene.n
    nr = cam_rec["width"] * cam_rec["height"] if pixels is None else len(pixels)
    px = None if pixels is None else np.ascontiguousarray(pixels, np.int32)
    g = {"positions": np.zeros((n, 3)), "scales": np.zeros((n, 3)), "quats": np.zeros((n, 4)),
         "densities": np.zeros(n), "rgb": np.zeros((n, 3))}
    ps = params_struct(prm)
    cs = camera_struct(cam_rec)
    lib().orc_backward(C.byref(ps), C.byref(h.s), _p(_f64(rgb)), _p(_f32(bmin)), _p(_f32(bmax)), C.byref(cs),
                       C.c_int64(nr), _p(px), _p(_f64(ray_colors)), _p(_f64(dLdC)), _p(g["positions"]),
                       _p(g["scales"]), _p(g["quats"]), _p(g["densities"]), _p(g["rgb"]), _rec_ptr(rh),
                       C.c_int(nthreads(threads)))
    return g


def full_backward(prm, scene, eye, bmin, bmax, cam_rec, ray_colors, dLdC, pixels=None, exact_sh=False,
                  threads=None, record=None):
    """Gradients of all nine boundary inputs, as RenderOptixFunction.backward assembles them
    (render_optix.py:158-190): the ray-caster terms plus the SH/SG chain and the position term
    through the view direction."""
    rgb = sh_forward(scene, eye, threads)
    g = backward(prm, scene, rgb, bmin, bmax, cam_rec, ray_colors, dLdC, pixels, threads, record)
    s = sh_backward(scene, eye, g["rgb"], exact_sh, threads)
    return {"positions": g["positions"] + s["positions"], "scales": g["scales"], "quats": g["quats"],
            "densities": g["densities"], "sh": s["sh"], "sg": s["sg"], "sharpness": s["sharpness"],
            "lobe_axis": s["lobe_axis"], "rgb": g["rgb"]}


def hits_fp32(mu, Mp, valid, o, d, tenter, texit, dt, spb, max_slabs, max_pairs, threads=None):
    n = mu.shape[0]
    nr = o.shape[0]
    pairs = np.full((nr, max_pairs, 2), -1, np.int32)
    counts = np.empty(nr, np.int64)
    lib().orc_hits_fp32(C.c_int64(n), _p(_f32(mu)), _p(_f32(Mp)), _p(np.ascontiguousarray(valid, np.int32)),
                        C.c_int64(nr), _p(_f32(o)), _p(_f32(d)), _p(_f32(tenter)), _p(_f32(texit)),
                        C.c_float(dt), C.c_int32(spb), C.c_int32(max_slabs), C.c_int32(max_pairs), _p(pairs),
                        _p(counts), C.c_int(nthreads(threads)))
    return pairs, counts


def intersect(scene, o, d, sigma_thr=0.1):
    """FP64 predicate for every (ray, Gaussian): dict of r1, r2, disc, b, c (nrays, n) and valid."""
    h = _SceneHolder(scene)
    o = _f64(o).reshape(-1, 3)
    d = _f64(d).reshape(-1, 3)
    nr, n = o.shape[0], scene.n
    out = np.empty((nr, n, 5))
    valid = np.empty((nr, n), np.int32)
    lib().orc_intersect(C.byref(h.s), C.c_float(sigma_thr), C.c_int64(nr), _p(o), _p(d), _p(out), _p(valid))
    return {"r1": out[..., 0], "r2": out[..., 1], "disc": out[..., 2], "b": out[..., 3], "c": out[..., 4],
            "valid": valid.astype(bool)}


# ------------------------------------------------------------------ FP32 CPU baseline ---------
# liboracle_rg32.so: the same source as the oracle built with real = float, without the oracle's
# fragility bookkeeping, for x86-64-v3 (AVX2 + FMA).  A speed reference, not a checker.

_LIB32_PATH = os.path.join(_HERE, "_build", "liboracle_rg32.so")
_lib32 = None


def cpu_model():
    """(model name, has AVX2 and FMA) of this host from /proc/cpuinfo."""
    name, flags = "unknown", set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and name == "unknown":
                    name = line.split(":", 1)[1].strip()
                elif line.startswith("flags") and not flags:
                    flags = set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return name, {"avx2", "fma"} <= flags


def baseline_lib():
    global _lib32
    if _lib32 is None:
        if not cpu_model()[1]:
            raise RuntimeError("the FP32 CPU baseline is built for x86-64-v3; this host lacks AVX2/FMA")
        if not os.path.exists(_LIB32_PATH):
            build()
        _lib32 = C.CDLL(_LIB32_PATH)
        assert _lib32.orc_baseline_real_bytes() == 4
    return _lib32


def baseline_step(prm, scene, rgb, bmin, bmax, cam_rec, pixels, target, threads=None):
    """One FP32 training pass of the ray caster over the given rays of one view (geometry + BVH,
    forward, L2 gradient against `target`, backward).  Returns the colours, the gradients and the
    seconds of the three phases."""
    h = _SceneHolder(scene)
    n = scene.n
    px = np.ascontiguousarray(pixels, np.int32)
    nr = len(px)
    f = lambda *shape: np.zeros(shape, np.float32)  # noqa: E731
    out = f(nr, 3)
    g = {"positions": f(n, 3), "scales": f(n, 3), "quats": f(n, 4), "densities": f(n), "rgb": f(n, 3)}
    times = np.zeros(3)
    ps = params_struct(prm)
    cs = camera_struct(cam_rec)
    baseline_lib().orc_baseline_step(
        C.byref(ps), C.byref(h.s), _p(_f32(rgb)), _p(_f32(bmin)), _p(_f32(bmax)), C.byref(cs), C.c_int64(nr), _p(px),
        _p(_f32(target)), _p(out), _p(g["positions"]), _p(g["scales"]), _p(g["quats"]), _p(g["densities"]),
        _p(g["rgb"]), _p(times), C.c_int(nthreads(threads)))
    return {"rgb": out, "grads": g, "times": tuple(float(t) for t in times)}
