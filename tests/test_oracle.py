"""Pins the CPU FP64 oracle against mathematics that is independent of its code: closed forms,
invariants, quadrature, finite differences and brute force.  The reference repository holds no
tests, fixtures or golden vectors for this path, so these checks (and the committed golden
vectors in tests/golden/) are what the oracle is trusted on.
"""

import math

import numpy as np
import pytest

from oracle import oracle as O
from ray_gauss_amd import scenes
from ray_gauss_amd.cameras import look_at_camera
from ray_gauss_amd.scenes import RenderParams, Scene


def make_scene(pos, scales, quats, dens, sh=None, sg=None, sharp=None, axis=None, deg=0, L=0):
    n = len(pos)
    K = (deg + 1) ** 2
    f = lambda a, s: np.ascontiguousarray(np.asarray(a, np.float64).reshape(s), np.float32)  # noqa: E731
    return Scene(f(pos, (n, 3)), f(scales, (n, 3)), f(quats, (n, 4)), f(dens, (n,)),
                 f(np.zeros((n, 3, K)) if sh is None else sh, (n, 3, K)),
                 f(np.zeros((n, 3, L)) if sg is None else sg, (n, 3, L)),
                 f(np.zeros((n, L)) if sharp is None else sharp, (n, L)),
                 f(np.zeros((n, L, 3)) if axis is None else axis, (n, L, 3)))


# ------------------------------------------------------------------------ SH / SG -------------

def real_sh_closed_form(x, y, z):
    """The 16 real SH of degree <= 3 written from their definitions (normalisation constants
    sqrt((2l+1)/(4 pi) (l-|m|)!/(l+|m|)!) with the Condon-Shortley phase), in the reference's order
    m = -l..l (utils/sh_utils.cu:60-117)."""
    pi = math.pi
    Y = [0.5 / math.sqrt(pi)]
    c1 = math.sqrt(3 / (4 * pi))
    Y += [-c1 * y, c1 * z, -c1 * x]
    Y += [0.5 * math.sqrt(15 / pi) * x * y, -0.5 * math.sqrt(15 / pi) * y * z,
          0.25 * math.sqrt(5 / pi) * (2 * z * z - x * x - y * y), -0.5 * math.sqrt(15 / pi) * x * z,
          0.25 * math.sqrt(15 / pi) * (x * x - y * y)]
    Y += [-0.25 * math.sqrt(35 / (2 * pi)) * y * (3 * x * x - y * y), 0.5 * math.sqrt(105 / pi) * x * y * z,
          -0.25 * math.sqrt(21 / (2 * pi)) * y * (4 * z * z - x * x - y * y),
          0.25 * math.sqrt(7 / pi) * z * (2 * z * z - 3 * x * x - 3 * y * y),
          -0.25 * math.sqrt(21 / (2 * pi)) * x * (4 * z * z - x * x - y * y),
          0.25 * math.sqrt(105 / pi) * z * (x * x - y * y), -0.25 * math.sqrt(35 / (2 * pi)) * x * (x * x - 3 * y * y)]
    return np.array(Y)


def sh_basis_via_oracle(dirs):
    """Evaluates the oracle's SH basis: one Gaussian per direction, eye at the origin, one unit
    coefficient per basis function (channel = basis index mod 3 is irrelevant: use red)."""
    n = len(dirs)
    out = np.empty((n, 16))
    for k in range(16):
        sh = np.zeros((n, 3, 16))
        sh[:, 0, k] = 1.0
        sc = make_scene(dirs * 2.0, np.ones((n, 3)), np.tile([1, 0, 0, 0], (n, 1)), np.ones(n), sh=sh, deg=3)
        out[:, k] = O.sh_forward(sc, np.zeros(3, np.float32))[:, 0] - 0.5
    return out


def sphere_quadrature(nt=24, nphi=48):
    xg, wg = np.polynomial.legendre.leggauss(nt)
    phi = (np.arange(nphi) + 0.5) * 2 * np.pi / nphi
    ct, ph = np.meshgrid(xg, phi, indexing="ij")
    st = np.sqrt(1 - ct ** 2)
    dirs = np.stack([st * np.cos(ph), st * np.sin(ph), ct], -1).reshape(-1, 3)
    w = (wg[:, None] * np.full(nphi, 2 * np.pi / nphi)[None, :]).reshape(-1)
    return dirs, w


def test_sh_basis_closed_form():
    rng = np.random.default_rng(0)
    d = rng.standard_normal((64, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    got = sh_basis_via_oracle(d.astype(np.float32).astype(np.float64))
    # directions are float32 positions / (x2) renormalised in FP64: compare on those directions
    d32 = (d * 2.0).astype(np.float32).astype(np.float64)
    d32 /= np.linalg.norm(d32, axis=1, keepdims=True)
    ref = np.array([real_sh_closed_form(*v) for v in d32])
    assert np.abs(got - ref).max() < 1e-7


def test_sh_orthonormal():
    dirs, w = sphere_quadrature()
    Y = sh_basis_via_oracle(dirs)
    # float32 positions perturb the directions slightly: loose but discriminating tolerance
    G = (Y * w[:, None]).T @ Y
    assert np.abs(G - np.eye(16)).max() < 1e-5


def _sh_scene(rng, n=20, L=3):
    sh = rng.normal(0, 0.3, (n, 3, 16))
    sg = rng.normal(0, 0.3, (n, 3, L))
    sharp = rng.uniform(1, 10, (n, L))
    ax = rng.standard_normal((n, L, 3))
    ax /= np.linalg.norm(ax, axis=2, keepdims=True)
    pos = rng.uniform(-1, 1, (n, 3))
    return make_scene(pos, np.ones((n, 3)), np.tile([1, 0, 0, 0], (n, 1)), np.ones(n), sh, sg, sharp, ax, 3, L)


def _fd(f, x, h):
    return (f(x + h) - f(x - h)) / (2 * h)


def test_sh_backward_matches_fd_exact_mode():
    rng = np.random.default_rng(1)
    sc = _sh_scene(rng)
    eye = np.array([3.0, -2.0, 0.5], np.float32)
    dL = rng.normal(0, 1, (sc.n, 3))
    g = O.sh_backward(sc, eye, dL, exact=True)

    def loss(s):
        return float((O.sh_forward(s, eye) * dL).sum())

    # the FD steps go through float32 inputs, so use steps large enough for FP32 resolution and
    # compare at the 1e-4 level
    for name, arr_name, idx in (("sh", "sh", (3, 1, 9)), ("sg", "sg", (5, 2, 1)), ("sharpness", "sharpness", (7, 2)),
                                ("lobe_axis", "lobe_axis", (2, 0, 1)), ("positions", "positions", (4, 1)),
                                ("positions", "positions", (11, 0))):
        base = getattr(sc, arr_name)

        def f(v):
            a = base.copy()
            a[idx] = v
            s = Scene(**{**sc.__dict__, arr_name: a})
            return loss(s)

        x0 = float(base[idx])
        h = 1e-2 * max(1.0, abs(x0))
        num = (8 * (f(x0 + h) - f(x0 - h)) - (f(x0 + 2 * h) - f(x0 - 2 * h))) / (12 * h)
        ana = g[name][idx]
        assert abs(num - ana) <= 2e-4 * max(1.0, abs(ana)), (name, idx, num, ana)


def test_sh_backward_reference_quirk_is_exactly_the_y9_term():
    """backward_sh writes d/dy of Y9 as C3[0]*(3xx - yy) (utils/sh_utils.cu:254); the exact
    derivative is C3[0]*3*(xx - yy).  The two modes differ by exactly that term chained to mu."""
    rng = np.random.default_rng(2)
    sc = _sh_scene(rng, n=30, L=2)
    eye = np.array([0.3, 0.2, -2.5], np.float32)
    dL = rng.normal(0, 1, (sc.n, 3))
    gr = O.sh_backward(sc, eye, dL, exact=False)
    ge = O.sh_backward(sc, eye, dL, exact=True)
    for k in ("sh", "sg", "sharpness", "lobe_axis"):
        assert np.array_equal(gr[k], ge[k])
    d0 = sc.positions.astype(np.float64) - eye.astype(np.float64)
    r = np.linalg.norm(d0, axis=1)
    v = d0 / r[:, None]
    C30 = -0.25 * math.sqrt(35 / (2 * math.pi))
    # difference in dL/ddir.y = sum_c dL_c * sh9_c * C30 * ((3xx - yy) - 3(xx - yy)) = ... * 2yy
    ddy = (dL * sc.sh[:, :, 9]).sum(1) * C30 * 2 * v[:, 1] ** 2
    ddir = np.zeros_like(d0)
    ddir[:, 1] = ddy
    proj = (ddir - v * (v * ddir).sum(1, keepdims=True)) / r[:, None]
    assert np.abs((gr["positions"] - ge["positions"]) - proj).max() < 1e-10
    assert np.abs(ddy).max() > 1e-3  # the quirk is actually exercised


# ------------------------------------------------------------------------ geometry -----------

def test_aabb_support_function():
    rng = np.random.default_rng(3)
    sc = scenes.scene_uniform(50, seed=3)
    box = O.aabb(sc)
    from ray_gauss_amd.scenes import _quat_to_R

    R = _quat_to_R(sc.quats.astype(np.float64))
    dens = sc.densities.astype(np.float64)
    delta = np.sqrt(np.maximum(0, np.log((dens / 0.1) ** 2)))
    for g in range(sc.n):
        u = rng.standard_normal((4000, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        pts = sc.positions[g] + (R[g] @ (delta[g] * sc.scales[g].astype(np.float64) * u).T).T
        assert (pts >= box[g, :3] - 1e-12).all() and (pts <= box[g, 3:] + 1e-12).all()
        # tight: the support point along each axis reaches the face
        for a in range(3):
            e = np.zeros(3)
            e[a] = 1
            # support point u* = S R^T e / |S R^T e|
            v = delta[g] * sc.scales[g] * (R[g].T @ e)
            if np.linalg.norm(v) == 0:
                continue
            us = v / np.linalg.norm(v) * delta[g] * sc.scales[g]
            top = sc.positions[g][a] + (R[g] @ us)[a]
            assert abs(top - box[g, 3 + a]) < 1e-9


def test_ray_generation_closed_form():
    cam = look_at_camera([0.0, -4.0, 0.0], width=65, height=33, fovx_deg=50.0).launch_record()
    o, d = O.rays(cam)
    assert np.abs(np.linalg.norm(d.astype(np.float64), axis=1) - 1).max() < 3e-7
    # centre pixel of an odd-sized image looks along W
    c = 16 * 65 + 32
    assert np.abs(d[c] - cam["w"]).max() < 1e-6
    # horizontal angle of the last column centre: atan(tan(fov/2) (1 - 1/W))
    r = 16 * 65 + 64
    ang = math.atan2(float(np.dot(d[r], cam["u"])), float(np.dot(d[r], cam["w"])))
    assert abs(ang - math.atan(float(cam["tanx"]) * (1 - 1 / 65))) < 1e-6
    assert (o == cam["eye"]).all()


def test_predicate_against_dense_sampling():
    rng = np.random.default_rng(4)
    n = 40
    pos = rng.uniform(-0.5, 0.5, (n, 3))
    sc = make_scene(pos, np.exp(rng.normal(-2.0, 0.4, (n, 3))), scenes._unit_quats(rng, n), rng.uniform(0.2, 3, n))
    o = np.tile([0.1, -3.0, 0.2], (200, 1))
    tgt = rng.uniform(-0.6, 0.6, (200, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    I = O.intersect(sc, o, d)
    from ray_gauss_amd.scenes import _quat_to_R

    R = _quat_to_R(sc.quats.astype(np.float64))
    delta = np.sqrt(np.log((sc.densities.astype(np.float64) / 0.1) ** 2))
    valid_dense = np.zeros((200, n), bool)
    ts = np.linspace(0, 6, 60001)
    for g in range(n):
        if sc.densities[g] <= 0.1:
            continue
        M = (R[g] / (sc.scales[g] * delta[g])).T  # rows col_k/s'_k
        for r in range(200):
            x = o[r] + ts[:, None] * d[r] - sc.positions[g]
            q = ((x @ M.T) ** 2).sum(1)
            inside = q < 1.0
            valid_dense[r, g] = inside.any()
            if I["valid"][r, g] and inside.any():
                tin, tout = ts[inside][0], ts[inside][-1]
                assert abs(tin - I["r1"][r, g]) < 2e-4 and abs(tout - I["r2"][r, g]) < 2e-4
    # the predicate agrees with sampling except for grazing rays (disc within 1e-3)
    mism = valid_dense != I["valid"]
    assert (np.abs(I["disc"][mism]) < 1e-3).all()
    assert valid_dense.sum() > 50


def test_predicate_quirk_origin_inside():
    """A ray whose origin is inside the bounding ellipsoid hits only when b > 0 (the centre lies
    ahead): forward/gaussians_aabb.cu:374-399 with q = b - sqrt(disc) for b <= 0."""
    sc = make_scene([[0, 0, 0]], [[1, 1, 1]], [[1, 0, 0, 0]], [1.0])
    delta = math.sqrt(math.log((1.0 / float(np.float32(0.1))) ** 2))  # threshold enters as float32
    o = np.array([[-0.2, 0.0, 0.0], [0.2, 0.0, 0.0]])
    d = np.array([[1.0, 0, 0], [1.0, 0, 0]])
    I = O.intersect(sc, o, d)
    assert I["valid"][0, 0] and I["b"][0, 0] > 0
    assert abs(I["r1"][0, 0] - (-delta + 0.2)) < 1e-12 and abs(I["r2"][0, 0] - (delta + 0.2)) < 1e-12
    # centre behind: disc > 0 but roots come out reversed -> never overlaps any [tmin,tmax]
    assert I["b"][1, 0] < 0 and I["r1"][1, 0] > I["r2"][1, 0]
    # |sigma| <= thr: no bounding ellipsoid
    sc2 = make_scene([[0, 0, 0], [0, 0, 0]], [[1, 1, 1]] * 2, [[1, 0, 0, 0]] * 2, [0.1, -0.05])
    I2 = O.intersect(sc2, np.array([[-5.0, 0, 0]]), np.array([[1.0, 0, 0]]))
    assert not I2["valid"].any()


# ------------------------------------------------------------------------ volume render ------

def _render_one(sc, params, cam, pixels=None):
    rgb = O.sh_forward(sc, cam["eye"])
    bmin, bmax = O.scene_box(sc, params.sigma_thr)
    return O.render(params, sc, rgb, bmin, bmax, cam, pixels), rgb


def test_single_gaussian_telescoping_and_integral():
    """One Gaussian: C = rgb (1 - T) + bg T with T = exp(-dt sum_i sigma_i) exactly, and for a
    fine grid dt sum_i sigma_i approaches the integral of sigma w over the part of the ray where
    sigma w > thr (closed form with erf)."""
    s = 0.2
    sigma = 2.0
    sh = np.zeros((1, 3, 1))
    sh[0, :, 0] = [1.0, -0.5, 0.3]
    sc = make_scene([[0, 0, 0]], [[s, s, s]], [[1, 0, 0, 0]], [sigma], sh=sh)
    cam = look_at_camera([0.0, -3.0, 0.0], width=1, height=1, fovx_deg=10.0).launch_record()
    for bg in (False, True):
        params = RenderParams(dt=1e-4, t_eps=0.0, white_background=bg)
        out, rgb = _render_one(sc, params, cam)
        T = 1.0 - out["opacity"][0]
        C = out["rgb"][0]
        bgv = 1.0 if bg else 0.0
        assert np.abs(C - (rgb[0] * (1 - T) + bgv * T)).max() < 1e-12
        # integral: w(t) = exp(-(t-tc)^2/(2 s^2)), contributing where sigma w > 0.1, i.e. |t-tc| < s*delta
        delta = math.sqrt(2 * math.log(sigma / 0.1))
        integ = sigma * s * math.sqrt(2 * math.pi) * math.erf(delta / math.sqrt(2))
        assert abs(-math.log(T) - integ) < 2e-3 * integ


def test_empty_and_missing_rays_are_background():
    sc = make_scene([[0, 0, 0]], [[0.05] * 3], [[1, 0, 0, 0]], [0.05])  # sigma below threshold
    cam = look_at_camera([0.0, -3.0, 0.0], width=8, height=8).launch_record()
    for bg in (False, True):
        out, _ = _render_one(sc, RenderParams(white_background=bg), cam)
        assert np.all(out["rgb"] == (1.0 if bg else 0.0))
        assert np.all(out["K"] == 0)


def test_early_exit_at_slab_granularity():
    """A dense Gaussian drives T below T_eps inside a slab; the march stops after that slab."""
    sh = np.zeros((1, 3, 1))
    sc = make_scene([[0, 0, 0]], [[0.3] * 3], [[1, 0, 0, 0]], [200.0], sh=sh)
    cam = look_at_camera([0.0, -3.0, 0.0], width=1, height=1, fovx_deg=10.0).launch_record()
    p_exit = RenderParams(dt=0.01, t_eps=1e-3)
    p_full = RenderParams(dt=0.01, t_eps=0.0)
    a, _ = _render_one(sc, p_exit, cam)
    b, _ = _render_one(sc, p_full, cam)
    Ta, Tb = 1 - a["opacity"][0], 1 - b["opacity"][0]
    assert Ta < 1e-3 and Tb < Ta  # stopped past the threshold, the full march goes further
    assert a["K"][0] < b["K"][0]
    assert a["K"][0] % 1 == 0


def test_interval_oracle_equals_literal_transcription():
    wl = scenes.workload("small4k")
    cam = wl.cameras[0].launch_record()
    rgb = O.sh_forward(wl.scene, cam["eye"])
    bmin, bmax = O.scene_box(wl.scene)
    px = np.arange(0, 4096, 13).astype(np.int32)
    a = O.render(wl.params, wl.scene, rgb, bmin, bmax, cam, px)
    b = O.render_literal(wl.params, wl.scene, rgb, bmin, bmax, cam, px)
    assert np.array_equal(a["K"], b["K"])
    assert np.abs(a["rgb"] - b["rgb"]).max() < 1e-14
    assert a["maxhits"].max() < 1024  # the reference's max_prim_slice is never reached


# ------------------------------------------------------------------------ backward vs FD -----

def _fd_scene(seed=5, n=60, L=2):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-0.4, 0.4, (n, 3))
    sc = make_scene(pos, np.exp(rng.normal(-1.8, 0.3, (n, 3))), scenes._unit_quats(rng, n), rng.uniform(0.5, 3.0, n),
                    sh=rng.normal(0, 0.3, (n, 3, 16)), sg=rng.normal(0, 0.3, (n, 3, L)),
                    sharp=rng.uniform(1, 5, (n, L)), axis=scenes._unit_vectors(rng, (n, L)), deg=3, L=L)
    return sc


def _loss_fn(sc, params, cam, dL, pixels):
    rgb = O.sh_forward(sc, cam["eye"])
    bmin = np.array([-2, -2, -2], np.float32)
    bmax = np.array([2, 2, 2], np.float32)
    out = O.render(params, sc, rgb, bmin, bmax, cam, pixels)
    return float((out["rgb"] * dL).sum()), out


def test_backward_matches_central_fd():
    """Oracle backward (the reference's per-hit formulas, backward/gaussians_aabb.cu:628-763 and
    utils/sh_utils.cu:121-315) against FP64 central finite differences of the oracle forward.
    Densities large enough that the threshold region is well inside the sampled chords; samples
    whose threshold decision changes under the perturbation are excluded by checking fragility
    and K on both sides."""
    sc = _fd_scene()
    params = RenderParams(dt=0.004, t_eps=0.0, white_background=True)
    cam = look_at_camera([0.3, -2.5, 0.4], width=16, height=16, fovx_deg=30.0).launch_record()
    px = np.arange(256, dtype=np.int32)
    rng = np.random.default_rng(6)
    dL = rng.normal(0, 1, (256, 3))
    L0, out0 = _loss_fn(sc, params, cam, dL, px)
    bmin = np.array([-2, -2, -2], np.float32)
    bmax = np.array([2, 2, 2], np.float32)
    g = O.full_backward(params, sc, cam["eye"], bmin, bmax, cam, out0["rgb"], dL, px, exact_sh=True)
    checked = 0
    for name, shape_idx in (("positions", 3), ("scales", 3), ("quats", 4), ("densities", 1), ("sh", 48), ("sg", 6),
                            ("sharpness", 2), ("lobe_axis", 6)):
        arr = getattr(sc, name)
        gg = g[name]
        # pick the parameters with the largest gradients (those actually seen by rays)
        order = np.argsort(-np.abs(gg).reshape(-1))[:10]
        for flat in order:
            idx = np.unravel_index(flat, arr.shape)
            x0 = float(arr[idx])
            h = 5e-4 * max(abs(x0), 0.1)
            vals = []
            for sgn in (1, -1):
                a = arr.copy()
                a[idx] = np.float32(x0 + sgn * h)
                hh = float(a[idx]) - x0
                s2 = Scene(**{**sc.__dict__, name: a})
                Lv, o = _loss_fn(s2, params, cam, dL, px)
                vals.append((Lv, hh, o))
            (Lp, hp, op), (Lm, hm, om) = vals
            same = all(np.array_equal(o_[k], out0[k]) for o_ in (op, om) for k in ("K", "NC"))
            if not same:
                continue
            num = (Lp - Lm) / (hp - hm)
            ana = gg[idx]
            scale = max(abs(ana), 1e-3 * np.abs(gg).max())
            assert abs(num - ana) <= 1e-3 * scale, (name, idx, num, ana)
            checked += 1
    assert checked >= 30


# ---- random generator of jitter / rnd_sample (forward/gaussians_aabb.h:63-94) ----------------

def test_tea_is_the_published_cipher():
    """With 32 cycles and an all-zero key and block, TEA's published test vector is 0x41ea3a0a
    for the first word (Wheeler & Needham's cipher; the reference's tea<N> uses the same rounds
    with a fixed key and N cycles)."""
    assert O.tea(0, 0, 32, [0, 0, 0, 0]) == 0x41EA3A0A


def test_tea4_uses_the_reference_key():
    key = [0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E]
    for v0, v1 in [(0, 0), (7, 123456), (2 ** 32 - 1, 99)]:
        assert O.tea4(v0, v1) == O.tea(v0, v1, 4, key)
    assert len({O.tea4(5, px) for px in range(1000)}) == 1000  # distinct seeds per pixel


def test_lcg_closed_form():
    """x_{k+1} = 1664525 x_k + 1013904223 mod 2^32 (Numerical Recipes' generator); the output is
    the low 24 bits, and rnd = output / 2^24 is exact in float32."""
    ints, floats = O.lcg(12345, 50)
    x = 12345
    for k in range(50):
        x = (1664525 * x + 1013904223) % 2 ** 32
        assert ints[k] == x & 0xFFFFFF
        assert floats[k] == np.float32(ints[k]) / np.float32(2 ** 24)
    assert floats.min() >= 0.0 and floats.max() < 1.0


def test_jitter_moves_rays_within_the_pixel():
    wl = scenes.workload("small4k")
    cam = wl.cameras[0].launch_record()
    o0, d0 = O.rays(cam)
    o1, d1 = O.rays_rng(cam, None, jitter=0, iteration=3)
    assert np.array_equal(d0, d1)  # no jitter: the generator is seeded but unused
    o2, d2 = O.rays_rng(cam, None, jitter=1, iteration=3)
    o3, d3 = O.rays_rng(cam, None, jitter=1, iteration=4)
    assert not np.array_equal(d2, d0) and not np.array_equal(d2, d3)
    # the jittered sample point stays inside the pixel: [x + 1/3, x + 2/3)
    W, H = cam["width"], cam["height"]
    u = np.asarray(cam["u"]); v = np.asarray(cam["v"]); w = np.asarray(cam["w"])
    dd = d2.astype(np.float64)
    a = dd @ u / (dd @ w) / cam["tanx"]
    b = dd @ v / (dd @ w) / cam["tany"]
    fx = (a + 1) / 2 * W - np.tile(np.arange(W), H)
    fy = (b + 1) / 2 * H - np.repeat(np.arange(H), W)
    assert fx.min() > 0.333 - 1e-3 and fx.max() < 0.667 + 1e-3
    assert fy.min() > 0.333 - 1e-3 and fy.max() < 0.667 + 1e-3


def test_sample_offsets_keep_the_single_gaussian_closed_form():
    """With random (rnd_sample = 1) or midpoint (2) sample offsets the samples move, but one
    Gaussian's image still telescopes to C = rgb (1 - T); random offsets depend on the iteration
    that seeds the generator, midpoint offsets do not."""
    wl = scenes.workload("small4k")
    cam = wl.cameras[0].launch_record()
    sc = wl.scene
    one = scenes.Scene(positions=sc.positions[:1] * 0, scales=np.full((1, 3), 0.2, np.float32),
                       quats=np.array([[1, 0, 0, 0]], np.float32), densities=np.array([3.0], np.float32),
                       sh=sc.sh[:1], sg=sc.sg[:1], sharpness=sc.sharpness[:1], lobe_axis=sc.lobe_axis[:1])
    px = np.array([32 * 64 + 32, 30 * 64 + 33], np.int32)
    for mode in (1, 2):
        prm = scenes.RenderParams(**{**wl.params.__dict__, "rnd_sample": mode, "iteration": 11})
        bmin, bmax = O.scene_box(one, prm.sigma_thr)
        rgb = O.sh_forward(one, cam["eye"], 1)
        out = O.render(prm, one, rgb, bmin, bmax, cam, px, 1)
        T = 1.0 - out["opacity"]
        assert T.max() < 0.99
        np.testing.assert_allclose(out["rgb"], rgb[0][None, :] * (1.0 - T[:, None]), rtol=1e-12, atol=1e-12)
        prm2 = scenes.RenderParams(**{**prm.__dict__, "iteration": 12})
        out2 = O.render(prm2, one, rgb, bmin, bmax, cam, px, 1)
        base = O.render(scenes.RenderParams(**{**prm.__dict__, "rnd_sample": 0}), one, rgb, bmin, bmax, cam, px, 1)
        assert not np.allclose(out["rgb"], base["rgb"], rtol=0, atol=1e-9)
        assert np.array_equal(out["rgb"], out2["rgb"]) == (mode == 2)


# ---- spherical-Gaussian colour term and depth: closed forms ----------------------------------

def test_sg_term_closed_form():
    """rgb - 0.5 - SH = sum_l a[c,l] exp(lambda_l (nu_l . v - 1)) per channel, v the unit view
    direction (utils/sh_utils.cu:43-54), with layouts sg (n,3,L) = [g, channel, lobe], sharpness
    (n,L), lobe_axis (n,L,3).  Lobe axes at v, -v and perpendicular to v make the three terms
    exactly a, a exp(-2 lambda) and a exp(-lambda); every (channel, lobe) amplitude and every
    lobe's sharpness is distinct, so a transposed index or a wrong exponent changes the result."""
    eye = np.array([0.5, -2.0, 0.25], np.float32)
    pos = np.array([[0.25, 1.0, -0.5]], np.float32)
    v = pos[0].astype(np.float64) - eye.astype(np.float64)
    v /= np.linalg.norm(v)
    perp = np.cross(v, [0.0, 0.0, 1.0])
    perp /= np.linalg.norm(perp)
    axes = np.stack([v, -v, perp])[None]  # (1, L=3, 3)
    lam = np.array([[1.5, 0.75, 2.5]])
    a = np.array([[[0.11, -0.23, 0.37], [0.41, 0.05, -0.29], [-0.17, 0.31, 0.13]]])  # [g, c, l]
    sh = np.zeros((1, 3, 16))
    sh[0, :, 0] = [0.2, -0.1, 0.05]
    sc = make_scene(pos, [[1, 1, 1]], [[1, 0, 0, 0]], [1.0], sh=sh, sg=a, sharp=lam, axis=axes, deg=3, L=3)
    got = O.sh_forward(sc, eye)[0] - 0.5 - 0.28209479177387814 * sc.sh[0, :, 0].astype(np.float64)
    # the float32 axes are not exactly +-v / perpendicular: evaluate nu.v with them
    ax32 = sc.lobe_axis[0].astype(np.float64)
    lam32 = sc.sharpness[0].astype(np.float64)
    a32 = sc.sg[0].astype(np.float64)
    dots = ax32 @ v
    assert abs(dots[0] - 1) < 1e-6 and abs(dots[1] + 1) < 1e-6 and abs(dots[2]) < 1e-6
    special = np.exp(lam32 * (np.array([1.0, -1.0, 0.0]) - 1.0))  # 1, exp(-2 lam), exp(-lam)
    np.testing.assert_allclose(got, a32 @ special, rtol=0, atol=2e-6)
    exact = a32 @ np.exp(lam32 * (dots - 1.0))
    np.testing.assert_allclose(got, exact, rtol=0, atol=1e-12)
    # the transposed layout a[g, l, c] gives a different colour
    assert np.abs(a32.T @ special - got).max() > 1e-2


def test_depth_single_gaussian_analytic_sequence():
    """depth = sum_i T_i alpha_i t_i (cuda_gui/gaussians_aabb.cu:400-417) for one Gaussian centred
    on the optical axis, from the analytic density sequence sigma_i = sigma
    exp(-(t_i - t_c)^2 / (2 s_y^2)) (zero where sigma_i <= thr) on the FP32 sample grid
    t_i = tb + i dt, tb stepping from the scene-box entry by float32(dt * 8); t_c = |mu - eye|.
    The quaternion (1/2, 1/2, 1/2, 1/2) (exactly unit in float32) maps local axis 0 to world y,
    the viewing direction, so s_y = s_0; the transposed rotation would pick s_2 instead.  The
    oracle evaluates the weights through the general ellipsoid machinery, so the check pins it,
    the slab loop and the depth sum.  As the density goes to zero, depth / opacity tends to t_c
    by symmetry."""
    s3 = [0.1, 0.15, 0.2]
    cam = look_at_camera([0.0, -3.0, 0.0], width=1, height=1, fovx_deg=10.0).launch_record()
    for sigma in (0.6, 5.0):
        sc = make_scene([[0, 0, 0]], [s3], [[0.5, 0.5, 0.5, 0.5]], [sigma])
        params = RenderParams(dt=0.002, t_eps=0.0)
        out, _ = _render_one(sc, params, cam)
        bmin, bmax = O.scene_box(sc, params.sigma_thr)
        te, tx = O.ray_interval(cam, bmin, bmax, params.t_near, params.t_far)
        o, d = O.rays(cam)
        assert np.array_equal(d[0], np.array([0, 1, 0], np.float32))
        tc = float(np.linalg.norm(-o[0].astype(np.float64)))
        sy = float(sc.scales[0, 0])
        dt = np.float32(params.dt)
        slab = np.float32(dt * np.float32(8))
        sig32 = float(sc.densities[0])
        T, depth = 1.0, 0.0
        tb = np.float32(te[0])
        while tb < tx[0]:
            for i in range(8):
                t = float(np.float32(tb + np.float32(np.float32(i) * dt)))
                si = sig32 * math.exp(-((t - tc) ** 2) / (2 * sy ** 2))
                si = si if si > float(np.float32(0.1)) else 0.0
                alpha = 1.0 - math.exp(-si * float(dt))
                depth += T * alpha * t
                T *= 1.0 - alpha
            tb = np.float32(tb + slab)
        assert 1 - T > 1e-3
        assert abs(out["depth"][0] - depth) < 1e-11 * max(1.0, depth), (sigma, out["depth"][0], depth)
        assert abs(out["opacity"][0] - (1 - T)) < 1e-11
    # low-density limit: the expected termination distance is the centre
    sc = make_scene([[0, 0, 0]], [s3], [[1, 0, 0, 0]], [0.2])
    out, _ = _render_one(sc, RenderParams(dt=0.0005, t_eps=0.0), cam)
    assert out["opacity"][0] > 0
    assert abs(out["depth"][0] / out["opacity"][0] - 3.0) < 2e-3


# ---- fragility: measured FP32 margin, explicit per-ray bounds, decision-matched mode ----------

def cpu_record(scene, thr=0.1):
    """FP32 per-Gaussian records computed on the host the way the GPU preprocessing forms them
    (rows of M' = col_k(R) / (s_k delta), h = ln|sigma / thr|); the oracle's self-tests use these
    where the GPU tests use the kernels' own exported records."""
    f = np.float32
    q = scene.quats.astype(f)
    r, i, j, k = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    two, one = f(2), f(1)
    c0 = np.stack([one - two * (j * j + k * k), two * (i * j + r * k), two * (i * k - r * j)], 1)
    c1 = np.stack([two * (i * j - r * k), one - two * (i * i + k * k), two * (j * k + r * i)], 1)
    c2 = np.stack([two * (i * k + r * j), two * (j * k - r * i), one - two * (i * i + j * j)], 1)
    ratio = scene.densities.astype(f) / f(thr)
    with np.errstate(divide="ignore", invalid="ignore"):
        l2 = np.log(ratio * ratio).astype(f)
        valid = (l2 > 0) & np.isfinite(l2)
        delta = np.where(valid, np.sqrt(np.where(valid, l2, f(1))), f(0)).astype(f)
        isp = np.where(valid[:, None], one / (scene.scales.astype(f) * delta[:, None]), f(0)).astype(f)
    Mp = np.concatenate([c0 * isp[:, :1], c1 * isp[:, 1:2], c2 * isp[:, 2:3]], 1).astype(f)
    h = np.where(valid, f(0.5) * l2, f(0)).astype(f)
    return {"mu": scene.positions.astype(f), "Mp": Mp, "valid": valid.astype(np.int32), "h": h}


def _crop_case(name):
    wl = scenes.workload(name)
    v = {"small4k": 0, "blobs100k": 0, "unbounded1m_8views": 3}[name]
    cam = wl.cameras[v].launch_record()
    if name == "small4k":
        px = np.arange(0, 4096, 2, dtype=np.int32)
    else:
        x0, y0 = {"blobs100k": (384, 384), "unbounded1m_8views": (600, 380)}[name]
        px = scenes.crop_pixels(wl.cameras[v], x0, y0, 32, 32)
    return wl, cam, px


@pytest.mark.parametrize("name", ["small4k", "blobs100k", "unbounded1m_8views"])
def test_fp32_margin_covers_measured_error(name):
    """The FP32 error of sigma*w within 1 % of the threshold, for the reference's own FP32
    evaluation and for the HIP path's closed form, measured against FP64 on seeded crops: the
    fragility margin EPS_W is at least twice the largest seen (the 3M-Gaussian workload measures
    7.8e-5 and 1.5e-4, the largest of all)."""
    wl, cam, px = _crop_case(name)
    bmin, bmax = O.scene_box(wl.scene, wl.params.sigma_thr)
    e = O.sigmaw_fp32_error(wl.params, wl.scene, bmin, bmax, cam, px)
    assert e["samples"] > 1000
    assert 0 < e["reference"] <= O.EPS_W / 2 and 0 < e["closed_form"] <= O.EPS_W / 2, e


@pytest.mark.parametrize("name", ["small4k", "blobs100k"])
def test_decision_matched_hit_sets_equal_fp32_brute_force(name):
    """In the decision-matched mode each slab's hit set is the FP32 record predicate's: with no
    early exit, the oracle's per-ray slab-hit count K equals the FP32 brute force over all
    Gaussians (the near-miss candidates cover every FP32-only hit)."""
    wl, cam, px = _crop_case(name)
    prm = scenes.RenderParams(**{**wl.params.__dict__, "t_eps": 0.0})
    rec = cpu_record(wl.scene, prm.sigma_thr)
    bmin, bmax = O.scene_box(wl.scene, prm.sigma_thr)
    rgb = O.sh_forward(wl.scene, cam["eye"])
    out = O.render(prm, wl.scene, rgb, bmin, bmax, cam, px, record=rec)
    o, d = O.rays(cam, px)
    te, tx = O.ray_interval(cam, bmin, bmax, prm.t_near, prm.t_far, px)
    _, counts = O.hits_fp32(rec["mu"], rec["Mp"], rec["valid"], o, d, te, tx, np.float32(prm.dt),
                            prm.samples_per_slab, 100000, 1)
    assert counts.sum() > 1000
    assert np.array_equal(out["K"], counts)


@pytest.mark.parametrize("name", ["blobs100k", "unbounded1m_8views"])
def test_fp32_decisions_stay_within_the_fragile_bounds(name):
    """An FP32-decision render (the oracle's decision-matched mode, deciding like the HIP path)
    against the exact-decision render: rays with no decision within EPS_W of the threshold agree
    to FP64 rounding, and EVERY ray lies within its explicit bound (1e-4 plus, per near-threshold
    decision, 4 thr (1 + EPS_W) dt max(cmax, 1)); gradients with no row masked agree to 1e-3 of
    each tensor's largest entry for the 100k-Gaussian crop and to 5e-3 for the 1M-Gaussian crop (its
    FP32 decisions move the scale gradient by 1.6e-3: no FP32 implementation, the reference's
    included, can meet 1e-3 against exact decisions there, which is why the GPU's gradients are
    held to 1e-3 against the decision-matched mode)."""
    wl, cam, px = _crop_case(name)
    prm = wl.params
    rec = cpu_record(wl.scene, prm.sigma_thr)
    bmin, bmax = O.scene_box(wl.scene, prm.sigma_thr)
    rgb = O.sh_forward(wl.scene, cam["eye"])
    ex = O.render(prm, wl.scene, rgb, bmin, bmax, cam, px)
    dm = O.render(prm, wl.scene, rgb, bmin, bmax, cam, px, record=rec)
    err = np.abs(dm["rgb"] - ex["rgb"]).max(axis=1)
    safe = ex["fragile"] == 0
    assert err[safe].max() <= 1e-10
    bound = O.fragile_bound(prm, ex)
    assert np.isfinite(bound).all()
    assert (err <= bound).all(), (err - bound).max()
    if name == "unbounded1m_8views":
        assert (err > 1e-10).sum() > 0  # the FP32 decisions do flip some of the fragile rays
    dL = np.random.default_rng(4).normal(0, 1, (len(px), 3))
    g_ex = O.full_backward(prm, wl.scene, cam["eye"], bmin, bmax, cam, ex["rgb"], dL, px)
    g_dm = O.full_backward(prm, wl.scene, cam["eye"], bmin, bmax, cam, ex["rgb"], dL, px, record=rec)
    for k in ("positions", "scales", "quats", "densities", "sh", "sg", "sharpness", "lobe_axis"):
        e = np.abs(g_dm[k] - g_ex[k]).max() / np.abs(g_ex[k]).max()
        assert e <= (1e-3 if name == "blobs100k" else 5e-3), (k, e)
