"""Host object around one library handle: owns the device workspace of the ray caster and
exposes forward / backward / batched / fused-train calls on torch tensors.

All calls are enqueued on torch's current HIP stream; none synchronises the host except
``stats()`` and ``scene_box(host=True)``.
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .cameras import launch_record_from


@dataclass
class TuneConfig:
    """Launch tuning: interval-list depth per ray, initial slabs per gather, whether a wavefront's
    lanes walk the LBVH as one packet, the forward kernel's own (smaller) list depth, and the
    words per ray in which a forward records its hit lists for the backward over the same rays
    (0: the backward walks the tree again; 1536 words = 6 KB of device memory per ray), and the
    tree built at a rebuild: 0, the LBVH in Morton order (about 9 ms at 1M Gaussians with the
    preprocessing, cheap enough to rebuild on every call as the reference does); 1, a binned-SAH
    tree (about 70 ms, 11 % faster forward at 1M Gaussians), for loops that build once and then
    refit every step.  ray_chunk: 0 launches the forward as one wavefront per 8x8 tile; a multiple
    of 64 makes it persistent, wavefronts taking that many rays at a time from per-XCD work queues
    and stealing from the others' when their own is empty (64: 3 % faster at 1M Gaussians);
    refill_lanes: in that mode the forward hands new rays to idle lanes once that many are idle
    (0: only when all its lanes are done; refilling was measured 2.5x slower, see the kernel);
    bwd_ray_chunk: the same for the backward (persistent measured 9 % slower at 1M Gaussians);
    bwd_window: the backward's depth synchronisation, in slabs: lanes within this distance of the
    wave's laggard prepare their next slab together and merge their gradients (0.5 is best at 1M
    Gaussians, 1.0 on the 300k-Gaussian batch of uneven views); fwd_jump: after a round in which no
    lane of a wavefront had a hit, the forward's lanes step together over the slabs empty for
    them instead of one slab per round (2 % faster at 1M Gaussians; the backward always does)."""

    hit_depth: int = 16
    seg_slabs: int = 32
    packet: bool = True
    fwd_hit_depth: int = 12
    replay_words: int = 1536
    tree_builder: int = 0
    ray_chunk: int = 64
    refill_lanes: int = 0
    bwd_ray_chunk: int = 0
    bwd_window: float = 0.5
    fwd_jump: bool = True


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _camera_struct(cam) -> _lib.CameraS:
    rec = cam if isinstance(cam, dict) else launch_record_from(cam)
    c = _lib.CameraS()
    for k in ("eye", "u", "v", "w"):
        getattr(c, k)[:] = [float(x) for x in np.asarray(rec[k], np.float32).reshape(3)]
    c.tanx = float(np.float32(rec["tanx"]))
    c.tany = float(np.float32(rec["tany"]))
    c.width = int(rec["width"])
    c.height = int(rec["height"])
    return c


def _f32c(t, shape=None):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("ray caster inputs must be device tensors")
    t = t.detach()
    if t.dtype != torch.float32:
        raise TypeError(f"expected float32, got {t.dtype}")
    t = t.contiguous()
    if shape is not None and tuple(t.shape) != tuple(shape):
        t = t.reshape(shape)
    return t


class RayCaster:
    """One handle of the gfx950 ray caster.

    params: an object with dt, samples_per_slab, t_eps, sigma_thr, t_near, t_far,
    white_background (e.g. :class:`ray_gauss_amd.scenes.RenderParams`).
    """

    def __init__(self, params, tune: TuneConfig | None = None, exact_sh_grad: bool = False, max_hits_per_slab: int = 0):
        self._L = _lib.load()
        self.params = params
        self.tune = tune or TuneConfig()
        self.exact_sh_grad = bool(exact_sh_grad)
        self.max_hits_per_slab = int(max_hits_per_slab)
        h = C.c_void_p()
        cfg = self._config()
        rc = self._L.rg_create(C.byref(h), C.byref(cfg))
        if rc != 0:
            raise _lib.RayGaussError(f"rg_create failed ({rc})")
        self._h = h
        self.last_token = 0  # hit-list record token of the last forward (0: none recorded)
        self._inputs = None
        self._cams = []
        self.generation = 0
        self.n = 0
        self.sh_degree = 0
        self.num_sg = 0

    def _config(self) -> _lib.Config:
        p = self.params
        c = _lib.Config()
        c.dt = float(p.dt)
        c.samples_per_slab = int(p.samples_per_slab)
        c.t_eps = float(p.t_eps)
        c.sigma_thr = float(p.sigma_thr)
        c.t_near = float(p.t_near)
        c.t_far = float(p.t_far) if math.isfinite(p.t_far) else float("inf")
        c.white_background = int(bool(p.white_background))
        c.exact_sh_grad = int(self.exact_sh_grad)
        c.hit_depth = int(self.tune.hit_depth)
        c.seg_slabs = int(self.tune.seg_slabs)
        c.max_hits_per_slab = int(self.max_hits_per_slab)
        c.packet_traversal = int(bool(self.tune.packet))
        c.fwd_hit_depth = int(self.tune.fwd_hit_depth)
        c.replay_words = int(getattr(self.tune, "replay_words", 0))
        c.tree_builder = int(getattr(self.tune, "tree_builder", 0))
        c.ray_chunk = int(getattr(self.tune, "ray_chunk", 0))
        c.refill_lanes = int(getattr(self.tune, "refill_lanes", 0))
        c.bwd_ray_chunk = int(getattr(self.tune, "bwd_ray_chunk", 0))
        c.bwd_window = float(getattr(self.tune, "bwd_window", 0.5))
        c.fwd_round_per_slab = 0 if getattr(self.tune, "fwd_jump", True) else 1
        c.jitter = int(getattr(p, "jitter", 0))
        c.rnd_sample = int(getattr(p, "rnd_sample", 0))
        c.iteration = int(getattr(p, "iteration", 0) or 0) & 0xFFFFFFFF
        return c

    def set_tune(self, tune: TuneConfig):
        self.tune = tune
        cfg = self._config()
        _lib.check(self._L.rg_set_config(self._h, C.byref(cfg)), self._h)

    def close(self):
        if getattr(self, "_h", None):
            self._L.rg_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- Gaussians / views -------
    def prepare(self, positions, scales, quats, densities, sh, sg=None, sharpness=None, lobe_axis=None,
                rebuild: bool = False):
        """Preprocess the Gaussians and build (rebuild=True / size change) or refit the LBVH.
        sh: (N,3,K); sg: (N,3,L); sharpness: (N,L); lobe_axis: (N,L,3)."""
        n = positions.shape[0]
        K = sh.shape[2]
        deg = int(round(math.sqrt(K))) - 1
        if (deg + 1) ** 2 != K or deg > 3:
            raise ValueError("color_features must have (deg+1)^2 coefficients, deg <= 3")
        L = 0 if sg is None else int(sg.shape[2]) if sg.dim() == 3 else 0
        t = {
            "positions": _f32c(positions, (n, 3)), "scales": _f32c(scales, (n, 3)), "quats": _f32c(quats, (n, 4)),
            "densities": _f32c(densities, (n,)), "sh": _f32c(sh, (n, 3, K)),
            "sg": _f32c(sg, (n, 3, L)) if L else None, "sharpness": _f32c(sharpness, (n, L)) if L else None,
            "lobe_axis": _f32c(lobe_axis, (n, L, 3)) if L else None,
        }
        g = _lib.Gaussians()
        g.n = n
        g.sh_degree = deg
        g.num_sg = L
        for k, v in t.items():
            setattr(g, k, None if v is None else v.data_ptr())
        _lib.check(self._L.rg_prepare(self._h, C.byref(g), int(bool(rebuild)), _stream()), self._h)
        if self._inputs is None or n != self.n:
            self._cams = []
        self._inputs = t
        self.n, self.sh_degree, self.num_sg = n, deg, L
        self.generation += 1
        if self._cams:
            # refit keeps the colour tables current for the stored views
            pass

    def set_views(self, cams):
        """cams: list of reference-style camera objects or FP32 launch records."""
        arr = (_lib.CameraS * len(cams))(*[_camera_struct(c) for c in cams])
        _lib.check(self._L.rg_set_views(self._h, len(cams), arr, _stream()), self._h)
        self._cams = [c if isinstance(c, dict) else launch_record_from(c) for c in cams]

    def set_view_colors(self, view: int, rgb):
        """Replace view `view`'s per-Gaussian colours with rgb (n,3) computed by the caller."""
        if rgb.numel() != 3 * self.n:
            raise ValueError(f"set_view_colors: expected {self.n} x 3 colours, got shape {tuple(rgb.shape)}")
        rgb = _f32c(rgb, (self.n, 3))
        _lib.check(self._L.rg_set_view_colors(self._h, int(view), C.c_void_p(rgb.data_ptr()), _stream()), self._h)
        self._keep_colors = rgb

    @property
    def views(self):
        return self._cams

    # ---------------------------------------------------------------- single view -------------
    def forward(self, view: int = 0, pixels=None, depth: bool = False, opacity: bool = False):
        rec = self._cams[view]
        nr = rec["width"] * rec["height"] if pixels is None else int(pixels.shape[0])
        dev = self._inputs["positions"].device
        out = torch.empty((nr, 3), dtype=torch.float32, device=dev)
        d = torch.empty(nr, dtype=torch.float32, device=dev) if depth else None
        o = torch.empty(nr, dtype=torch.float32, device=dev) if opacity else None
        px = None if pixels is None else pixels.to(device=dev, dtype=torch.int32).contiguous()
        tok = C.c_uint64(0)
        _lib.check(self._L.rg_forward(self._h, view, nr, _ptr(px), _ptr(out), _ptr(d), _ptr(o), C.byref(tok),
                                      _stream()), self._h)
        self.last_token = int(tok.value)
        return out, d, o

    def backward(self, view: int, pixels, ray_colors, dLdC, replay_token: int = 0):
        """Backward of forward(view, pixels); replay_token = the forward's ``last_token`` lets the
        backward read that forward's recorded hit lists instead of walking the tree (0: walk)."""
        rec = self._cams[view]
        nr = rec["width"] * rec["height"] if pixels is None else int(pixels.shape[0])
        dev = self._inputs["positions"].device
        px = None if pixels is None else pixels.to(device=dev, dtype=torch.int32).contiguous()
        rc = ray_colors.detach().contiguous().float()
        dl = dLdC.detach().contiguous().float()
        if rc.shape != (nr, 3) or dl.shape != (nr, 3):
            raise ValueError("ray_colors and dL/dC must be (num_rays, 3)")
        _lib.check(self._L.rg_backward(self._h, view, nr, _ptr(px), _ptr(rc), _ptr(dl), int(replay_token),
                                       _stream()), self._h)

    # ---------------------------------------------------------------- batches -----------------
    @staticmethod
    def _batch(ray_view, ray_pixel):
        b = _lib.RayBatch()
        b.num_rays = int(ray_view.shape[0])
        b.ray_view = ray_view.data_ptr()
        b.ray_pixel = ray_pixel.data_ptr()
        return b

    def forward_batch(self, ray_view, ray_pixel, depth=False, opacity=False):
        nr = int(ray_view.shape[0])
        dev = ray_view.device
        out = torch.empty((nr, 3), dtype=torch.float32, device=dev)
        d = torch.empty(nr, dtype=torch.float32, device=dev) if depth else None
        o = torch.empty(nr, dtype=torch.float32, device=dev) if opacity else None
        b = self._batch(ray_view, ray_pixel)
        tok = C.c_uint64(0)
        _lib.check(self._L.rg_forward_batch(self._h, C.byref(b), _ptr(out), _ptr(d), _ptr(o), C.byref(tok),
                                            _stream()), self._h)
        self.last_token = int(tok.value)
        return out, d, o

    def backward_batch(self, ray_view, ray_pixel, ray_colors, dLdC, replay_token: int = 0):
        b = self._batch(ray_view, ray_pixel)
        _lib.check(self._L.rg_backward_batch(self._h, C.byref(b), _ptr(ray_colors.contiguous()),
                                             _ptr(dLdC.contiguous()), int(replay_token), _stream()), self._h)

    def train_batch(self, ray_view, ray_pixel, target, loss_kind=0, out_rgb=None, out_loss=None):
        b = self._batch(ray_view, ray_pixel)
        _lib.check(self._L.rg_train_batch(self._h, C.byref(b), _ptr(target.contiguous()), int(loss_kind),
                                          _ptr(out_rgb), _ptr(out_loss), _stream()), self._h)

    # ---------------------------------------------------------------- gradients ---------------
    def zero_grads(self):
        _lib.check(self._L.rg_zero_grads(self._h, _stream()), self._h)

    def finish_grads(self):
        """Gradients of the nine boundary inputs (original order), accumulated since the last call."""
        t = self._inputs
        n, L = self.n, self.num_sg
        K = (self.sh_degree + 1) ** 2
        dev = t["positions"].device
        z = lambda *s: torch.empty(s, dtype=torch.float32, device=dev)  # noqa: E731
        g = {"positions": z(n, 3), "scales": z(n, 3), "quats": z(n, 4), "densities": z(n), "sh": z(n, 3, K),
             "sg": z(n, 3, L) if L else None, "sharpness": z(n, L) if L else None,
             "lobe_axis": z(n, L, 3) if L else None}
        s = _lib.Grads()
        for k, v in g.items():
            setattr(s, k, None if v is None else v.data_ptr())
        _lib.check(self._L.rg_finish_grads(self._h, C.byref(s), _stream()), self._h)
        return g

    # ---------------------------------------------------------------- inspection --------------
    def scene_box(self):
        dev = self._inputs["positions"].device
        out = torch.empty(6, dtype=torch.float32, device=dev)
        _lib.check(self._L.rg_scene_box(self._h, _ptr(out), _stream()), self._h)
        return out

    def stats(self) -> dict:
        s = _lib.Stats()
        _lib.check(self._L.rg_get_stats(self._h, C.byref(s), _stream()), self._h)
        return {k: int(getattr(s, k)) for k, _ in s._fields_}

    def export_predicate(self):
        dev = self._inputs["positions"].device
        n = self.n
        mu = torch.empty((n, 3), dtype=torch.float32, device=dev)
        Mp = torch.empty((n, 9), dtype=torch.float32, device=dev)
        valid = torch.empty(n, dtype=torch.int32, device=dev)
        _lib.check(self._L.rg_export_predicate(self._h, _ptr(mu), _ptr(Mp), _ptr(valid), None, _stream()), self._h)
        return mu, Mp, valid

    def export_record(self) -> dict:
        """The FP32 per-Gaussian records the kernels decide with (original order), as numpy
        arrays: mu (n,3), Mp (n,9), valid (n), h (n) = ln|sigma / thr|."""
        dev = self._inputs["positions"].device
        n = self.n
        mu = torch.empty((n, 3), dtype=torch.float32, device=dev)
        Mp = torch.empty((n, 9), dtype=torch.float32, device=dev)
        valid = torch.empty(n, dtype=torch.int32, device=dev)
        h = torch.empty(n, dtype=torch.float32, device=dev)
        _lib.check(self._L.rg_export_predicate(self._h, _ptr(mu), _ptr(Mp), _ptr(valid), _ptr(h), _stream()), self._h)
        return {"mu": mu.cpu().numpy(), "Mp": Mp.cpu().numpy(), "valid": valid.cpu().numpy(), "h": h.cpu().numpy()}

    def debug_hits(self, view, pixels, max_slabs, max_pairs):
        dev = self._inputs["positions"].device
        px = pixels.to(device=dev, dtype=torch.int32).contiguous()
        nr = int(px.shape[0])
        pairs = torch.full((nr, max_pairs, 2), -1, dtype=torch.int32, device=dev)
        counts = torch.empty(nr, dtype=torch.int64, device=dev)
        ro = torch.empty((nr, 3), dtype=torch.float32, device=dev)
        rd = torch.empty((nr, 3), dtype=torch.float32, device=dev)
        te = torch.empty(nr, dtype=torch.float32, device=dev)
        tx = torch.empty(nr, dtype=torch.float32, device=dev)
        _lib.check(self._L.rg_debug_hits(self._h, view, nr, _ptr(px), max_slabs, max_pairs, _ptr(pairs), _ptr(counts),
                                         _ptr(ro), _ptr(rd), _ptr(te), _ptr(tx), _stream()), self._h)
        return pairs, counts, ro, rd, te, tx

    def alloc_epoch(self) -> int:
        """Changes whenever the handle (re)allocates device memory (graphs over it are then stale)."""
        return int(self._L.rg_alloc_epoch(self._h))

    def tree_cost(self) -> float:
        """Surface-area cost of the current tree (sum of node box areas over the root's);
        synchronises the stream."""
        c = C.c_double()
        _lib.check(self._L.rg_tree_cost(self._h, C.byref(c), _stream()), self._h)
        return float(c.value)

    def device_bytes(self) -> dict:
        """Device memory the handle holds: {"total": bytes, "records": the hit-list records'
        share (rays x replay_words x 4 B + 12 B per ray)}."""
        rec = C.c_int64()
        total = int(self._L.rg_device_bytes(self._h, C.byref(rec)))
        return {"total": total, "records": int(rec.value)}

    def check_bvh(self) -> int:
        e = C.c_int64()
        _lib.check(self._L.rg_debug_check_bvh(self._h, C.byref(e), _stream()), self._h)
        return int(e.value)
