"""Multi-view batching: ray batches over views of uneven size, a hipGraph-captured train step
(refit + colour + forward + loss/backward + SH/SG backward), and the launch
autotuner.

The graph holds only kernel launches and device-to-device copies (no host transfers, no
memsets): camera upload happens once, outside the graph, in ``set_views``.  Parameters must be
updated in place between replays (as an optimiser does); the gradients are returned in the
graph's static output tensors.  A graph bakes in device addresses: the trainer captures again
whenever the Gaussian count, a parameter or ray tensor, or the handle's own workspace
(``RayCaster.alloc_epoch``) has changed since the capture.
"""

from __future__ import annotations

import itertools
import json
import os
import time

import numpy as np
import torch

from .caster import RayCaster, TuneConfig

HERE = os.path.dirname(os.path.abspath(__file__))
TUNE_FILE = os.path.join(HERE, "_build", "autotune.json")


def tile_order_pixels(width: int, height: int) -> np.ndarray:
    """All pixels of a view, enumerated in 8x8 tiles (one tile per wavefront)."""
    tx, ty = (width + 7) // 8, (height + 7) // 8
    t = np.arange(tx * ty * 64)
    tile, lane = t // 64, t % 64
    px = (tile % tx) * 8 + lane % 8
    py = (tile // tx) * 8 + lane // 8
    ok = (px < width) & (py < height)
    return (py * width + px)[ok].astype(np.int32)


def tile_order_subset(width: int, pixels: np.ndarray) -> np.ndarray:
    """Reorders an arbitrary pixel subset by 8x8 tile, then within the tile."""
    px = np.asarray(pixels, np.int64)
    x, y = px % width, px // width
    tx = (width + 7) // 8
    key = ((y // 8) * tx + (x // 8)) * 64 + (y % 8) * 8 + (x % 8)
    return px[np.argsort(key, kind="stable")].astype(np.int32)


def make_ray_batch(cams, pixels=None, device="cuda"):
    """(ray_view, ray_pixel) int32 device tensors over all views; pixels[v] = None means every
    pixel of view v.  Returns also per-view offsets."""
    views, pix, offs = [], [], [0]
    for v, c in enumerate(cams):
        w, h = int(c["width"]), int(c["height"])
        p = tile_order_pixels(w, h) if (pixels is None or pixels[v] is None) else tile_order_subset(w, pixels[v])
        pix.append(p)
        views.append(np.full(len(p), v, np.int32))
        offs.append(offs[-1] + len(p))
    rv = torch.from_numpy(np.concatenate(views)).to(device)
    rp = torch.from_numpy(np.concatenate(pix)).to(device)
    return rv, rp, offs


class BatchedTrainer:
    """One training pass over a fixed ray batch: refit, colour, forward + loss + backward (a
    forward launch, then a backward launch that forms the loss gradient), SH/SG backward.  With
    use_graph=True the pass is captured in a HIP graph on first call and replayed afterwards."""

    def __init__(self, caster: RayCaster, params: dict, cams, ray_view, ray_pixel, target, loss_kind: int = 0,
                 use_graph: bool = True):
        self.rc = caster
        self.p = params
        self.cams = cams
        self.rv, self.rp, self.target = ray_view, ray_pixel, target
        self.loss_kind = int(loss_kind)
        self.use_graph = use_graph
        self.loss = torch.zeros(1, dtype=torch.float32, device=ray_view.device)
        self.graph = None
        self.graph_key = None
        self.grads = None
        L = params["sg"].shape[2] if params.get("sg") is not None else 0
        self.L = L
        self._prepare(rebuild=True)
        self.rc.set_views(cams)

    def _prepare(self, rebuild=False):
        p, L = self.p, self.L
        self.rc.prepare(p["positions"], p["scales"], p["quats"], p["densities"], p["sh"], p["sg"] if L else None,
                        p["sharpness"] if L else None, p["lobe_axis"] if L else None, rebuild=rebuild)

    def _pass(self):
        self._prepare(rebuild=False)
        self.rc.train_batch(self.rv, self.rp, self.target, self.loss_kind, None, self.loss)
        return self.rc.finish_grads()

    def _key(self):
        """Everything a captured graph depends on: sizes, tensor addresses, workspace epoch."""
        ptrs = tuple((k, None if v is None else (v.data_ptr(), tuple(v.shape))) for k, v in sorted(self.p.items()))
        return (self.rc.n, ptrs, self.rv.data_ptr(), self.rp.data_ptr(), int(self.rv.shape[0]),
                self.target.data_ptr(), self.rc.alloc_epoch())

    def step(self):
        if not self.use_graph:
            self.grads = self._pass()
            return self.grads
        if self.graph is not None and self.graph_key != self._key():
            self.graph = None  # stale addresses: capture again
        if self.graph is None:
            # warm up on a side stream, then capture
            s = torch.cuda.Stream()
            s.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(s):
                self._pass()
            torch.cuda.current_stream().wait_stream(s)
            torch.cuda.synchronize()
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph):
                self.grads = self._pass()
            self.graph_key = self._key()
        self.graph.replay()
        return self.grads

    def rebuild(self, params: dict | None = None):
        """Rebuild the tree outside the graph, optionally over new parameter tensors (after
        densification).  With the same Gaussian count and tensors the handle's buffers keep their
        addresses and the graph stays valid; otherwise the next step captures again."""
        if params is not None:
            self.p = params
            self.L = params["sg"].shape[2] if params.get("sg") is not None else 0
        self._prepare(rebuild=True)
        self.rc.set_views(self.cams)
        if self.graph is not None and self.graph_key != self._key():
            self.graph = None


def autotune(caster: RayCaster, ray_view, ray_pixel, target, depths=(8, 16, 32), segs=(8, 32, 64),
             packets=(True, False), reps: int = 2, key: str | None = None, save: bool = True):
    """Times the train step for every (hit_depth, seg_slabs, packet) and keeps the fastest.
    Returns (best TuneConfig, table of (config, ms))."""
    table = []
    loss = torch.zeros(1, dtype=torch.float32, device=ray_view.device)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for d, s, c in itertools.product(depths, segs, packets):
        cfg = TuneConfig(d, s, packet=c)
        caster.set_tune(cfg)
        caster.train_batch(ray_view, ray_pixel, target, 0, None, loss)
        caster.zero_grads()
        ms = []
        for _ in range(reps):
            e0.record()
            caster.train_batch(ray_view, ray_pixel, target, 0, None, loss)
            e1.record()
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1))
        caster.zero_grads()
        table.append(((d, s, c), float(min(ms))))
    best = min(table, key=lambda r: r[1])[0]
    cfg = TuneConfig(best[0], best[1], packet=best[2])
    caster.set_tune(cfg)
    if save and key is not None:
        try:
            data = json.load(open(TUNE_FILE)) if os.path.exists(TUNE_FILE) else {}
            data[key] = {"hit_depth": best[0], "seg_slabs": best[1], "packet": bool(best[2]), "time": time.time()}
            os.makedirs(os.path.dirname(TUNE_FILE), exist_ok=True)
            json.dump(data, open(TUNE_FILE, "w"), indent=1)
        except OSError:
            pass
    return cfg, table


def tune_key(n: int, rays: int) -> str:
    """Bucket of (Gaussians, rays) the autotuner's choice is stored under."""
    return f"n{int(round(np.log2(max(n, 1))))}_r{int(round(np.log2(max(rays, 1))))}"


def load_tuned(n: int, rays: int) -> TuneConfig | None:
    try:
        d = json.load(open(TUNE_FILE))[tune_key(n, rays)]
        return TuneConfig(d["hit_depth"], d["seg_slabs"], packet=bool(d["packet"]))
    except Exception:
        return None
