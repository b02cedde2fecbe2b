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

import dataclasses
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
    use_graph=True the pass is captured in a HIP graph on first call and replayed afterwards.

    Rebuild policy: each step refits the tree, whose boxes grow as the Gaussians move away from
    the positions it was built for.  Every `check_every` steps the trainer compares the tree's
    surface-area cost (RayCaster.tree_cost) with its value just after the last rebuild and
    rebuilds, outside the graph, once the ratio exceeds `rebuild_ratio` (None: never)."""

    def __init__(self, caster: RayCaster, params: dict, cams, ray_view, ray_pixel, target, loss_kind: int = 0,
                 use_graph: bool = True, rebuild_ratio: float | None = 1.5, check_every: int = 10):
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
        self.rebuild_ratio = rebuild_ratio
        self.check_every = max(1, int(check_every))
        self.steps = 0
        self.rebuilds = 0  # made by the policy
        self._prepare(rebuild=True)
        self.rc.set_views(cams)
        self.base_cost = self.rc.tree_cost() if rebuild_ratio is not None else None
        self.cost = self.base_cost

    def _check_tree(self):
        self.steps += 1
        if self.rebuild_ratio is None or self.steps % self.check_every:
            return
        self.cost = self.rc.tree_cost()
        if self.cost > self.rebuild_ratio * self.base_cost:
            self.rebuild()
            self.rebuilds += 1

    def _prepare(self, rebuild=False):
        p, L = self.p, self.L
        self.rc.prepare(p["positions"], p["scales"], p["quats"], p["densities"], p["sh"], p["sg"] if L else None,
                        p["sharpness"] if L else None, p["lobe_axis"] if L else None, rebuild=rebuild)

    def _pass(self):
        self._prepare(rebuild=False)
        self.rc.train_batch(self.rv, self.rp, self.target, self.loss_kind, None, self.loss)
        return self.rc.finish_grads()

    def _key(self):
        """Everything a captured graph depends on: sizes, tensor addresses, workspace epoch and the
        launch configuration (baked into the captured launches)."""
        ptrs = tuple((k, None if v is None else (v.data_ptr(), tuple(v.shape))) for k, v in sorted(self.p.items()))
        return (self.rc.n, ptrs, self.rv.data_ptr(), self.rp.data_ptr(), int(self.rv.shape[0]),
                self.target.data_ptr(), self.rc.alloc_epoch(), dataclasses.astuple(self.rc.tune))

    def step(self):
        if not self.use_graph:
            self.grads = self._pass()
            self._check_tree()
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
        self._check_tree()
        return self.grads

    def rebuild(self, params: dict | None = None):
        """Rebuild the tree outside the graph, optionally over new parameter tensors (after
        densification).  With the same Gaussian count and tensors the handle's buffers keep their
        addresses and an LBVH's graph stays valid; a binned-SAH rebuild (whose refit runs one
        launch per tree level) or new sizes or tensors change the key, and the next step captures
        again."""
        if params is not None:
            self.p = params
            self.L = params["sg"].shape[2] if params.get("sg") is not None else 0
        self._prepare(rebuild=True)
        self.rc.set_views(self.cams)
        if self.rebuild_ratio is not None:
            self.base_cost = self.cost = self.rc.tree_cost()
        if self.graph is not None and self.graph_key != self._key():
            self.graph = None


# Launch parameters the autotuner sweeps, with the values it tries (hit-list replay and the tree
# builder trade memory and build time, not speed alone, and are left as configured).
TUNE_AXES = {
    "hit_depth": (8, 16, 32),
    "seg_slabs": (8, 32, 64),
    "packet": (True, False),
    "fwd_hit_depth": (6, 12, 16),
    "ray_chunk": (0, 64, 128),
    "bwd_ray_chunk": (0, 64),
    "bwd_window": (0.5, 1.0, 2.0),
    "fwd_jump": (True, False),
}


def autotune(caster: RayCaster, ray_view, ray_pixel, target, axes: dict | None = None, reps: int = 2,
             passes: int = 1, key: str | None = None, save: bool = True):
    """Coordinate search over the launch parameters of the train step (forward + L2 loss +
    backward over the batch): starting from the caster's current TuneConfig, each axis in turn is
    swept with the others held at the best values so far, `passes` times over all axes; each
    configuration is run once untimed, then timed `reps` times with HIP events (the minimum
    counts).  Returns (best TuneConfig, table of (config fields, ms)) and leaves the caster on the
    best configuration."""
    axes = TUNE_AXES if axes is None else axes
    table, seen = [], {}
    loss = torch.zeros(1, dtype=torch.float32, device=ray_view.device)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(cfg: TuneConfig) -> float:
        k = tuple(sorted(dataclasses.asdict(cfg).items()))
        if k in seen:
            return seen[k]
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
        seen[k] = float(min(ms))
        table.append(({f: getattr(cfg, f) for f in axes}, seen[k]))
        return seen[k]

    best = caster.tune
    best_ms = timed(best)
    for _ in range(passes):
        for field, values in axes.items():
            for v in values:
                cfg = dataclasses.replace(best, **{field: v})
                ms = timed(cfg)
                if ms < best_ms:
                    best, best_ms = cfg, ms
    caster.set_tune(best)
    if save and key is not None:
        try:
            data = json.load(open(TUNE_FILE)) if os.path.exists(TUNE_FILE) else {}
            data[key] = {**dataclasses.asdict(best), "ms": best_ms, "time": time.time()}
            os.makedirs(os.path.dirname(TUNE_FILE), exist_ok=True)
            json.dump(data, open(TUNE_FILE, "w"), indent=1)
        except OSError:
            pass
    return best, table


def tune_key(n: int, rays: int) -> str:
    """Bucket of (Gaussians, rays) the autotuner's choice is stored under."""
    return f"n{int(round(np.log2(max(n, 1))))}_r{int(round(np.log2(max(rays, 1))))}"


def load_tuned(n: int, rays: int) -> TuneConfig | None:
    try:
        d = json.load(open(TUNE_FILE))[tune_key(n, rays)]
        return TuneConfig(**{f.name: d[f.name] for f in dataclasses.fields(TuneConfig) if f.name in d})
    except Exception:
        return None
