"""Benchmark of the gfx950 differentiable ray caster: train rays/s (forward + backward).

A step is one training pass of the hot path over one batch: LBVH refit + per-Gaussian
preprocessing, per-view SH/SG colour, forward and backward of every ray of every view (L2 loss
against a fixed synthetic target), and the SH/SG backward.  Inputs are resident in HBM before
the timed region.

Two launch structures are timed on the same step:
  batched  -- one forward launch and one backward launch (which forms the loss gradient itself)
              over the rays of all views, captured with the rest of the step in one HIP graph:
              the number reported as `value`;
  per_view -- the plain path: one forward launch and one backward launch per view.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu]
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0


def tile_order_pixels(w, h):
    import numpy as np

    tx, ty = (w + 7) // 8, (h + 7) // 8
    t = np.arange(tx * ty * 64)
    tile, lane = t // 64, t % 64
    px = (tile % tx) * 8 + lane % 8
    py = (tile // tx) * 8 + lane // 8
    ok = (px < w) & (py < h)
    return (py * w + px)[ok].astype(np.int32)


def algorithmic_bytes(R, P):
    """Per-ray and per-(ray, Gaussian) HBM bytes of forward + backward: per ray the colour
    written / re-read and dL/dC read (36 B); per distinct pair the forward read of the Gaussian
    (mu, s, q, sigma, rgb: 56 B), the backward re-read (56 B) and the gradient write (14 fp32 =
    56 B)."""
    return 36.0 * R + 168.0 * P


def backward_bytes(R, P):
    """The backward launch's share of algorithmic_bytes: colour re-read + dL/dC read per ray
    (24 B), Gaussian re-read + gradient write per pair (112 B)."""
    return 24.0 * R + 112.0 * P


def measured_traffic(workload, kernel="k_backward"):
    """HBM-side bytes per launch of `kernel` from the rocprofv3 summary committed for this workload
    (FETCH_SIZE + WRITE_SIZE, each scaled by its reading for a device copy of known size); None
    when no summary exists.  The summary names the commands it was made with."""
    path = os.path.join(ROOT, "profiles", f"k_backward_{workload}.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(f"traffic_bytes_per_launch {kernel}:"):
                    return float(line.split(":")[1]), os.path.relpath(path, ROOT)
    except OSError:
        pass
    return None, None


def forward_bytes(R, P):
    """The forward launch's share: colour write per ray (12 B), Gaussian read per pair (56 B)."""
    return 12.0 * R + 56.0 * P


def step_bytes(R, P, n, K, L, V):
    """Whole-step algorithmic bytes: rays and pairs, preprocessing (geometry read, box write),
    SH/SG forward (coefficients read once, V colour tables written) and SH/SG backward (V colour
    gradients read, coefficients read, coefficient gradients and direction term written)."""
    coef = 4.0 * (3 * K + 7 * L)
    return (algorithmic_bytes(R, P) + n * (44.0 + 24.0) + n * (coef + 12.0 * V)
            + n * (12.0 * V + coef + coef + 12.0))


def cpu_baseline(wl, threads=None, runs=5):
    """The FP32 build of the CPU transcription (oracle/liboracle_rg32.so: the oracle's algorithm
    with float arithmetic, AVX2/FMA, OpenMP over rays, without the oracle's fragility bookkeeping)
    on a bounded sample: a 192x192 crop at the centre of view 0, one training pass (forward, L2
    gradient, backward) per run, one warm-up run, then the median of `runs`.  The per-call CPU BVH
    build is timed apart and left out of rays/s (amortised over a whole view it is negligible)."""
    import numpy as np

    from oracle import oracle as O
    from ray_gauss_amd import scenes

    cam_obj = wl.cameras[0]
    cam = cam_obj.launch_record()
    w, h = cam["width"], cam["height"]
    c = 192
    px = scenes.crop_pixels(cam_obj, w // 2 - c // 2, h // 2 - c // 2, c, c)
    n_threads = O.nthreads(threads)
    rgb = O.sh_forward(wl.scene, cam["eye"], n_threads)
    bmin, bmax = O.scene_box(wl.scene, wl.params.sigma_thr)
    target = np.random.default_rng(0).uniform(0, 1, (len(px), 3))
    times = []
    for _ in range(runs + 1):
        times.append(O.baseline_step(wl.params, wl.scene, rgb, bmin, bmax, cam, px, target, n_threads)["times"])
    times = np.array(times[1:])
    ray_s = np.median(times[:, 1] + times[:, 2])
    model, _ = O.cpu_model()
    return {"value": len(px) / ray_s, "unit": "rays/s", "cores": n_threads, "kind": "port", "cpu": model,
            "sample": f"{c}x{c} crop ({len(px)} rays) at the centre of view 0; FP32 build of the transcription; "
                      f"forward + L2 gradient + backward, median of {runs} after 1 warm-up: {ray_s:.2f} s "
                      f"(forward {np.median(times[:, 1]):.2f} s, backward {np.median(times[:, 2]):.2f} s); "
                      f"CPU BVH build {np.median(times[:, 0]):.2f} s per call, excluded"}


# learning rates of the full-step benchmark (the timing does not depend on them)
FULL_STEP_LR = {"positions": 1.6e-4, "rgb": 2.5e-3, "sh_rest": 1.25e-4, "log_scales": 5e-3, "quats": 1e-3,
                "densities": 5e-2, "sg": 2.5e-3, "sharpness_raw": 1e-2, "lobe_axis": 1e-3}


def full_step(args):
    """The complete training step of the 3M-Gaussian workload: per step mu += N(0, 1e-3) and
    log s += N(0, 1e-2), the reference's activations (exp scales, normalised quaternions and lobe
    axes, softplus sharpness, densities as they are: classes/point_cloud.py:689-739) with their
    autograd, the graph-captured refit + render + loss + backward, one Adam launch over the nine
    parameter tensors and the post-step projections (scripts/train.py:121-124, 168-183)."""
    import numpy as np
    import torch

    from ray_gauss_amd import scenes
    from ray_gauss_amd.batch import BatchedTrainer, make_ray_batch
    from ray_gauss_amd.caster import RayCaster, TuneConfig
    from ray_gauss_amd.optim import FusedAdam, project_params, step_all

    dev = torch.device("cuda:0")
    wl = scenes.workload(args.workload)
    sc = wl.scene
    cams = [c.launch_record() for c in wl.cameras]
    f32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(dev)  # noqa: E731
    raw = {"positions": f32(sc.positions), "rgb": f32(sc.sh[:, :, 0]), "sh_rest": f32(sc.sh[:, :, 1:]),
           "log_scales": f32(np.log(sc.scales)), "quats": f32(sc.quats), "densities": f32(sc.densities),
           "sg": f32(sc.sg), "sharpness_raw": f32(np.log(np.expm1(sc.sharpness))), "lobe_axis": f32(sc.lobe_axis)}
    for t in raw.values():
        t.requires_grad_(True)

    def activate():
        return {"positions": raw["positions"], "scales": torch.exp(raw["log_scales"]),
                "quats": raw["quats"] / torch.norm(raw["quats"], dim=1)[:, None], "densities": raw["densities"],
                "sh": torch.cat([raw["rgb"][:, :, None], raw["sh_rest"]], dim=2), "sg": raw["sg"],
                "sharpness": torch.nn.functional.softplus(raw["sharpness_raw"]),
                "lobe_axis": raw["lobe_axis"] / torch.norm(raw["lobe_axis"], dim=2)[:, :, None]}

    with torch.no_grad():
        buf = {k: v.detach().clone().contiguous() for k, v in activate().items()}
    rv, rp, _ = make_ray_batch(cams, wl.pixels, dev)
    R = int(rv.shape[0])
    target = torch.rand((R, 3), generator=torch.Generator(device=dev).manual_seed(0), device=dev)
    tune = TuneConfig(args.hit_depth, args.seg_slabs, packet=not args.no_packet, fwd_hit_depth=args.fwd_hit_depth,
                      replay_words=args.replay_words, tree_builder=args.tree_builder,
                      ray_chunk=args.ray_chunk, refill_lanes=args.refill_lanes,
                      bwd_ray_chunk=args.bwd_ray_chunk, bwd_window=args.bwd_window,
                      fwd_jump=not args.fwd_round_per_slab)
    trainer = BatchedTrainer(RayCaster(wl.params, tune), buf, cams, rv, rp, target, loss_kind=0,
                             use_graph=not args.no_graph)
    opts = [FusedAdam([raw[k]], lr=FULL_STEP_LR[k]) for k in raw]
    gen = torch.Generator(device=dev).manual_seed(1)
    filter_3d = torch.full((sc.n, 1), math.log(1e-4), device=dev)

    def step():
        with torch.no_grad():
            raw["positions"].add_(torch.randn(raw["positions"].shape, generator=gen, device=dev), alpha=1e-3)
            raw["log_scales"].add_(torch.randn(raw["log_scales"].shape, generator=gen, device=dev), alpha=1e-2)
        act = activate()
        with torch.no_grad():
            for k, v in act.items():
                buf[k].copy_(v)
        g = trainer.step()
        names = [k for k in act if g.get(k) is not None]
        torch.autograd.backward([act[k] for k in names], [g[k] for k in names])
        step_all(opts)
        project_params(densities=raw["densities"].data, log_scales=raw["log_scales"].data, filter_3d=filter_3d,
                       max_scale=float("inf"), quaternions=raw["quats"].data, lobe_axis=raw["lobe_axis"].data)
        for o in opts:
            o.zero_grad(set_to_none=True)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    free0, total = torch.cuda.mem_get_info()
    peak_used = total - free0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        free, _ = torch.cuda.mem_get_info()
        peak_used = max(peak_used, total - free)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    n_params = sum(t.numel() for t in raw.values())
    print(json.dumps({
        "metric": "full train step (refit + fwd + bwd + Adam + projections), rays/s",
        "value": R / dt, "unit": "rays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "gaussians": sc.n, "views": len(cams), "rays_per_step": R,
                   "sg_lobes": sc.num_sg, "parameters": n_params},
        "device_memory_used_peak_GB": peak_used / 1e9,
        "adam_bytes_per_step": 28 * n_params,
    }))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="unbounded1m_8views")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-per-view", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="run the batched step eagerly instead of as a HIP graph")
    ap.add_argument("--hit-depth", type=int, default=16)
    ap.add_argument("--seg-slabs", type=int, default=32)
    ap.add_argument("--fwd-hit-depth", type=int, default=12, help="list depth of the forward kernel (6, 12, 16, 32)")
    ap.add_argument("--replay-words", type=int, default=1536,
                    help="words per ray of the forward's hit-list records replayed by the backward (0: off)")
    ap.add_argument("--no-packet", action="store_true", help="per-lane LBVH traversal")
    ap.add_argument("--tree-builder", type=int, default=1, help="0: LBVH (Morton order), 1: binned SAH")
    ap.add_argument("--ray-chunk", type=int, default=64,
                    help="forward launch: 0 = one wavefront per 8x8 tile (grid launch); a multiple of 64 = "
                         "persistent wavefronts taking that many rays at a time from per-XCD work queues")
    ap.add_argument("--bwd-ray-chunk", type=int, default=0, help="backward launch, as --ray-chunk")
    ap.add_argument("--bwd-window", type=float, default=0.5,
                    help="backward depth synchronisation in slabs (lanes within it of the laggard step together)")
    ap.add_argument("--fwd-round-per-slab", action="store_true",
                    help="forward: one round per slab (no joint step over slabs empty for every lane)")
    ap.add_argument("--refill-lanes", type=int, default=0,
                    help="persistent forward: hand new rays to idle lanes once this many are idle (0: whole chunks)")
    ap.add_argument("--autotune", action="store_true",
                    help="pick the launch configuration by the autotuner's coordinate search before timing "
                         "(the table is reported)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 5 device copies of 256 MiB (reference kernel for rocprof byte counters)")
    ap.add_argument("--full-step", action="store_true",
                    help="time the complete training step incl. Adam (use with --workload unbounded3m_4views)")
    args = ap.parse_args()
    if args.full_step:
        if args.gpus != 1:
            print(json.dumps({"error": "this build runs on one GPU", "n_gpus": args.gpus}))
            sys.exit(2)
        full_step(args)
        return
    if args.calibrate_copy:
        import torch

        x = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel() * 4, "copy_bytes_written": y.numel() * 4, "copies": 5}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU scaling is design-only "
                                   "(one replica per GPU, gradient all-reduce over RCCL)", "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    from ray_gauss_amd import scenes
    from ray_gauss_amd.caster import RayCaster, TuneConfig

    dev = torch.device("cuda:0")
    wl = scenes.workload(args.workload)
    sc = wl.scene
    cams = [c.launch_record() for c in wl.cameras]
    t = {k: torch.from_numpy(np.ascontiguousarray(getattr(sc, k))).to(dev) for k in
         ("positions", "scales", "quats", "densities", "sh", "sg", "sharpness", "lobe_axis")}
    L = sc.num_sg

    # ray batch: every pixel of every view (or the workload's subsets), 8x8 tiles for coherence
    views, pixels = [], []
    for v, (c, px) in enumerate(zip(cams, wl.pixels)):
        p = tile_order_pixels(c["width"], c["height"]) if px is None else np.asarray(px, np.int32)
        pixels.append(p)
        views.append(np.full(len(p), v, np.int32))
    ray_view = torch.from_numpy(np.concatenate(views)).to(dev)
    ray_pixel = torch.from_numpy(np.concatenate(pixels)).to(dev)
    R = int(ray_view.shape[0])
    gen = torch.Generator(device=dev).manual_seed(0)
    target = torch.rand((R, 3), generator=gen, device=dev)
    targets_pv = [target[int(sum(len(p) for p in pixels[:v])):][:len(pixels[v])] for v in range(len(cams))]
    px_dev = [torch.from_numpy(p).to(dev) for p in pixels]

    tune = TuneConfig(args.hit_depth, args.seg_slabs, packet=not args.no_packet, fwd_hit_depth=args.fwd_hit_depth,
                      replay_words=args.replay_words, tree_builder=args.tree_builder,
                      ray_chunk=args.ray_chunk, refill_lanes=args.refill_lanes,
                      bwd_ray_chunk=args.bwd_ray_chunk, bwd_window=args.bwd_window,
                      fwd_jump=not args.fwd_round_per_slab)
    rc = None  # the per-view caster, made after the batched run so that its memory is not counted there

    def prepare(rebuild=False):
        rc.prepare(t["positions"], t["scales"], t["quats"], t["densities"], t["sh"], t["sg"] if L else None,
                   t["sharpness"] if L else None, t["lobe_axis"] if L else None, rebuild=rebuild)

    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    torch.cuda.synchronize()
    free0, hbm_total = torch.cuda.mem_get_info()

    def step_per_view():
        prepare()
        rc.set_views(cams)
        for v in range(len(cams)):
            rgb, _, _ = rc.forward(v, px_dev[v])
            dL = 2.0 * (rgb - targets_pv[v])
            rc.backward(v, px_dev[v], rgb, dL, rc.last_token)
        rc.finish_grads()

    # ---- batched: refit + colour + forward + loss/backward + SH/SG backward, one HIP graph ----
    from ray_gauss_amd.batch import BatchedTrainer

    rc_b = RayCaster(wl.params, tune)
    trainer = BatchedTrainer(rc_b, t, cams, ray_view, ray_pixel, target, loss_kind=0, use_graph=not args.no_graph)
    tune_table = None
    if args.autotune:
        from ray_gauss_amd.batch import autotune

        tune, tune_table = autotune(rc_b, ray_view, ray_pixel, target, reps=2, save=False)
    for _ in range(args.warmup):
        trainer.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        trainer.step()
    torch.cuda.synchronize()
    dt_b = (time.perf_counter() - t0) / args.steps
    # the two ray kernels timed alone with HIP events on the stream they are launched on (the
    # caster launches on torch's current stream)
    f_ms, b_ms = [], []
    for _ in range(2):
        ev[0].record()
        rgb_b, _, _ = rc_b.forward_batch(ray_view, ray_pixel)
        ev[1].record()
        torch.cuda.synchronize()
        f_ms.append(ev[0].elapsed_time(ev[1]))
        st = rc_b.stats()
        dL_b = 2.0 * (rgb_b - target)
        torch.cuda.synchronize()
        ev[0].record()
        rc_b.backward_batch(ray_view, ray_pixel, rgb_b, dL_b, rc_b.last_token)
        ev[1].record()
        torch.cuda.synchronize()
        b_ms.append(ev[0].elapsed_time(ev[1]))
        rc_b.zero_grads()
    fk_ms, bk_ms = float(np.mean(f_ms)), float(np.mean(b_ms))
    # device memory of the batched step: the caster's own buffers (tree, records, accumulators),
    # torch's peak (parameters, targets, colours, gradients) and the device-wide use
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    hb = rc_b.device_bytes()
    memory = {"handle_bytes": hb["total"], "record_bytes": hb["records"],
              "torch_peak_bytes": int(torch.cuda.max_memory_allocated()),
              "device_used_bytes": int(hbm_total - free1), "device_used_before_bytes": int(hbm_total - free0),
              "device_total_bytes": int(hbm_total)}
    del trainer, rgb_b, dL_b

    # ---- per-view plain launches ----
    dt_pv = None
    if not args.no_per_view:
        rc = RayCaster(wl.params, tune)
        prepare(rebuild=True)
        rc.set_views(cams)
        for _ in range(args.warmup):
            step_per_view()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step_per_view()
        torch.cuda.synchronize()
        dt_pv = (time.perf_counter() - t0) / args.steps

    # tree builds (outside the timed step, which refits the tree it was set up with): wall time
    # of rg_prepare(rebuild) incl. preprocessing, second of two calls per builder
    build_ms = {}
    for name, builder in (("lbvh", 0), ("sah", 1)):
        tb = TuneConfig(args.hit_depth, args.seg_slabs, packet=not args.no_packet, fwd_hit_depth=args.fwd_hit_depth,
                        replay_words=0, tree_builder=builder)
        rc_t = RayCaster(wl.params, tb)
        for _ in range(2):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rc_t.prepare(t["positions"], t["scales"], t["quats"], t["densities"], t["sh"], t["sg"] if L else None,
                         t["sharpness"] if L else None, t["lobe_axis"] if L else None, rebuild=True)
            torch.cuda.synchronize()
            build_ms[name] = (time.perf_counter() - t0) * 1e3
        del rc_t

    P, K = st["pairs"], st["slab_hits"]
    B_bwd = backward_bytes(R, P)
    achieved = B_bwd / (bk_ms * 1e-3) / 1e9
    # hit-list records: written by the forward, read back by the backward (4 B a word, each way)
    B_rec = 8.0 * st["replay_words"]
    B_step = step_bytes(R, P, sc.n, (sc.sh_degree + 1) ** 2, L, len(cams)) + B_rec
    traffic, traffic_src = measured_traffic(args.workload)
    out = {
        "metric": "train rays/s (fwd+bwd) at 1M Gaussians, batched vs per-view; HBM GB/s",
        "value": R / dt_b,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt_b * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "gaussians": sc.n, "views": len(cams),
                   "sizes": sorted({(c["width"], c["height"]) for c in cams})[:4], "rays_per_step": R,
                   "sh_degree": sc.sh_degree, "sg_lobes": L, "dt": wl.params.dt, "t_eps": wl.params.t_eps,
                   "hit_depth": tune.hit_depth, "seg_slabs": tune.seg_slabs,
                   "packet_traversal": tune.packet, "fwd_hit_depth": tune.fwd_hit_depth,
                   "replay_words": tune.replay_words, "tree_builder": ("lbvh", "sah")[tune.tree_builder],
                   "ray_chunk": tune.ray_chunk, "refill_lanes": tune.refill_lanes,
                   "bwd_ray_chunk": tune.bwd_ray_chunk, "bwd_window": tune.bwd_window,
                   "fwd_jump": tune.fwd_jump},
        "tree_build_ms": build_ms,
        "per_view_rays_per_s": (R / dt_pv) if dt_pv else None,
        "batched_over_per_view": (dt_pv / dt_b) if dt_pv else None,
        "counters": {"K_slab_hits": K, "P_pairs": P, "fallback_slabs": st["fallback_slabs"],
                     "gathers": st["gathers"], "replay_words": st["replay_words"],
                     "replay_truncated_rays": st["replay_truncated"]},
        "roofline": {"bound": "hbm", "kernel": "k_backward", "kernel_ms": bk_ms, "achieved": achieved,
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
                     "traffic_source": traffic_src, "algorithmic_bytes": B_bwd},
        "forward_kernel": {"kernel": "k_forward", "kernel_ms": fk_ms,
                           "achieved_GBps": forward_bytes(R, P) / (fk_ms * 1e-3) / 1e9},
        "step_hbm_GBps": B_step / dt_b / 1e9,
        "step_bytes": {"total": B_step, "records": B_rec},
        "memory": memory,
    }
    if tune_table is not None:
        out["autotune"] = [{"config": c, "ms": round(ms, 2)} for c, ms in tune_table]
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(wl)
        except Exception as e:  # pragma: no cover
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
