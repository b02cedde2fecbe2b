"""Batched multi-view path, train step (forward + loss/backward), hipGraph replay and the drop-in autograd boundary
on the GPU, each checked against the per-view path or the CPU oracle."""

import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from ray_gauss_amd import scenes
from ray_gauss_amd.batch import BatchedTrainer, autotune, make_ray_batch
from ray_gauss_amd.cameras import look_at_camera, orbit_eye
from ray_gauss_amd.caster import TuneConfig

from gpu_helpers import check_image, gpu_box, max_rel, oracle_forward, prepared_caster, to_dev

pytestmark = pytest.mark.gpu


def _uneven_views(rng, count=3):
    cams, pixels = [], []
    for _ in range(count):
        w = int(rng.integers(40, 120))
        h = int(round(0.75 * w))
        cams.append(look_at_camera(orbit_eye(4.0, rng.uniform(0, 360), rng.uniform(-20, 40)), width=w,
                                   height=h).launch_record())
        npx = int(rng.uniform(0.25, 1.0) * w * h)
        pixels.append(np.sort(rng.permutation(w * h)[:npx]).astype(np.int32))
    return cams, pixels


@pytest.fixture(scope="module")
def blobs(gpu_device):
    sc = scenes.scene_blobs(20000, 16, seed=3, num_sg=4)
    prm = scenes.RenderParams(dt=0.01, t_eps=1e-4, white_background=True)
    cams, pixels = _uneven_views(np.random.default_rng(5))
    rc, t = prepared_caster(sc, prm, cams, gpu_device)
    rv, rp, offs = make_ray_batch(cams, pixels, gpu_device)
    return sc, prm, cams, pixels, rc, t, rv, rp, offs


def test_batch_forward_equals_per_view(blobs):
    sc, prm, cams, pixels, rc, t, rv, rp, offs = blobs
    out, _, _ = rc.forward_batch(rv, rp)
    for v in range(len(cams)):
        pv = rp[offs[v]:offs[v + 1]]
        ref, _, _ = rc.forward(v, pv)
        assert torch.equal(out[offs[v]:offs[v + 1]], ref)


def test_batch_forward_parity_with_oracle(blobs):
    sc, prm, cams, pixels, rc, t, rv, rp, offs = blobs
    out, _, _ = rc.forward_batch(rv, rp)
    box = gpu_box(rc)
    for v in range(len(cams)):
        pv = rp[offs[v]:offs[v + 1]].cpu().numpy()[::5]
        sel = np.arange(offs[v], offs[v + 1])[::5]
        check_image(rc, sc, prm, cams[v], out.cpu().numpy()[sel], pv, box=box, max_flagged=0.02)


def test_fused_train_equals_forward_backward(blobs):
    sc, prm, cams, pixels, rc, t, rv, rp, offs = blobs
    gen = torch.Generator(device=rv.device).manual_seed(0)
    target = torch.rand((rv.shape[0], 3), generator=gen, device=rv.device)
    rgb, _, _ = rc.forward_batch(rv, rp)
    rc.backward_batch(rv, rp, rgb, 2.0 * (rgb - target), rc.last_token)
    g_sep = rc.finish_grads()
    loss = torch.zeros(1, device=rv.device)
    out = torch.empty_like(rgb)
    rc.train_batch(rv, rp, target, 0, out, loss)
    g_fused = rc.finish_grads()
    torch.cuda.synchronize()
    assert torch.equal(out, rgb)
    assert abs(loss.item() - float(((rgb - target) ** 2).sum())) <= 1e-4 * float(((rgb - target) ** 2).sum())
    for k in ("positions", "scales", "quats", "densities", "sh", "sg", "sharpness", "lobe_axis"):
        assert max_rel(g_fused[k].cpu().numpy(), g_sep[k].cpu().numpy()) <= 1e-5, k


@pytest.mark.parametrize("words", [0, 16, 64, 4096])
def test_replayed_hit_lists_give_the_same_step(blobs, gpu_device, words):
    """train_batch's backward reads the hit lists its forward recorded (words per ray); rays whose
    records run out continue with their own traversal, and without records the backward walks
    the tree: the image and loss are identical, the gradients equal up to the order of atomic
    sums."""
    sc, prm, cams, pixels, rc0, t0, rv, rp, offs = blobs
    gen = torch.Generator(device=gpu_device).manual_seed(2)
    target = torch.rand((rv.shape[0], 3), generator=gen, device=gpu_device)
    res = {}
    for w in (0, words):
        rc, _ = prepared_caster(sc, prm, cams, gpu_device, TuneConfig(replay_words=w))
        out = torch.empty((rv.shape[0], 3), device=gpu_device)
        loss = torch.zeros(1, device=gpu_device)
        rc.train_batch(rv, rp, target, 0, out, loss)
        st = rc.stats()
        res[w] = (out.cpu().numpy(), float(loss.item()), {k: v.cpu().numpy() for k, v in rc.finish_grads().items()
                                                          if v is not None}, st)
    a, b = res[0], res[words]
    assert np.array_equal(a[0], b[0]) and abs(a[1] - b[1]) <= 1e-6 * abs(a[1])
    for k, v in a[2].items():
        assert max_rel(b[2][k], v) <= 1e-5, k
    if words:
        assert b[3]["replay_words"] > 0
        if words == 16:
            assert b[3]["replay_truncated"] > 0
        if words == 4096:
            assert b[3]["replay_truncated"] == 0
    else:
        assert a[3]["replay_words"] == 0


def test_replay_only_for_the_recorded_rays(blobs, gpu_device):
    """forward_batch records and returns a token; backward_batch replays only when given that
    token with nothing changed in between: after another recording forward or set_views the
    token is stale and the backward walks the tree, giving the same gradients as without
    records.  A caller that rewrites its ray arrays in place between the two calls still gets
    the walked gradients for the rewritten rays (each output row's records name their ray)."""
    sc, prm, cams, pixels, rc0, t0, rv, rp, offs = blobs
    gen = torch.Generator(device=gpu_device).manual_seed(3)
    target = torch.rand((rv.shape[0], 3), generator=gen, device=gpu_device)
    rv2, rp2 = rv.flip(0).contiguous(), rp.flip(0).contiguous()

    def run(words, between, rewrite=False):
        rc, _ = prepared_caster(sc, prm, cams, gpu_device, TuneConfig(replay_words=words))
        rv_, rp_ = rv.clone(), rp.clone()
        rgb, _, _ = rc.forward_batch(rv_, rp_)
        tok = rc.last_token
        assert (tok != 0) == (words > 0)
        between(rc)
        if rewrite:  # rows 0, 3, 6, ... now name the next pixel of their own view
            npx = torch.tensor([c["width"] * c["height"] for c in cams], device=rv.device)[rv.long()]
            rp_[::3] = ((rp[::3].long() + 1) % npx[::3]).int()
        rc.backward_batch(rv_, rp_, rgb, 2.0 * (rgb - target), tok)
        st = rc.stats()
        assert st["dropped"] == 0 and st["bad_rays"] == 0
        return {k: v.cpu().numpy() for k, v in rc.finish_grads().items() if v is not None}

    ref = run(0, lambda rc: None)
    for between in (lambda rc: None, lambda rc: rc.forward_batch(rv2, rp2), lambda rc: rc.set_views(cams)):
        g = run(512, between)
        for k, v in ref.items():
            assert max_rel(g[k], v) <= 1e-5, k
    ref_rw = run(0, lambda rc: None, rewrite=True)
    g_rw = run(512, lambda rc: None, rewrite=True)
    for k, v in ref_rw.items():
        assert max_rel(g_rw[k], v) <= 1e-5, k
    assert max_rel(ref_rw["densities"], ref["densities"]) > 1e-3  # the rewrite does change the rays


def test_batch_backward_parity_with_oracle(blobs):
    sc, prm, cams, pixels, rc, t, rv, rp, offs = blobs
    rgb, _, _ = rc.forward_batch(rv, rp)
    tok = rc.last_token
    box = gpu_box(rc)
    rec = rc.export_record()
    rng = np.random.default_rng(9)
    dL = np.zeros((rv.shape[0], 3))
    total = None
    for v in range(len(cams)):
        pv = rp[offs[v]:offs[v + 1]].cpu().numpy()
        ref = oracle_forward(sc, prm, cams[v], box, pv, record=rec)
        ok = (ref["fragile"] == 0)[:, None]
        assert ok.mean() >= 0.98
        d = rng.normal(0, 1, (len(pv), 3)) * ok
        dL[offs[v]:offs[v + 1]] = d
        og = O.full_backward(prm, sc, cams[v]["eye"], box[0], box[1], cams[v],
                             rgb[offs[v]:offs[v + 1]].cpu().numpy().astype(np.float64), d, pv, record=rec)
        total = og if total is None else {k: total[k] + og[k] for k in og}
    rc.backward_batch(rv, rp, rgb, torch.from_numpy(dL.astype(np.float32)).to(rv.device), tok)
    assert rc.stats()["dropped"] == 0
    g = rc.finish_grads()
    for k in ("positions", "scales", "quats", "densities", "sh", "sg", "sharpness", "lobe_axis"):
        e = max_rel(g[k].cpu().numpy(), total[k])
        assert e <= 1e-3, (k, e)


def test_graph_replay_matches_eager(blobs, gpu_device):
    sc, prm, cams, pixels, rc0, t0, rv, rp, offs = blobs
    rc, t = prepared_caster(sc, prm, cams, gpu_device)
    gen = torch.Generator(device=gpu_device).manual_seed(1)
    target = torch.rand((rv.shape[0], 3), generator=gen, device=gpu_device)
    eager = BatchedTrainer(rc, t, cams, rv, rp, target, use_graph=False)
    ge = {k: (v.clone() if v is not None else None) for k, v in eager.step().items()}
    rc2, t2 = prepared_caster(sc, prm, cams, gpu_device)
    tr = BatchedTrainer(rc2, t2, cams, rv, rp, target, use_graph=True)
    for rep in range(3):
        g = tr.step()
        torch.cuda.synchronize()
        _ = (target * 2).sum().item()  # eager work and a host sync between replays
        for k in ("positions", "scales", "quats", "densities", "sh", "sg"):
            assert max_rel(g[k].cpu().numpy(), ge[k].cpu().numpy()) <= 1e-5, (rep, k)


@pytest.mark.parametrize("words", [0, 1536])
def test_joint_step_over_empty_slabs_changes_nothing(blobs, gpu_device, words):
    """After a round in which no lane had a hit, the lanes step together over the slabs empty for
    them (always in the backward; in the forward unless fwd_jump is off): the slabs with hits and
    their order per ray are unchanged, so the image and loss are identical and the gradients equal
    up to the order of atomic sums."""
    sc, prm, cams, pixels, rc0, t0, rv, rp, offs = blobs
    gen = torch.Generator(device=gpu_device).manual_seed(4)
    target = torch.rand((rv.shape[0], 3), generator=gen, device=gpu_device)
    res = {}
    for jump in (False, True):
        rc, _ = prepared_caster(sc, prm, cams, gpu_device, TuneConfig(replay_words=words, fwd_jump=jump))
        out = torch.empty((rv.shape[0], 3), device=gpu_device)
        loss = torch.zeros(1, device=gpu_device)
        rc.train_batch(rv, rp, target, 0, out, loss)
        st = rc.stats()
        res[jump] = (out.cpu().numpy(), float(loss.item()),
                     {k: v.cpu().numpy() for k, v in rc.finish_grads().items() if v is not None}, st)
    a, b = res[False], res[True]
    assert np.array_equal(a[0], b[0]) and abs(a[1] - b[1]) <= 1e-6 * abs(a[1])
    assert a[3]["slab_hits"] == b[3]["slab_hits"] and a[3]["pairs"] == b[3]["pairs"]
    assert a[3]["dropped"] == 0 and b[3]["dropped"] == 0
    for k, v in a[2].items():
        assert max_rel(b[2][k], v) <= 1e-5, k


def test_autotune_picks_a_valid_config(blobs):
    sc, prm, cams, pixels, rc, t, rv, rp, offs = blobs
    """The coordinate search times its starting point and then each axis's other values (one new
    configuration per axis here), leaves the caster on the fastest, and that configuration renders
    the same image."""
    target = torch.zeros((rv.shape[0], 3), device=rv.device)
    rc.set_tune(TuneConfig())
    ref, _, _ = rc.forward_batch(rv, rp)
    axes = {"hit_depth": (8, 16), "packet": (True, False), "ray_chunk": (0, 64), "bwd_ray_chunk": (0, 64)}
    cfg, table = autotune(rc, rv, rp, target, axes=axes, reps=1, save=False)
    assert len(table) == 5 and all(ms > 0 for _, ms in table)
    assert len({tuple(sorted(c.items())) for c, _ in table}) == 5
    assert rc.tune == cfg and cfg.hit_depth in (8, 16) and cfg.ray_chunk in (0, 64)
    best = min(table, key=lambda r: r[1])[0]
    assert all(getattr(cfg, k) == v for k, v in best.items())
    out, _, _ = rc.forward_batch(rv, rp)
    assert (out - ref).abs().max().item() <= 1e-5
    rc.set_tune(TuneConfig())


def test_dropin_autograd_boundary(gpu_device):
    """RenderOptixFunction.apply in the reference's calling convention, with the activations of
    classes/point_cloud.py:689-739 in front of it, back-propagated by torch autograd."""
    from ray_gauss_amd import optix_utils as u_ox
    from ray_gauss_amd import render as r_ox

    wl = scenes.workload("small4k")
    sc = wl.scene
    dev = gpu_device
    log_s = torch.log(torch.from_numpy(sc.scales)).to(dev).requires_grad_(True)
    q_raw = (torch.from_numpy(sc.quats) * 1.7).to(dev).requires_grad_(True)
    pos = torch.from_numpy(sc.positions).to(dev).requires_grad_(True)
    dens = torch.from_numpy(sc.densities).to(dev).requires_grad_(True)
    rgb_dc = torch.from_numpy(sc.sh[:, :, 0].copy()).to(dev).requires_grad_(True)
    sh_rest = torch.from_numpy(sc.sh[:, :, 1:].copy()).to(dev).requires_grad_(True)
    empty = torch.zeros((sc.n, 3, 0), device=dev)
    ctx = u_ox.create_context()
    settings = r_ox.RenderOptixSettings(ctx, False, None, None, None, None, None, wl.cameras[0], 1024, 0, 0, 0, 0,
                                        (1, 1), False, None, params=wl.params)
    color = torch.cat([rgb_dc[:, :, None], sh_rest], dim=2)
    q = q_raw / torch.norm(q_raw, dim=1)[:, None]
    img = r_ox.RenderOptixFunction.apply(pos, torch.exp(log_s), q, dens, color, empty, torch.zeros((sc.n, 0), device=dev),
                                         torch.zeros((sc.n, 0, 3), device=dev), settings)
    assert img.shape == (64 * 64, 3)
    gen = torch.Generator(device=dev).manual_seed(0)
    target = torch.rand(img.shape, generator=gen, device=dev)
    loss = ((img - target) ** 2).sum()
    loss.backward()
    # same image and gradients through the caster directly, on the same activated tensors
    from ray_gauss_amd.caster import RayCaster

    with torch.no_grad():
        s_act = torch.exp(log_s)
        q_act = q_raw / torch.norm(q_raw, dim=1)[:, None]
        c_act = torch.cat([rgb_dc[:, :, None], sh_rest], dim=2)
    rc = RayCaster(wl.params)
    rc.prepare(pos.detach(), s_act, q_act, dens.detach(), c_act, rebuild=True)
    rc.set_views([wl.cameras[0].launch_record()])
    rgb, _, _ = rc.forward(0)
    assert torch.equal(rgb, img.detach())
    rc.backward(0, None, rgb, 2 * (rgb - target), rc.last_token)
    g = rc.finish_grads()
    assert max_rel(dens.grad.cpu().numpy(), g["densities"].cpu().numpy()) <= 1e-5
    assert max_rel(pos.grad.cpu().numpy(), g["positions"].cpu().numpy()) <= 1e-5
    # chain through exp(log s): d/dlog s = s * d/ds
    assert max_rel(log_s.grad.cpu().numpy(), (g["scales"] * s_act).cpu().numpy()) <= 1e-5
    # chain through q / |q|: (I - q^ q^T) / |q| applied to d/dq^
    qn = torch.norm(q_raw.detach(), dim=1, keepdim=True)
    gq = g["quats"]
    dq_raw = (gq - q_act * (q_act * gq).sum(1, keepdim=True)) / qn
    assert max_rel(q_raw.grad.cpu().numpy(), dq_raw.cpu().numpy()) <= 1e-5
    assert max_rel(rgb_dc.grad.cpu().numpy(), g["sh"][:, :, 0].cpu().numpy()) <= 1e-5


def test_dropin_jitter_and_random_samples(gpu_device):
    """The drop-in accepts the reference's jitter / rnd_sample settings; the image depends on the
    iteration that seeds the generator, and its gradient equals the caster's for the same
    settings."""
    from ray_gauss_amd import render as r_ox
    from ray_gauss_amd.caster import RayCaster

    wl = scenes.workload("small4k")
    t = to_dev(wl.scene, gpu_device)
    imgs = []
    for it in (5, 6):
        settings = r_ox.RenderOptixSettings(None, False, None, None, None, None, None, wl.cameras[0], 1024, 0, it,
                                            1, 1, (1, 1), False, None, params=wl.params)
        imgs.append(r_ox.RenderOptixFunction.apply(t["positions"], t["scales"], t["quats"], t["densities"], t["sh"],
                                                   t["sg"], t["sharpness"], t["lobe_axis"], settings))
    assert not torch.equal(imgs[0], imgs[1])
    prm = scenes.RenderParams(**{**wl.params.__dict__, "jitter": 1, "rnd_sample": 1, "iteration": 6})
    rc = RayCaster(prm)
    rc.prepare(t["positions"], t["scales"], t["quats"], t["densities"], t["sh"], rebuild=True)
    rc.set_views([wl.cameras[0].launch_record()])
    rgb, _, _ = rc.forward(0)
    assert torch.equal(rgb, imgs[1])


@pytest.mark.parametrize("chunk,refill", [(64, 0), (128, 0), (256, 0), (64, 32), (64, 8)])
def test_persistent_kernels_equal_grid_launch(blobs, gpu_device, chunk, refill):
    """Persistent wavefronts taking `chunk` rays at a time from per-XCD work queues (with work
    stealing), optionally refilling idle forward lanes: every ray is processed exactly once with
    the same slabs and hit lists, so images are bit-identical to the grid launch and gradients
    equal up to the order of atomic sums -- for a separate forward / backward (records replayed)
    and for the fused train step."""
    sc, prm, cams, pixels, rc0, t0, rv, rp, offs = blobs
    gen = torch.Generator(device=gpu_device).manual_seed(7)
    target = torch.rand((rv.shape[0], 3), generator=gen, device=gpu_device)
    res = {}
    for ch, rf in ((0, 0), (chunk, refill)):
        rc, _ = prepared_caster(sc, prm, cams, gpu_device, TuneConfig(ray_chunk=ch, refill_lanes=rf, bwd_ray_chunk=ch))
        rgb, depth, opac = rc.forward_batch(rv, rp, depth=True, opacity=True)
        st = rc.stats()
        rc.backward_batch(rv, rp, rgb, 2.0 * (rgb - target), rc.last_token)
        assert rc.stats()["dropped"] == 0
        g = {k: v.cpu().numpy() for k, v in rc.finish_grads().items() if v is not None}
        out = torch.empty_like(rgb)
        loss = torch.zeros(1, device=gpu_device)
        rc.train_batch(rv, rp, target, 0, out, loss)
        gt = {k: v.cpu().numpy() for k, v in rc.finish_grads().items() if v is not None}
        res[ch] = (rgb, depth, opac, st, g, out, float(loss.item()), gt)
    a, b = res[0], res[chunk]
    for i in (0, 1, 2, 5):
        assert torch.equal(a[i], b[i]), i
    # the same slab hits and pairs; with whole tiles per wavefront also the same gathers and records
    for k in ("slab_hits", "pairs") + (("gathers", "replay_words") if refill == 0 else ()):
        assert a[3][k] == b[3][k], (k, a[3][k], b[3][k])
    assert abs(a[6] - b[6]) <= 1e-6 * abs(a[6])
    for k, v in a[4].items():
        assert max_rel(b[4][k], v) <= 1e-5, k
        assert max_rel(b[7][k], a[7][k]) <= 1e-5, k


def test_persistent_full_view_and_pixel_lists(gpu_device):
    """The persistent kernels over the implicit full-view enumeration (8x8 tiles with padding) and
    over a pixel list give the grid launch's images."""
    wl = scenes.workload("small4k")
    cam = wl.cameras[0].launch_record()
    px = torch.arange(3, 4096, 5, dtype=torch.int32)
    out = {}
    for ch in (0, 128):
        rc, _ = prepared_caster(wl.scene, wl.params, [cam], gpu_device,
                                TuneConfig(ray_chunk=ch, refill_lanes=16 if ch else 0, bwd_ray_chunk=ch))
        out[ch] = (rc.forward(0)[0], rc.forward(0, px)[0])
    assert torch.equal(out[0][0], out[128][0]) and torch.equal(out[0][1], out[128][1])
