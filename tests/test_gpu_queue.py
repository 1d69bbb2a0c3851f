"""GPU parity of the persistent work-queue paths: the GI batch (GI ray per pixel + sun ray per GI hit)
against the oracle, and the uneven view + probe batch against per-request plain launches.  Hit records
must be bit-identical whatever the tile shape, chunk size or sorting."""
import numpy as np
import pytest
import torch

import vftrace as vf
import oracle as O
from helpers import camera, config1_camera, hit_diff, world

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


@pytest.fixture(scope="module")
def ctx():
    c = vf.Context(0)
    w, a, _ = world(1)
    c.upload(w)
    c.set_scene(vf.default_scene(a["world_bricks"], a["world_win"]))
    yield c
    c.close()


def dev_hits(n):
    return torch.zeros(max(n, 1) * 20, dtype=torch.uint8, device=DEV)


def gbuffer(ctx, pu, W, H):
    hits = dev_hits(W * H)
    gbp = torch.zeros(W * H * 4, dtype=torch.float32, device=DEV)
    gbn = torch.zeros(W * H, dtype=torch.int32, device=DEV)
    ctx.trace_primary(pu, W, H, hits, gbp, gbn)
    ctx.sync()
    return gbp, gbn


@pytest.mark.parametrize("tile,chunk,sort,refill", [((8, 8), 64, False, 32), ((16, 8), 256, True, 32),
                                                    ((16, 16), 1024, True, 48), ((8, 8), 512, True, 0),
                                                    ((8, 8), 64, False, 0), ((16, 16), 64, False, -1)])
def test_gi_batch_matches_oracle(ctx, tile, chunk, sort, refill):
    _, a, ow = world(1)
    pu, _ = config1_camera(frame=6)
    W, H = 640, 360
    gbp, gbn = gbuffer(ctx, pu, W, H)
    gi = dev_hits(W * H)
    sun = dev_hits(W * H)
    ctx.trace_gi_batch(pu, W, H, 640.0, gbp, gbn, gi, sun, ctx.queue_cfg(tile, chunk, sort, refill=refill))
    ctx.sync()
    scene = vf.default_scene(a["world_bricks"])
    og, osun = O.gi_batch(ow, W, H, 6, gbp.cpu().numpy().reshape(-1, 4), gbn.cpu().numpy().view(np.uint32),
                          np.array(scene.sun_dir, np.float32), 640.0)
    g = vf.hits_to_numpy(gi)[:W * H]
    s = vf.hits_to_numpy(sun)[:W * H]
    assert hit_diff(g, og).size == 0
    assert hit_diff(s, osun).size == 0
    assert 0.05 < g["hit"].mean() < 0.95 and s["hit"].sum() > 0


@pytest.mark.parametrize("refill", [0, 32, -1])
def test_gi_batch_counters_match_oracle(ctx, refill):
    _, a, ow = world(1)
    pu, _ = config1_camera(frame=2)
    W, H = 320, 180
    gbp, gbn = gbuffer(ctx, pu, W, H)
    gi, sun = dev_hits(W * H), dev_hits(W * H)
    ctx.set_counting(True)
    try:
        ctx.counters(reset=True)
        ctx.trace_gi_batch(pu, W, H, 640.0, gbp, gbn, gi, sun, ctx.queue_cfg((16, 8), 128, True, refill=refill))
        c = ctx.counters(reset=True)
    finally:
        ctx.set_counting(False)
    scene = vf.default_scene(a["world_bricks"])
    _, _, oc = O.gi_batch(ow, W, H, 2, gbp.cpu().numpy().reshape(-1, 4), gbn.cpu().numpy().view(np.uint32),
                          np.array(scene.sun_dir, np.float32), 640.0, return_counts=True)
    assert c == oc


def _requests(ctx, sizes, probes):
    reqs, outs, passes = [], [], []
    rng = np.random.default_rng(5)
    for (w, h) in sizes:
        pos = [rng.uniform(20, 230), rng.uniform(70, 120), rng.uniform(20, 230)]
        pu, _ = camera(pos, rng.uniform(0, 6.28), rng.uniform(-0.6, 0.1))
        out = dev_hits(w * h)
        reqs.append(ctx.view_request(pu, w, h, out))
        outs.append(out)
        passes.append((pu, w, h))
    for (origin, spacing, dims, nr, frame, reach) in probes:
        out = dev_hits(int(np.prod(dims)) * nr)
        reqs.append(ctx.probe_request(origin, spacing, dims, nr, frame, reach, out))
        outs.append(out)
        passes.append((origin, spacing, dims, nr, frame, reach))
    return reqs, outs, passes


def test_uneven_batch_matches_plain_launches(ctx):
    sizes = [(64, 64), (256, 144), (130, 75), (480, 270)]
    probes = [((-64.5, -32.25, -64.75), 20.0, (8, 6, 8), 256, 3, 256.0)]
    reqs, outs, passes = _requests(ctx, sizes, probes)
    for cfg in (ctx.queue_cfg((8, 8), 64), ctx.queue_cfg((16, 16), 1024), ctx.queue_cfg((16, 8), 256),
                ctx.queue_cfg((8, 8), 64, refill=0), ctx.queue_cfg((16, 16), 512, refill=0),
                ctx.queue_cfg((8, 8), 64, refill=-1), ctx.queue_cfg((16, 8), 256, refill=-1),
                ctx.queue_cfg((16, 8), 64, refill=-1, traversal=1)):
        for o in outs:
            o.zero_()
        ctx.trace_batch(reqs, cfg)
        ctx.sync()
        times = ctx.batch_times(len(reqs))
        assert all(t > 0 for t in times), times
        for k, p in enumerate(passes):
            if k < len(sizes):
                pu, w, h = p
                ref = dev_hits(w * h)
                ctx.trace_primary(pu, w, h, ref)
                n = w * h
            else:
                origin, spacing, dims, nr, frame, reach = p
                n = int(np.prod(dims)) * nr
                ref = dev_hits(n)
                ctx.trace_probes(origin, spacing, dims, nr, frame, reach, ref)
            ctx.sync()
            assert torch.equal(outs[k][:n * 20], ref[:n * 20]), (k, cfg.tile_w, cfg.chunk)


def test_batch_and_gi_batch_in_a_graph(ctx):
    sizes = [(96, 64), (200, 120)]
    probes = [((10.5, 60.5, 10.5), 20.0, (4, 2, 4), 64, 1, 128.0)]
    reqs, outs, _ = _requests(ctx, sizes, probes)
    pu, _ = config1_camera(frame=3)
    W, H = 160, 96
    gbp, gbn = gbuffer(ctx, pu, W, H)
    gi, sun = dev_hits(W * H), dev_hits(W * H)
    s = torch.cuda.Stream()
    ctx.set_stream(s)
    try:
        cfg = ctx.queue_cfg((16, 8), 128, True)
        ctx.trace_batch(reqs, cfg)          # eager run first: sizes the queue buffers
        ctx.trace_gi_batch(pu, W, H, 640.0, gbp, gbn, gi, sun, cfg)
        ctx.sync()
        eager = [o.clone() for o in outs] + [gi.clone(), sun.clone()]
        ctx.graph_begin()
        ctx.trace_batch(reqs, cfg)
        ctx.trace_gi_batch(pu, W, H, 640.0, gbp, gbn, gi, sun, cfg)
        g = ctx.graph_end()
        for rep in range(3):
            for o in outs + [gi, sun]:
                o.zero_()
            torch.cuda.synchronize()
            ctx.graph_launch(g)
            ctx.sync()
            for e, o in zip(eager, outs + [gi, sun]):
                assert torch.equal(e, o), rep
        g.close()
    finally:
        ctx.set_stream(None)


def test_autotuner_picks_a_measured_config(ctx):
    pu, _ = config1_camera(frame=1)
    W, H = 256, 144
    gbp, gbn = gbuffer(ctx, pu, W, H)
    gi, sun = dev_hits(W * H), dev_hits(W * H)
    best, table = ctx.autotune_gi_batch(pu, W, H, 640.0, gbp, gbn, gi, sun)
    assert len(table) == 2 * 3 * 5 * 2 + 1   # + the queue-free configuration
    fastest = min(table, key=lambda r: r[3])
    assert (best.tile_w, best.tile_h) == fastest[0] and best.chunk == fastest[1] and bool(best.sort) == fastest[2]
    assert all(r[3] > 0 for r in table)


def test_full_config5_batch_matches_plain_launches_and_oracle(ctx):
    """The uneven batch at its full size: 16 seeded views drawn from {256^2, 512x288, 960x540, 1024^2} and a
    32x16x32 probe grid x 256 spherical-Fibonacci rays (about 13M rays) in one launch, in chunk mode and in
    refill mode and in the flat one-launch mode.  Every request's hit records equal its own plain launch byte for byte, and every 97th ray of
    every request equals the oracle's traceVolumeS on the same ray."""
    import math
    _, a, ow = world(1)
    rng = np.random.default_rng(1337)
    sizes = [(256, 256), (512, 288), (960, 540), (1024, 1024)]
    views = []
    for _ in range(16):
        vw, vh = sizes[int(rng.integers(0, 4))]
        pos = [float(rng.uniform(20, 236)), float(rng.uniform(75, 120)), float(rng.uniform(20, 236))]
        vp, vr, vu, vfw = vf.fly_camera(pos, float(rng.uniform(0, 2 * math.pi)), float(rng.uniform(-0.7, 0.1)))
        assert ow.voxel(*[int(math.floor(c)) for c in vp]) == 0   # cameras in air: rays start at the camera
        views.append((vw, vh, vf.make_pass(vp, vr, vu, vfw, 70.0, 0), (vp, vr, vu, vfw)))
    views.sort(key=lambda v: v[0] * v[1])
    assert {(v[0], v[1]) for v in views} >= {(1024, 1024)}
    dims, nr, frame, reach, spacing = (32, 16, 32), 256, 7, 256.0, 20.0
    origin = (128.0 - 310.0 + 0.5, 64.0 - 150.0 + 0.5, 128.0 - 310.0 + 0.5)
    n_probe = int(np.prod(dims)) * nr
    outs = [dev_hits(vw * vh) for vw, vh, _, _ in views] + [dev_hits(n_probe)]
    reqs = [ctx.view_request(vp, vw, vh, o) for (vw, vh, vp, _), o in zip(views, outs)]
    reqs.append(ctx.probe_request(origin, spacing, dims, nr, frame, reach, outs[-1]))
    ref = [dev_hits(vw * vh) for vw, vh, _, _ in views] + [dev_hits(n_probe)]
    for (vw, vh, vp, _), r in zip(views, ref):
        ctx.trace_primary(vp, vw, vh, r)
    ctx.trace_probes(origin, spacing, dims, nr, frame, reach, ref[-1])
    ctx.sync()
    for cfg in (ctx.queue_cfg((16, 8), 512, refill=0), ctx.queue_cfg((8, 8), 256, refill=32),
                ctx.queue_cfg((16, 16), 64, refill=-1), ctx.queue_cfg((16, 8), 64, refill=-1, traversal=1)):
        for o in outs:
            o.zero_()
        ctx.trace_batch(reqs, cfg)
        ctx.sync()
        times = ctx.batch_times(len(reqs))
        assert all(t > 0 for t in times), times
        for k, (o, r) in enumerate(zip(outs, ref)):
            assert torch.equal(o, r), (k, cfg.refill)
    # every 97th ray of each request against the oracle
    L = O.lib()
    jit = np.zeros(2, np.float32)
    tmp = np.zeros(3, np.float32)
    for (vw, vh, pu, (vp, vr, vu, vfw)), o in zip(views, outs):
        idx = np.arange(0, vw * vh, 97)
        rays = np.zeros((idx.size, 6), np.float32)
        R, U, F = (np.ascontiguousarray(x, np.float32) for x in (vr, vu, vfw))
        for j, i in enumerate(idx):
            L.vfo_primary_dir(R.ctypes.data, U.ctypes.data, F.ctypes.data, float(pu.tan_half_fov),
                              vw, vh, int(i % vw), int(i // vw), jit.ctypes.data, tmp.ctypes.data)
            rays[j, :3], rays[j, 3:] = vp, tmp
        oh, _ = ow.trace_rays(rays, 1e30, 0)
        g = vf.hits_to_numpy(o)[idx]
        assert hit_diff(g, oh).size == 0, (vw, vh)
    dirs = O.probe_dirs(nr, frame)
    idx = np.arange(0, n_probe, 97)
    p, i = idx // nr, idx % nr
    pc = np.stack([p % dims[0], (p // dims[0]) % dims[1], p // (dims[0] * dims[1])], 1).astype(np.float32)
    o3 = np.array(origin, np.float32)[None, :] + np.float32(spacing) * pc
    oh, _ = ow.trace_rays(np.concatenate([o3, dirs[i]], 1).astype(np.float32), reach, 5)
    assert hit_diff(vf.hits_to_numpy(outs[-1])[idx], oh).size == 0
    assert 0.05 < oh["hit"].mean() < 0.95
