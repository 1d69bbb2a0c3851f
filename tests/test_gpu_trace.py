"""GPU parity of the traced passes against the CPU oracle, through the C ABI.

Bar: hit records (cell, t, material, face, hit flag) bit-identical to the oracle's FP32 restatement
of traceVolumeS / brickVoxels (reference shaders/common.glsl:474-658) on the same inputs; G-buffer
position bits identical; shadowRaw halves identical; GI radiance within rtol 2e-4 / atol 2e-5 (the
sky term goes through the device's powf / expf, which are not correctly rounded).
"""
import numpy as np
import pytest
import torch

import vftrace as vf
import oracle as O
from helpers import camera, config1_camera, dense_world, hit_diff, random_rays, world

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


@pytest.fixture(scope="module", params=[1, 2], ids=["plain", "small_frame"])
def ctx(request):
    """Every test runs with the plain traversal and with the small-frame one (vf_set_traversal 1 / 2), which
    skips the walks its 2^3-block test rules out: both must give the oracle's hits."""
    c = vf.Context(0)
    c.set_traversal(request.param)
    yield c
    c.close()


def dev_hits(n):
    return torch.zeros(max(n, 1) * 20, dtype=torch.uint8, device=DEV)


def upload(ctx, w, a):
    ctx.upload(w)
    scene = vf.default_scene(a["world_bricks"], a["world_win"])
    ctx.set_scene(scene)
    return scene


def trace_rays_gpu(ctx, rays, t_cap=1e30, skip_mat=0):
    r = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).to(DEV)
    h = dev_hits(len(rays))
    ctx.trace_rays(r, len(rays), h, t_cap, skip_mat)
    ctx.sync()
    return vf.hits_to_numpy(h)[:len(rays)]


def primary_gpu(ctx, pu, W, H):
    hits = dev_hits(W * H)
    gbp = torch.zeros(W * H * 4, dtype=torch.float32, device=DEV)
    gbn = torch.zeros(W * H, dtype=torch.int32, device=DEV)
    ctx.trace_primary(pu, W, H, hits, gbp, gbn)
    ctx.sync()
    return hits, gbp, gbn


def test_primary_config1_full_res(ctx):
    w, a, ow = world(1)
    upload(ctx, w, a)
    pu, cam = config1_camera()
    W, H = 640, 360
    hits, gbp, gbn = primary_gpu(ctx, pu, W, H)
    oh, opos, onrm, _ = ow.pass_primary(cam, W, H)
    g = vf.hits_to_numpy(hits)
    bad = hit_diff(g, oh)
    assert bad.size == 0, f"{bad.size} primary hits differ, first {bad[:5]}: {g[bad[:3]]} vs {oh[bad[:3]]}"
    assert 0.3 < g["hit"].mean() <= 1.0
    assert np.array_equal(gbp.cpu().numpy().reshape(-1, 4).view(np.uint32), opos.view(np.uint32))
    assert np.array_equal(gbn.cpu().numpy().view(np.uint32), onrm)


def test_shadow_config1(ctx):
    w, a, ow = world(1)
    scene = upload(ctx, w, a)
    pu, cam = config1_camera()
    W, H = 640, 360
    _, gbp, gbn = primary_gpu(ctx, pu, W, H)
    sh = torch.zeros(W * H, dtype=torch.int32, device=DEV)
    shh = dev_hits(W * H)
    ctx.trace_shadow(W, H, gbp, gbn, sh, shh)
    ctx.sync()
    _, opos, onrm, _ = ow.pass_primary(cam, W, H)
    osh, ohits, _ = ow.pass_shadow(np.array(scene.sun_dir, np.float32), W, H, opos, onrm)
    g = vf.hits_to_numpy(shh)
    bad = hit_diff(g, ohits)
    assert bad.size == 0, f"{bad.size} shadow hits differ"
    gsh = sh.cpu().numpy().view(np.uint32)
    vis = (gsh & 0xFFFF).astype(np.uint16).view(np.float16)
    tb = (gsh >> 16).astype(np.uint16).view(np.float16)
    assert np.array_equal(vis, osh[:, 0].astype(np.float16))
    assert np.array_equal(tb, osh[:, 1].astype(np.float16))
    assert 0.0 < (vis == 0).mean() < 1.0


@pytest.mark.parametrize("size", [(640, 360), (320, 180), (1280, 720)])
def test_fused_frame_config1_matches_oracle(ctx, size):
    """The fused primary + sun pass on the fixed world (config 1's frame and the sizes around it): hits,
    G-buffer, shadowRaw and sun-ray hits equal the oracle's two passes."""
    w, a, ow = world(1)
    scene = upload(ctx, w, a)
    pu, cam = config1_camera()
    W, H = size
    hits, gbp, gbn = dev_hits(W * H), torch.zeros(W * H * 4, dtype=torch.float32, device=DEV), \
        torch.zeros(W * H, dtype=torch.int32, device=DEV)
    sh, shh = torch.zeros(W * H, dtype=torch.int32, device=DEV), dev_hits(W * H)
    ctx.trace_primary_shadow(pu, W, H, hits, gbp, gbn, sh, shh)
    ctx.sync()
    oh, opos, onrm, _ = ow.pass_primary(cam, W, H)
    osh, ohits, _ = ow.pass_shadow(np.array(scene.sun_dir, np.float32), W, H, opos, onrm)
    assert hit_diff(vf.hits_to_numpy(hits), oh).size == 0
    assert np.array_equal(gbp.cpu().numpy().reshape(-1, 4).view(np.uint32), opos.view(np.uint32))
    assert np.array_equal(gbn.cpu().numpy().view(np.uint32), onrm)
    assert hit_diff(vf.hits_to_numpy(shh), ohits).size == 0
    gsh = sh.cpu().numpy().view(np.uint32)
    assert np.array_equal((gsh & 0xFFFF).astype(np.uint16).view(np.float16), osh[:, 0].astype(np.float16))
    assert np.array_equal((gsh >> 16).astype(np.uint16).view(np.float16), osh[:, 1].astype(np.float16))


@pytest.mark.parametrize("half_res,frame", [(True, 0), (True, 5), (False, 2)])
def test_gi_config1(ctx, half_res, frame):
    w, a, ow = world(1)
    scene = upload(ctx, w, a)
    pu, cam = config1_camera(frame)
    W, H = 320, 180
    _, gbp, gbn = primary_gpu(ctx, pu, W, H)
    OW, OH = ((W + 1) // 2, (H + 1) // 2) if half_res else (W, H)
    rad = torch.zeros(OW * OH * 4, dtype=torch.float32, device=DEV)
    bh = dev_hits(OW * OH)
    ctx.trace_gi(pu, W, H, gbp, gbn, rad, bh, half_res=half_res, gi_reach=48.0)
    ctx.sync()
    _, opos, onrm, _ = ow.pass_primary(cam, W, H)
    orad, obounce, _ = ow.pass_gi(scene, frame, W, H, opos, onrm, half_res=half_res, gi_reach=48.0)
    g = vf.hits_to_numpy(bh)
    bad = hit_diff(g, obounce)
    assert bad.size == 0, f"{bad.size} GI bounce hits differ"
    gr = rad.cpu().numpy().reshape(-1, 4)
    np.testing.assert_allclose(gr, orad, rtol=2e-4, atol=2e-5)
    assert gr[:, :3].max() > 0.0


@pytest.mark.parametrize("skip_mat", [0, 5])
def test_explicit_rays_config1(ctx, skip_mat):
    w, a, ow = world(1)
    upload(ctx, w, a)
    # origins inside and around the world box; 10 % with tiny / zero direction components
    rays = random_rays(200_000, [-40, -20, -40], [296, 150, 296], seed=3 + skip_mat)
    g = trace_rays_gpu(ctx, rays, 1e30, skip_mat)
    oh, _ = ow.trace_rays(rays, 1e30, skip_mat)
    bad = hit_diff(g, oh)
    assert bad.size == 0, f"{bad.size} ray hits differ, first {bad[:4]}"


@pytest.mark.parametrize("t_cap", [0.0, 5.0, 48.0, 640.0])
def test_explicit_rays_tcap(ctx, t_cap):
    w, a, ow = world(1)
    upload(ctx, w, a)
    rays = random_rays(50_000, [0, 40, 0], [256, 128, 256], seed=11)
    g = trace_rays_gpu(ctx, rays, t_cap, 0)
    oh, _ = ow.trace_rays(rays, t_cap, 0)
    assert hit_diff(g, oh).size == 0


def test_counters_match_oracle(ctx):
    w, a, ow = world(1)
    upload(ctx, w, a)
    rays = random_rays(40_000, [0, 0, 0], [256, 128, 256], seed=5)
    ctx.set_counting(True)
    try:
        ctx.counters(reset=True)
        trace_rays_gpu(ctx, rays, 1e30, 5)
        c = ctx.counters(reset=True)
    finally:
        ctx.set_counting(False)
    _, oc = ow.trace_rays(rays, 1e30, 5)
    assert c == oc


def test_probe_rays(ctx):
    w, a, ow = world(1)
    upload(ctx, w, a)
    dims, nr, origin, spacing, frame = (4, 3, 4), 256, (10.5, 60.25, 12.75), 20.0, 7
    h = dev_hits(int(np.prod(dims)) * nr)
    ctx.trace_probes(origin, spacing, dims, nr, frame, 256.0, h)
    ctx.sync()
    g = vf.hits_to_numpy(h)[:int(np.prod(dims)) * nr]
    dirs = O.probe_dirs(nr, frame)
    rays = []
    for p in range(int(np.prod(dims))):
        px, py, pz = p % dims[0], (p // dims[0]) % dims[1], p // (dims[0] * dims[1])
        o = np.array([origin[0] + np.float32(spacing) * np.float32(px), origin[1] + np.float32(spacing) * np.float32(py),
                      origin[2] + np.float32(spacing) * np.float32(pz)], np.float32)
        for i in range(nr):
            rays.append(np.concatenate([o, dirs[i]]))
    oh, _ = ow.trace_rays(np.array(rays, np.float32), 256.0, 5)
    assert hit_diff(g, oh).size == 0


def test_brickdist_and_skytop_match_oracle(ctx):
    w, a, ow = world(1)
    upload(ctx, w, a)
    d, r = ctx.read_brickdist()
    assert np.array_equal(d, ow.d[0]) and np.array_equal(r, ow.d[2])
    assert np.array_equal(ctx.read_skytop(), ow.skytop)


def test_brickdist_incremental_update(ctx):
    """GpuWorld::sync-style edit: clear a box of bricks and fill another, rebuild dirty +- 16."""
    w, a, _ = world(1)
    a = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in a.items()}
    upload(ctx, w, a)
    wb = a["world_bricks"]
    grid = a["grid"]
    idx, val = [], []
    for z in range(10, 14):
        for y in range(2, 6):
            for x in range(5, 9):
                i = (z * wb[1] + y) * wb[0] + x
                idx.append(i); val.append(0)
    for z in range(20, 22):
        for y in range(12, 14):
            for x in range(20, 23):
                i = (z * wb[1] + y) * wb[0] + x
                idx.append(i); val.append(9)
    ctx.update(idx, val)
    grid[np.array(idx)] = np.array(val, np.uint32)
    lo, hi = [5 - 16, 2 - 16, 10 - 16], [22 + 16, 13 + 16, 21 + 16]
    dlo, dhi = [5, 2, 10], [22, 13, 21]
    ctx.build_brickdist(lo, hi, dlo, dhi)
    d, r = ctx.read_brickdist()
    full = O.OracleWorld(a)
    assert np.array_equal(d, full.d[0])
    assert np.array_equal(r, full.d[2])
    # the traversal sees the edit
    rays = random_rays(30_000, [0, 0, 0], [256, 128, 256], seed=9)
    g = trace_rays_gpu(ctx, rays, 1e30, 0)
    oh, _ = full.trace_rays(rays, 1e30, 0)
    assert hit_diff(g, oh).size == 0


def test_ring_world_primary_and_rays(ctx):
    """Endless-world shape in miniature: 3x3 chunks in a 4x4-chunk ring grid (negative absolute bricks)."""
    w, a, ow = world(2, world_bricks=(64, 32, 64), chunks=3, trees_per_chunk=60)
    assert a["world_win"][3] == 1 and a["world_win"][0] < 0
    upload(ctx, w, a)
    gy = w.ground(-37, 21)
    pu, cam = camera([-37.4, gy + 17.0, 21.6], 2.1, -0.12)
    W, H = 320, 200
    hits, gbp, _ = primary_gpu(ctx, pu, W, H)
    oh, opos, _, _ = ow.pass_primary(cam, W, H)
    g = vf.hits_to_numpy(hits)
    assert hit_diff(g, oh).size == 0
    assert np.array_equal(gbp.cpu().numpy().reshape(-1, 4).view(np.uint32), opos.view(np.uint32))
    win = a["world_win"]
    lo = [win[0] * 8 - 30, -10, win[2] * 8 - 30]
    hi = [(win[0] + 64) * 8 + 30, 266, (win[2] + 64) * 8 + 30]
    rays = random_rays(100_000, lo, hi, seed=21)
    for skip in (0, 5):
        assert hit_diff(trace_rays_gpu(ctx, rays, 1e30, skip), ow.trace_rays(rays, 1e30, skip)[0]).size == 0
    d, r = ctx.read_brickdist()
    assert np.array_equal(d, ow.d[0]) and np.array_equal(r, ow.d[2])


def test_analytic_scenes(ctx):
    """Empty world, one voxel, a water slab over a floor (skipMat semantics), a full solid world."""
    dims = (64, 32, 48)   # z, y, x
    rays = random_rays(20_000, [-8, -8, -8], [56, 40, 72], seed=4)
    empty = np.zeros(dims, np.uint8)
    one = empty.copy(); one[20, 13, 30] = 7
    water = empty.copy(); water[:, 0:8, :] = 3; water[:, 8:14, :] = 5
    solid = np.full(dims, 2, np.uint8)
    for vox in (empty, one, water, solid):
        w, a, ow = dense_world(vox)
        upload(ctx, w, a)
        for skip in (0, 5):
            g = trace_rays_gpu(ctx, rays, 1e30, skip)
            oh, _ = ow.trace_rays(rays, 1e30, skip)
            assert hit_diff(g, oh).size == 0
    # the water slab: a ray straight down stops at the water with skipMat 0, at the floor with 5
    w, a, ow = dense_world(water)
    upload(ctx, w, a)
    down = np.array([[10.5, 30.0, 10.5, 0.0, -1.0, 0.0]], np.float32)
    g0 = trace_rays_gpu(ctx, down, 1e30, 0)[0]
    g5 = trace_rays_gpu(ctx, down, 1e30, 5)[0]
    assert g0["hit"] and g0["mat"] == 5 and tuple(g0["cell"]) == (10, 13, 10) and g0["face"] == 3
    assert g5["hit"] and g5["mat"] == 3 and tuple(g5["cell"]) == (10, 7, 10) and g5["face"] == 3
    assert g0["t"] == np.float32(16.0) and g5["t"] == np.float32(22.0)


def test_frame_graph_replays_match_eager(ctx):
    """A frame recorded with vf_graph_begin/end replays to the eager result, several times, with eager
    work and synchronisations between the replays; calls that synchronise are refused while recording."""
    w, a, ow = world(1)
    upload(ctx, w, a)
    pu, _ = config1_camera()
    W, H = 320, 180
    s = torch.cuda.Stream()
    ctx.set_stream(s)
    try:
        e_hits, e_gbp, e_gbn = primary_gpu(ctx, pu, W, H)
        e_sh = torch.zeros(W * H, dtype=torch.int32, device=DEV)
        ctx.trace_shadow(W, H, e_gbp, e_gbn, e_sh)
        ctx.sync()
        hits = dev_hits(W * H)
        gbp = torch.zeros(W * H * 4, dtype=torch.float32, device=DEV)
        gbn = torch.zeros(W * H, dtype=torch.int32, device=DEV)
        sh = torch.zeros(W * H, dtype=torch.int32, device=DEV)
        ctx.graph_begin()
        with pytest.raises(vf.VfError):
            ctx.sync()
        ctx.trace_primary(pu, W, H, hits, gbp, gbn)
        ctx.trace_shadow(W, H, gbp, gbn, sh)
        g = ctx.graph_end()
        for k in range(3):
            hits.zero_(); gbp.zero_(); gbn.zero_(); sh.zero_()
            torch.cuda.synchronize()
            ctx.graph_launch(g)
            ctx.sync()
            assert torch.equal(hits, e_hits) and torch.equal(gbp, e_gbp) and torch.equal(sh, e_sh), k
            rays = torch.from_numpy(random_rays(1000, [0, 60, 0], [256, 128, 256], seed=k)).to(DEV)
            tmp = dev_hits(1000)
            ctx.trace_rays(rays, 1000, tmp)    # eager work between replays
            ctx.sync()
        g.close()
    finally:
        ctx.set_stream(None)


def test_fused_primary_shadow_equals_two_passes(ctx):
    w, a, ow = world(2, world_bricks=(64, 32, 64), chunks=3, trees_per_chunk=60)
    upload(ctx, w, a)
    gy = w.ground(-37, 21)
    pu, _ = camera([-37.4, gy + 17.0, 21.6], 4.1, -0.08)
    W, H = 333, 197
    hits, gbp, gbn = primary_gpu(ctx, pu, W, H)
    sh = torch.zeros(W * H, dtype=torch.int32, device=DEV)
    shh = dev_hits(W * H)
    ctx.trace_shadow(W, H, gbp, gbn, sh, shh)
    f_hits, f_gbp, f_gbn = dev_hits(W * H), torch.zeros_like(gbp), torch.zeros_like(gbn)
    f_sh, f_shh = torch.zeros_like(sh), dev_hits(W * H)
    ctx.trace_primary_shadow(pu, W, H, f_hits, f_gbp, f_gbn, f_sh, f_shh)
    ctx.sync()
    for x, y in ((hits, f_hits), (gbp, f_gbp), (gbn, f_gbn), (sh, f_sh), (shh, f_shh)):
        assert torch.equal(x, y)


def _camera_spots(w):
    """A camera inside solid terrain and one inside water of the fixed world."""
    solid = water = None
    for x in range(0, 256):
        for z in range(0, 256):
            gy = w.ground(x, z)
            if solid is None and x >= 60 and z >= 60 and gy > 12 and w.voxel(x, gy - 6, z) not in (0, 5):
                solid = [x + 0.37, gy - 5.6, z + 0.61]
            if water is None and w.voxel(x, 39, z) == 5 and w.voxel(x, 38, z) == 5:   # water below level 40
                water = [x + 0.52, 38.55, z + 0.33]
            if solid is not None and water is not None:
                return solid, water
    return solid, water


def test_camera_inside_solid_or_water_matches_oracle(ctx):
    """gbuffer.comp:233-266: from inside solid terrain each ray starts where it leaves the solid (rays
    that only leave the world are misses); from inside water the ray passes through water."""
    w, a, ow = world(1)
    upload(ctx, w, a)
    solid, water = _camera_spots(w)
    assert solid is not None and water is not None
    assert ow.voxel_mat_at(solid) not in (0, 5) and ow.voxel_mat_at(water) == 5
    W, H = 320, 180
    for pos, yaw, pitch in ((solid, 0.7, -0.2), (solid, 2.5, 0.6), (water, 1.9, -0.1)):
        pu, cam = camera(pos, yaw, pitch)
        hits, gbp, gbn = primary_gpu(ctx, pu, W, H)
        oh, opos, onrm, _ = ow.pass_primary(cam, W, H)
        g = vf.hits_to_numpy(hits)
        bad = hit_diff(g, oh)
        assert bad.size == 0, (pos, bad.size, g[bad[:3]], oh[bad[:3]])
        assert np.array_equal(gbp.cpu().numpy().reshape(-1, 4).view(np.uint32), opos.view(np.uint32))
        assert np.array_equal(gbn.cpu().numpy().view(np.uint32), onrm)
        assert g["hit"].mean() > 0.01
        # the fused kernel takes the same pre-step
        f_hits, f_gbp, f_gbn = dev_hits(W * H), torch.zeros_like(gbp), torch.zeros_like(gbn)
        sh = torch.zeros(W * H, dtype=torch.int32, device=DEV)
        ctx.trace_primary_shadow(pu, W, H, f_hits, f_gbp, f_gbn, sh)
        ctx.sync()
        assert torch.equal(hits, f_hits) and torch.equal(gbp, f_gbp) and torch.equal(gbn, f_gbn)
    # inside solid the pre-step matters: some rays start past tSkip > 0
    pu, cam = camera(solid, 0.7, -0.2)
    oh, opos, _, _ = ow.pass_primary(cam, W, H)
    assert np.any((oh["hit"] == 1) & (oh["t"] > 1.0))


def test_batch_view_from_inside_solid_equals_plain_launch(ctx):
    w, a, _ = world(1)
    upload(ctx, w, a)
    solid, water = _camera_spots(w)
    views = [(solid, 0.7, -0.2, 200, 120), (water, 1.9, -0.1, 128, 96), ([60.5, 100.5, 60.5], 0.4, -0.3, 96, 64)]
    reqs, outs, plain = [], [], []
    for pos, yaw, pitch, vw, vh in views:
        pu, _ = camera(pos, yaw, pitch)
        o = dev_hits(vw * vh)
        reqs.append(ctx.view_request(pu, vw, vh, o))
        outs.append(o)
        ref = dev_hits(vw * vh)
        ctx.trace_primary(pu, vw, vh, ref)
        plain.append(ref)
    for cfg in (ctx.queue_cfg((8, 8), 64), ctx.queue_cfg((16, 8), 128, refill=0)):
        for o in outs:
            o.zero_()
        ctx.trace_batch(reqs, cfg)
        ctx.sync()
        for k, (o, ref) in enumerate(zip(outs, plain)):
            assert torch.equal(o, ref), (k, cfg.refill)


def _scattered_leaves():
    """A stone floor (y < 8) under 3000 single leaf voxels (palette 7, kind 3) at y 12..55."""
    rng = np.random.default_rng(8)
    vox = np.zeros((64, 64, 64), np.uint8)   # (z, y, x)
    vox[:, 0:8, :] = 3
    idx = rng.integers(0, 64, (3000, 3))
    idx[:, 1] = rng.integers(12, 56, 3000)
    vox[idx[:, 0], idx[:, 1], idx[:, 2]] = 7
    return dense_world(vox)


def _forest(ctx):
    w, a, ow = world(2, world_bricks=(64, 32, 64), chunks=3, trees_per_chunk=60)
    scene = upload(ctx, w, a)
    gy = w.ground(-37, 21)
    pos = [-37.4, gy + 17.0, 21.6]
    return w, a, ow, scene, pos


@pytest.mark.parametrize("leaf", [0.6, 0.0])
def test_shadow_through_leaves_matches_oracle(ctx, leaf):
    """shadow_trace.comp with FOLIAGE_LIGHT on a forest: visibility leafLight^k through k leaves (exact
    FP16 halves), tFirst and the first segment's hit bit-identical; leafLight 0 is the hard shadow."""
    w, a, ow, scene, pos = _forest(ctx)
    W, H = 320, 200
    for yaw, pitch in ((4.1, -0.08), (1.0, 0.25)):
        pu, cam = camera(pos, yaw, pitch)
        pu.look_p[3] = leaf
        _, gbp, gbn = primary_gpu(ctx, pu, W, H)
        sh = torch.zeros(W * H, dtype=torch.int32, device=DEV)
        shh = dev_hits(W * H)
        ctx.trace_shadow(W, H, gbp, gbn, sh, shh, pass_ubo=pu)
        ctx.sync()
        _, opos, onrm, _ = ow.pass_primary(cam, W, H)
        osh, ohits, _ = ow.pass_shadow(np.array(scene.sun_dir, np.float32), W, H, opos, onrm, leaf_light=leaf)
        assert hit_diff(vf.hits_to_numpy(shh), ohits).size == 0
        gsh = sh.cpu().numpy().view(np.uint32)
        vis = (gsh & 0xFFFF).astype(np.uint16).view(np.float16)
        tb = (gsh >> 16).astype(np.uint16).view(np.float16)
        assert np.array_equal(vis, osh[:, 0].astype(np.float16))
        assert np.array_equal(tb, osh[:, 1].astype(np.float16))
    # scattered single leaves over a floor: many sun rays cross one to three of them
    w2, a2, ow2 = _scattered_leaves()
    scene2 = upload(ctx, w2, a2)
    pu, cam = camera([32.5, 9.5, 32.5], 0.3, -0.45)
    pu.look_p[3] = leaf
    _, gbp, gbn = primary_gpu(ctx, pu, W, H)
    sh = torch.zeros(W * H, dtype=torch.int32, device=DEV)
    shh = dev_hits(W * H)
    ctx.trace_shadow(W, H, gbp, gbn, sh, shh, pass_ubo=pu)
    ctx.sync()
    _, opos, onrm, _ = ow2.pass_primary(cam, W, H)
    osh, ohits, _ = ow2.pass_shadow(np.array(scene2.sun_dir, np.float32), W, H, opos, onrm, leaf_light=leaf)
    assert hit_diff(vf.hits_to_numpy(shh), ohits).size == 0
    gsh = sh.cpu().numpy().view(np.uint32)
    vis = (gsh & 0xFFFF).astype(np.uint16).view(np.float16)
    assert np.array_equal(vis, osh[:, 0].astype(np.float16))
    partial = ((vis > 0) & (vis < 1)).mean()
    assert (partial > 0.01) == (leaf > 0.0), partial


def test_back_lit_foliage_matches_oracle(ctx):
    """FOLIAGE_SSS, the reference default (shadow_trace.comp:35-46, lighting.comp:66-71): on the forest with the
    camera facing the sun, leaf texels facing away from the sun (gbNormal.a = 0.8) trace the sun through the
    leaf from foliageSunOrigin.  The fused pass, the G-buffer + shadow pair and the SVGF sun cone match the
    oracle bit for bit, G-buffer alpha included, and such rays are more than 1 % of the frame."""
    w, a, ow, scene, pos = _forest(ctx)
    sun = np.array(scene.sun_dir, np.float32)
    W, H = 320, 200
    yaw = float(np.arctan2(sun[0], sun[2]))
    views = [(pos, -0.05), (pos, 0.2), None, ([32.5, 9.5, 32.5], 0.1)]
    lit_some = False
    for v in views:
        if v is None:   # single leaves over a floor: back-lit leaves with the sun through 0..5 further leaves
            w, a, ow = _scattered_leaves()
            scene = upload(ctx, w, a)
            continue
        pu, cam = camera(v[0], yaw, v[1], frame=3)
        hits, gbp, gbn = dev_hits(W * H), torch.zeros(W * H * 4, dtype=torch.float32, device=DEV), \
            torch.zeros(W * H, dtype=torch.int32, device=DEV)
        sh, shh = torch.zeros(W * H, dtype=torch.int32, device=DEV), dev_hits(W * H)
        ctx.trace_primary_shadow(pu, W, H, hits, gbp, gbn, sh, shh)
        sh2, shh2 = torch.zeros(W * H, dtype=torch.int32, device=DEV), dev_hits(W * H)
        ctx.trace_shadow(W, H, gbp, gbn, sh2, shh2, pass_ubo=pu)
        cone, coneh = torch.zeros(W * H, dtype=torch.float32, device=DEV), dev_hits(W * H)
        ctx.trace_sun_cone(pu, W, H, gbp, gbn, cone, coneh)
        ctx.sync()
        oh, opos, onrm, _ = ow.pass_primary(cam, W, H)
        assert hit_diff(vf.hits_to_numpy(hits), oh).size == 0
        assert np.array_equal(gbn.cpu().numpy().view(np.uint32), onrm)
        osh, ohits, ocnt = ow.pass_shadow(sun, W, H, opos, onrm, leaf_light=0.6)
        want = ((osh[:, 0].astype(np.float16).view(np.uint16).astype(np.uint32)) |
                (osh[:, 1].astype(np.float16).view(np.uint16).astype(np.uint32) << 16))
        for g, gh in ((sh, shh), (sh2, shh2)):
            assert np.array_equal(g.cpu().numpy().view(np.uint32), want)
            assert hit_diff(vf.hits_to_numpy(gh), ohits).size == 0
        oc, och = ow.pass_sun_cone(sun, scene.sun_angle, 0.6, 3, W, H, opos, onrm)
        assert np.array_equal(cone.cpu().numpy(), oc)
        assert hit_diff(vf.hits_to_numpy(coneh), och).size == 0
        nb = onrm.view(np.uint32)
        N = np.stack([((nb >> (8 * k)) & 255).astype(np.float32) / 255 * 2 - 1 for k in range(3)], 1)
        back_lit = (opos[:, 3] >= 0) & ((nb >> 24) == 204) & ((N * sun).sum(1) <= 0)
        assert back_lit.mean() > 0.01, back_lit.mean()
        assert (osh[back_lit, 1] == 0).all()
        lit_some |= bool((osh[back_lit, 0] > 0).any())
    assert lit_some


def test_sun_cone_matches_oracle(ctx):
    """lighting.comp's cone-jittered, stochastic sun sample: bit-identical to the oracle."""
    w, a, ow = _scattered_leaves()
    scene = upload(ctx, w, a)
    W, H = 320, 200
    for frame in (0, 7):
        pu, cam = camera([32.5, 9.5, 32.5], 0.3, -0.45, frame=frame)
        _, gbp, gbn = primary_gpu(ctx, pu, W, H)
        out = torch.zeros(W * H, dtype=torch.float32, device=DEV)
        hh = dev_hits(W * H)
        ctx.trace_sun_cone(pu, W, H, gbp, gbn, out, hh)
        ctx.sync()
        _, opos, onrm, _ = ow.pass_primary(cam, W, H)
        o, oh = ow.pass_sun_cone(np.array(scene.sun_dir, np.float32), scene.sun_angle, 0.6, frame, W, H, opos, onrm)
        assert np.array_equal(out.cpu().numpy(), o)
        assert hit_diff(vf.hits_to_numpy(hh), oh).size == 0
        assert 0.0 < (o == 0).mean() < 1.0


def test_world_loaded_from_a_save_traces_identically(ctx, tmp_path):
    """The config-1 world written as a .vfworld save (the reference's format), read back and uploaded:
    the primary pass equals the oracle's on the original world."""
    from vftrace import vfworld as V
    w, a, ow = world(1)
    path = tmp_path / "cfg1.vfworld"
    V.write_vfworld(path, a)
    back = V.arrays_from_save(V.read_vfworld(path))
    ctx.upload_arrays(back)
    ctx.set_scene(vf.default_scene(back["world_bricks"], back["world_win"]))
    pu, cam = config1_camera()
    W, H = 320, 180
    hits, gbp, gbn = primary_gpu(ctx, pu, W, H)
    oh, opos, onrm, _ = ow.pass_primary(cam, W, H)
    assert hit_diff(vf.hits_to_numpy(hits), oh).size == 0
    assert np.array_equal(gbp.cpu().numpy().reshape(-1, 4).view(np.uint32), opos.view(np.uint32))
