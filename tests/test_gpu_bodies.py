"""GPU parity of the traced passes over a scene with bodies (traceSceneS / traceObjectsA / traceObject,
reference shaders/common.glsl:664-793; GI_OBJECTS, common.glsl:780, 2267-2271), through the C ABI.

Bar: as for the world (hit records, G-buffer bits, shadowRaw halves, body indices bit-identical to the
oracle; GI radiance within rtol 2e-4 / atol 2e-5), with the BVH built by the library equal to the
oracle's restatement of build_bvh (src/gfx/gpu_world.rs:550-600).
"""
import numpy as np
import pytest
import torch

import oracle as O
import vftrace as vf
from vftrace import bodies as B
from helpers import camera, config1_camera, hit_diff, random_bodies, world

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")
_SCENES = {}


@pytest.fixture(scope="module")
def ctx():
    c = vf.Context(0)
    yield c
    c.close()


def dev_hits(n):
    return torch.zeros(max(n, 1) * 20, dtype=torch.uint8, device=DEV)


def body_scene(seed=1, n_random=20):
    """The config-1 world with the seven demo-shaped bodies and n_random small ones in the config-1
    camera's view: (arrays, headers, grid, oracle world, vols, poses)."""
    key = (seed, n_random)
    if key not in _SCENES:
        w, a, _ = world(1)
        rng = np.random.default_rng(seed)
        _, cam = config1_camera()
        pos, fwd = np.asarray(cam["pos"], np.float64), np.asarray(cam["fwd"], np.float64)
        vols, poses = B.demo_bodies((0.0, 0.0), lambda x, z: 0.0)
        placed = []
        for k, v in enumerate(vols):
            c = pos + fwd * (40.0 + 18.0 * k) + rng.uniform(-12, 12, 3)
            placed.append(B.pose(c, v.shape, float(rng.uniform(0, 6.28)), float(rng.uniform(-0.6, 0.6))))
        rv, rp = random_bodies(rng, n_random, pos + fwd * 15.0 - 30.0, pos + fwd * 150.0 + 30.0)
        vols, poses = vols + rv, placed + rp
        a2, hd, g = B.attach(a, vols, poses)
        ow = O.OracleWorld(a2)
        ow.set_bodies(hd, g)
        _SCENES[key] = (a2, hd, g, ow, vols, poses)
    return _SCENES[key]


def upload(ctx, seed=1, n_random=20, num_objects=None):
    a2, hd, g, ow, vols, poses = body_scene(seed, n_random)
    ctx.upload_arrays(a2)
    ctx.objects_upload(hd, g)
    scene = vf.default_scene(a2["world_bricks"], a2["world_win"])
    scene.num_objects = len(hd) if num_objects is None else num_objects
    ctx.set_scene(scene)
    return scene, ow, hd


def gbuffer_gpu(ctx, pu, W, H, shadow=False):
    hits = dev_hits(W * H)
    gbp = torch.zeros(W * H * 4, dtype=torch.float32, device=DEV)
    gbn = torch.zeros(W * H, dtype=torch.int32, device=DEV)
    body = torch.full((W * H,), 7, dtype=torch.int32, device=DEV)
    sh = torch.zeros(W * H, dtype=torch.int32, device=DEV) if shadow else None
    shh = dev_hits(W * H) if shadow else None
    ctx.trace_gbuffer(pu, W, H, hits, gbp, gbn, body, sh, shh)
    ctx.sync()
    return hits, gbp, gbn, body, sh, shh


def shadow_halves(sh):
    raw = sh.cpu().numpy().view(np.uint32)
    return np.stack([(raw & 0xFFFF).astype(np.uint16).view(np.float16), (raw >> 16).astype(np.uint16).view(np.float16)], 1)


def test_library_bvh_equals_the_oracle_builder(ctx):
    _, ow, hd = upload(ctx)
    nodes, prims = ctx.objects_read_bvh()
    assert np.array_equal(nodes.view(np.float32).reshape(-1, 8).view(np.uint32), ow.bvh_nodes.view(np.uint32))
    assert np.array_equal(prims, ow.bvh_prims)
    assert len(prims) == len(hd)


@pytest.mark.parametrize("view", [0, 1])
def test_gbuffer_with_bodies_matches_oracle(ctx, view):
    _, ow, _ = upload(ctx)
    pu, cam = config1_camera() if view == 0 else camera([60.5, 100.25, 90.75], 1.1, -0.15)
    W, H = 480, 270
    hits, gbp, gbn, body, _, _ = gbuffer_gpu(ctx, pu, W, H)
    oh, opos, onrm, _ = ow.pass_primary(cam, W, H)
    obody = ow.last_body
    g = vf.hits_to_numpy(hits)
    bad = hit_diff(g, oh)
    assert bad.size == 0, f"{bad.size} hits differ, first {bad[:5]}: {g[bad[:3]]} vs {oh[bad[:3]]}"
    assert np.array_equal(gbp.cpu().numpy().reshape(-1, 4).view(np.uint32), opos.view(np.uint32))
    assert np.array_equal(gbn.cpu().numpy().view(np.uint32), onrm)
    assert np.array_equal(body.cpu().numpy(), obody)
    assert (obody >= 0).sum() > 2000, "the view shows too few body pixels"
    assert np.all((g["body"] == 1) == (obody >= 0))
    # vf_trace_primary gives the same hits and G-buffer without the body index
    h2 = dev_hits(W * H)
    p2 = torch.zeros_like(gbp)
    n2 = torch.zeros_like(gbn)
    ctx.trace_primary(pu, W, H, h2, p2, n2)
    ctx.sync()
    assert torch.equal(h2, hits) and torch.equal(p2, gbp) and torch.equal(n2, gbn)


def test_fused_and_separate_shadow_with_bodies(ctx):
    scene, ow, _ = upload(ctx)
    pu, cam = config1_camera()
    W, H = 400, 240
    hits, gbp, gbn, body, sh, shh = gbuffer_gpu(ctx, pu, W, H, shadow=True)
    oh, opos, onrm, _ = ow.pass_primary(cam, W, H)
    assert hit_diff(vf.hits_to_numpy(hits), oh).size == 0
    osh, osh_hits, _ = ow.pass_shadow(np.array(scene.sun_dir, np.float32), W, H, opos, onrm)
    want = np.stack([osh[:, 0].astype(np.float16), osh[:, 1].astype(np.float16)], 1)
    assert np.array_equal(shadow_halves(sh).view(np.uint16), want.view(np.uint16))
    g = vf.hits_to_numpy(shh)
    assert hit_diff(g, osh_hits).size == 0
    assert (g["body"] == 1).sum() > 200, "too few sun rays blocked by bodies"
    # the fused entry point of the world-only API and the separate pass agree
    f_hits, f_gbp, f_gbn = dev_hits(W * H), torch.zeros_like(gbp), torch.zeros_like(gbn)
    f_sh = torch.zeros_like(sh)
    ctx.trace_primary_shadow(pu, W, H, f_hits, f_gbp, f_gbn, f_sh)
    s_sh = torch.zeros_like(sh)
    s_hits = dev_hits(W * H)
    ctx.trace_shadow(W, H, gbp, gbn, s_sh, s_hits, pass_ubo=pu)
    ctx.sync()
    assert torch.equal(f_hits, hits) and torch.equal(f_sh, sh) and torch.equal(s_sh, sh) and torch.equal(s_hits, shh)


def test_sun_cone_with_bodies(ctx):
    scene, ow, _ = upload(ctx)
    W, H = 320, 180
    for frame in (0, 3):
        pu, cam = config1_camera(frame)
        _, gbp, gbn, _, _, _ = gbuffer_gpu(ctx, pu, W, H)
        out = torch.zeros(W * H, dtype=torch.float32, device=DEV)
        hh = dev_hits(W * H)
        ctx.trace_sun_cone(pu, W, H, gbp, gbn, out, hh)
        ctx.sync()
        _, opos, onrm, _ = ow.pass_primary(cam, W, H)
        o, oh = ow.pass_sun_cone(np.array(scene.sun_dir, np.float32), scene.sun_angle, 0.6, frame, W, H, opos, onrm)
        assert np.array_equal(out.cpu().numpy(), o)
        assert hit_diff(vf.hits_to_numpy(hh), oh).size == 0


@pytest.mark.parametrize("half_res", [True, False])
def test_gi_with_bodies(ctx, half_res):
    scene, ow, _ = upload(ctx)
    pu, cam = config1_camera(2)
    W, H = 320, 180
    _, gbp, gbn, _, _, _ = gbuffer_gpu(ctx, pu, W, H)
    OW, OH = ((W + 1) // 2, (H + 1) // 2) if half_res else (W, H)
    rad = torch.zeros(OW * OH * 4, dtype=torch.float32, device=DEV)
    bh = dev_hits(OW * OH)
    ctx.trace_gi(pu, W, H, gbp, gbn, rad, bh, half_res=half_res, gi_reach=48.0)
    ctx.sync()
    _, opos, onrm, _ = ow.pass_primary(cam, W, H)
    orad, obounce, _ = ow.pass_gi(scene, 2, W, H, opos, onrm, half_res=half_res, gi_reach=48.0)
    g = vf.hits_to_numpy(bh)
    assert hit_diff(g, obounce).size == 0
    assert (g["body"] == 1).sum() > 5, "too few bounce rays meet bodies"
    np.testing.assert_allclose(rad.cpu().numpy().reshape(-1, 4), orad, rtol=2e-4, atol=2e-5)


@pytest.mark.parametrize("bounces", [2, 3])
def test_gi_path_with_bodies(ctx, bounces):
    scene, ow, _ = upload(ctx)
    pu, cam = config1_camera()
    W, H = 160, 90
    _, gbp, gbn, _, _, _ = gbuffer_gpu(ctx, pu, W, H)
    out = torch.zeros(W * H * 4, dtype=torch.float32, device=DEV)
    ctx.trace_gi_path(pu, W, H, gbp, gbn, out, frame=1, spp=2, bounces=bounces)
    ctx.sync()
    _, opos, onrm, _ = ow.pass_primary(cam, W, H)
    o = ow.pass_gi_path(scene, W, H, opos, onrm, 1, 2, 0x68e31da4, bounces=bounces)
    np.testing.assert_allclose(out.cpu().numpy().reshape(-1, 4), o, rtol=2e-4, atol=2e-5)


def test_probes_with_bodies(ctx):
    _, ow, _ = upload(ctx)
    _, cam = config1_camera()
    pos, fwd = np.asarray(cam["pos"], np.float32), np.asarray(cam["fwd"], np.float32)
    origin = (pos + fwd * 60.0 - 24.0).astype(np.float32)
    dims, nr, spacing, frame, reach = (4, 3, 4), 128, 16.0, 3, 96.0
    h = dev_hits(int(np.prod(dims)) * nr)
    ctx.trace_probes(origin, spacing, dims, nr, frame, reach, h)
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
    rays = np.array(rays, np.float32)
    oh, _ = ow.trace_rays(rays, reach, 5)
    # GI_ADD_OBJECTS (probe_trace.comp:47): bodies nearer than min(the world hit or the reach, 16)
    n_body = 0
    for k in range(len(rays)):
        lim = min(float(oh["t"][k]) if oh["hit"][k] else reach, 16.0)
        b = ow.objects_only(rays[k, :3], rays[k, 3:], lim)
        if b.hit:
            n_body += 1
            oh[k] = ((b.cell[0], b.cell[1], b.cell[2]), b.t, b.mat, O.face_code(b.n), 1, 1)
    assert hit_diff(g, oh).size == 0
    assert n_body > 10


def test_camera_in_solid_and_in_a_body(ctx):
    _, ow, hd = upload(ctx)
    w, _, _ = world(1)
    W, H = 240, 135
    # a camera inside solid terrain: the exitSolid pre-step, then world and bodies from there
    gx, gz = 100, 120
    spots = [np.array([gx + 0.5, w.ground(gx, gz) - 3.5, gz + 0.5], np.float32)]
    # a camera inside the first body's box
    spots.append(((hd["aabb_min"][0, :3] + hd["aabb_max"][0, :3]) * 0.5).astype(np.float32))
    for pos in spots:
        for yaw, pitch in ((0.4, 0.1), (2.2, -0.3)):
            pu, cam = camera(pos, yaw, pitch)
            hits, gbp, gbn, body, _, _ = gbuffer_gpu(ctx, pu, W, H)
            oh, opos, onrm, _ = ow.pass_primary(cam, W, H)
            assert hit_diff(vf.hits_to_numpy(hits), oh).size == 0
            assert np.array_equal(gbp.cpu().numpy().reshape(-1, 4).view(np.uint32), opos.view(np.uint32))
            assert np.array_equal(gbn.cpu().numpy().view(np.uint32), onrm)
            assert np.array_equal(body.cpu().numpy(), ow.last_body)


def test_moving_bodies_and_graph_replay(ctx):
    """update_objects: new poses and a rebuilt BVH; a recorded frame graph sees them on replay."""
    a2, hd, g, ow, vols, poses = body_scene()
    upload(ctx)
    pu, cam = config1_camera()
    W, H = 256, 144
    hits = dev_hits(W * H)
    gbp = torch.zeros(W * H * 4, dtype=torch.float32, device=DEV)
    gbn = torch.zeros(W * H, dtype=torch.int32, device=DEV)
    body = torch.zeros(W * H, dtype=torch.int32, device=DEV)
    ctx.set_stream(torch.cuda.Stream())
    try:
        ctx.graph_begin()
        ctx.trace_gbuffer(pu, W, H, hits, gbp, gbn, body)
        graph = ctx.graph_end()
        rng = np.random.default_rng(9)
        for step in range(3):
            moved = [B.pose(np.asarray(p[1]) + r @ (np.array(v.shape[::-1], np.float32) * 0.5) + rng.uniform(-6, 6, 3),
                            v.shape, float(rng.uniform(0, 6.28)), float(rng.uniform(-0.5, 0.5)))
                     for v, p, r in zip(vols, poses, [p[0] for p in poses])]
            hd2, g2, _ = B.pack(vols, moved, a2["atlas"].size // 512 - int((g >= 256).sum()))
            assert np.array_equal(g2, g)
            ctx.objects_set_headers(hd2)
            ctx.graph_launch(graph)
            ctx.sync()
            ow2 = O.OracleWorld(a2)
            ow2.set_bodies(hd2, g2)
            oh, opos, onrm, _ = ow2.pass_primary(cam, W, H)
            assert hit_diff(vf.hits_to_numpy(hits), oh).size == 0, step
            assert np.array_equal(body.cpu().numpy(), ow2.last_body)
            nodes, prims = ctx.objects_read_bvh()
            assert np.array_equal(prims, ow2.bvh_prims)
        graph.close()
    finally:
        ctx.set_stream(None)


def test_body_grid_update_and_scene_switch(ctx):
    a2, hd, g, ow, vols, poses = body_scene()
    upload(ctx)
    pu, cam = config1_camera()
    W, H = 256, 144
    # empty every brick of the first three bodies (sync_objects)
    idx = np.concatenate([np.arange(hd["dims_bricks"][k, 3], hd["dims_bricks"][k, 3] + np.prod(hd["dims_bricks"][k, :3]))
                          for k in range(3)]).astype(np.uint32)
    ctx.objects_update_grid(idx, np.zeros_like(idx))
    g2 = g.copy()
    g2[idx] = 0
    ow2 = O.OracleWorld(a2)
    ow2.set_bodies(hd, g2)
    hits, _, _, body, _, _ = gbuffer_gpu(ctx, pu, W, H)
    oh, _, _, _ = ow2.pass_primary(cam, W, H)
    assert hit_diff(vf.hits_to_numpy(hits), oh).size == 0
    assert not np.isin(body.cpu().numpy(), [0, 1, 2]).any()
    # num_objects = 0: the world alone (the world-only kernels; body index -1)
    scene = vf.default_scene(a2["world_bricks"], a2["world_win"])
    ctx.set_scene(scene)
    hits, _, _, body, _, _ = gbuffer_gpu(ctx, pu, W, H)
    w1, a1, ow1 = world(1)
    oh1, _, _, _ = ow1.pass_primary(cam, W, H)
    assert hit_diff(vf.hits_to_numpy(hits), oh1).size == 0
    assert (body.cpu().numpy() == -1).all()
    # a scene whose num_objects differs from the uploaded bodies is refused, by the passes and the batches
    _, gbp, gbn, _, _, _ = gbuffer_gpu(ctx, pu, W, H)
    gi_h, sun_h = dev_hits(W * H), dev_hits(W * H)
    scene.num_objects = len(hd) + 1
    ctx.set_scene(scene)
    with pytest.raises(vf.VfError) as e:
        gbuffer_gpu(ctx, pu, W, H)
    assert "num_objects" in str(e.value)
    with pytest.raises(vf.VfError) as e:
        ctx.trace_gi_batch(pu, W, H, 48.0, gbp, gbn, gi_h, sun_h)
    assert "num_objects" in str(e.value)
    with pytest.raises(vf.VfError) as e:
        ctx.trace_batch([ctx.view_request(pu, W, H, gi_h)])
    assert "num_objects" in str(e.value)
    scene.num_objects = len(hd)
    ctx.set_scene(scene)


def test_caller_bvh_checks(ctx):
    _, ow, hd = upload(ctx)
    nodes = np.zeros(len(ow.bvh_nodes), vf.BVH_NODE_DTYPE)
    nodes.view(np.float32).reshape(-1, 8)[:] = ow.bvh_nodes
    ctx.objects_set_bvh(nodes, ow.bvh_prims)   # the oracle's own tree is accepted
    bad = nodes.copy()
    bad["first"][0] = len(bad)                 # children out of range
    with pytest.raises(vf.VfError):
        ctx.objects_set_bvh(bad, ow.bvh_prims)
    bad = nodes.copy()
    leaf = np.nonzero(bad["count"] > 0)[0][0]
    bad["first"][leaf] = len(hd)               # prims out of range
    with pytest.raises(vf.VfError):
        ctx.objects_set_bvh(bad, ow.bvh_prims)
    pr = ow.bvh_prims.copy()
    pr[0] = len(hd) + 3                         # a prim naming no body
    with pytest.raises(vf.VfError):
        ctx.objects_set_bvh(nodes, pr)
    # a degenerate chain of 30 interior nodes (node 0, then 1, 3, 5, ...; each one's right child a leaf):
    # deeper than the walk's 24-entry stack
    n = 30
    chain = np.zeros(2 * n + 1, vf.BVH_NODE_DTYPE)
    chain["min"] = -1e30
    chain["max"] = 1e30
    chain["count"] = 1                          # leaves of prim 0 ...
    for k in range(n):
        node = 0 if k == 0 else 2 * k - 1
        chain["count"][node] = 0                # ... except the interior chain
        chain["first"][node] = 2 * k + 1
    with pytest.raises(vf.VfError) as e:
        ctx.objects_set_bvh(chain, np.zeros(1, np.uint32))
    assert "stack" in str(e.value)
    ctx.objects_set_bvh(nodes, ow.bvh_prims)


_GI_BATCH_ORACLE = {}


@pytest.mark.parametrize("refill", [32, 0, -1])
def test_gi_batch_with_bodies(ctx, refill):
    """vf_trace_gi_batch over a scene with bodies, every queue mode: GI rays meet the bodies nearer than
    min(world hit or reach, 16), a body hit's sun ray leaves along the body's rotated normal, a sun ray that
    missed the world meets any body (GI_OBJECTS, common.glsl:780, 2267-2271)."""
    scene, ow, _ = upload(ctx)
    pu, _ = config1_camera(3)
    W, H, reach = 320, 180, 96.0
    _, gbp, gbn, _, _, _ = gbuffer_gpu(ctx, pu, W, H)
    gi, sun = dev_hits(W * H), dev_hits(W * H)
    ctx.trace_gi_batch(pu, W, H, reach, gbp, gbn, gi, sun, ctx.queue_cfg((16, 8), 128, True, refill=refill))
    ctx.sync()
    if "o" not in _GI_BATCH_ORACLE:
        _GI_BATCH_ORACLE["o"] = O.gi_batch(ow, W, H, 3, gbp.cpu().numpy().reshape(-1, 4), gbn.cpu().numpy().view(np.uint32),
                                           np.array(scene.sun_dir, np.float32), reach, bodies=True)
    og, osun = _GI_BATCH_ORACLE["o"]
    g, s = vf.hits_to_numpy(gi)[:W * H], vf.hits_to_numpy(sun)[:W * H]
    assert hit_diff(g, og).size == 0
    assert hit_diff(s, osun).size == 0
    assert (g["body"] == 1).sum() > 20, "too few GI rays meet bodies"
    assert ((s["body"] == 1) & (s["hit"] == 1)).sum() > 0, "no sun ray is blocked by a body"


def test_uneven_batch_with_bodies_matches_plain_launches(ctx):
    """vf_trace_batch over a scene with bodies: each view equals its plain primary launch (traceSceneS from
    the camera, with the exitSolid pre-step for a camera in terrain), each probe grid its plain probe
    launch (GI_ADD_OBJECTS), in the persistent and the chunk modes."""
    _, ow, hd = upload(ctx)
    w, _, _ = world(1)
    pu0, cam = config1_camera()
    pos = np.asarray(cam["pos"], np.float32)
    fwd = np.asarray(cam["fwd"], np.float32)
    gx, gz = 100, 120
    in_solid = np.array([gx + 0.5, w.ground(gx, gz) - 3.5, gz + 0.5], np.float32)
    in_body = ((hd["aabb_min"][0, :3] + hd["aabb_max"][0, :3]) * 0.5).astype(np.float32)
    views = [(pu0, 320, 180), (camera(in_solid, 0.4, 0.1)[0], 96, 64), (camera(in_body, 2.2, -0.3)[0], 130, 75)]
    origin = (pos + fwd * 60.0 - 24.0).astype(np.float32)
    probes = [(origin, 16.0, (4, 3, 4), 128, 3, 96.0)]
    reqs, outs = [], []
    for (pu, vw, vh) in views:
        outs.append(dev_hits(vw * vh))
        reqs.append(ctx.view_request(pu, vw, vh, outs[-1]))
    for (o, sp, dims, nr, fr, rc) in probes:
        outs.append(dev_hits(int(np.prod(dims)) * nr))
        reqs.append(ctx.probe_request(o, sp, dims, nr, fr, rc, outs[-1]))
    refs = []
    for (pu, vw, vh) in views:
        r = dev_hits(vw * vh)
        ctx.trace_primary(pu, vw, vh, r)
        refs.append((r, vw * vh))
    for (o, sp, dims, nr, fr, rc) in probes:
        n = int(np.prod(dims)) * nr
        r = dev_hits(n)
        ctx.trace_probes(o, sp, dims, nr, fr, rc, r)
        refs.append((r, n))
    ctx.sync()
    body_hits = sum(int((vf.hits_to_numpy(r)[:n]["body"] == 1).sum()) for r, n in refs)
    assert body_hits > 50, body_hits
    for cfg in (ctx.queue_cfg((8, 8), 64), ctx.queue_cfg((16, 16), 512, refill=0), ctx.queue_cfg((8, 8), 64, refill=-1),
                ctx.queue_cfg((8, 8), 64, refill=-1, traversal=1)):
        for o in outs:
            o.zero_()
        ctx.trace_batch(reqs, cfg)
        ctx.sync()
        for k, (r, n) in enumerate(refs):
            assert torch.equal(outs[k][:n * 20], r[:n * 20]), (k, cfg.refill)
