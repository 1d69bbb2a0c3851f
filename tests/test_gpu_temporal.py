"""GPU parity of the temporal consumers (include/vf_temporal.h) against the CPU oracle, through the C ABI:
gbuffer.comp's motion vectors, gi.comp's reprojected accumulation, lighting.comp's shadow history and the
a-trous levels of atrous.comp.

Bar: bit-identical outputs (half-float history bits, float moments and shadow history, guide words,
motion and delta halves).  The inputs are random well-formed frames at ragged and 1080p sizes, and a
chain of traced frames of a moving camera (G-buffer, motion, GI, sun cone, accumulation, filter), in
which the oracle carries its own histories from frame to frame."""
import numpy as np
import pytest
import torch

import oracle as O
import vftrace as vf
from helpers import camera, world
from temporal_helpers import random_frame

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


@pytest.fixture(scope="module")
def ctx():
    c = vf.Context(0)
    yield c
    c.close()


def dev(a, dtype=None):
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint16:
        a = a.view(np.int16)
    elif a.dtype == np.uint32:
        a = a.view(np.int32)
    return torch.from_numpy(a).to(DEV)


def host(t, dtype):
    return t.cpu().numpy().view(dtype)


def first_diff(a, b):
    bad = np.nonzero(a.reshape(a.shape[0], -1) != b.reshape(b.shape[0], -1))
    return None if bad[0].size == 0 else (bad[0][0], bad[1][0], bad[0].size)


def gi_acc_gpu(ctx, pu, w, h, fr, bicubic):
    hw, hh = O.half_dims(w, h)
    hist = torch.zeros(hh * hw * 4, dtype=torch.int16, device=DEV)
    mom = torch.zeros(hh * hw * 4, dtype=torch.float32, device=DEV)
    guide = torch.zeros(hh * hw * 4, dtype=torch.int32, device=DEV)
    ins = [dev(fr[k]) for k in ("gb_pos", "gb_normal", "gb_pos_prev", "gb_normal_prev", "gb_motion", "gb_delta",
                                "indirect", "hist_prev", "mom_prev")]
    ctx.gi_accumulate(pu, w, h, *ins, hist, mom, guide, bicubic=bicubic)
    ctx.sync()
    return (host(hist, np.uint16).reshape(hh, hw, 4), host(mom, np.float32).reshape(hh, hw, 4),
            host(guide, np.uint32).reshape(hh, hw, 4))


def pass_of(fr, frame):
    pu = vf.PassUbo()
    pu.frame_idx = frame
    for i in range(3):
        pu.shift_p[i] = float(fr["shift"][i])
    return pu


@pytest.mark.parametrize("w,h,frame,bicubic", [(333, 187, 0, True), (333, 187, 1, False), (1920, 1080, 2, True),
                                               (64, 9, 3, True), (1, 1, 0, True), (2, 3, 5, False)])
def test_gi_accumulate_matches_oracle(ctx, w, h, frame, bicubic):
    fr = random_frame(w, h, seed=w * 7 + h + frame)
    pu = pass_of(fr, frame)
    g = gi_acc_gpu(ctx, pu, w, h, fr, bicubic)
    o = O.gi_accumulate(w, h, frame, fr["shift"], fr["gb_pos"], fr["gb_normal"], fr["gb_pos_prev"],
                        fr["gb_normal_prev"], fr["gb_motion"], fr["gb_delta"], fr["indirect"], fr["hist_prev"],
                        fr["mom_prev"], bicubic=bicubic)
    for name, a, b in zip(("hist", "mom", "guide"), g, o):
        d = first_diff(a.view(np.uint32) if a.dtype == np.float32 else a, b.view(np.uint32) if b.dtype == np.float32 else b)
        assert d is None, f"{name}: first difference at texel {d[0]} word {d[1]}, {d[2]} words differ"
    if w * h > 100:   # the random frame reuses history for a good share of texels
        assert (g[0][..., 3].view(np.float16) > 1).mean() > 0.1


@pytest.mark.parametrize("w,h", [(333, 187), (1920, 1080), (7, 5)])
def test_light_accumulate_matches_oracle(ctx, w, h):
    fr = random_frame(w, h, seed=w + h)
    pu = pass_of(fr, 0)
    out = torch.zeros(h * w * 4, dtype=torch.float32, device=DEV)
    ins = [dev(fr[k]) for k in ("gb_pos", "gb_normal", "gb_pos_prev", "gb_normal_prev", "gb_motion", "gb_delta",
                                "shadow", "light_hist_prev")]
    ctx.light_accumulate(pu, w, h, *ins, out)
    ctx.sync()
    g = host(out, np.float32).reshape(h, w, 4)
    o = O.light_accumulate(w, h, fr["shift"], fr["gb_pos"], fr["gb_normal"], fr["gb_pos_prev"], fr["gb_normal_prev"],
                           fr["gb_motion"], fr["gb_delta"], fr["shadow"], fr["light_hist_prev"])
    d = first_diff(g.view(np.uint32), o.view(np.uint32))
    assert d is None, f"first difference at pixel {d[0]} word {d[1]}, {d[2]} words differ"


@pytest.mark.parametrize("w,h,levels,histfix", [(333, 187, 3, True), (333, 187, 5, False), (1920, 1080, 3, True),
                                                (9, 7, 5, True), (1, 1, 1, True)])
def test_atrous_matches_oracle(ctx, w, h, levels, histfix):
    fr = random_frame(w, h, seed=3 * w + h, max_count=8.0)      # many fresh texels: the history fix runs
    o_h, o_m, o_g = O.gi_accumulate(w, h, 1, fr["shift"], fr["gb_pos"], fr["gb_normal"], fr["gb_pos_prev"],
                                    fr["gb_normal_prev"], fr["gb_motion"], fr["gb_delta"], fr["indirect"],
                                    fr["hist_prev"], fr["mom_prev"])
    hw, hh = O.half_dims(w, h)
    f0 = torch.zeros(hh * hw * 4, dtype=torch.int16, device=DEV)
    f1 = torch.zeros(hh * hw * 4, dtype=torch.int16, device=DEV)
    ctx.gi_atrous(w, h, levels, dev(fr["gb_pos"]), dev(fr["gb_normal"]), dev(o_m), dev(o_h), dev(o_g), f0, f1,
                  histfix=histfix)
    ctx.sync()
    e0, e1 = O.gi_atrous(w, h, levels, fr["gb_pos"], fr["gb_normal"], o_m, o_h, o_g, histfix=histfix)
    for name, a, b in (("filt0", host(f0, np.uint16).reshape(hh, hw, 4), e0),
                       ("filt1", host(f1, np.uint16).reshape(hh, hw, 4), e1)):
        if levels == 1 and name == "filt1":
            continue
        d = first_diff(a, b)
        assert d is None, f"{name}: first difference at texel {d[0]} half {d[1]}, {d[2]} halves differ"


def test_motion_matches_oracle_world_and_bodies(ctx):
    from test_gpu_bodies import upload as upload_bodies
    scene, ow, hd = upload_bodies(ctx)
    W, H = 320, 180
    pu0, _ = camera([37.3, 96.4, 41.7], 0.8, -0.35)
    pu, _ = camera([38.1, 96.0, 42.9], 0.83, -0.33, frame=1)
    vf.set_prev_camera(pu, pu0)
    pu.jitter[0], pu.jitter[1] = 0.25, -0.125
    # the bodies moved since the last frame: previous pose = a shifted and turned copy
    hd2 = hd.copy()
    rng = np.random.default_rng(4)
    for i in range(len(hd2)):
        o2w = hd2[i]["obj_to_world"].reshape(4, 4).T.copy()
        o2w[:3, 3] -= rng.uniform(-1.5, 1.5, 3).astype(np.float32)
        c, s = np.cos(0.05), np.sin(0.05)
        o2w[:3, :3] = o2w[:3, :3] @ np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]], np.float32)
        hd2[i]["prev_obj_to_world"] = o2w.T.reshape(-1)
    ctx.objects_set_headers(hd2)
    hits = torch.zeros(W * H * 20, dtype=torch.uint8, device=DEV)
    gbp = torch.zeros(W * H * 4, dtype=torch.float32, device=DEV)
    gbn = torch.zeros(W * H, dtype=torch.int32, device=DEV)
    body = torch.zeros(W * H, dtype=torch.int32, device=DEV)
    ctx.trace_gbuffer(pu, W, H, hits, gbp, gbn, body)
    mot = torch.zeros(W * H, dtype=torch.int32, device=DEV)
    dl = torch.zeros(W * H * 4, dtype=torch.int16, device=DEV)
    ctx.gbuffer_motion(pu, W, H, gbp, mot, dl, body=body)
    ctx.sync()
    b = host(body, np.int32).reshape(H, W)
    assert (b >= 0).sum() > 200, "bodies should cover part of the view"
    pos = host(gbp, np.float32).reshape(H, W, 4)
    om, od = O.gbuffer_motion(W, H, pu, pos, b, hd2)
    assert first_diff(host(mot, np.uint32).reshape(H, W), om) is None
    assert first_diff(host(dl, np.uint16).reshape(H, W, 4), od) is None
    # without the body indices every hit is the static world
    ctx.gbuffer_motion(pu, W, H, gbp, mot, dl)
    ctx.sync()
    om2, od2 = O.gbuffer_motion(W, H, pu, pos)
    assert first_diff(host(mot, np.uint32).reshape(H, W), om2) is None and np.all(host(dl, np.uint16) == 0)
    scene.num_objects = 0
    ctx.set_scene(scene)
    ctx.objects_upload(np.zeros(0, O.OBJ_HEADER_DTYPE), np.zeros(0, np.uint32))


def test_traced_frame_chain(ctx):
    """Six frames of a moving camera over the config-1 world: the G-buffer, motion vectors, half-res GI,
    the cone-jittered sun ray, then gi.comp accumulation, lighting.comp history and 3 a-trous levels on
    the GPU; the oracle runs the temporal passes on the same traced inputs with its own histories."""
    w, a, ow = world(1)
    ctx.upload(w)
    scene = vf.default_scene(a["world_bricks"], a["world_win"])
    ctx.set_scene(scene)
    W, H = 400, 224
    hw, hh = O.half_dims(W, H)
    zeros = lambda n, dt: torch.zeros(n, dtype=dt, device=DEV)
    g_hist = [zeros(hh * hw * 4, torch.int16) for _ in range(2)]
    g_mom = [zeros(hh * hw * 4, torch.float32) for _ in range(2)]
    g_lh = [zeros(H * W * 4, torch.float32) for _ in range(2)]
    g_pos = [zeros(W * H * 4, torch.float32) for _ in range(2)]
    g_n = [zeros(W * H, torch.int32) for _ in range(2)]
    g_pos[1][3::4] = -1.0                                        # the cleared previous G-buffer (w = -1)
    guide = zeros(hh * hw * 4, torch.int32)
    f0, f1 = zeros(hh * hw * 4, torch.int16), zeros(hh * hw * 4, torch.int16)
    mot, dl = zeros(W * H, torch.int32), zeros(W * H * 4, torch.int16)
    rad, sun = zeros(hh * hw * 4, torch.float32), zeros(W * H, torch.float32)
    hits = zeros(W * H * 20, torch.uint8)
    o_hist = np.zeros((hh, hw, 4), np.uint16)
    o_mom = np.zeros((hh, hw, 4), np.float32)
    o_lh = np.zeros((H, W, 4), np.float32)
    prev = None
    for f in range(6):
        pu, _ = camera([37.3 + 0.4 * f, 96.4 - 0.1 * f, 41.7 + 0.3 * f], 0.8 + 0.01 * f, -0.35, frame=f)
        vf.set_prev_camera(pu, prev if prev is not None else pu)
        cur, old = f & 1, (f + 1) & 1
        ctx.trace_gbuffer(pu, W, H, hits, g_pos[cur], g_n[cur], None)
        ctx.gbuffer_motion(pu, W, H, g_pos[cur], mot, dl)
        ctx.trace_gi(pu, W, H, g_pos[cur], g_n[cur], rad, None, half_res=True, gi_reach=48.0)
        ctx.trace_sun_cone(pu, W, H, g_pos[cur], g_n[cur], sun)
        ctx.gi_accumulate(pu, W, H, g_pos[cur], g_n[cur], g_pos[old], g_n[old], mot, dl, rad, g_hist[old], g_mom[old],
                          g_hist[cur], g_mom[cur], guide)
        ctx.light_accumulate(pu, W, H, g_pos[cur], g_n[cur], g_pos[old], g_n[old], mot, dl, sun, g_lh[old], g_lh[cur])
        ctx.gi_atrous(W, H, 3, g_pos[cur], g_n[cur], g_mom[cur], g_hist[cur], guide, f0, f1)
        ctx.sync()
        pos, nrm = host(g_pos[cur], np.float32).reshape(H, W, 4), host(g_n[cur], np.uint32).reshape(H, W)
        ppos, pnrm = host(g_pos[old], np.float32).reshape(H, W, 4), host(g_n[old], np.uint32).reshape(H, W)
        m = host(mot, np.uint32).reshape(H, W)
        d = host(dl, np.uint16).reshape(H, W, 4)
        om, od = O.gbuffer_motion(W, H, pu, pos)
        assert first_diff(m, om) is None and first_diff(d, od) is None, f
        r = host(rad, np.float32).reshape(hh, hw, 4)
        s = host(sun, np.float32).reshape(H, W)
        o_hist, o_mom, o_guide = O.gi_accumulate(W, H, f, (0, 0, 0), pos, nrm, ppos, pnrm, m, d, r, o_hist, o_mom)
        o_lh = O.light_accumulate(W, H, (0, 0, 0), pos, nrm, ppos, pnrm, m, d, s, o_lh)
        e0, _ = O.gi_atrous(W, H, 3, pos, nrm, o_mom, o_hist, o_guide)
        for name, g, o in (("hist", host(g_hist[cur], np.uint16).reshape(hh, hw, 4), o_hist),
                           ("mom", host(g_mom[cur], np.uint32).reshape(hh, hw, 4), o_mom.view(np.uint32)),
                           ("guide", host(guide, np.uint32).reshape(hh, hw, 4), o_guide),
                           ("light", host(g_lh[cur], np.uint32).reshape(H, W, 4), o_lh.view(np.uint32)),
                           ("atrous", host(f0, np.uint16).reshape(hh, hw, 4), e0)):
            dd = first_diff(g, o)
            assert dd is None, f"frame {f} {name}: first difference at {dd[0]} word {dd[1]}, {dd[2]} differ"
        prev = pu
    cnt = o_hist[..., 3].view(np.float16).astype(np.float32)
    assert cnt.max() == 6.0 and (cnt >= 4).mean() > 0.3         # history survived the camera motion


def test_argument_errors(ctx):
    pu = vf.PassUbo()
    t = torch.zeros(64, dtype=torch.float32, device=DEV)
    with pytest.raises(vf.VfError):
        ctx.gi_accumulate(pu, 4, 4, t, t, t, t, t, t, t, t, t, t, t, t)     # previous == current
    with pytest.raises(vf.VfError):
        ctx.light_accumulate(pu, 4, 4, t, t, t, t, t, t, t, t, t)
    with pytest.raises(vf.VfError):
        ctx.gi_atrous(4, 4, 0, t, t, t, t, t, t, torch.zeros_like(t))
    with pytest.raises(vf.VfError):
        ctx.gi_atrous(4, 4, 3, t, t, t, t, t, t, t)
    with pytest.raises(vf.VfError):
        ctx.gbuffer_motion(pu, 0, 4, t, t, t)
