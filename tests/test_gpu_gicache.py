"""GPU parity of the two-bounce GI path and the learned GI radiance cache, through the C ABI.

* traced parts (plain two-bounce GI, training samples, the cached render's traced head) against the C
  oracle: positions / normals / directions / tags bit-exact, radiance within rtol 2e-4 (the sky term's
  powf / expf differ in the last bits between the device and host libraries);
* the network (encoding, inference, gradient, Adam) against oracle/mlp.py: the encoding within 2e-6,
  inference and gradients against the oracle's BF16-emulating network (same rounding points, FP64 sums)
  to FP32-accumulation error, the gradient's direction against the plain FP64 network, Adam bit for bit;
* the cache's behaviour: warm-up frames are the traced path exactly, and a trained cache lowers the
  error against a 64-spp traced reference below that of the same path without it.
"""
import numpy as np
import pytest
import torch

import mlp as M
import oracle as O
import vftrace as vf
from helpers import config1_camera, world

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


@pytest.fixture(scope="module")
def net(ctx):
    g = vf.GiCache(ctx, warmup_frames=0, train_samples=4096, train_steps=1)
    yield g
    g.close()


def to_dev(recs):
    return torch.from_numpy(np.ascontiguousarray(recs).view(np.uint8).copy()).to(DEV)


def gbuffer(ctx, pu, W, H):
    hits = torch.zeros(W * H * 20, dtype=torch.uint8, device=DEV)
    gbp = torch.zeros(W * H * 4, dtype=torch.float32, device=DEV)
    gbn = torch.zeros(W * H, dtype=torch.int32, device=DEV)
    ctx.trace_primary(pu, W, H, hits, gbp, gbn)
    ctx.sync()
    return gbp, gbn, gbp.cpu().numpy().reshape(-1, 4), gbn.cpu().numpy().view(np.uint32)


def gpu_encode(net, recs):
    n = len(recs)
    enc = torch.zeros(n * 96, dtype=torch.float32, device=DEV)
    net.encode(to_dev(recs), n, enc)
    torch.cuda.synchronize()
    return enc.cpu().numpy().reshape(n, 96).astype(np.float64)


def gpu_infer(net, recs):
    n = len(recs)
    out = torch.zeros(n * 4, dtype=torch.int16, device=DEV)
    net.infer(to_dev(recs), n, out)
    torch.cuda.synchronize()
    return out.cpu().numpy().view(np.float16).reshape(n, 4).astype(np.float64)


def test_encoding_matches_oracle(net):
    recs = M.random_records(1000, 1)
    recs["pos"][:300] -= 300.0   # ring worlds have negative absolute coordinates
    got = gpu_encode(net, recs)
    want = M.encode(recs, 8)
    # bands after the first of each group of four come from double-angle steps (error < 2e-6)
    assert np.abs(got - want).max() < 2e-6
    first = [24 * a + 2 * k + c for a in range(3) for k in (0, 4, 8) for c in (0, 1)]
    assert np.abs(got[:, first] - want[:, first]).max() < 5e-7


def test_inference_matches_emulated_oracle(net):
    params = M.init_params(7)
    net.set_params(params)
    recs = M.random_records(1000, 2)
    got = gpu_infer(net, recs)
    X = gpu_encode(net, recs)
    _, y = M.forward(params, X, emulate=True)
    assert np.all(got[:, 3] == 0.0)
    err = np.abs(got[:, :3] - y)
    assert np.all(err <= 2e-3 + 2e-3 * np.abs(y)), err.max()
    assert np.abs(y).max() > 0.05   # a non-trivial output


@pytest.mark.parametrize("n", [4096, 40000])
def test_gradient_matches_oracle(net, n):
    """Gradient check: the matrix-core backward pass against the CPU network (40000 samples also
    exercises wavefronts that own several 64-sample chunks)."""
    params = M.init_params(9)
    net.set_params(params)
    recs = M.random_records(n, 3)
    grad = torch.zeros(M.N_PARAMS, dtype=torch.float32, device=DEV)
    loss = torch.zeros(1, dtype=torch.float32, device=DEV)
    net.grad(to_dev(recs), n, grad, loss)
    torch.cuda.synchronize()
    g = grad.cpu().numpy().astype(np.float64)
    X = gpu_encode(net, recs)
    lo, go, _ = M.loss_grad(params, recs, 8, 0.01, emulate=True, X=X)
    l64, g64, _ = M.loss_grad(params, recs, 8, 0.01, emulate=False)
    assert abs(float(loss.item()) - lo) < 1e-4 * lo
    for l in range(6):
        for a, b in ((M.w_off(l), M.b_off(l)), (M.b_off(l), M.b_off(l) + M.OUT_DIMS[l])):
            e = np.linalg.norm(g[a:b] - go[a:b]) / max(np.linalg.norm(go[a:b]), 1e-30)
            assert e < 2e-3, (l, a, e)
            cos = g[a:b] @ g64[a:b] / (np.linalg.norm(g[a:b]) * np.linalg.norm(g64[a:b]) + 1e-30)
            assert cos > 0.99, (l, a, cos)
    # padded input rows see zero features
    assert not np.any(g[M.w_off(0) + 78 * 64:M.b_off(0)])


def test_adam_matches_oracle_and_updates_the_network(net):
    p = net.params
    w0 = M.init_params(11)
    net.set_params(w0)
    rng = np.random.default_rng(12)
    m = np.zeros_like(w0)
    v = np.zeros_like(w0)
    w = w0
    assert np.array_equal(net.get_render_params(), w0)
    we = w0
    ema = np.float32(p.weight_ema)
    assert ema > 0
    for step in (1, 2):
        gh = (rng.normal(size=M.N_PARAMS) * 0.05).astype(np.float32)
        net.adam(torch.from_numpy(gh).to(DEV))
        w, m, v = M.adam(w, m, v, gh, step, p.lr, p.beta1, p.beta2, p.eps)
        got = net.get_params()
        assert np.array_equal(got, w), (step, np.abs(got - w).max())
        # the render's weights: we = ema * we + (1 - ema) * w, FP32 in that order
        we = M.weight_ema(we, w, ema)
        assert np.array_equal(net.get_render_params(), we), step
    assert not np.array_equal(we, w)
    # the BF16 copies the kernels read were rewritten: inference runs the render's weights
    recs = M.random_records(500, 4)
    _, y = M.forward(we, gpu_encode(net, recs), emulate=True)
    assert np.all(np.abs(gpu_infer(net, recs)[:, :3] - y) <= 2e-3 + 2e-3 * np.abs(y))


@pytest.mark.parametrize("bounces", [2, 3])
def test_plain_two_bounce_gi_matches_oracle(ctx, bounces):
    _, a, ow = world(1)
    pu, _ = config1_camera(frame=2)
    W, H = 160, 90
    gbp, gbn, gp, gn = gbuffer(ctx, pu, W, H)
    out = torch.zeros(W * H * 4, dtype=torch.float32, device=DEV)
    ctx.trace_gi_path(pu, W, H, gbp, gbn, out, frame=3, spp=2, salt=0x1234, bounces=bounces)
    ctx.sync()
    got = out.cpu().numpy().reshape(-1, 4)
    want = ow.pass_gi_path(vf.default_scene(a["world_bricks"]), W, H, gp, gn, 3, 2, 0x1234, bounces=bounces)
    assert np.array_equal(got[:, 3], want[:, 3])
    assert np.allclose(got, want, rtol=2e-4, atol=2e-6)
    assert got[:, :3].max() > 0.1 and (got[:, 3] == 1.0).mean() > 0.5


@pytest.mark.parametrize("terminate,bounces", [(0, 2), (1, 2), (0, 3), (1, 3)])
def test_training_samples_match_oracle(ctx, terminate, bounces):
    _, a, ow = world(1)
    pu, _ = config1_camera(frame=1)
    W, H = 160, 90
    gbp, gbn, gp, gn = gbuffer(ctx, pu, W, H)
    g = vf.GiCache(ctx, terminate=terminate, train_samples=8192, train_steps=2, bounces=bounces)
    try:
        out = torch.zeros(8192 * 64, dtype=torch.uint8, device=DEV)
        g.samples(pu, W, H, gbp, gbn, 5, out)
        ctx.sync()
        got = out.cpu().numpy().view(O.RECORD_DTYPE)
        want = ow.gicache_samples(vf.default_scene(a["world_bricks"]), list(pu.cam_pos), W, H, gp, gn, 5, 8192,
                                  terminate, g.params.salt, bounces=bounces)
    finally:
        g.close()
    assert np.array_equal(got["tag"], want["tag"])
    for f in ("pos", "n", "view"):
        assert np.array_equal(got[f], want[f]), f
    assert np.allclose(got["rgb"], want["rgb"], rtol=2e-4, atol=2e-6)
    assert got["tag"].mean() > (0.3 if terminate == 0 else 0.05)


def _oracle_render(ow, scene, net_params, X_of, pre, queries, spp, npix):
    valid = queries["tag"] != 0xFFFFFFFF
    q = queries[valid]
    _, y = M.forward(net_params, X_of(q), emulate=True)
    yh = np.maximum(y.astype(np.float16).astype(np.float64), 0.0)
    tail = np.zeros((npix * spp, 3))
    tail[q["tag"]] = q["rgb"].astype(np.float64) * yh
    out = (pre[:, :3].astype(np.float64) + tail.reshape(npix, spp, 3).sum(1)) / spp
    out[pre[:, 3] == 0.0] = 0.0
    return out, int(valid.sum())


@pytest.mark.parametrize("terminate,spp,bounces", [(1, 1, 2), (1, 2, 2), (0, 1, 2), (1, 1, 3)])
def test_cached_render_matches_oracle(ctx, terminate, spp, bounces):
    _, a, ow = world(1)
    scene = vf.default_scene(a["world_bricks"])
    pu, _ = config1_camera(frame=4)
    W, H = 128, 72
    gbp, gbn, gp, gn = gbuffer(ctx, pu, W, H)
    g = vf.GiCache(ctx, terminate=terminate, spp=spp, warmup_frames=0, train_samples=4096, train_steps=1,
                   bounces=bounces)
    try:
        out = torch.zeros(W * H * 4, dtype=torch.float32, device=DEV)
        g.frame(pu, W, H, gbp, gbn, out, 7)
        st = g.stats()
        params = g.get_render_params()
        assert not np.array_equal(params, g.get_params())   # the exponential mean, one step behind
        got = out.cpu().numpy().reshape(-1, 4)
        if terminate == 1:
            pre, queries = ow.gicache_prepass(scene, W, H, gp, gn, 7, spp, g.params.salt, bounces=bounces)
        else:
            npix = W * H
            pre = np.zeros((npix, 4), np.float32)
            pre[:, 3] = (gp[:, 3] >= 0.0).astype(np.float32)
            queries = np.zeros(npix, O.RECORD_DTYPE)
            queries["tag"] = 0xFFFFFFFF
            sel = np.nonzero(gp[:, 3] >= 0.0)[0]
            Nq = np.stack([((gn[sel] >> (8 * k)) & 255).astype(np.float32) / np.float32(255.0) * np.float32(2.0)
                           - np.float32(1.0) for k in range(3)], 1)
            # view = normalize(cam - P) in FP32, the kernel's operation order
            v = np.asarray(pu.cam_pos, np.float32)[None, :] - gp[sel, :3]
            ln = np.sqrt((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2])
            queries["tag"][sel] = sel
            queries["pos"][sel] = gp[sel, :3]
            queries["n"][sel] = Nq
            queries["view"][sel] = v / ln[:, None]
            queries["rgb"][sel] = 1.0
            spp = 1
        want, nq = _oracle_render(ow, scene, params, lambda q: gpu_encode(g, q), pre, queries, spp, W * H)
    finally:
        g.close()
    assert st["fallback"] == 0 and st["n_queries"] == nq and st["frames"] == 1
    assert np.array_equal(got[:, 3], pre[:, 3])
    # BF16 activations: an FP32-vs-FP64 difference of ~1e-7 can round a hidden unit to the neighbouring
    # BF16 value (~5e-5 of the units), which moves that output by up to ~1e-3; everything else agrees to
    # FP16 output rounding.
    err = np.abs(got[:, :3] - want)
    assert np.all(err <= 2e-3 + 2e-3 * np.abs(want)), err.max()
    assert np.mean(err <= 2e-4 + 3e-3 * np.abs(want)) > 0.99


@pytest.mark.parametrize("bounces", [2, 3])
def test_warmup_frames_render_the_traced_path(ctx, bounces):
    _, a, _ = world(1)
    pu, _ = config1_camera(frame=6)
    W, H = 128, 72
    gbp, gbn, _, _ = gbuffer(ctx, pu, W, H)
    g = vf.GiCache(ctx, warmup_frames=2, spp=1, train_samples=4096, train_steps=1, bounces=bounces)
    try:
        out = torch.zeros(W * H * 4, dtype=torch.float32, device=DEV)
        ref = torch.zeros_like(out)
        for f in range(2):
            g.frame(pu, W, H, gbp, gbn, out, f)
            ctx.trace_gi_path(pu, W, H, gbp, gbn, ref, frame=f, spp=1, salt=g.params.salt, bounces=bounces)
            ctx.sync()
            assert g.stats()["fallback"] == 1
            assert torch.equal(out, ref), f
        g.frame(pu, W, H, gbp, gbn, out, 2)
        assert g.stats()["fallback"] == 0
    finally:
        g.close()


def _mse(x, ref, mask):
    d = (x[mask, :3] - ref[mask, :3]).astype(np.float64)
    rel = d * d / (ref[mask, :3].astype(np.float64) ** 2 + 1e-2)
    return float((d * d).mean()), float(rel.mean())


@pytest.mark.parametrize("terminate", [1, 0])
def test_trained_cache_beats_the_path_without_it(ctx, terminate):
    """What the network gains: after online training the cached GI is closer to a 64-spp traced reference
    than the same frame's traced path without the cache (terminate 1: the same bounce-1 rays, bounce 2
    traced once; terminate 0: one traced path per pixel)."""
    pu, _ = config1_camera(frame=0)
    W, H = 240, 135
    gbp, gbn, gp, _ = gbuffer(ctx, pu, W, H)
    ref = torch.zeros(W * H * 4, dtype=torch.float32, device=DEV)
    ctx.trace_gi_path(pu, W, H, gbp, gbn, ref, frame=0, spp=64, salt=0x5EED)
    g = vf.GiCache(ctx, terminate=terminate, warmup_frames=4)
    try:
        out = torch.zeros_like(ref)
        frames = 160
        for f in range(frames):
            g.frame(pu, W, H, gbp, gbn, out, f)
        st = g.stats()
        plain = torch.zeros_like(ref)
        ctx.trace_gi_path(pu, W, H, gbp, gbn, plain, frame=frames - 1, spp=1, salt=g.params.salt)
        ctx.sync()
    finally:
        g.close()
    mask = gp[:, 3] >= 0.0
    r = ref.cpu().numpy().reshape(-1, 4)
    mc, rc = _mse(out.cpu().numpy().reshape(-1, 4), r, mask)
    mp, rp = _mse(plain.cpu().numpy().reshape(-1, 4), r, mask)
    print(f"terminate {terminate}: cache mse {mc:.5f} relmse {rc:.4f}; traced 1spp mse {mp:.5f} relmse {rp:.4f}; "
          f"stats {st}")
    assert st["fallback"] == 0 and st["steps"] == frames * 4
    assert mc < mp and rc < rp
