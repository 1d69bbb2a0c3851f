"""The allocation-free uneven batch (vf_trace_batch_noalloc) against vf_trace_batch: two calls in a row, the
second smaller, byte-identical hit records and completion stamps for every request; recorded into a graph
before any batch ever ran on the context (the case it exists for); refused beyond its request limit."""
import numpy as np
import pytest
import torch

import vftrace as vf
from helpers import camera, world

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def dev_hits(n):
    return torch.zeros(max(n, 1) * 20, dtype=torch.uint8, device=DEV)


def _batch(ctx, sizes, probes, seed):
    rng = np.random.default_rng(seed)
    reqs, outs = [], []
    for (w, h) in sizes:
        pos = [rng.uniform(20, 230), rng.uniform(70, 120), rng.uniform(20, 230)]
        pu, _ = camera(pos, rng.uniform(0, 6.28), rng.uniform(-0.6, 0.1))
        out = dev_hits(w * h)
        reqs.append(ctx.view_request(pu, w, h, out))
        outs.append(out)
    for (origin, spacing, dims, nr, frame, reach) in probes:
        out = dev_hits(int(np.prod(dims)) * nr)
        reqs.append(ctx.probe_request(origin, spacing, dims, nr, frame, reach, out))
        outs.append(out)
    return reqs, outs


def _fresh_ctx():
    c = vf.Context(0)
    _, a, _ = world(1)
    c.upload(world(1)[0])
    c.set_scene(vf.default_scene(a["world_bricks"], a["world_win"]))
    return c


CALLS = [
    ([(480, 270), (256, 144), (130, 75)], [((-64.5, -32.25, -64.75), 20.0, (8, 6, 8), 256, 3, 256.0)], 11),
    ([(64, 64)], [((10.5, 60.5, 10.5), 20.0, (4, 2, 4), 64, 1, 128.0)], 12),
]


def test_noalloc_batch_matches_the_allocating_batch():
    ctx = _fresh_ctx()
    try:
        for cfg in (ctx.queue_cfg((8, 8), 64), ctx.queue_cfg((16, 16), 512, refill=0),
                    ctx.queue_cfg((8, 8), 64, refill=-1)):
            for sizes, probes, seed in CALLS:   # the second call is the smaller one
                reqs, outs = _batch(ctx, sizes, probes, seed)
                ctx.trace_batch_noalloc(reqs, cfg)
                ctx.sync()
                t_noalloc = ctx.batch_times(len(reqs))
                got = [o.clone() for o in outs]
                for o in outs:
                    o.zero_()
                ctx.trace_batch(reqs, cfg)
                ctx.sync()
                t_alloc = ctx.batch_times(len(reqs))
                for k, (g, o) in enumerate(zip(got, outs)):
                    assert torch.equal(g, o), (k, cfg.tile_w, cfg.chunk, seed)
                assert all(t > 0 for t in t_noalloc) and all(t > 0 for t in t_alloc), (t_noalloc, t_alloc)
                assert int(vf.hits_to_numpy(got[0])["hit"].sum()) > 0
    finally:
        ctx.close()


def test_noalloc_batch_recorded_before_any_batch_ran():
    ctx = _fresh_ctx()   # no batch has run on it: the allocating entry point could not be recorded yet
    s = torch.cuda.Stream()
    try:
        sizes, probes, seed = CALLS[1]
        reqs, outs = _batch(ctx, sizes, probes, seed)
        cfg = ctx.queue_cfg((16, 8), 128)
        ctx.set_stream(s)
        ctx.graph_begin()
        ctx.trace_batch_noalloc(reqs, cfg)
        g = ctx.graph_end()
        replays = []
        for rep in range(3):
            for o in outs:
                o.zero_()
            torch.cuda.synchronize()
            ctx.graph_launch(g)
            ctx.sync()
            replays.append([o.clone() for o in outs])
            assert all(t > 0 for t in ctx.batch_times(len(reqs))), rep
        g.close()
        ctx.set_stream(None)
        for o in outs:
            o.zero_()
        ctx.trace_batch(reqs, cfg)
        ctx.sync()
        for rep, r in enumerate(replays):
            for k, (a, b) in enumerate(zip(r, outs)):
                assert torch.equal(a, b), (rep, k)
    finally:
        ctx.set_stream(None)
        ctx.close()


def test_noalloc_batch_refuses_beyond_its_limit():
    ctx = _fresh_ctx()
    try:
        out = dev_hits(16)
        pu, _ = camera([30.0, 90.0, 30.0], 0.3, -0.2)
        reqs = [ctx.view_request(pu, 4, 4, out)] * (vf.VF_MAX_REQUESTS + 1)
        with pytest.raises(vf.VfError, match="requests required"):
            ctx.trace_batch_noalloc(reqs)
    finally:
        ctx.close()
