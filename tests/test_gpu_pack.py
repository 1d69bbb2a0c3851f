"""GPU parity of the device-side world pack (vf_world_pack: GpuWorld::upload's pack, gpu_world.rs:228-330)
against the oracle's restatement, through the C ABI: the encoded grid, the atlas slots, the brick masks
and each brick's slot bit-identical (bricks named twice, a full atlas with the majority rule, ring
worlds); and a world packed on the device traces, and builds its distance fields and sky-top map,
exactly like the same world packed on the host and uploaded."""
import numpy as np
import pytest
import torch

import oracle as O
import vftrace as vf
from helpers import random_rays

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = vf.Context(0)
    yield c
    c.close()


def _random_src(seed, wb, n_b):
    rng = np.random.default_rng(seed)
    n = int(np.prod(wb))
    bricks = np.where(rng.random((n_b, 512)) < 0.4, rng.integers(1, 40, (n_b, 512)), 0).astype(np.uint8)
    bricks[rng.random(n_b) < 0.2] = 254                       # TORCH_FLAME bricks
    bricks[rng.random(n_b) < 0.2, :300] = rng.integers(1, 40)  # majority candidates
    grid = np.where(rng.random(n) < 0.3, rng.integers(1, 40, n), 0).astype(np.uint32)
    det = rng.random(n) < 0.4
    grid[det] = 256 + rng.integers(0, n_b, int(det.sum()))     # bricks named more than once
    return grid, bricks


@pytest.mark.parametrize("cap_frac", [0.0, 1.5, 0.5])
def test_pack_matches_oracle(ctx, cap_frac):
    wb, n_b = (16, 12, 20), 3000
    grid, bricks = _random_src(int(cap_frac * 10), wb, n_b)
    pal = vf.GenWorld.generate(1, world_bricks=(4, 4, 4)).arrays()["palette"]
    need = len(np.unique(grid[grid >= 256]))
    cap = int(need * cap_frac)
    slot = ctx.pack_world(wb, (0, 0, 0, 0), grid, bricks, pal, atlas_slots=cap)
    out, atlas, masks, oslot = O.pack_world(grid, bricks, cap)
    g, a, m = ctx.world_read(atlas.shape[0])
    assert np.array_equal(g, out), f"{int((g != out).sum())} grid cells differ"
    assert np.array_equal(a, atlas) and np.array_equal(m, masks)
    assert np.array_equal(slot, oslot)
    if cap_frac == 0.5:
        assert (out[grid >= 256] < 256).any()   # the majority rule ran


def test_packed_world_traces_like_the_uploaded_one(ctx):
    dev = torch.device("cuda:0")
    for kind, kw in ((1, {}), (2, {"chunks": 3, "trees_per_chunk": 30, "world_bricks": (64, 32, 64)})):
        w = vf.GenWorld.generate(kind, **kw)
        a = w.arrays()
        grid, atlas = np.asarray(a["grid"], np.uint32), np.asarray(a["atlas"], np.uint8).reshape(-1, 512)
        rng = np.random.default_rng(kind)
        perm = rng.permutation(atlas.shape[0])
        src = grid.copy()
        det = grid >= 256
        src[det] = 256 + np.argsort(perm)[grid[det] - 256]
        lo = np.array([a["world_win"][0], 0, a["world_win"][2]]) * 8
        hi = lo + np.array(a["world_bricks"]) * 8
        rays = torch.from_numpy(random_rays(100_000, lo - 20, hi + 20, seed=kind)).to(dev)
        outs = []
        for mode in ("upload", "pack"):
            if mode == "upload":
                ctx.upload_arrays(a)
            else:
                slot = ctx.pack_world(a["world_bricks"], a["world_win"], src, atlas[perm], a["palette"])
                assert np.array_equal(slot, perm)
            ctx.set_scene(vf.default_scene(a["world_bricks"], a["world_win"]))
            h = torch.zeros(100_000 * 20, dtype=torch.uint8, device=dev)
            ctx.trace_rays(rays, 100_000, h)
            ctx.sync()
            n = int(np.prod(a["world_bricks"]))
            d = np.zeros(n, np.uint8)
            dry = np.zeros(n, np.uint8)
            vf.lib().vf_read_brickdist(ctx._h, d.ctypes.data, dry.ctypes.data)
            sky = np.zeros(int(a["world_bricks"][0] * 8 * a["world_bricks"][2] * 8), np.uint16)
            vf.lib().vf_read_skytop(ctx._h, sky.ctypes.data)
            outs.append((h.cpu().numpy(), d, dry, sky, ctx.world_read(atlas.shape[0])[0]))
        for x, y in zip(*outs):
            assert np.array_equal(x, y)
        assert vf.hits_to_numpy(torch.from_numpy(outs[0][0]))["hit"].mean() > 0.2


def test_pack_argument_errors(ctx):
    pal = np.zeros((256, 12), np.float32)
    with pytest.raises(vf.VfError):
        ctx.pack_world((4, 4, 4), (0, 0, 0, 0), np.full(64, 256 + 5, np.uint32), np.zeros((2, 512), np.uint8), pal)
    with pytest.raises(vf.VfError):
        ctx.pack_world((0, 4, 4), (0, 0, 0, 0), np.zeros(0, np.uint32), np.zeros((0, 512), np.uint8), pal)
