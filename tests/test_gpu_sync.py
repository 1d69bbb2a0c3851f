"""GPU parity of the device world syncs against the oracle's transcriptions, through the C ABI: the chunk sync
(vf_world_sync_chunks: GpuWorld::sync_chunks, gpu_world.rs:1194-1376) on the endless-world shape (a ring
world with negative absolute bricks) and the edit sync (vf_world_sync_edits: GpuWorld::sync, :985-1142) on
it and on the fixed world.  After each sync the encoded grid, the atlas slots handed out, their brick masks,
the coarse pyramid, every brick's slot and the free list bit-identical; the distance fields equal to the
oracle's record_distance_field over the same boxes; the sky-top map and traced rays equal to the oracle's
world.  Chunks: new and reused bricks, bricks named twice, a chunk named twice in one call (one sync per
ring slot), evicted chunks, released slots (non-world, repeated and negative ones ignored), a full atlas.
Edits: clustered and scattered cells, cells listed twice, bricks edited in place, bricks that became
uniform (compacted, their slots freed and reused), new bricks past a full atlas."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle as O
import vftrace as vf
from helpers import hit_diff, random_rays

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


@pytest.fixture(scope="module")
def ctx():
    c = vf.Context(0)
    yield c
    c.close()


def _endless(extra):
    w = vf.GenWorld.generate(2, world_bricks=(64, 32, 64), chunks=3, trees_per_chunk=30)
    a = w.arrays()
    grid = np.asarray(a["grid"], np.uint32)
    bricks = np.asarray(a["atlas"], np.uint8).reshape(-1, 512)
    return a, grid, bricks, bricks.shape[0] + extra


def _cells(rng, wby, pool):
    n = 16 * wby * 16
    c = np.where(rng.random(n) < 0.15, rng.integers(1, 40, n), 0).astype(np.uint32)
    det = rng.random(n) < 0.1
    c[det] = 256 + rng.choice(pool, int(det.sum()))
    return c


def _check(ctx, st, gslot, dslot):
    used, n_free, cap = ctx.atlas_state()
    assert (used, n_free, cap) == (st.used, len(st.free), st.cap)
    assert np.array_equal(dslot, gslot)
    g, at, m = ctx.world_read(st.cap)
    assert np.array_equal(g, st.grid), f"{int((g != st.grid).sum())} grid cells differ"
    assert np.array_equal(at[:st.used], st.atlas[:st.used]) and np.array_equal(m[:st.used], st.masks[:st.used])
    c1, c2 = ctx.read_coarse()
    assert np.array_equal(c1, st.c1) and np.array_equal(c2, st.c2)


def _fields_and_traces(ctx, a, ow, st, boxes, seed):
    ow.grid[:] = st.grid
    ow.atlas[:] = st.atlas.reshape(-1)
    ow.c1[:] = st.c1
    ow.c2[:] = st.c2
    for lo, hi in boxes:
        ow.rebuild_fields([lo[0] - 16, lo[1] - 16, lo[2] - 16], [hi[0] + 16, hi[1] + 16, hi[2] + 16], lo, hi)
    d, r = ctx.read_brickdist()
    assert np.array_equal(d, ow.d[0]) and np.array_equal(r, ow.d[2])
    kinds = np.ascontiguousarray(ow.palette[:, 6])
    O.lib().vfo_skytop(C.byref(ow.s), O._p(kinds), O._p(ow.skytop))
    sky = np.zeros_like(ow.skytop)
    vf.lib().vf_read_skytop(ctx._h, sky.ctypes.data)
    assert np.array_equal(sky, ow.skytop)
    win = a["world_win"]
    rays = random_rays(60_000, [win[0] * 8 - 30, -10, win[2] * 8 - 30], [(win[0] + 64) * 8 + 30, 266,
                                                                        (win[2] + 64) * 8 + 30], seed=seed)
    rt = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).to(DEV)
    h = torch.zeros(len(rays) * 20, dtype=torch.uint8, device=DEV)
    ctx.trace_rays(rt, len(rays), h)
    ctx.sync()
    g = vf.hits_to_numpy(h)[:len(rays)]
    assert hit_diff(g, ow.trace_rays(rays)[0]).size == 0


@pytest.mark.parametrize("extra", [3000, 40])
def test_sync_chunks_matches_oracle(ctx, extra):
    a, grid, bricks0, cap = _endless(extra)
    wb, win = a["world_bricks"], a["world_win"]
    assert win[3] == 1 and win[0] < 0
    dslot = ctx.pack_world(wb, win, grid, bricks0, a["palette"], atlas_slots=cap)
    st = O.GpuWorldState(wb, grid, bricks0, cap, win)
    assert np.array_equal(dslot, st.slot0)
    ctx.set_scene(vf.default_scene(wb, win))
    ow = O.OracleWorld({"world_bricks": wb, "world_win": win, "grid": st.grid, "atlas": st.atlas.reshape(-1),
                        "palette": a["palette"]})
    rng = np.random.default_rng(extra)
    n0 = bricks0.shape[0]
    new = rng.integers(0, 60, (5000, 512)).astype(np.uint8)
    new[:, :200] = 0
    new[rng.random(5000) < 0.1] = 254
    new[rng.random(5000) < 0.3, :320] = rng.integers(1, 60)          # majority candidates
    bricks = np.concatenate([bricks0, new])
    gslot = np.concatenate([st.slot0, np.full(5000, -1, np.int32)]).astype(np.int32)
    cx0, cz0 = win[0] // 16, win[2] // 16
    src3 = st.src.reshape(wb[2], wb[1], wb[0])
    nxt = n0
    for it in range(3):
        chunks = []
        for j in range(3):
            cx, cz = cx0 + int(rng.integers(0, 4)), cz0 + int(rng.integers(0, 4))
            if it == 1 and j == 0:
                chunks.append((cx, cz, None))                         # evicted, not refilled
                continue
            live = np.nonzero(gslot >= 0)[0]
            pool = np.concatenate([np.arange(nxt, nxt + 400), rng.choice(live, 100)])
            nxt += 400
            chunks.append((cx, cz, _cells(rng, wb[1], pool)))
        chunks.append((chunks[-1][0] + 4, chunks[-1][1], _cells(rng, wb[1], np.arange(nxt, nxt + 10))))   # same ring slot
        # bricks of the chunks about to be replaced or cleared, named nowhere else: the world drops them
        replaced = np.zeros(src3.shape, bool)
        for cx, cz, _ in chunks:
            tx, tz = (cx * 16) & (wb[0] - 1), (cz * 16) & (wb[2] - 1)
            replaced[tz:tz + 16, :, tx:tx + 16] = True
        gone = set((src3[replaced & (src3 >= 256)] - 256).tolist()) - set((src3[~replaced & (src3 >= 256)] - 256).tolist())
        gone = np.array(sorted(b for b in gone if gslot[b] >= 0), np.int64)
        back = gone if it == 2 else gone[:300] if extra > 100 else gone[:50]   # pops, then fresh slots
        released = gslot[back].tolist() + [-1, cap + 3, int(gslot[gone[0]]) if len(gone) else 0]
        gslot[gone] = -1
        gslot_ref, boxes = O.sync_chunks(st, chunks, bricks, gslot, released)
        dslot = ctx.sync_chunks(chunks, bricks, gslot, released)
        _check(ctx, st, gslot_ref, dslot)
        gslot = gslot_ref
        _fields_and_traces(ctx, a, ow, st, boxes, seed=it)
    if extra == 40:                                                   # the atlas filled: majority cells
        named = np.unique(st.src[st.src >= 256] - 256)
        assert st.used == cap and (gslot[named] < 0).any()


def _edits(rng, st, wb, pool_new, bricks):
    """A cluster of cells around a random brick and scattered single cells: air, materials, new bricks,
    bricks edited in place and bricks edited into one material."""
    n = int(np.prod(wb))
    c = rng.integers(0, [wb[0] - 6, wb[1] - 6, wb[2] - 6])
    z, y, x = np.meshgrid(np.arange(6) + c[2], np.arange(6) + c[1], np.arange(6) + c[0], indexing="ij")
    cells = np.concatenate([((z * wb[1] + y) * wb[0] + x).reshape(-1), rng.integers(0, n, 6)])
    cells = np.concatenate([cells, cells[:10]])                            # listed twice
    k = cells.size
    vals = np.where(rng.random(k) < 0.4, rng.integers(0, 40, k), 0).astype(np.uint32)
    new = rng.random(k) < 0.35
    vals[new] = 256 + rng.choice(pool_new, int(new.sum()), replace=False)
    for i in np.nonzero(rng.random(k) < 0.25)[0]:
        v = int(st.src[cells[i]])
        if v >= 256:                                                     # the world's brick there, edited
            if rng.random() < 0.4:
                bricks[v - 256] = rng.integers(0, 40)                    # into one material: compacted
            else:
                bricks[v - 256, rng.integers(0, 512, 20)] = rng.integers(1, 40)
            vals[i] = v
    return cells, vals


@pytest.mark.parametrize("kind", [1, 2])
def test_sync_edits_matches_oracle(ctx, kind):
    if kind == 2:
        a, grid, bricks0, cap = _endless(150)
    else:
        a = vf.GenWorld.generate(1, world_bricks=(192, 16, 96)).arrays()
        grid = np.asarray(a["grid"], np.uint32)
        bricks0 = np.asarray(a["atlas"], np.uint8).reshape(-1, 512)
        cap = bricks0.shape[0] + 150
    wb, win = a["world_bricks"], a["world_win"]
    dslot = ctx.pack_world(wb, win, grid, bricks0, a["palette"], atlas_slots=cap)
    st = O.GpuWorldState(wb, grid, bricks0, cap, win)
    assert np.array_equal(dslot, st.slot0)
    ctx.set_scene(vf.default_scene(wb, win))
    ow = O.OracleWorld({"world_bricks": wb, "world_win": win, "grid": st.grid.copy(), "atlas": st.atlas.reshape(-1),
                        "palette": a["palette"]})
    rng = np.random.default_rng(10 + kind)
    n0 = bricks0.shape[0]
    new = rng.integers(0, 40, (3000, 512)).astype(np.uint8)
    new[:, :150] = 0
    new[rng.random(3000) < 0.3, :320] = rng.integers(1, 40)               # majority candidates
    bricks = np.concatenate([bricks0, new])
    gslot = np.concatenate([st.slot0, np.full(3000, -1, np.int32)]).astype(np.int32)
    pool = np.arange(n0, n0 + 3000)
    n_compacted = 0
    for it in range(4):
        cells, vals = _edits(rng, st, wb, pool[it * 700:(it + 1) * 700], bricks)
        vref, gref, boxes = O.sync_edits(st, cells, vals, bricks, gslot)
        vdev, gdev = ctx.sync_edits(cells, vals, bricks, gslot)
        assert np.array_equal(vdev, vref)
        _check(ctx, st, gref, gdev)
        n_compacted += int(((vals >= 256) & (vref < 256)).sum())
        gslot = gref
        _fields_and_traces(ctx, a, ow, st, boxes, seed=20 + it)
    named = np.unique(st.src[st.src >= 256] - 256)
    assert n_compacted > 0 and (gslot[named] < 0).any()                   # compaction and a full atlas ran


def test_sync_chunks_argument_errors(ctx):
    a, grid, bricks0, cap = _endless(10)
    wb, win = a["world_bricks"], a["world_win"]
    ctx.pack_world(wb, win, grid, bricks0, a["palette"], atlas_slots=cap)
    before = ctx.atlas_state()
    cells = np.full(16 * wb[1] * 16, 256 + bricks0.shape[0], np.uint32)   # names a brick beyond the list
    with pytest.raises(vf.VfError):
        ctx.sync_chunks([(win[0] // 16, win[2] // 16, cells)], bricks0, np.arange(bricks0.shape[0], dtype=np.int32),
                        released=[0, 1, 2])
    assert ctx.atlas_state() == before                                   # nothing released on failure
    flat = vf.GenWorld.generate(1, world_bricks=(16, 8, 16)).arrays()
    ctx.upload_arrays(flat)                                               # not a ring world
    with pytest.raises(vf.VfError):
        ctx.sync_chunks([(0, 0, None)], np.zeros((0, 512), np.uint8), np.zeros(0, np.int32))


BRICK_RECORD = np.dtype([("vox", np.uint8, 512), ("gpu_slot", np.int32), ("mask_ok", np.uint8), ("mask", np.uint32, 16)],
                        align=True)   # a brick record as the caller's world holds it (voxels, slot, cached masks)


def test_brick_records_read_and_written_in_place(ctx):
    """The pack, the chunk sync and the edit sync given brick records (voxels and slot fields at the record
    stride) leave the same device world and slots as given packed arrays."""
    a, grid, bricks0, cap = _endless(2000)
    wb, win = a["world_bricks"], a["world_win"]
    rng = np.random.default_rng(7)
    n0 = bricks0.shape[0]
    bricks = np.concatenate([bricks0, rng.integers(0, 30, (1500, 512)).astype(np.uint8)])
    rec = np.zeros(bricks.shape[0], BRICK_RECORD)
    rec["vox"] = bricks
    rec["gpu_slot"] = -7
    cx0, cz0 = win[0] // 16, win[2] // 16
    chunks = [(cx0 + 1, cz0 + 2, _cells(rng, wb[1], np.arange(n0, n0 + 800))), (cx0 + 3, cz0, None)]
    cells = rng.integers(0, int(np.prod(wb)), 200)
    vals = np.where(rng.random(200) < 0.5, 256 + rng.integers(n0 + 800, n0 + 1500, 200), rng.integers(0, 30, 200)).astype(np.uint32)
    results = []
    for mode in ("packed", "records"):
        if mode == "packed":
            slot = ctx.pack_world(wb, win, grid, bricks, a["palette"], atlas_slots=cap)
            slot = ctx.sync_chunks(chunks, bricks, slot, released=[5, 9])
            v, slot = ctx.sync_edits(cells, vals, bricks, slot)
        else:
            r = rec.copy()
            slot = ctx.pack_world(wb, win, grid, r, a["palette"], atlas_slots=cap)
            assert np.array_equal(slot, r["gpu_slot"])
            ctx.sync_chunks(chunks, r, None, released=[5, 9])
            v, _ = ctx.sync_edits(cells, vals, r, None)
            slot = r["gpu_slot"].copy()
            assert np.array_equal(r["vox"], bricks) and not r["mask_ok"].any()   # only gpu_slot written
        g, at, m = ctx.world_read(cap)
        c1, c2 = ctx.read_coarse()
        d, dry = ctx.read_brickdist()
        results.append((slot, v, g, at, m, c1, c2, d, dry, ctx.atlas_state()))
    for x, y in zip(*results):
        assert np.array_equal(np.asarray(x), np.asarray(y))
    assert (results[0][0][n0:n0 + 800] >= 0).any()
