"""Pins of the chunk-sync oracle (oracle.sync_chunks: GpuWorld::sync_chunks, gpu_world.rs:1194-1376, with
alloc_slot :178-186 and encode_brick :191-214): hand-computed slot orders through the free list (released
slots pushed highest first, popped from the back), the release rules (world slots only, once), one sync per
ring slot, cleared evicted chunks, the full-atlas majority rule; a one-chunk ring world synced from empty
equals the whole-world pack (oracle.pack_world); and after random sequences of syncs the coarse pyramid
equals a full rebuild from the source grid, every world brick's slot holds its voxels and the slots in use
and the free list partition the slots handed out."""
import numpy as np
import pytest

import oracle as O


def _solid(m):
    return np.full(512, m, np.uint8)


def _chunk(wby, cells_at):
    c = np.zeros(16 * wby * 16, np.uint32)
    for (x, y, z), v in cells_at.items():
        c[(z * wby + y) * 16 + x] = v
    return c


def _state(wb, src_cells, bricks, cap):
    src = np.zeros(int(np.prod(wb)), np.uint32)
    g = src.reshape(wb[2], wb[1], wb[0])
    for (x, y, z), v in src_cells.items():
        g[z, y, x] = v
    return O.GpuWorldState(wb, src, bricks, cap)


def test_free_list_release_and_slot_order():
    wb = (32, 4, 32)
    bricks = np.stack([_solid(m) for m in (10, 11, 12, 13, 14, 15)])
    st = _state(wb, {(0, 0, 0): 256 + 0, (1, 0, 0): 256 + 1, (2, 0, 0): 256 + 2}, bricks, cap=8)
    assert st.slot0[:3].tolist() == [0, 1, 2] and st.used == 3
    gslot = np.array([0, -1, 2, -1, -1, -1], np.int32)   # brick 1 dropped by the world
    # 7 is not a world slot, 1 twice, -1 nothing: free list [1, 0] -> pops 0 then 1, then slot 3
    cells = _chunk(4, {(0, 0, 0): 256 + 3, (1, 0, 0): 256 + 4, (2, 0, 0): 256 + 3, (0, 1, 0): 256 + 5,
                       (3, 0, 0): 256 + 2, (4, 0, 0): 9})
    out, boxes = O.sync_chunks(st, [(1, 0, cells), (3, 0, _chunk(4, {(0, 0, 0): 256 + 1})), (0, 0, None)],
                               bricks, gslot, released=[1, 0, 7, -1, 1])
    assert out.tolist() == [0, -1, 2, 0, 1, 3]
    assert st.used == 4 and st.free == []
    g = st.grid.reshape(32, 4, 32)
    assert g[0, 0, 16] == 256 and g[0, 0, 17] == 257 and g[0, 0, 18] == 256 and g[0, 1, 16] == 259
    assert g[0, 0, 19] == 258 and g[0, 0, 20] == 9
    assert not g[:16, :, :16].any()                      # (0, 0) evicted: cleared
    assert (st.atlas[0] == 13).all() and (st.atlas[1] == 14).all() and (st.atlas[3] == 15).all()
    assert st.world_slot.tolist() == [1, 1, 1, 1, 0, 0, 0, 0]
    assert boxes == [([16, 0, 0], [31, 3, 15]), ([0, 0, 0], [15, 3, 15])]   # (3, 0) shares (1, 0)'s ring slot


def test_full_atlas_takes_the_majority_material():
    wb = (16, 4, 16)
    b_major = np.zeros(512, np.uint8)
    b_major[:300] = 7
    bricks = np.stack([_solid(3), _solid(4), b_major, _solid(254)])
    st = _state(wb, {(0, 0, 0): 256 + 0}, bricks, cap=2)
    cells = _chunk(4, {(0, 0, 0): 256 + 1, (1, 0, 0): 256 + 2, (2, 0, 0): 256 + 3, (3, 0, 0): 256 + 2})
    out, _ = O.sync_chunks(st, [(0, 0, cells)], bricks, np.array([-1, -1, -1, -1], np.int32), released=[0])
    # slot 0 released (brick 0 dropped) -> brick 1; slot 1 fresh -> brick 2; brick 3 finds none (TORCH_FLAME
    # never wins: majority 1 with 0 voxels -> 0)
    assert out.tolist() == [-1, 0, 1, -1]
    assert st.grid[:4].tolist() == [256, 257, 0, 257]
    bricks2 = np.stack([_solid(3), _solid(4), b_major, _solid(5)])
    st = _state(wb, {(0, 0, 0): 256 + 0}, bricks2, cap=1)
    out, _ = O.sync_chunks(st, [(0, 0, _chunk(4, {(0, 0, 0): 256 + 2, (1, 0, 0): 256 + 3}))], bricks2,
                           np.array([0, -1, -1, -1], np.int32))
    assert out.tolist() == [0, -1, -1, -1] and st.grid[:2].tolist() == [7, 5]


@pytest.mark.parametrize("seed", [0, 1])
def test_one_chunk_world_synced_from_empty_equals_the_pack(seed):
    rng = np.random.default_rng(seed)
    wb, n_b = (16, 6, 16), 400
    bricks = np.where(rng.random((n_b, 512)) < 0.5, rng.integers(1, 30, (n_b, 512)), 0).astype(np.uint8)
    bricks[rng.random(n_b) < 0.1] = 254
    cells = rng.integers(0, 30, 16 * 6 * 16).astype(np.uint32)
    det = rng.random(cells.size) < 0.4
    cells[det] = 256 + rng.integers(0, n_b, int(det.sum()))
    for cap in (1000, 150):
        st = O.GpuWorldState(wb, np.zeros(cells.size, np.uint32), bricks, cap)
        slot, _ = O.sync_chunks(st, [(0, 0, cells)], bricks, np.full(n_b, -1, np.int32))
        out, atlas, masks, pslot = O.pack_world(cells, bricks, cap)
        assert np.array_equal(st.grid, out) and np.array_equal(slot, pslot)
        k = int((pslot >= 0).sum())
        assert st.used == k
        assert np.array_equal(st.atlas[:k], atlas[:k]) and np.array_equal(st.masks[:k], masks[:k])


def _random_cells(rng, wby, live, new_ids):
    c = np.where(rng.random(16 * wby * 16) < 0.2, rng.integers(1, 30, 16 * wby * 16), 0).astype(np.uint32)
    pool = np.concatenate([live, new_ids]) if len(live) else new_ids
    det = rng.random(c.size) < 0.15
    c[det] = 256 + rng.choice(pool, int(det.sum()))
    return c


@pytest.mark.parametrize("seed,cap", [(0, 4000), (1, 700)])
def test_random_sync_sequences_keep_the_invariants(seed, cap):
    rng = np.random.default_rng(seed)
    wb = (64, 5, 64)
    n_b = 3000
    bricks = rng.integers(0, 20, (n_b, 512)).astype(np.uint8)
    bricks[rng.random(n_b) < 0.1] = 254
    src = np.zeros(int(np.prod(wb)), np.uint32)
    st = O.GpuWorldState(wb, src, bricks, cap)
    gslot = np.full(n_b, -1, np.int32)
    nxt = 0
    for it in range(6):
        chunks, released = [], []
        for _ in range(rng.integers(1, 5)):
            cx, cz = int(rng.integers(-6, 6)), int(rng.integers(-6, 6))
            if rng.random() < 0.3:
                chunks.append((cx, cz, None))
            else:
                new = np.arange(nxt, nxt + 60)
                nxt += 60
                live = np.nonzero(gslot >= 0)[0]
                chunks.append((cx, cz, _random_cells(rng, wb[1], live[:40], new)))
        if it % 2:   # the world drops bricks no cell names any more: their slots come back
            named = set((st.src[st.src >= 256] - 256).tolist())
            drop = np.array([b for b in np.nonzero(gslot >= 0)[0] if b not in named][:50], np.int64)
            released = gslot[drop].tolist() + [cap + 5, -3]
            gslot[drop] = -1
        gslot, _ = O.sync_chunks(st, chunks, bricks, gslot, released)
        ws = set(np.nonzero(st.world_slot)[0].tolist())
        assert not ws & set(st.free) and ws | set(st.free) == set(range(st.used))
        assert len(st.free) == len(set(st.free))
    c1, c2 = np.zeros_like(st.c1), np.zeros_like(st.c2)
    O.lib().vfo_coarse(O._p(st.src), (O.C.c_int32 * 3)(*wb), O._p(c1), O._p(c2))
    assert np.array_equal(st.c1, c1) and np.array_equal(st.c2, c2)
    g = st.grid
    for v in np.unique(g[g >= 256]):
        b = np.nonzero(gslot == v - 256)[0]
        assert len(b) >= 1
        assert np.array_equal(st.atlas[v - 256], np.where(bricks[b[0]] == 254, 0, bricks[b[0]]))
        assert np.array_equal(st.masks[v - 256], O.brick_mask(bricks[b[0]]))


# ---- edit sync (GpuWorld::sync) ------------------------------------------------------------------------

def test_edit_compaction_frees_a_slot_the_next_cell_pops():
    wb = (16, 4, 16)
    b_uni = _solid(5)
    bricks = np.stack([_solid(1), b_uni, _solid(2), _solid(3)])
    bricks[0, 7] = 9                                                 # brick 0 detailed
    bricks[2, 3] = 8
    bricks[3, 1] = 7
    st = _state(wb, {(0, 0, 0): 256 + 0, (1, 0, 0): 256 + 1}, bricks, cap=3)
    assert st.used == 2
    gslot = st.slot0.copy()                                          # [0, 1, -1, -1]
    # cell 1 (brick 1, now uniform 5) -> 5, slot 1 freed; cell 30 names new brick 2 -> pops slot 1; cell 40 names
    # new brick 3 -> fresh slot 2; cell 0 listed twice: its last value (air)
    vals, out, boxes = O.sync_edits(st, [40, 1, 0, 30, 0], [256 + 3, 256 + 1, 256 + 0, 256 + 2, 0], bricks, gslot)
    assert vals.tolist() == [256 + 3, 5, 0, 256 + 2, 0]
    assert out.tolist() == [0, -1, 1, 2]
    assert st.grid[[0, 1, 30, 40]].tolist() == [0, 5, 256 + 1, 256 + 2]
    assert st.free == [] and st.used == 3 and st.world_slot.tolist() == [1, 1, 1]
    assert np.array_equal(st.atlas[1], bricks[2]) and np.array_equal(st.atlas[2], bricks[3])
    assert boxes == [([0, 0, 0], [14, 2, 0])]   # cells 0, 1, 30 (x 14, y 1), 40 (x 8, y 2): one box


def test_edit_clusters_grow_then_merge():
    wb = (128, 2, 16)
    st = O.GpuWorldState(wb, np.zeros(128 * 2 * 16, np.uint32), np.zeros((1, 512), np.uint8), 4)
    assert O.edit_boxes(st, [0, 10, 45, 60]) == [([0, 0, 0], [10, 0, 0]), ([45, 0, 0], [60, 0, 0])]
    assert O.edit_boxes(st, [0, 10, 40, 45, 60]) == [([0, 0, 0], [60, 0, 0])]       # 40 <= 10 + 32: merged
    assert O.edit_boxes(st, [5, 128 * 2 * 15 + 5]) == [([5, 0, 0], [5, 0, 15])]      # z 15 within 16
    ring = O.GpuWorldState((32, 2, 32), np.zeros(32 * 2 * 32, np.uint32), np.zeros((1, 512), np.uint8), 4,
                           win=(-16, 0, -32, 1))
    # grid x 20 is absolute -12 (the window starts at -16: x in [-16, 16)), grid x 3 is absolute 3
    assert O.brick_coords(ring, 20) == [-12, 0, -32] and O.brick_coords(ring, 3) == [3, 0, -32]


@pytest.mark.parametrize("seed,cap", [(0, 3000), (1, 520)])
def test_random_edits_render_the_source_world(seed, cap):
    """After any edits, every cell renders what the World holds: uniform cells as given, detailed cells from
    their brick's atlas slot (TORCH_FLAME as air) or, with no slot, as its majority material; the coarse
    pyramid equals a full rebuild; slots in use and the free list partition the slots handed out."""
    rng = np.random.default_rng(seed)
    wb = (32, 8, 32)
    n = 32 * 8 * 32
    n_b = 4000
    bricks = rng.integers(0, 6, (n_b, 512)).astype(np.uint8)
    bricks[rng.random(n_b) < 0.05] = 254
    src = np.where(rng.random(n) < 0.2, rng.integers(1, 30, n), 0).astype(np.uint32)
    det = rng.random(n) < 0.05
    src[det] = 256 + np.arange(int(det.sum()))
    nxt = int(det.sum())
    st = O.GpuWorldState(wb, src, bricks, cap)
    gslot = np.concatenate([st.slot0, np.full(n_b - st.slot0.size, -1, np.int32)])[:n_b].astype(np.int32)
    for it in range(5):
        k = 300
        cells = rng.integers(0, n, k)
        vals = np.where(rng.random(k) < 0.5, rng.integers(0, 30, k), 0).astype(np.uint32)
        new = rng.random(k) < 0.3
        vals[new] = 256 + np.arange(nxt, nxt + int(new.sum()))
        nxt += int(new.sum())
        uni = rng.random(k) < 0.1                                      # bricks edited into one material
        for c in np.nonzero(uni)[0]:
            v = int(st.src[cells[c]])
            if v >= 256:
                bricks[v - 256] = rng.integers(0, 30)
                vals[c] = v
        out_vals, gslot, _ = O.sync_edits(st, cells, vals, bricks, gslot)
        ws = set(np.nonzero(st.world_slot)[0].tolist())
        assert not ws & set(st.free) and ws | set(st.free) == set(range(st.used))
    g, s = st.grid, st.src
    assert np.array_equal(g[s < 256], s[s < 256])
    for gi in np.nonzero(s >= 256)[0]:
        b = int(s[gi]) - 256
        if g[gi] >= 256:
            assert gslot[b] == g[gi] - 256
            assert np.array_equal(st.atlas[g[gi] - 256], np.where(bricks[b] == 254, 0, bricks[b]))
        else:
            assert gslot[b] < 0 and g[gi] == O.majority_material(bricks[b])
    c1, c2 = np.zeros_like(st.c1), np.zeros_like(st.c2)
    O.lib().vfo_coarse(O._p(st.src), (O.C.c_int32 * 3)(*wb), O._p(c1), O._p(c2))
    assert np.array_equal(st.c1, c1) and np.array_equal(st.c2, c2)
