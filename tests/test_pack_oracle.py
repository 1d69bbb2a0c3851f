"""Pins of the world-pack oracle (oracle.pack_world: GpuWorld::upload's CPU pack, gpu_world.rs:228-275 with
encode_brick :191-214 and gpu_voxels :58-67): hand-computed slot orders, shared slots for a brick named
twice, the full-atlas majority rule, TORCH_FLAME as air in the atlas but not in the masks, the
vectorised form against a line-by-line transcription of the reference's loop, and the synthetic worlds'
own packer (slots in grid order) recovered from shuffled bricks."""
import numpy as np
import pytest

import oracle as O
import vftrace as vf


def _brick(fill=0, **counts):
    v = np.full(512, fill, np.uint8)
    at = 0
    for m, c in counts.items():
        v[at:at + c] = int(m[1:])
        at += c
    return v


def test_slots_follow_first_cells_in_grid_order():
    bricks = np.stack([_brick(1), _brick(2), _brick(3)])
    grid = np.array([0, 256 + 2, 5, 256 + 0, 256 + 2, 256 + 1, 0], np.uint32)
    out, atlas, masks, slot = O.pack_world(grid, bricks)
    assert out.tolist() == [0, 256, 5, 257, 256, 258, 0]
    assert slot.tolist() == [1, 2, 0]
    assert (atlas[0] == 3).all() and (atlas[1] == 1).all() and (atlas[2] == 2).all()
    assert (masks == 0xFFFFFFFF).all()


def test_full_atlas_majority_rule():
    b0 = _brick(0, m7=300, m9=100)                   # 7 fills more than half? 300*2 >= 512 -> 7
    b1 = _brick(0, m254=256, m3=256)                 # TORCH_FLAME never wins: 3 (256*2 >= 512)
    b2 = _brick(0, m3=200, m4=200)                   # tie keeps 3, but 400 < 512 -> 0
    b3 = _brick(0, m1=10, m5=10)                     # starts from 1; 5 is not strictly more -> 1, 20 < 512 -> 0
    b4 = _brick(0, m255=400)                         # 255 counts
    bricks = np.stack([_brick(6), b0, b1, b2, b3, b4])
    grid = np.array([256 + 0, 256 + 1, 256 + 2, 256 + 3, 256 + 4, 256 + 5, 256 + 1], np.uint32)
    out, atlas, _, slot = O.pack_world(grid, bricks, atlas_slots=1)
    assert out.tolist() == [256, 7, 3, 0, 0, 255, 7]
    assert slot.tolist() == [0, -1, -1, -1, -1, -1] and atlas.shape == (1, 512)
    assert [O.majority_material(b) for b in (b0, b1, b2, b3, b4)] == [7, 3, 0, 0, 255]


def test_torch_flame_is_air_in_the_atlas_but_not_in_the_masks():
    b = _brick(0, m254=64, m2=8)
    out, atlas, masks, _ = O.pack_world(np.array([256], np.uint32), b[None])
    assert (atlas[0][:64] == 0).all() and (atlas[0][64:72] == 2).all()
    assert np.array_equal(masks[0], O.brick_mask(b)) and masks[0].any()
    assert not np.array_equal(masks[0], O.brick_mask(atlas[0]))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_vectorised_pack_equals_the_reference_loop(seed):
    rng = np.random.default_rng(seed)
    n_b = 200
    bricks = rng.integers(0, 12, (n_b, 512)).astype(np.uint8)
    bricks[rng.random(n_b) < 0.3] = rng.integers(1, 12)            # some single-material bricks
    grid = rng.integers(0, 20, 3000).astype(np.uint32)
    det = rng.random(3000) < 0.5
    grid[det] = 256 + rng.integers(0, n_b, int(det.sum()))          # repeats: shared slots
    for cap in (0, 50, 500):
        out, atlas, masks, slot = O.pack_world(grid, bricks, cap)
        lout, lslot, lcap = O.pack_world_loop(grid, bricks, cap)
        assert np.array_equal(out, lout) and np.array_equal(slot, lslot) and atlas.shape[0] == lcap


def test_generated_worlds_pack_their_slots_in_grid_order():
    """The synthetic world's packer numbers slots in grid order, so its bricks given back in any order
    with references pack to exactly its grid and atlas."""
    a = vf.GenWorld.generate(1, world_bricks=(16, 16, 16)).arrays()
    grid, atlas = np.asarray(a["grid"], np.uint32), np.asarray(a["atlas"], np.uint8).reshape(-1, 512)
    rng = np.random.default_rng(3)
    perm = rng.permutation(atlas.shape[0])        # brick i of the source = slot perm[i]
    inv = np.argsort(perm)
    src = grid.copy()
    det = grid >= 256
    src[det] = 256 + inv[grid[det] - 256]
    out, atlas2, _, slot = O.pack_world(src, atlas[perm])
    assert np.array_equal(out, grid)
    assert np.array_equal(atlas2, np.where(atlas == 254, 0, atlas))
    assert np.array_equal(slot, perm)
