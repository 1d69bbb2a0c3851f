"""CPU pins of the oracle's body traversal (traceObject / traceObjectsA / traceSceneS, reference
shaders/common.glsl:664-720, 787-793) and of its BVH builder (build_bvh, src/gfx/gpu_world.rs:550-600).

* the BVH: structural invariants of the reference's builder (boxes, leaf sizes, median split on the
  widest centroid axis, node order), and the BVH walk returns exactly what the reference's plain loop
  over every body returns (common.glsl:719, USE_OBJ_BVH off);
* a body posed by an integer translation, or by a quarter turn about y plus a translation, is hit where
  the same voxels placed in the world are hit (cell, material, world normal; t to a few ulps: both DDAs
  accumulate tMax += tDelta from their own box entry, so t may differ in its last bits): the body DDA
  then reduces to the world DDA, which is pinned on its own;
* an arbitrary pose against an FP64 march of the body's voxels along the FP64-transformed ray;
* the limit rules: the world hit in front hides a body, a body in front replaces the world hit, any-hit;
* GI_OBJECTS: a bounce ray meets bodies only within 16 voxels (common.glsl:780, 2267-2271).
"""
import numpy as np

import oracle as O
import vftrace as vf
from vftrace import bodies as B
from helpers import body_world, dense_world, random_bodies


def _nodes(ow):
    n = ow.bvh_nodes
    first = n[:, 3].copy().view(np.uint32)
    count = n[:, 7].copy().view(np.uint32)
    return n[:, 0:3], n[:, 4:7], first, count


def test_bvh_builder_invariants():
    rng = np.random.default_rng(5)
    for nb in (1, 2, 3, 7, 24, 61):
        vols, poses = random_bodies(rng, nb, (20, 20, 20), (200, 120, 200))
        _, hd, g, ow = body_world(np.zeros((8, 8, 8), np.uint8), vols, poses)
        lo, hi, first, count = _nodes(ow)
        prims = ow.bvh_prims
        assert sorted(prims.tolist()) == list(range(nb))
        amin = hd["aabb_min"][:, :3]
        amax = hd["aabb_max"][:, :3]
        cen = amin + amax
        leaves = []

        def walk(i):
            if count[i] == 0:
                left, right = int(first[i]), int(first[i]) + 1
                assert i < left and right < len(first)
                pl, pr = walk(left), walk(right)
                ps = pl + pr
                ext = cen[ps].max(0) - cen[ps].min(0)
                ax = 0
                for k in (1, 2):
                    if ext[k] > ext[ax]:
                        ax = k
                # median split on the widest centroid axis: the left half's centroids come first
                assert len(pl) == len(ps) // 2
                assert cen[pl, ax].max() <= cen[pr, ax].min()
            else:
                assert 1 <= count[i] <= 2
                ps = prims[first[i]:first[i] + count[i]].tolist()
                leaves.append((int(first[i]), int(count[i])))
            assert np.array_equal(lo[i], amin[ps].min(0)) and np.array_equal(hi[i], amax[ps].max(0))
            return ps

        walk(0)
        # the leaves, in walk order (left subtree first), cover the prims array front to back
        pos = 0
        for f, c in leaves:
            assert f == pos
            pos += c
        assert pos == nb
        # every node is reached once: the array holds the tree and nothing else
        assert len(first) == (1 if nb <= 2 else 2 * len(leaves) - 1)


def test_bvh_walk_equals_the_loop_over_every_body():
    rng = np.random.default_rng(11)
    vox = np.zeros((128, 96, 128), np.uint8)
    vox[:, :20, :] = 3   # ground
    vox[40:60, 20:50, 70:90] = 7
    vols, poses = random_bodies(rng, 40, (10, 25, 10), (118, 85, 118))
    _, hd, g, ow = body_world(vox, vols, poses)
    rays = []
    for _ in range(3000):
        o = rng.uniform([0, 20, 0], [128, 96, 128])
        d = rng.normal(size=3)
        if rng.random() < 0.1:
            d[rng.integers(0, 3)] = 0.0
        d /= np.linalg.norm(d)
        rays.append(np.concatenate([o, d]))
    rays = np.asarray(rays, np.float32)
    ow.use_bvh(True)
    h1, b1 = ow.trace_scene_rays(rays)
    ow.use_bvh(False)
    h2, b2 = ow.trace_scene_rays(rays)
    ow.use_bvh(True)
    # equal, except that two bodies hit at exactly the same t go to whichever is visited first (both the
    # reference's walks keep the first: traceObject needs oh.t < limit); such ties are rare
    diff = np.nonzero((h1.view(np.uint8).reshape(-1, 20) != h2.view(np.uint8).reshape(-1, 20)).any(1) | (b1 != b2))[0]
    for i in diff:
        assert b1[i] >= 0 and b2[i] >= 0 and h1["t"][i] == h2["t"][i], (i, h1[i], h2[i])
    assert len(diff) <= 5
    assert (b1 >= 0).sum() > 300, "too few body hits to pin anything"
    assert ((b1 < 0) & (h1["hit"] == 1)).sum() > 300
    assert np.all(h1["body"] == (b1 >= 0))


def _nice(rng, n, lo, hi):
    return (np.floor(rng.uniform(lo, hi, (n, 3)) * 64.0) / 64.0).astype(np.float32)


def _random_volume(rng, shape):
    v = np.where(rng.random(shape) < 0.12, rng.integers(1, 30, shape), 0).astype(np.uint8)
    v[v == 5] = 6
    v[:8, :8, :8] = 9          # a uniform brick
    v[8:16, 8:16, :8] = 0      # an empty brick
    return v


def test_translated_body_is_the_world_dda():
    rng = np.random.default_rng(3)
    v = _random_volume(rng, (24, 16, 32))   # z, y, x
    for off in ((16, 8, 24), (13, 5, 22)):
        world = np.zeros((64, 64, 64), np.uint8)
        world[off[2]:off[2] + 24, off[1]:off[1] + 16, off[0]:off[0] + 32] = v
        _, _, owA = dense_world(world)
        r = np.eye(3, dtype=np.float32)
        _, _, _, owB = body_world(np.zeros((64, 64, 64), np.uint8), [v], [(r, np.asarray(off, np.float32))])
        n_hit = n_exact = 0
        box_lo = np.asarray(off, np.float64)
        box_hi = box_lo + np.array([32, 16, 24])
        for o in _nice(rng, 600, 0, 64):
            d = (rng.uniform(box_lo, box_hi) - o).astype(np.float32)   # aimed at the body's box
            d /= np.float32(np.linalg.norm(d))
            ha, _ = owA.trace_one(o, d)
            hb = owB.trace_scene_one(o, d)
            assert ha.hit == hb.hit
            if not ha.hit:
                continue
            n_hit += 1
            assert hb.body == 1
            assert abs(ha.t - hb.t) <= 4e-7 * max(1.0, ha.t)
            n_exact += np.float32(ha.t).view(np.uint32) == np.float32(hb.t).view(np.uint32)
            assert ha.mat == hb.mat and tuple(ha.n) == tuple(hb.n)
            assert tuple(ha.cell) == tuple(np.asarray(hb.cell) + np.asarray(off))
        assert n_hit > 200 and n_exact > 0.5 * n_hit


def test_quarter_turn_body_is_the_world_dda():
    """rotY(90 deg): body (x, y, z) -> world (z + tx, y + ty, tz - x); the matrices hold 0 and +-1 only."""
    rng = np.random.default_rng(4)
    v = _random_volume(rng, (16, 24, 32))   # z, y, x: nx 32, ny 24, nz 16
    t = np.array([20.0, 10.0, 50.0], np.float32)
    r = np.array([[0, 0, 1], [0, 1, 0], [-1, 0, 0]], np.float32)
    world = np.zeros((64, 64, 64), np.uint8)
    nz, ny, nx = v.shape
    for xb in range(nx):
        # world voxel (tx + zb, ty + yb, tz - 1 - xb) <- v[zb, yb, xb]
        world[int(t[2]) - 1 - xb, int(t[1]):int(t[1]) + ny, int(t[0]):int(t[0]) + nz] = v[:, :, xb].T
    _, _, owA = dense_world(world)
    _, _, _, owB = body_world(np.zeros((64, 64, 64), np.uint8), [v], [(r, t)])
    n_hit = n_exact = 0
    for o in _nice(rng, 800, 0, 64):
        d = (rng.uniform([20, 10, 18], [36, 34, 50]) - o).astype(np.float32)   # aimed at the body's box
        d /= np.float32(np.linalg.norm(d))
        ha, _ = owA.trace_one(o, d)
        hb = owB.trace_scene_one(o, d)
        assert ha.hit == hb.hit
        if not ha.hit:
            continue
        n_hit += 1
        assert abs(ha.t - hb.t) <= 4e-7 * max(1.0, ha.t)
        n_exact += np.float32(ha.t).view(np.uint32) == np.float32(hb.t).view(np.uint32)
        assert ha.mat == hb.mat
        if any(ha.n):
            assert np.array_equal(owB.hit_normal(hb), np.asarray(ha.n, np.float32))
        else:   # the ray starts in a solid voxel: no face, normalize(0) is NaN for the body
            assert np.isnan(owB.hit_normal(hb)).all()
        xb, yb, zb = hb.cell
        assert tuple(ha.cell) == (int(t[0]) + zb, int(t[1]) + yb, int(t[2]) - 1 - xb)
    assert n_hit > 250 and n_exact > 0.5 * n_hit


def test_rotated_body_against_fp64_march():
    rng = np.random.default_rng(8)
    v = _random_volume(rng, (24, 24, 16))
    body_as_world = dense_world(v)[2]
    misses = 0
    total = 0
    for k in range(6):
        r, t = B.pose((40.0, 40.0, 40.0), v.shape, float(rng.uniform(0, 6.28)), float(rng.uniform(-1, 1)))
        _, _, _, ow = body_world(np.zeros((80, 80, 80), np.uint8), [v], [(r, t)])
        r64, t64 = r.astype(np.float64), t.astype(np.float64)
        for _ in range(150):
            o = rng.uniform(0, 80, 3).astype(np.float32)
            d = rng.normal(size=3).astype(np.float32)
            d /= np.float32(np.linalg.norm(d))
            h = ow.trace_scene_one(o, d)
            ro = r64.T @ (o.astype(np.float64) - t64)
            rd = r64.T @ d.astype(np.float64)
            hit, t_ref, cell, n, mat = body_as_world.march_f64(ro, rd, 1e30, 5)
            total += 1
            if bool(h.hit) != hit or (hit and (tuple(h.cell) != cell or h.mat != mat)):
                misses += 1
                continue
            if hit:
                assert abs(h.t - t_ref) <= 2e-3 * max(1.0, t_ref), (h.t, t_ref)
    assert misses <= 0.02 * total, f"{misses} of {total} rays disagree with the FP64 march"


def test_limits_and_any_hit():
    v = np.full((8, 8, 8), 6, np.uint8)
    r = np.eye(3, dtype=np.float32)
    for wall_x, want_body in ((20, False), (50, True)):
        world = np.zeros((64, 64, 64), np.uint8)
        world[:, :, wall_x] = 7
        _, _, _, ow = body_world(world, [v], [(r, np.array([32.0, 28.0, 28.0], np.float32))])
        o = np.array([2.25, 32.5, 31.75], np.float32)
        d = np.array([1.0, 0.0, 0.0], np.float32)
        h = ow.trace_scene_one(o, d)
        assert h.hit and (h.body == 1) == want_body
        assert abs(h.t - ((32.0 if want_body else wall_x) - 2.25)) < 1e-5
        # limit below the body's distance: nothing; above: the body
        assert not ow.objects_only(o, d, 29.0).hit
        hb = ow.objects_only(o, d, 31.0)
        assert hb.hit and hb.body == 1
        assert ow.objects_only(o, d, 1e30, any_hit=True).hit
        assert not ow.objects_only(o, -d, 1e30, any_hit=True).hit


def test_gi_bounce_meets_bodies_within_16_voxels():
    """An empty world: the bounce ray escapes to the sky; a body 10 voxels along it is a bounce hit (the
    path's light changes), a body 20 voxels along it is beyond GI_OBJECTS' 16 and changes nothing."""
    a_empty = vf.GenWorld.from_dense(np.zeros((64, 64, 64), np.uint8)).arrays()
    scene = vf.default_scene((8, 8, 8))
    v = np.full((8, 8, 8), 6, np.uint8)
    P = np.array([20.25, 30.5, 20.75], np.float32)
    N = np.array([0.0, 1.0, 0.0], np.float32)
    d = np.array([1.0, 0.0, 0.0], np.float32)
    ow0 = O.OracleWorld(a_empty)
    L0 = ow0.gi_path(scene, P, N, d, 1, last_stage=3)[0]
    out = {}
    for dist in (10.0, 20.0):
        t = np.array([P[0] + dist, 26.0, 16.0], np.float32)
        a2, hd, g = B.attach(a_empty, [v], [(np.eye(3, dtype=np.float32), t)])
        ow = O.OracleWorld(a2)
        ow.set_bodies(hd, g)
        out[dist] = ow.gi_path(scene, P, N, d, 1, last_stage=3)[0]
    assert np.array_equal(np.asarray(out[20.0]), np.asarray(L0))
    assert not np.array_equal(np.asarray(out[10.0]), np.asarray(L0))
