"""Pins of the CPU oracle (no GPU).  Each test checks the oracle against something it does not share
code with: the reference's own known answers, brute-force closed forms, an independent FP64
march, or invariants that a dropped term / wrong sign / swapped index would break."""
import ctypes
import math

import numpy as np
import pytest

import oracle as O
import vftrace as vf
from helpers import config1_camera, dense_world, random_rays, world


# --- the reference's own known-answer tests ------------------------------------------------------
def test_mask_bits_reference_vector():
    """sparse_volume.rs:660-668: voxel (7,7,7) -> m[15] == 1<<31, all other words 0."""
    vox = np.zeros(512, np.uint8)
    vox[7 * 64 + 7 * 8 + 7] = 1
    m = O.brick_mask(vox)
    assert m[15] == 1 << 31
    assert (m[:15] == 0).all()


def test_raycast_hits_floor_reference_case():
    """sparse_volume.rs:669-679: floor y 0..9 in 64^3; ray (20.5, 40, 20.5) down -> cell [20,9,20],
    normal [0,1,0], |t - 30| < 1e-3.  Checked for the reference's CPU raycast and for traceVolumeS."""
    vox = np.zeros((64, 64, 64), np.uint8)
    vox[:, 0:10, :] = 3
    _, _, ow = dense_world(vox)
    hit, t, cell, n, mat = ow.raycast_cpu([20.5, 40.0, 20.5], [0.0, -1.0, 0.0], 100.0)
    assert hit and cell == (20, 9, 20) and n == (0, 1, 0) and abs(t - 30.0) < 1e-3 and mat == 3
    h, _ = ow.trace_one([20.5, 40.0, 20.5], [0.0, -1.0, 0.0])
    assert h.hit and tuple(h.cell) == (20, 9, 20) and tuple(h.n) == (0, 1, 0) and abs(h.t - 30.0) < 1e-3


def test_product_packer_masks_match_reference_layout():
    """The generator's 64-byte masks (uploaded as brickMasks) follow brick_mask's bit order."""
    _, a, _ = world(1)
    atlas = a["atlas"].reshape(-1, 512)
    masks = a["masks"].reshape(-1, 16)
    rng = np.random.default_rng(0)
    for s in rng.choice(atlas.shape[0], 200, replace=False):
        assert np.array_equal(masks[s], O.brick_mask(atlas[s]))


# --- coarse pyramid, distance field, sky-top: closed forms ------------------------------------
def _grid3(a):
    wb = a["world_bricks"]
    return a["grid"].reshape(wb[2], wb[1], wb[0])


def test_coarse_pyramid_closed_form():
    _, a, ow = world(1)
    g = _grid3(a) != 0
    wb = a["world_bricks"]
    for cell, arr, dims in ((4, ow.c1, ow.c1d), (16, ow.c2, ow.c2d)):
        exp = np.zeros((dims[2], dims[1], dims[0]), np.uint8)
        for z in range(dims[2]):
            for y in range(dims[1]):
                for x in range(dims[0]):
                    exp[z, y, x] = g[z * cell:(z + 1) * cell, y * cell:(y + 1) * cell, x * cell:(x + 1) * cell].any()
        assert np.array_equal(arr.reshape(exp.shape), exp)
    assert np.array_equal(ow.c1, a["coarse1"]) and np.array_equal(ow.c2, a["coarse2"])
    assert wb == (32, 16, 32)


def test_brickdist_separable_equals_closed_form():
    """brickdist.comp's 3 separable passes == min(16, L-inf distance to an occupied brick)."""
    _, a, ow = world(1)
    wb = a["world_bricks"]
    assert np.array_equal(ow.d[0], O.brickdist_brute(a["grid"], a["atlas"], wb, False, False))
    assert np.array_equal(ow.d[2], O.brickdist_brute(a["grid"], a["atlas"], wb, False, True))
    # spot-check the closed form itself with numpy on a few bricks
    g = _grid3(a) != 0
    occ = np.argwhere(g)
    rng = np.random.default_rng(1)
    for _ in range(50):
        z, y, x = rng.integers(0, wb[2]), rng.integers(0, wb[1]), rng.integers(0, wb[0])
        d = np.abs(occ - np.array([z, y, x])).max(1).min() if len(occ) else 16
        assert ow.d[0][(z * wb[1] + y) * wb[0] + x] == min(16, d)


def test_brickdist_ring_wrap_and_water():
    """Ring mode wraps x/z (above/below = empty); the dry field counts water-only bricks as empty."""
    rng = np.random.default_rng(2)
    vox = np.zeros((64 * 8, 4 * 8, 64 * 8), np.uint8)
    for _ in range(40):
        bz, by, bx = rng.integers(0, 64), rng.integers(0, 4), rng.integers(0, 64)
        vox[bz * 8:(bz + 1) * 8, by * 8:by * 8 + 3, bx * 8:(bx + 1) * 8] = rng.choice([3, 5])
    vox[0:8, 0:8, 0:8] = 0
    vox[2, 3, 4] = 7               # a single voxel at the ring's corner brick
    vox[8 * 63 + 1, 2, 8 * 63] = 5  # a water-only detailed brick at the opposite corner
    w, a, ow = dense_world(vox, world_win=(-16, 0, 8, 1))
    wb = a["world_bricks"]
    assert np.array_equal(ow.d[0], O.brickdist_brute(a["grid"], a["atlas"], wb, True, False))
    assert np.array_equal(ow.d[2], O.brickdist_brute(a["grid"], a["atlas"], wb, True, True))
    # the corner brick's wrapped neighbour (63, 0, 63) is 1 away through the wrap
    assert ow.d[0][(63 * wb[1] + 0) * wb[0] + 63] <= 1


def test_brickdist_incremental_equals_full():
    """record_distance_field over dirty +- 16 after an edit == a full rebuild (triangle inequality)."""
    _, a0, _ = world(1)
    a = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in a0.items()}
    ow = O.OracleWorld(a)
    wb = a["world_bricks"]
    g = ow.grid.reshape(wb[2], wb[1], wb[0])
    g[3:7, 1:4, 10:12] = 0
    g[20:23, 12:14, 2:5] = 9
    ow.rebuild_fields([10 - 16, 1 - 16, 3 - 16], [11 + 16, 13 + 16, 22 + 16], [2, 1, 3], [11, 13, 22])
    full = O.OracleWorld(a)
    assert np.array_equal(ow.d[0], full.d[0]) and np.array_equal(ow.d[2], full.d[2])


def test_skytop_closed_form():
    """skyTop = 1 + highest voxel whose kind is not water (1) or foliage (3)."""
    vox = np.zeros((32, 32, 32), np.uint8)
    vox[:, 0:5, :] = 3
    vox[:, 5:9, 0:16] = 5          # water above the ground on the left half
    vox[10:14, 12:20, 20:24] = 7   # a leaf clump
    vox[10:12, 20:21, 22:24] = 2   # dirt on top of the clump
    _, a, ow = dense_world(vox)
    kinds = ow.palette[:, 6]
    exp = np.zeros((32, 32), np.uint16)
    for z in range(32):
        for x in range(32):
            ys = [y for y in range(32) if vox[z, y, x] and kinds[vox[z, y, x]] not in (1.0, 3.0)]
            exp[z, x] = (max(ys) + 1) if ys else 0
    assert np.array_equal(ow.skytop.reshape(32, 32), exp)


# --- traversal --------------------------------------------------------------------------------
def _same_hits(a, b):
    return (a["hit"] == b["hit"]) & (a["cell"] == b["cell"]).all(1) & (a["face"] == b["face"]) & (a["mat"] == b["mat"])


@pytest.mark.parametrize("kind", [1, 2])
def test_skips_equal_plain_brick_dda(kind):
    """The reference states that the distance-field / coarse skips leave the output identical to the
    plain brick-by-brick DDA (common.glsl:598-601): same voxel, face and material; t within 1 ulp-ish."""
    if kind == 1:
        _, a, ow = world(1)
        rays = random_rays(60_000, [-30, -10, -30], [286, 140, 286], seed=7)
    else:
        _, a, ow = world(2, world_bricks=(64, 32, 64), chunks=3, trees_per_chunk=60)
        win = a["world_win"]
        rays = random_rays(60_000, [win[0] * 8, 0, win[2] * 8], [(win[0] + 64) * 8, 256, (win[2] + 64) * 8], seed=8)
    plain = O.OracleWorld(a, use_dist=False, coarse_skip=False)
    coarse = O.OracleWorld(a, use_dist=False, coarse_skip=True)
    for skip in (0, 5):
        h, c = ow.trace_rays(rays, 1e30, skip)
        hp, cp = plain.trace_rays(rays, 1e30, skip)
        hc, _ = coarse.trace_rays(rays, 1e30, skip)
        assert _same_hits(h, hp).all() and _same_hits(hc, hp).all()
        m = hp["hit"] == 1
        rel = np.abs(h["t"][m] - hp["t"][m]) / np.maximum(hp["t"][m], 1e-3)
        assert rel.max() < 2e-6
        assert c["outer"] < cp["outer"]   # the skips do skip


def _config1_frame_rays(ow, scene_sun):
    """The config-1 frame at 640x360: its primary rays and the sun rays of its lit texels."""
    pu, cam = config1_camera()
    W, H = 640, 360
    _, pos, nrm, _ = ow.pass_primary(cam, W, H)
    dirs = O.primary_dirs(cam, W, H)
    prim = np.concatenate([np.broadcast_to(np.asarray(cam["pos"], np.float32), dirs.shape), dirs], 1)
    N = np.stack([((nrm >> (8 * k)) & 255) / np.float32(255.0) * 2 - 1 for k in range(3)], 1).astype(np.float32)
    sun = np.asarray(scene_sun, np.float32)
    lit = (pos[:, 3] >= 0) & ((N @ sun) > 0)
    so = pos[lit, :3] + N[lit] * np.float32(0.03)
    return prim.astype(np.float32), np.concatenate([so, np.broadcast_to(sun, so.shape)], 1).astype(np.float32)


def test_small_frame_mask_test_never_skips_a_hit():
    """The library's small-frame traversal skips a detailed brick's walk when the brick's mask has no solid
    voxel in the rows and columns of the box from the walk's entry voxel to the voxel at the brick DDA's
    exit time (taken 0.05 voxel further).  Restated in the oracle: every hit equals the reference
    traversal's on the config-1 frame's primary and sun rays, on random rays with tiny and zero direction
    components, and on the ring world; the test does skip walks; and a planted mistake in it (the entry
    voxel left out of the box, or the exit coordinate taken 0.6 voxel short) changes hits."""
    _, a, ow = world(1)
    sun = vf.default_scene((32, 16, 32)).sun_dir
    prim, sunr = _config1_frame_rays(ow, sun)
    cases = [(prim, 0), (sunr, 5), (random_rays(60_000, [-30, -10, -30], [286, 140, 286], seed=17), 0),
             (random_rays(60_000, [0, 30, 0], [256, 128, 256], seed=18), 5)]
    skipped_total = walked_total = 0
    for rays, skip in cases:
        ref, _ = ow.trace_rays(rays, 1e30, skip)
        got, skipped, walked = ow.trace_rays_box(rays, 1e30, skip)
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))
        skipped_total += skipped
        walked_total += walked
    assert skipped_total > 0.2 * (skipped_total + walked_total)
    _, a2, ow2 = world(2, world_bricks=(64, 32, 64), chunks=3, trees_per_chunk=60)
    win = a2["world_win"]
    rays = random_rays(60_000, [win[0] * 8, 0, win[2] * 8], [(win[0] + 64) * 8, 256, (win[2] + 64) * 8], seed=19)
    for skip in (0, 5):
        ref, _ = ow2.trace_rays(rays, 1e30, skip)
        got, _, _ = ow2.trace_rays_box(rays, 1e30, skip)
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))
    # planted mistakes must change hits
    ref, _ = ow.trace_rays(prim, 1e30, 0)
    for margin, mode in ((0.05, 1), (-0.6, 0)):
        bad, _, _ = ow.trace_rays_box(prim, 1e30, 0, margin=margin, mode=mode)
        assert (bad.view(np.uint8).reshape(-1, 20) != ref.view(np.uint8).reshape(-1, 20)).any(1).sum() > 100


def test_fp64_remarch_agrees():
    """Independent FP64 voxel march: same voxel except at near-ties, |t32 - t64| <= 1e-5 relative."""
    _, a, ow = world(1)
    rays = random_rays(4000, [0, 60, 0], [256, 128, 256], seed=12, tiny_frac=0.0)
    hits, _ = ow.trace_rays(rays, 1e30, 0)
    mismatch = 0
    worst = 0.0
    for i in range(len(rays)):
        hit, t64, cell, n, mat = ow.march_f64(rays[i, :3].astype(np.float64), rays[i, 3:].astype(np.float64))
        if hit != bool(hits["hit"][i]) or (hit and tuple(hits["cell"][i]) != cell):
            mismatch += 1
            continue
        if hit:
            assert vf_face(n) == hits["face"][i] and mat == hits["mat"][i]
            worst = max(worst, abs(float(hits["t"][i]) - t64) / max(t64, 1e-6))
    assert mismatch <= 2, mismatch
    assert worst <= 1e-5, worst


def vf_face(n):
    for a in range(3):
        if n[a]:
            return 2 * a + (1 if n[a] > 0 else 2)
    return 0


def test_hit_invariants():
    """The hit voxel is solid for the ray's skipMat; the voxel the ray came from (cell + normal) is
    not solid; a hit lies on the ray (the hit point is inside or on the cell)."""
    _, a, ow = world(1)
    rays = random_rays(3000, [0, 50, 0], [256, 128, 256], seed=13, tiny_frac=0.0)
    for skip in (0, 5):
        hits, _ = ow.trace_rays(rays, 1e30, skip)
        for i in np.nonzero(hits["hit"])[0][:1500]:
            c = hits["cell"][i]
            m = ow.voxel(*c)
            assert m != 0 and m != skip and m == hits["mat"][i]
            f = hits["face"][i]
            if f:
                ax, sgn = (f - 1) // 2, (1 if f % 2 == 1 else -1)
                prev = c.copy()
                prev[ax] += sgn
                pm = ow.voxel(*prev)
                assert pm == 0 or pm == skip or (skip == 5 and pm == 5)
            p = rays[i, :3].astype(np.float64) + rays[i, 3:].astype(np.float64) * float(hits["t"][i])
            assert np.all(p >= c - 1e-3) and np.all(p <= c + 1 + 1e-3)


def test_tcap_brick_entry_rule():
    """t > tCap is tested once per brick (common.glsl:574): a hit can lie past tCap, but never past the
    far side of the brick the ray entered at t <= tCap (at most one brick diagonal, ~13.9 voxels)."""
    _, a, ow = world(1)
    rays = random_rays(20_000, [0, 40, 0], [256, 128, 256], seed=14)
    for cap in (0.0, 3.0, 20.0):
        h, _ = ow.trace_rays(rays, cap, 0)
        m = h["hit"] == 1
        assert np.all(h["t"][m] <= cap + 8.0 * math.sqrt(3.0) + 1e-3)
        h_inf, _ = ow.trace_rays(rays, 1e30, 0)
        close = (h_inf["hit"] == 1) & (h_inf["t"] <= cap)
        assert np.all(h["hit"][close] == 1)


# --- ray generators ---------------------------------------------------------------------------
def test_sincos_accuracy():
    xs = np.linspace(0.0, 700.0, 20001, dtype=np.float32)
    s = np.array([O.lib().vfo_sin(float(x)) for x in xs])
    c = np.array([O.lib().vfo_cos(float(x)) for x in xs])
    assert np.max(np.abs(s - np.sin(xs.astype(np.float64)))) < 2e-6
    assert np.max(np.abs(c - np.cos(xs.astype(np.float64)))) < 2e-6


def test_pcg_is_a_permutation_sample():
    """pcgHash is RXS-M-XS (a bijection on u32): no collisions on 2^16 consecutive inputs and an
    avalanche of roughly half the bits.  Exact values are pinned only GPU-vs-oracle."""
    vals = np.array([O.pcg(i) for i in range(1 << 16)], np.uint64)
    assert len(np.unique(vals)) == 1 << 16
    flips = [bin(O.pcg(i) ^ O.pcg(i ^ 1)).count("1") for i in range(2000)]
    assert 12 < np.mean(flips) < 20


def test_primary_dirs_pinhole_geometry():
    _, cam = config1_camera()
    W, H = 64, 36
    d = O.primary_dirs(cam, W, H)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    fwd = np.asarray(cam["fwd"], np.float64)
    # pixel (W/2 - 0.5, ...) straddles the centre: the mean of the 4 centre pixels points forward
    c = d[[(H // 2 - 1) * W + W // 2 - 1, (H // 2 - 1) * W + W // 2, (H // 2) * W + W // 2 - 1, (H // 2) * W + W // 2]]
    m = c.mean(0)
    assert np.dot(m / np.linalg.norm(m), fwd) > 0.9999
    # row 0 is the bottom of the image: its rays point below the top row's
    assert d[0] @ np.asarray(cam["up"]) < d[(H - 1) * W] @ np.asarray(cam["up"])
    # vertical angle between the bottom and top pixel rows ~ the 70 deg FOV minus one pixel
    top, bot = d[(H - 1) * W + W // 2], d[W // 2]
    ang = math.degrees(math.acos(float(np.clip(top @ bot, -1, 1))))
    exp = 2 * math.degrees(math.atan(math.tan(math.radians(35)) * (H - 1) / H))
    assert abs(ang - exp) < 0.2


def test_cosine_hemisphere_distribution():
    n = np.array([0.0039215684, 1.0, 0.0039215684], np.float32)
    out = np.zeros(3, np.float32)
    u = np.zeros(2, np.float32)
    cosines = []
    for k in range(4000):
        O.lib().vfo_ld_sample2(k % 97, k // 97, k, 0x68e31da4, u.ctypes.data)
        O.lib().vfo_cosine_hemisphere_u(n.ctypes.data, u.ctypes.data, out.ctypes.data)
        assert abs(np.linalg.norm(out) - 1.0) < 1e-5
        cosines.append(float(out @ (n / np.linalg.norm(n))))
    cosines = np.array(cosines)
    assert cosines.min() > -1e-6
    assert abs(cosines.mean() - 2.0 / 3.0) < 0.02     # E[cos] of a cosine-weighted hemisphere


def _ld(px, py, k, salt):
    out = np.zeros(2, np.float32)
    O.lib().vfo_ld_sample2(int(px), int(py), int(k) & 0xFFFFFFFF, int(salt) & 0xFFFFFFFF, out.ctypes.data)
    return out


def _r2_constants():
    """Roberts' R2 sequence step (1/g, 1/g^2), g the real root of x^3 = x + 1 (the plastic number), as float32."""
    g = max(r.real for r in np.roots([1.0, 0.0, -1.0, -1.0]) if abs(r.imag) < 1e-12)
    assert abs(g ** 3 - g - 1.0) < 1e-14
    return np.array([1.0 / g, 1.0 / (g * g)], np.float32)


def _r2_sequence(u0, ks, alpha):
    x = u0[None, :] + (ks & 0xFFFF).astype(np.float32)[:, None] * alpha[None, :]   # float32 product, then sum
    return (x - np.floor(x)).astype(np.float32)                                     # GLSL fract


def test_ld_sample2_is_the_r2_sequence():
    """ldSample2 (common.glsl:1284-1288): sample 0 of a pixel is its rotation, a multiple of 2^-24 per axis; sample
    k is fract(rotation + float(k & 0xFFFF) * (1/g, 1/g^2)) bit for bit in float32, with g the plastic number
    computed here from its defining cubic; k wraps at 2^16.  A constant one float32 ulp off fails."""
    alpha = _r2_constants()
    ks = np.arange(1, 3000, dtype=np.int64)
    for px, py, salt in ((0, 0, 0), (17, 931, 0x68e31da4), (1919, 1079, 0x5EED), (5, 7, 123456789)):
        u0 = _ld(px, py, 0, salt)
        assert np.all(u0 * np.float32(16777216.0) == np.floor(u0 * np.float32(16777216.0)))
        got = np.array([_ld(px, py, k, salt) for k in ks])
        assert np.array_equal(got, _r2_sequence(u0, ks, alpha)), (px, py)
        assert np.array_equal(_ld(px, py, 65536 + 77, salt), got[76])
        for ax in (0, 1):
            bad = alpha.copy()
            bad[ax] = np.nextafter(bad[ax], np.float32(1.0))
            assert not np.array_equal(got, _r2_sequence(u0, ks, bad))


def _fibonacci_set(n, golden):
    i = np.arange(n, dtype=np.float64)
    z = 1.0 - (2.0 * i + 1.0) / n
    rr = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    ph = i * golden
    return np.stack([np.cos(ph) * rr, z, np.sin(ph) * rr], 1)


def test_probe_dirs_are_the_rotated_fibonacci_set():
    """probeRayDir (probes.glsl:155-170): direction i of n is the spherical-Fibonacci point z = 1 - (2i+1)/n,
    phi = i * golden angle (pi (3 - sqrt 5)), turned by one rotation per frame; so every pair's dot product
    equals that of the closed-form set, index for index.  A one-digit slip in the golden angle or the
    z spacing fails."""
    golden = np.pi * (3.0 - np.sqrt(5.0))
    # the float32 angle i * golden is off by up to ~ulp(i * golden): 6e-6 at n = 64, 2.5e-5 at n = 256
    for n, tol in ((64, 1e-5), (256, 4e-5)):
        b = _fibonacci_set(n, golden)
        for frame in (0, 3, 11):
            d = O.probe_dirs(n, frame).astype(np.float64)
            assert np.abs(d @ d.T - b @ b.T).max() < tol, (n, frame)
        wrong_golden = _fibonacci_set(n, 2.39996423)
        assert np.abs(d @ d.T - wrong_golden @ wrong_golden.T).max() > 2 * tol
        i = np.arange(n)
        wz = np.stack([np.cos(i * golden) * np.sqrt(1 - (1 - 2 * i / n) ** 2), 1 - 2 * i / n,
                       np.sin(i * golden) * np.sqrt(1 - (1 - 2 * i / n) ** 2)], 1)
        assert np.abs(d @ d.T - wz @ wz.T).max() > 1e-3


def _sky_closed_form(scene, rd, model, glow_exp=8.0, ph_exp=0.7, sunset_k=9.0):
    rd = np.asarray(rd, np.float64)
    sun = np.asarray(scene.sun_dir, np.float64)
    h = min(max(rd[1] * 0.5 + 0.5, 0.0), 1.0)
    ph = h ** ph_exp
    mu = max(float(rd @ sun), 0.0)
    glow = mu ** glow_exp
    tint = np.array([1.0, 0.9, 0.7])
    if model == 0:
        sky = np.array([0.75, 0.82, 0.90]) * (1 - ph) + np.array([0.25, 0.45, 0.85]) * ph
        out = sky + tint * (glow * 0.25)
    else:
        sky = np.array(scene.sky_horizon, np.float64) * (1 - ph) + np.array(scene.sky_zenith, np.float64) * ph
        sky = sky + np.array(scene.sunset_rgb, np.float64) * (mu ** 3 * np.exp(-max(rd[1], 0.0) * 5.0) +
                                                             0.25 * np.exp(-abs(rd[1]) * sunset_k))
        srgb = np.array(scene.sun_rgb, np.float64)
        out = sky + srgb / max(srgb[0], srgb[1], 1e-3) * tint * (glow * 0.25)
    return np.minimum(out * scene.sky_scale, 3.0)


def test_sky_color_closed_forms():
    """skyColor(rd, 0) (common.glsl:1082-1099) as a GI ray's escape radiance in an empty world (giEscape 1,
    clamped to 3): both sky models against their closed forms at the zenith, near the horizon, toward the sun
    and around it; a changed exponent (glow 8 -> 16, horizon ramp 0.7 -> 0.8, sunset falloff 9 -> 8) fails."""
    vox = np.zeros((64, 64, 64), np.uint8)
    _, a, ow = dense_world(vox)
    P = np.array([30.5, 10.0, 30.5], np.float32)
    N = np.array([0.0, 1.0, 0.0], np.float32)
    L = np.zeros(3, np.float32)
    rng = np.random.default_rng(2)
    for model in (0, 1):
        scene = vf.default_scene((8, 8, 8), sky_model=model)
        sun = np.array(scene.sun_dir, np.float32)
        dirs = [np.array([0.0, 1.0, 0.0]), sun, np.array([0.6, 0.03, -0.8]), np.array([-0.99, 0.02, 0.1])]
        for _ in range(12):
            v = rng.normal(size=3)
            v[1] = abs(v[1]) + 0.01
            dirs.append(v)
        t = sun + np.float32(0.2) * np.array([sun[2], 0.0, -sun[0]], np.float32)
        dirs.append(t)
        s = ow.sky(scene)
        worst = {"glow": 0.0, "ph": 0.0, "sunset": 0.0}
        for d in dirs:
            d = (np.asarray(d, np.float64) / np.linalg.norm(d)).astype(np.float32)
            O.lib().vfo_gi_path(ctypes.byref(ow.s), ctypes.byref(s), P.ctypes.data, N.ctypes.data, d.ctypes.data,
                                L.ctypes.data, None, None)
            want = _sky_closed_form(scene, d, model)
            assert np.allclose(L, want, rtol=2e-5, atol=1e-6), (model, d, L, want)
            for k, kw in (("glow", {"glow_exp": 16.0}), ("ph", {"ph_exp": 0.8}), ("sunset", {"sunset_k": 8.0})):
                worst[k] = max(worst[k], float(np.abs(L - _sky_closed_form(scene, d, model, **kw)).max()))
        assert worst["glow"] > 1e-2 and worst["ph"] > 1e-2, worst
        if model == 1:
            assert worst["sunset"] > 1e-3, worst


def test_probe_dirs_spherical_fibonacci():
    d = O.probe_dirs(256, 3)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=2e-5)
    assert np.linalg.norm(d.mean(0)) < 0.02
    d0 = O.probe_dirs(256, 0)
    # the rotation preserves the pairwise angles of the unrotated set
    assert np.allclose(d0 @ d0.T, d @ d.T, atol=5e-5)


# --- GI path ----------------------------------------------------------------------------------
def test_gi_open_sky_and_water_semantics():
    """An empty world: every GI ray escapes to the sky (sky colour x escape 1 at the top).  A bounce
    ray that meets water is a miss (common.glsl:2295)."""
    scene = vf.default_scene((8, 8, 8))
    vox = np.zeros((64, 64, 64), np.uint8)
    _, a, ow = dense_world(vox)
    s = ow.sky(scene)
    L = np.zeros(3, np.float32)
    P = np.array([30.5, 10.0, 30.5], np.float32)
    N = np.array([0.0, 1.0, 0.0], np.float32)
    d = np.array([0.0, 1.0, 0.0], np.float32)
    O.lib().vfo_gi_path(ctypes.byref(ow.s), ctypes.byref(s), P.ctypes.data, N.ctypes.data, d.ctypes.data,
                        L.ctypes.data, None, None)
    # classic sky straight up: mix(a, b, 1) + glow
    mu = max(float(np.dot(d, scene.sun_dir)), 0.0)
    exp = np.array([0.25, 0.45, 0.85]) + np.array([1.0, 0.9, 0.7]) * (mu ** 8 * 0.25)
    assert np.allclose(L, exp, rtol=1e-5)
    vox[:, 0:20, :] = 5
    _, a, ow = dense_world(vox)
    s = ow.sky(scene)
    down = np.array([0.0, -1.0, 0.0], np.float32)
    P2 = np.array([30.5, 30.0, 30.5], np.float32)
    O.lib().vfo_gi_path(ctypes.byref(ow.s), ctypes.byref(s), P2.ctypes.data, N.ctypes.data, down.ctypes.data,
                        L.ctypes.data, None, None)
    # water is a miss: escape downward inside the world -> giEscape ends at the water surface column
    assert np.all(np.isfinite(L))


def _gbuffer_points(n):
    """Non-sky G-buffer points (position, quantised normal) of the fixed world's test camera."""
    _, a, ow = world(1)
    _, cam = config1_camera()
    hits, pos, nrm, _ = ow.pass_primary(cam, 96, 54)
    idx = np.nonzero(pos[:, 3] >= 0.0)[0][:n]
    N = np.stack([((nrm[idx] >> (8 * k)) & 255).astype(np.float32) / np.float32(255.0) * np.float32(2.0)
                  - np.float32(1.0) for k in range(3)], 1).astype(np.float32)
    return ow, a, pos[idx, :3].copy(), N


def test_two_bounce_path_without_bounce_light_is_the_one_bounce_path():
    """giScale 0 zeroes the throughput of bounce 2 (common.glsl:2277, :2339): the two-bounce path then
    returns exactly what the one-bounce path (pinned above) returns."""
    ow, a, P, N = _gbuffer_points(300)
    scene = vf.default_scene(a["world_bricks"])
    scene.gi_scale = 0.0
    rng = np.random.default_rng(4)
    for i in range(len(P)):
        seed = int(rng.integers(0, 2 ** 32))
        d, seed = O.cosine_hemisphere(N[i], seed)
        L3, _, _ = ow.gi_path(scene, P[i], N[i], d, seed, last_stage=3)
        L1, _, _ = ow.gi_path(scene, P[i], N[i], d, seed, last_stage=1)
        assert np.array_equal(L3, L1), i


def test_two_bounce_path_is_its_head_plus_tail():
    """Stopping a path where bounce 2 starts and continuing it from that point by hand (bounce 2 +
    its sun ray, no trace-on: giPathImpl(..., 1, false)) gives the whole path: L = pre + thr * rest.
    This is the split the learned cache's targets and queries rely on."""
    ow, a, P, N = _gbuffer_points(400)
    scene = vf.default_scene(a["world_bricks"])
    rng = np.random.default_rng(5)
    tails = 0
    for i in range(len(P)):
        seed0 = int(rng.integers(0, 2 ** 32))
        d, seed = O.cosine_hemisphere(N[i], seed0)
        L, _, _ = ow.gi_path(scene, P[i], N[i], d, seed, last_stage=3)
        Lh, s_after, t = ow.gi_path(scene, P[i], N[i], d, seed, last_stage=3, tail=True)
        assert s_after == seed   # nothing drawn before bounce 2
        if not t["valid"]:
            assert np.array_equal(L, Lh)
            continue
        tails += 1
        d2, s2 = O.cosine_hemisphere(t["n"], seed)
        Lr, _, _ = ow.gi_path(scene, t["p"], t["n"], d2, s2, last_stage=1, trace_on=False)
        want = Lh.astype(np.float64) + t["thr"].astype(np.float64) * Lr
        assert np.allclose(L, want, rtol=2e-6, atol=1e-7), (i, L, want)
        # the throughput is albedo x giScale of the bounce-1 hit
        assert np.all(t["thr"] >= 0.0) and np.all(t["thr"] <= 1.0)
    assert tails > 40


def test_three_bounce_path_is_quadratic_in_the_bounce_weight():
    """The three-bounce path's light is A + g*B + g^2*C in the bounce weight g = giScale (the throughput picks
    up albedo x g at each of bounce 1's and bounce 2's hits, and nothing else depends on g).  Pinned from three
    evaluations: g = 0 is the one-bounce path exactly; B equals the two-bounce path's own bounce-2 term (so the
    continuation draws bounce 2's direction from the same seed state, starts at the same point and carries the
    same throughput); C >= 0 and is nonzero for many paths (bounce 3 adds light)."""
    ow, a, P, N = _gbuffer_points(1500)
    rng = np.random.default_rng(9)
    scenes = {}
    for g in (0.0, 1.0, 2.0):
        sc = vf.default_scene(a["world_bricks"])
        sc.gi_scale = g
        scenes[g] = sc
    third = 0
    for i in range(len(P)):
        seed0 = int(rng.integers(0, 2 ** 32))
        d, seed = O.cosine_hemisphere(N[i], seed0)
        L0, _ = ow.gi_path3(scenes[0.0], P[i], N[i], d, seed)
        L1p, _, _ = ow.gi_path(scenes[0.0], P[i], N[i], d, seed, last_stage=1)
        assert np.array_equal(L0, L1p), i
        A = L0.astype(np.float64)
        y1 = ow.gi_path3(scenes[1.0], P[i], N[i], d, seed)[0].astype(np.float64) - A
        y2 = ow.gi_path3(scenes[2.0], P[i], N[i], d, seed)[0].astype(np.float64) - A
        C = (y2 - 2.0 * y1) / 2.0
        B = y1 - C
        B2 = ow.gi_path(scenes[1.0], P[i], N[i], d, seed, last_stage=3)[0].astype(np.float64) - A
        scale = 1e-5 * (1.0 + np.abs(y2).max())
        assert np.allclose(B, B2, atol=scale), (i, B, B2)
        assert np.all(C >= -scale), (i, C)
        third += int(C.max() > 10 * scale)
    assert third > 30, third


# --- camera inside solid terrain (gbuffer.comp:233-266, exitSolid common.glsl:990-1016) ------------
def _solid_world():
    """64^3 voxels: a solid block [8,40) x [8,24) x [8,40) of material 3 (whole bricks inside, partial ones
    at its faces) with a pocket of water at [24,28)^3 and a cave of air at [12,16) x [12,16) x [30,34)."""
    vox = np.zeros((64, 64, 64), np.uint8)   # indexed (x, y, z); from_dense takes (z, y, x)
    vox[8:40, 8:24, 8:40] = 3
    vox[24:28, 16:20, 24:28] = 5
    vox[12:16, 12:16, 30:34] = 0
    return vox


def _march_exit_f64(vox, ro, rd):
    """First non-solid voxel along the ray by an FP64 voxel march: its entry t; -2 out of the box."""
    ro, rd = np.asarray(ro, np.float64), np.asarray(rd, np.float64)
    sd = np.where(np.abs(rd) < 1e-5, 1e-5, rd)
    v = np.floor(ro).astype(np.int64)
    st = np.where(sd > 0, 1, -1)
    tmax = ((v + (st > 0)) - ro) / sd
    tdel = np.abs(1.0 / sd)
    t = 0.0
    for _ in range(4096):
        if np.any(v < 0) or np.any(v >= 64):
            return -2.0
        m = vox[v[0], v[1], v[2]]
        if m == 0 or m == 5:
            return t
        a = int(np.argmin(tmax))
        t = tmax[a]
        v[a] += st[a]
        tmax[a] += tdel[a]
    return -1.0


def test_exit_solid_closed_forms():
    vox = _solid_world()
    _, _, ow = dense_world(vox.transpose(2, 1, 0))
    # axis-aligned from inside the block: the distance to its face
    assert ow.exit_solid([20.25, 10.5, 20.5], [1.0, 0.0, 0.0]) == pytest.approx(40.0 - 20.25, abs=1e-5)
    assert ow.exit_solid([20.25, 10.5, 20.5], [0.0, 1.0, 0.0]) == pytest.approx(24.0 - 10.5, abs=1e-5)
    assert ow.exit_solid([20.25, 10.5, 20.5], [0.0, 0.0, -1.0]) == pytest.approx(20.5 - 8.0, abs=1e-5)
    # into the water pocket: water is not solid
    assert ow.exit_solid([26.5, 12.5, 26.5], [0.0, 1.0, 0.0]) == pytest.approx(16.0 - 12.5, abs=1e-5)
    # a camera in air: 0; the material queries
    assert ow.exit_solid([50.5, 50.5, 50.5], [0.3, 0.4, 0.5]) == 0.0
    assert ow.voxel_mat_at([20.2, 10.1, 20.9]) == 3 and ow.voxel_mat_at([26.5, 17.5, 26.5]) == 5
    assert ow.voxel_mat_at([50.5, 50.5, 50.5]) == 0 and ow.voxel_mat_at([-1.0, 5.0, 5.0]) == 0


def test_exit_solid_matches_fp64_voxel_march():
    """exitSolid crosses whole uniform bricks in one step; the voxel-by-voxel FP64 march must find the
    same exit distance (and -2 when the ray leaves the world while still in solid)."""
    vox = _solid_world()
    vox[:, 0:8, :] = 2   # a solid floor reaching the world's edge: rays down leave the world in solid
    _, _, ow = dense_world(vox.transpose(2, 1, 0))
    rng = np.random.default_rng(21)
    n_out = 0
    for _ in range(600):
        ro = rng.uniform([8.0, 1.0, 8.0], [40.0, 24.0, 40.0])
        if vox[int(ro[0]), int(ro[1]), int(ro[2])] in (0, 5):
            continue
        d = rng.normal(size=3)
        if rng.uniform() < 0.15:
            d[rng.integers(0, 3)] = 0.0
        d /= np.linalg.norm(d)
        ro32, d32 = ro.astype(np.float32), d.astype(np.float32)
        got = ow.exit_solid(ro32, d32)
        want = _march_exit_f64(vox, ro32, d32)
        if want == -2.0:
            n_out += 1
            assert got == -2.0
        else:
            assert got == pytest.approx(want, rel=1e-5, abs=1e-4), (ro32, d32)
    assert n_out > 10


# --- sun rays through leaves (FOLIAGE_LIGHT, sunVisibilityT common.glsl:1964-1981) ----------------------
def _leaf_column_world(k, blocker=False):
    """64^3 voxels: a floor of stone at y < 8 and, above (20, *, 20), k leaf voxels stacked from y = 12
    (with a stone voxel on top when `blocker`).  Indexed (x, y, z); from_dense takes (z, y, x)."""
    vox = np.zeros((64, 64, 64), np.uint8)
    vox[:, 0:8, :] = 3
    for j in range(k):
        vox[20, 12 + 2 * j, 20] = 7          # leaves (palette kind 3), a gap between them
    if blocker:
        vox[20, 12 + 2 * k, 20] = 3
    return vox


def test_sun_visibility_through_leaves_closed_form():
    """Straight up through k separated leaf voxels: deterministic visibility leafLight^k (k <= 5), 0 from
    the sixth leaf on or at any non-leaf blocker; tFirst = distance to the first leaf; leafLight 0 is the
    hard shadow ray."""
    o = np.array([20.5, 8.03, 20.5], np.float32)
    up = np.array([0.0, 1.0, 0.0], np.float32)
    for k in range(0, 8):
        _, _, ow = dense_world(_leaf_column_world(k).transpose(2, 1, 0))
        v, tf, _ = ow.sun_visibility(o, up, 0.6)
        want = 0.6 ** k if k <= 5 else 0.0
        assert v == pytest.approx(want, rel=1e-6), (k, v)
        assert tf == (pytest.approx(12.0 - 8.03, abs=1e-5) if k > 0 else 0.0)
        v0, tf0, _ = ow.sun_visibility(o, up, 0.0)
        assert v0 == (0.0 if k > 0 else 1.0) and tf0 == tf
    _, _, ow = dense_world(_leaf_column_world(2, blocker=True).transpose(2, 1, 0))
    assert ow.sun_visibility(o, up, 0.6)[0] == 0.0
    # stochastic: pass-or-block with probability 0.6 per leaf
    _, _, ow = dense_world(_leaf_column_world(2).transpose(2, 1, 0))
    hits = [ow.sun_visibility(o, up, 0.6, seed=s, stochastic=True)[0] for s in range(4000)]
    assert set(hits) <= {0.0, 1.0}
    assert abs(np.mean(hits) - 0.36) < 0.03


def test_sample_cone_stays_in_the_sun_disk():
    rng = np.random.default_rng(6)
    lib = O.lib()
    for _ in range(20):
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        ang = float(rng.uniform(0.01, 0.2))
        r2 = []
        seed = ctypes.c_uint32(int(rng.integers(0, 2 ** 32)))
        out = np.zeros(3, np.float32)
        for _ in range(400):
            lib.vfo_sample_cone(d.ctypes.data, ctypes.c_float(ang), ctypes.byref(seed), out.ctypes.data)
            c = float(np.clip(np.dot(out, d), -1.0, 1.0))
            th = math.acos(c)
            assert th <= math.atan(ang) + 1e-5
            r2.append(math.tan(th) ** 2)
        # r = ang * sqrt(u): uniform over the disk, E[r^2] = ang^2 / 2
        assert abs(np.mean(r2) / (ang * ang) - 0.5) < 0.06


# --- back-lit foliage (FOLIAGE_SSS: shadow_trace.comp:35-46, lighting.comp:66-71, common.glsl:1953-1989) ---
def _nbytes(n, alpha):
    b = [int(math.floor(min(max(c * 0.5 + 0.5, 0.0), 1.0) * 255.0 + 0.5)) for c in n]
    return np.uint32(b[0] | (b[1] << 8) | (b[2] << 16) | (alpha << 24))


def _leaf_slabs_world(ks, y0=20):
    """64^3 voxels: stone floor at y < 8; column j (x = 4 + 4j, z = 20) holds a slab of ks[j] leaf voxels
    from y0 up.  Indexed (x, y, z)."""
    vox = np.zeros((64, 64, 64), np.uint8)
    vox[:, 0:8, :] = 3
    for j, k in enumerate(ks):
        vox[4 + 4 * j, y0:y0 + k, 20] = 7
    return vox


LEAF_A, BLADE_A = 204, 178   # unorm8(0.8), unorm8(0.7): gbNormal.a of a leaf voxel and of a grass blade


def test_back_lit_leaf_closed_form():
    """A G-buffer texel on the underside of a slab of k leaf voxels, the sun straight up: the texel faces
    away from the sun, so only FOLIAGE_SSS traces.  A leaf voxel's ray starts at its centre stepped 0.9
    toward the sun (foliageSunOrigin: P - N*0.5 + sun*0.9), inside the second layer, so it crosses k - 1
    leaves: visibility leafLight^(k-1), 0 from the sixth leaf on; a blade starts just off the surface
    (P + sun*0.05), inside the first layer: leafLight^k.  No blocker distance is stored.  An opaque
    texel (alpha 255) and a sun below the horizon get no ray."""
    ks = [1, 2, 3, 4, 5, 6, 7]
    _, _, ow = dense_world(_leaf_slabs_world(ks).transpose(2, 1, 0))
    up = np.array([0.0, 1.0, 0.0], np.float32)
    down_n = (0.0, -1.0, 0.0)
    cases = []   # (P, normal bytes, expected visibility)
    for j, k in enumerate(ks):
        P = (4 + 4 * j + 0.5, 20.0, 20.5)
        m = k - 1
        cases.append((P, _nbytes(down_n, LEAF_A), float(np.float32(0.6) ** m) if m <= 5 else 0.0))
        cases.append((P, _nbytes(down_n, BLADE_A), float(np.float32(0.6) ** k) if k <= 5 else 0.0))
        cases.append((P, _nbytes(down_n, 255), 0.0))
    n = len(cases)
    pos = np.array([[*c[0], 5.0] for c in cases], np.float32)
    nrm = np.array([c[1] for c in cases], np.uint32)
    sh, _, cnt = ow.pass_shadow(up, n, 1, pos, nrm, leaf_light=0.6)
    for i, c in enumerate(cases):
        assert sh[i, 0] == pytest.approx(c[2], rel=1e-6), (i, c, sh[i])
        assert sh[i, 1] == 0.0
    assert cnt["rays"] == 2 * len(ks)   # one ray per foliage texel, however many leaves it crosses
    # a lone leaf voxel's side face, the sun grazing it from behind: a ray (nothing beyond: lit) only while the
    # sun is above the horizon
    _, _, ow = dense_world(_leaf_slabs_world([1]).transpose(2, 1, 0))
    side = _nbytes((-1.0, 0.0, 0.0), LEAF_A)
    P = np.array([[4.0, 20.5, 20.5, 5.0]], np.float32)
    for sy, want in ((-0.1, 0.0), (0.1, 1.0)):
        sun = np.array([1.0, sy, 0.0], np.float32)
        sun /= np.linalg.norm(sun)
        sh, _, cnt = ow.pass_shadow(sun, 1, 1, P, np.array([side], np.uint32), leaf_light=0.6)
        assert sh[0, 0] == want and cnt["rays"] == (1 if sy > 0 else 0), (sy, sh)


def test_back_lit_leaf_mistakes_are_caught():
    """The closed form separates foliageSunOrigin from its plausible misreadings: the lit-face origin
    P + N*0.03 or a flipped N*0.5 start below or inside the first layer and cross all k leaves."""
    k = 3
    _, _, ow = dense_world(_leaf_slabs_world([k]).transpose(2, 1, 0))
    up = np.array([0.0, 1.0, 0.0], np.float32)
    P = np.array([4.5, 20.0, 20.5], np.float32)
    N = np.array([128 / 255 * 2 - 1, -1.0, 128 / 255 * 2 - 1], np.float32)
    right = ow.sun_visibility((P - N * np.float32(0.5)) + up * np.float32(0.9), up, 0.6)[0]
    assert right == pytest.approx(0.36, rel=1e-6)
    for wrong_o in (P + N * np.float32(0.03), (P + N * np.float32(0.5)) + up * np.float32(0.9)):
        assert ow.sun_visibility(wrong_o, up, 0.6)[0] == pytest.approx(0.216, rel=1e-6)


def test_back_lit_leaf_sun_cone_mean():
    """lighting.comp's stochastic sample from the underside of a wide slab of 3 leaf voxels: each leaf
    passes the ray with probability 0.6, so the mean over many texels is 0.6^2 (the origin is in the
    second layer); texels that are not foliage stay 0."""
    vox = np.zeros((64, 64, 64), np.uint8)
    vox[:, 0:8, :] = 3
    vox[8:56, 20:23, 8:56] = 7
    _, _, ow = dense_world(vox.transpose(2, 1, 0))
    sun = np.array([0.05, 1.0, 0.03], np.float32)
    sun /= np.linalg.norm(sun)
    rng = np.random.default_rng(3)
    W, H = 64, 64
    xz = rng.uniform(16, 48, (W * H, 2)).astype(np.float32)
    pos = np.stack([xz[:, 0], np.full(W * H, 20.0, np.float32), xz[:, 1], np.full(W * H, 5.0, np.float32)], 1)
    nrm = np.full(W * H, _nbytes((0.0, -1.0, 0.0), LEAF_A), np.uint32)
    out, _ = ow.pass_sun_cone(sun, 0.02, 0.6, 7, W, H, pos, nrm)
    assert set(np.unique(out)) <= {0.0, 1.0}
    assert abs(out.mean() - 0.36) < 0.03
    nrm[:] = _nbytes((0.0, -1.0, 0.0), 255)
    out, _ = ow.pass_sun_cone(sun, 0.02, 0.6, 7, W, H, pos, nrm)
    assert not out.any()


def test_gbuffer_alpha_marks_foliage():
    """gbuffer.comp:526 (WATER_CLEAR on): gbNormal.a = 0.8 (byte 204) on a hit whose material is palette
    kind 3, 1.0 (255) on every other hit, 0 on the sky."""
    _, a, ow = world(1)
    pu, cam = config1_camera()
    hits, pos, nrm, _ = ow.pass_primary(cam, 160, 90)
    kinds = ow.palette[hits["mat"], 6]
    alpha = nrm >> 24
    hit = hits["hit"] == 1
    assert (alpha[~hit] == 0).all()
    assert (alpha[hit & (kinds == 3)] == LEAF_A).all() and (alpha[hit & (kinds != 3)] == 255).all()
    vox = _leaf_slabs_world([2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2])
    _, _, ow2 = dense_world(vox.transpose(2, 1, 0))
    cam2 = {"pos": [30.5, 40.0, 20.7], "right": [1.0, 0.0, 0.0], "up": [0.0, 0.0, 1.0], "fwd": [0.0, -1.0, 0.0],
            "tan_half_fov": 0.7}
    h2, _, n2, _ = ow2.pass_primary(cam2, 64, 64)
    leaf = (h2["hit"] == 1) & (h2["mat"] == 7)
    assert leaf.sum() > 20 and ((n2 >> 24)[leaf] == LEAF_A).all()
    assert ((n2 >> 24)[(h2["hit"] == 1) & (h2["mat"] == 3)] == 255).all()
