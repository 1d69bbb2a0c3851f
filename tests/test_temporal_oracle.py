"""Pins of the temporal-pass oracle (oracle/vfo_temporal.c) by what the mathematics fixes, not by its own
formulas: exp against float64, half conversions against numpy, a still camera's running mean and history
count, whole- and half-texel reprojection (the textbook Catmull-Rom weights -1/16, 9/16, 9/16, -1/16),
tap rejection by plane, normal, floating-origin shift and surface delta, the firefly cap and anti-lag
clamp on hand-computed cases, the guide bits, and the a-trous filter's B3-spline variance factor
(sum h^2)^2, identity on constant fields and its edge stops (gi.comp, lighting.comp, atrous.comp)."""
import ctypes

import numpy as np
import pytest

import oracle as O
from temporal_helpers import (DOWN, EAST, UP, f16_bits, f16_vals, gi_hist, gi_mom, motion_bits, pack_normal,
                              plane_gbuffer, zero_delta)

W, H = 128, 64           # full res; half res 64 x 32 (powers of two: (px + 0.5) / res * res is exact)
HW, HH = 64, 32
Z3 = (0.0, 0.0, 0.0)


def test_exp_and_powers():
    L = O.lib()
    x = np.concatenate([np.linspace(-86, 88, 200001), np.linspace(-1, 1, 20001)]).astype(np.float32)
    y = np.array([L.vfo_exp(float(v)) for v in x], np.float64)
    ref = np.exp(x.astype(np.float64))
    ulp = np.spacing(ref.astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(y - ref) / ulp) < 2.0
    assert L.vfo_exp(0.0) == 1.0 and L.vfo_exp(-90.0) == 0.0 and L.vfo_exp(float("-inf")) == 0.0
    assert np.isnan(L.vfo_exp(float("nan"))) and L.vfo_exp(100.0) == float("inf")
    # repeated squaring: relative error at most (1 + 2 + ... + 2^(k-1)) half-ulps for k squarings
    for v in np.float32([0.0, 0.5, 0.9, 0.999, 1.0, 1.0001, 0.7071]):
        v = float(v)
        assert abs(L.vfo_pow16(v) - v ** 16) <= 15 * 2 ** -24 * max(v ** 16, 1e-30)
        assert abs(L.vfo_pow64(v) - v ** 64) <= 63 * 2 ** -24 * max(v ** 64, 1e-30)


def test_half_conversions_match_numpy():
    L = O.lib()
    rng = np.random.default_rng(0)
    vals = np.concatenate([rng.normal(0, 1, 20000), rng.normal(0, 3000, 5000), rng.uniform(-7e4, 7e4, 2000),
                           np.exp2(rng.uniform(-27, -13, 5000)) * rng.choice([-1, 1], 5000),
                           [0.0, -0.0, 65504, 65519.99, 65520, 1e9, -1e9, 2 ** -24, 2 ** -25, 3 * 2 ** -26,
                            2 ** -25 * 1.0000001, 5.960464477539063e-08, 6.097555160522461e-05]]).astype(np.float32)
    # halfway cases between neighbouring halves (round to even)
    h = rng.integers(0, 0x7bff, 3000).astype(np.uint16)
    mid = ((h.view(np.float16).astype(np.float64) + (h + 1).view(np.float16).astype(np.float64)) / 2).astype(np.float32)
    vals = np.concatenate([vals, mid])
    got = np.array([L.vfo_f32_to_f16(float(v)) for v in vals], np.uint16)
    with np.errstate(over="ignore"):
        want = vals.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)
    allh = np.arange(0, 0x10000, dtype=np.uint32).astype(np.uint16)
    finite = (allh & 0x7c00) != 0x7c00
    back = np.array([L.vfo_f16_to_f32(int(b)) for b in allh[finite]], np.float32)
    assert np.array_equal(back.view(np.uint32), allh[finite].view(np.float16).astype(np.float32).view(np.uint32))


def _still(indirect_rgb, hist_prev, mom_prev, frame=0, bicubic=True, prev_gb=None, motion=(0.0, 0.0), delta=None,
           shift=Z3, cur_gb=None):
    pos, nrm = cur_gb if cur_gb is not None else plane_gbuffer(W, H)
    ppos, pnrm = prev_gb if prev_gb is not None else (pos, nrm)
    ind = np.zeros((HH, HW, 4), np.float32)
    ind[..., :3] = indirect_rgb
    ind[..., 3] = 1.0
    return O.gi_accumulate(W, H, frame, shift, pos, nrm, ppos, pnrm, motion_bits(W, H, *motion),
                           delta if delta is not None else zero_delta(W, H), ind, hist_prev, mom_prev, bicubic=bicubic)


def test_gi_still_camera_running_mean_and_count():
    """A still camera: frame k's history is the mean of the first k samples (w = 1/(count+1)) and the
    count runs 1, 2, ..., 48, then stays 49 (kMaxHistory 48)."""
    rng = np.random.default_rng(1)
    hist = np.zeros((HH, HW, 4), np.uint16)
    mom = np.zeros((HH, HW, 4), np.float32)
    samples = []
    for k in range(52):
        # colours of equal luminance (no firefly cap, anti-lag band never binds)
        c = rng.uniform(0.2, 0.8, 3)
        c *= 0.5 / (c @ [0.2126, 0.7152, 0.0722])
        samples.append(c)
        hist, mom, _ = _still(c.astype(np.float32), hist, mom, frame=k)
        cnt = f16_vals(hist[..., 3])
        assert np.all(cnt == min(k + 1, 49)), (k, np.unique(cnt))
        if k < 48:
            want = np.mean(samples, axis=0)
            got = f16_vals(hist[5:-5, 5:-5, :3]).reshape(-1, 3)
            assert np.abs(got - want).max() < 2.5e-3, k
    assert np.allclose(mom[3:-3, 3:-3, 0], 0.5, atol=1e-3)


def test_gi_whole_texel_reprojection_reads_the_neighbour():
    """motion = +1 half-res texel in x: the history comes from texel x+1; the last column has none."""
    rng = np.random.default_rng(2)
    field = rng.uniform(0, 1, (HH, HW, 3)).astype(np.float32)
    hist = gi_hist(HW, HH, field, 10.0)
    mom = gi_mom(HW, HH, 0.0, 100.0, 0.0)              # sigma 10: the anti-lag band never binds
    h, m, _ = _still(np.zeros(3, np.float32), hist, mom, motion=(1.0 / HW, 0.0))
    got = f16_vals(h)
    src = f16_vals(hist)[:, 1:, :3]
    assert np.allclose(got[:, :-1, :3], src * (10.0 / 11.0), rtol=1e-3, atol=1e-4)
    assert np.all(got[:, :-1, 3] == 11.0)
    assert np.all(got[:, -1, :3] == 0.0) and np.all(got[:, -1, 3] == 1.0)
    assert np.allclose(m[:, :-1, 1], 100.0 * 10.0 / 11.0, rtol=1e-6)
    # and -1 texel in y: from row y-1
    h, _, _ = _still(np.zeros(3, np.float32), hist, mom, motion=(0.0, -1.0 / HH))
    assert np.allclose(f16_vals(h)[1:, :, :3], f16_vals(hist)[:-1, :, :3] * (10.0 / 11.0), rtol=1e-3, atol=1e-4)
    assert np.all(f16_vals(h)[0, :, 3] == 1.0)


def test_gi_half_texel_bilinear_and_catmull_rom():
    """Half a texel in x: bilinear = the two taps' mean; Catmull-Rom = (-1, 9, 9, -1)/16 over x-1..x+2,
    clamped to the central taps' range."""
    rng = np.random.default_rng(3)
    row = rng.uniform(0, 1, (HW, 3)).astype(np.float32)
    field = np.broadcast_to(row, (HH, HW, 3))
    hist = gi_hist(HW, HH, field, 10.0)
    mom = gi_mom(HW, HH, 0.0, 100.0, 0.0)
    hv = f16_vals(hist)[0, :, :3].astype(np.float64)
    x = np.arange(1, HW - 2)
    lin = 0.5 * (hv[x] + hv[x + 1])
    cr = (-hv[x - 1] + 9 * hv[x] + 9 * hv[x + 1] - hv[x + 2]) / 16
    cr = np.clip(cr, np.minimum(hv[x], hv[x + 1]), np.maximum(hv[x], hv[x + 1]))
    h0, _, _ = _still(np.zeros(3, np.float32), hist, mom, motion=(0.5 / HW, 0.0), bicubic=False)
    h1, _, _ = _still(np.zeros(3, np.float32), hist, mom, motion=(0.5 / HW, 0.0), bicubic=True)
    g0 = f16_vals(h0)[4:-4, x, :3]
    g1 = f16_vals(h1)[4:-4, x, :3]
    assert np.allclose(g0, lin * 10 / 11, rtol=1e-3, atol=2e-4)
    assert np.allclose(g1, cr * 10 / 11, rtol=1e-3, atol=2e-4)
    assert np.abs(g1 - g0).max() > 0.01                  # the two filters differ on this field


def test_gi_tap_rejection():
    hist = gi_hist(HW, HH, (0.7, 0.7, 0.7), 10.0)
    mom = gi_mom(HW, HH, 0.7, 100.0, 0.7)
    ind = np.full(3, 0.1, np.float32)
    pos, nrm = plane_gbuffer(W, H)

    def count(**kw):
        h, _, _ = _still(ind, hist, mom, **kw)
        return f16_vals(h)[2:-2, 2:-2, 3]

    assert np.all(count() == 11.0)
    assert np.all(count(prev_gb=plane_gbuffer(W, H, normal=DOWN)) == 1.0)          # normal dot <= 0.85
    assert np.all(count(prev_gb=plane_gbuffer(W, H, normal=EAST)) == 1.0)
    # plane distance: tol = 0.5 + t*0.004 = 0.54 at t = 10 (N.y = 1 + 1/255 after decoding)
    assert np.all(count(prev_gb=plane_gbuffer(W, H, y=0.53)) == 11.0)
    assert np.all(count(prev_gb=plane_gbuffer(W, H, y=0.55)) == 1.0)
    # a floating-origin rebase: old positions + shiftP
    assert np.all(count(prev_gb=plane_gbuffer(W, H, y=-3.0), shift=(0.0, 3.0, 0.0)) == 11.0)
    assert np.all(count(prev_gb=plane_gbuffer(W, H, y=-3.0), shift=(0.0, -3.0, 0.0)) == 1.0)
    # a moving surface: P - delta is where it was
    d = np.zeros((H, W, 4), np.float32)
    d[..., 1] = 2.0
    assert np.all(count(prev_gb=plane_gbuffer(W, H, y=-2.0), delta=f16_bits(d)) == 11.0)
    assert np.all(count(prev_gb=plane_gbuffer(W, H, y=-2.0)) == 1.0)
    # previous hit distance 0 is not history (w <= 0), current w == 0 is geometry
    p0 = pos.copy()
    p0[..., 3] = 0.0
    assert np.all(count(prev_gb=(p0, nrm)) == 1.0)
    assert np.all(count(cur_gb=(p0, nrm), prev_gb=(pos, nrm)) == 11.0)


def test_gi_rotating_subtexel_and_its_normal():
    """frame & 3 picks the sampled sub-texel (off = (f & 1, (f >> 1) & 1)); its normal is the one tested."""
    pos, nrm = plane_gbuffer(W, H)
    n2 = nrm.copy()
    n2[0::2, 1::2] = pack_normal(DOWN)                  # sub-texel (1, 0) of every 2x2 block faces down
    hist = gi_hist(HW, HH, (0.5, 0.5, 0.5), 10.0)
    mom = gi_mom(HW, HH, 0.5, 100.0, 0.5)
    for f in range(8):
        h, _, g = _still(np.full(3, 0.5, np.float32), hist, mom, frame=f, cur_gb=(pos, n2), prev_gb=(pos, nrm))
        want = 1.0 if (f & 3) == 1 else 11.0
        assert np.all(f16_vals(h)[..., 3] == want), f
        assert np.all(g[..., 3] == pack_normal(UP))       # the guide is the representative (x*2, y*2)


def test_gi_firefly_cap_and_anti_lag():
    # firefly: count 10 >= 4, m1 0.2, sigma 0.1 -> cap = 0.2 + 4*0.1 + 0.05 = 0.65
    hist = gi_hist(HW, HH, (0.2, 0.2, 0.2), 10.0)
    mom = gi_mom(HW, HH, 0.2, 0.05, 0.2)
    h, m, _ = _still(np.full(3, 5.0, np.float32), hist, mom)
    got = f16_vals(h)[3:-3, 3:-3, 0]
    assert np.allclose(got, (10 * 0.2 + 0.65) / 11, rtol=2e-3)
    assert np.allclose(m[3:-3, 3:-3, 0], (10 * 0.2 + 0.65) / 11, rtol=1e-5)
    assert np.allclose(m[3:-3, 3:-3, 2], 0.75 * 0.2 + 0.25 * 0.65, rtol=1e-5)        # fast, alpha 0.25
    # below 4 frames of history there is no cap: (3*0.2 + 5)/4 (sigma 10: the anti-lag band is wide)
    h, _, _ = _still(np.full(3, 5.0, np.float32), gi_hist(HW, HH, (0.2, 0.2, 0.2), 3.0), gi_mom(HW, HH, 0.2, 100, 0.2))
    assert np.allclose(f16_vals(h)[3:-3, 3:-3, 0], (3 * 0.2 + 5.0) / 4, rtol=2e-3)
    # anti-lag: history luma 1, sigma 0, fast history 0.1 and a 0.1 sample -> band 0.02, slow = 0.12
    hist = gi_hist(HW, HH, (1.0, 1.0, 1.0), 10.0)
    mom = gi_mom(HW, HH, 0.5, 0.25, 0.1)
    h, _, _ = _still(np.full(3, 0.1, np.float32), hist, mom)
    assert np.allclose(f16_vals(h)[3:-3, 3:-3, :3], (10 * 0.12 + 0.1) / 11, rtol=2e-3)


def test_gi_guide_and_sky():
    pos, nrm = plane_gbuffer(W, H)
    nrm = (nrm & np.uint32(0x00ffffff)) | (np.arange(W, dtype=np.uint32)[None, :] << 24)   # tag the alpha byte
    pos[0:2, 0:2, 3] = -1.0                               # block (0, 0): all sky
    pos[2:4, 0:2, 3] = -1.0
    pos[3, 1, 3] = 5.0                                    # block (0, 1): only texel (1, 3) is geometry
    hist = gi_hist(HW, HH, (0.5, 0.5, 0.5), 10.0)
    mom = gi_mom(HW, HH, 0.5, 0.3, 0.5)
    h, m, g = _still(np.full(3, 0.5, np.float32), hist, mom, cur_gb=(pos, nrm))
    assert np.all(g[0, 0] == np.r_[pos[0, 0, :3].view(np.uint32), 0]) and np.all(h[0, 0] == 0) and np.all(m[0, 0] == 0)
    assert np.all(g[1, 0, :3] == pos[3, 1, :3].view(np.uint32)) and g[1, 0, 3] == nrm[3, 1]
    assert np.all(g[5, 7, :3] == pos[10, 14, :3].view(np.uint32)) and g[5, 7, 3] == nrm[10, 14]
    assert f16_vals(h[1, 0, 3]) > 0


def _light(shadow, hist_prev, motion=(0.0, 0.0), prev_gb=None, cur_gb=None):
    pos, nrm = cur_gb if cur_gb is not None else plane_gbuffer(W, H)
    ppos, pnrm = prev_gb if prev_gb is not None else (pos, nrm)
    s = np.broadcast_to(np.asarray(shadow, np.float32), (H, W))
    return O.light_accumulate(W, H, Z3, pos, nrm, ppos, pnrm, motion_bits(W, H, *motion), zero_delta(W, H), s,
                              hist_prev)


def test_light_history():
    hist = np.zeros((H, W, 4), np.float32)
    for k in range(51):
        hist = _light(0.3, hist)
        assert np.all(hist[..., 2] == min(k + 1, 49))
        assert np.allclose(hist[..., 0], 0.3, atol=1e-6) and np.allclose(hist[..., 3], 0.3, atol=1e-6)
    # sky pixels: (1, 1, 0, 0)
    pos, nrm = plane_gbuffer(W, H)
    pos[7, 9, 3] = -0.5
    out = _light(0.0, hist, cur_gb=(pos, nrm))
    assert np.all(out[7, 9] == [1, 1, 0, 0])
    # whole-pixel reprojection: +1 pixel in x reads x+1
    rng = np.random.default_rng(5)
    hp = np.zeros((H, W, 4), np.float32)
    hp[..., 0] = rng.random((H, W))
    hp[..., 1] = 100.0
    hp[..., 2] = 10.0
    hp[..., 3] = hp[..., 0]
    out = _light(0.0, hp, motion=(1.0 / W, 0.0))
    assert np.allclose(out[:, :-1, 0], hp[:, 1:, 0] * 10 / 11, rtol=1e-6, atol=1e-7)
    assert np.all(out[:, -1, 2] == 1.0)
    # anti-lag: mean 1, sigma 0, count 10, a 0 sample: fast = 0.75, band 0.02 -> slow 0.77
    hp = np.zeros((H, W, 4), np.float32)
    hp[..., 0], hp[..., 1], hp[..., 2], hp[..., 3] = 1.0, 1.0, 10.0, 1.0
    out = _light(0.0, hp)
    assert np.allclose(out[..., 0], 0.77 * 10 / 11, rtol=1e-6)
    assert np.allclose(out[..., 1], 10 / 11, rtol=1e-6) and np.allclose(out[..., 3], 0.75, rtol=1e-6)


def _guide_of(pos, nrm):
    hist = gi_hist(HW, HH, (0, 0, 0), 0.0)
    _, _, g = _still(np.zeros(3, np.float32), hist, gi_mom(HW, HH, 0, 0, 0), cur_gb=(pos, nrm))
    return g


def test_atrous_b3_variance_factor_and_constant_field():
    """A converged constant field on one plane: colour unchanged; the variance shrinks by the B3 spline's
    (sum h^2)^2 = 0.2734375^2 at an interior texel (edge stops ~1 on a plane)."""
    pos, nrm = plane_gbuffer(W, H, t=50.0)
    g = _guide_of(pos, nrm)
    c = np.array([0.3, 0.5, 0.2], np.float32)
    lum = float(c @ [0.2126, 0.7152, 0.0722])
    v = 0.04
    hist = gi_hist(HW, HH, c, 48.0)
    mom = gi_mom(HW, HH, lum, lum * lum + v, lum)
    out = O.gi_atrous_level(W, H, 0, 1, pos, nrm, mom, hist, g)
    val = f16_vals(out)
    assert np.allclose(val[..., :3], f16_vals(hist)[..., :3], rtol=2e-3)
    inner = val[4:-4, 4:-4, 3]
    assert np.allclose(inner, v * 0.2734375 ** 2, rtol=0.02)
    # levels 1, 2 with larger steps leave the constant colour alone too, and shrink the variance again
    f0, f1 = O.gi_atrous(W, H, 3, pos, nrm, mom, hist, g)
    assert np.allclose(f16_vals(f0)[..., :3], c, rtol=3e-3)
    assert f16_vals(f0)[12:-12, 12:-12, 3].max() < inner.max() * 0.2734375 ** 2 * 1.1


def test_atrous_edge_stops():
    pos, nrm = plane_gbuffer(W, H, t=50.0)
    pos2 = pos.copy()
    pos2[:, W // 2:, 1] = 10.0                        # right half: a parallel plane 10 voxels up
    nrm2 = nrm.copy()
    nrm2[H // 2:, :] = pack_normal(EAST)               # bottom half: perpendicular normal
    a, b = np.array([1.0, 0.0, 0.0], np.float32), np.array([0.0, 0.0, 1.0], np.float32)
    col = np.zeros((HH, HW, 3), np.float32)
    col[:, :HW // 2] = a
    col[:, HW // 2:] = b
    hist = gi_hist(HW, HH, col, 48.0)
    mom = gi_mom(HW, HH, 0.1, 0.2, 0.1)               # wide variance: the luminance stop is loose
    for p, n, sel in [(pos2, nrm, np.s_[4:-4, HW // 2 - 1]), (pos, nrm2, None)]:
        g = _guide_of(p, n)
        out = f16_vals(O.gi_atrous_level(W, H, 0, 1, p, n, mom, hist, g))
        if sel is not None:   # plane step: left of the edge stays colour a
            assert np.allclose(out[sel][:, :3], a, atol=2e-3)
            assert np.allclose(out[4:-4, HW // 2][:, :3], b, atol=2e-3)
    # normal step: colour a above, b below, the halves exchange nothing (pow(dot, 64) ~ 0)
    col2 = np.zeros((HH, HW, 3), np.float32)
    col2[:HH // 2] = a
    col2[HH // 2:] = b
    hist2 = gi_hist(HW, HH, col2, 48.0)
    g = _guide_of(pos, nrm2)
    out = f16_vals(O.gi_atrous_level(W, H, 0, 1, pos, nrm2, mom, hist2, g))
    assert np.all(out[HH // 2 - 1, 4:-4, 2] == 0.0) and np.all(out[HH // 2, 4:-4, 0] == 0.0)
    # luminance stop: zero temporal variance (m2 = m1^2) -> lumPhi = 1e-4, different-luma taps get 0
    col3 = np.where((np.arange(HW) % 2 == 0)[None, :, None], a, 2 * a).astype(np.float32)
    col3 = np.broadcast_to(col3, (HH, HW, 3))
    hist3 = gi_hist(HW, HH, col3, 48.0)
    mom3 = gi_mom(HW, HH, 0.5, 0.25, 0.5)
    g = _guide_of(pos, nrm)
    out = f16_vals(O.gi_atrous_level(W, H, 0, 1, pos, nrm, mom3, hist3, g))
    assert np.allclose(out[4:-4, 4:-4, :3], f16_vals(hist3)[4:-4, 4:-4, :3], rtol=1e-3)


def test_atrous_history_fix():
    """A freshly revealed texel (history 1) among converged ones (48) on the same surface borrows their
    GI: with the fix its output is ~ the neighbours' colour, without it is pulled toward its own sample."""
    pos, nrm = plane_gbuffer(W, H, t=50.0)
    g = _guide_of(pos, nrm)
    y = np.array([0.2, 0.2, 0.2], np.float32)
    col = np.broadcast_to(y, (HH, HW, 3)).copy()
    lens = np.full((HH, HW), 48.0, np.float32)
    col[16, 30] = [3.0, 3.0, 3.0]
    lens[16, 30] = 1.0
    v = np.zeros((HH, HW, 4), np.float32)
    v[..., :3], v[..., 3] = col, lens
    hist = f16_bits(v)
    mom = gi_mom(HW, HH, 0.2, 0.05, 0.2)
    on = f16_vals(O.gi_atrous_level(W, H, 0, 1, pos, nrm, mom, hist, g, histfix=True))
    off = f16_vals(O.gi_atrous_level(W, H, 0, 1, pos, nrm, mom, hist, g, histfix=False))
    assert abs(on[16, 30, 0] - 0.2) < 0.01
    assert off[16, 30, 0] > 0.3


@pytest.mark.parametrize("levels", [1, 3, 5])
def test_atrous_ping_pong(levels):
    """gi_atrous = the levels in sequence with steps 1, 2, 4, ...; level i writes filt[i & 1]."""
    from temporal_helpers import random_frame
    fr = random_frame(W, H, 9)
    h, m, g = O.gi_accumulate(W, H, 3, fr["shift"], fr["gb_pos"], fr["gb_normal"], fr["gb_pos_prev"],
                              fr["gb_normal_prev"], fr["gb_motion"], fr["gb_delta"], fr["indirect"], fr["hist_prev"],
                              fr["mom_prev"])
    f0, f1 = O.gi_atrous(W, H, levels, fr["gb_pos"], fr["gb_normal"], m, h, g)
    cur = None
    for i in range(levels):
        cur = O.gi_atrous_level(W, H, i, 1 << i, fr["gb_pos"], fr["gb_normal"], m, h, g, in_cv=cur)
    assert np.array_equal(cur, f0 if levels % 2 else f1)


# ---- gbuffer.comp motion vectors ---------------------------------------------------------------------
def _cam(pos, yaw, pitch, fov=70.0):
    import vftrace as vf
    p, r, u, f = vf.fly_camera(pos, yaw, pitch)
    return vf.make_pass(p, r, u, f, fov)


def _project(pu, P, W_, H_, prev=True):
    """Float64 pinhole projection of world points into a camera of the pass (uv in [0, 1])."""
    g = lambda k: np.asarray(getattr(pu, ("prev_" if prev else "") + k), np.float64)
    c, r, u, f = g("cam_pos"), g("cam_right"), g("cam_up"), g("cam_fwd")
    thf = pu.prev_tan_half_fov if prev else pu.tan_half_fov
    dp = P - c
    z = dp @ f
    return np.stack([(dp @ r) / (z * thf * W_ / H_) * 0.5 + 0.5, (dp @ u) / (z * thf) * 0.5 + 0.5], -1), z


def test_motion_vectors_by_projection():
    """World points: motion = (projection into last frame's camera) - (sample / res) for a camera that
    moved and turned; a still camera gives ~0; a point behind the previous camera gives 0; the sky moves
    with the rotation only."""
    import vftrace as vf
    W_, H_ = 96, 64
    rng = np.random.default_rng(7)
    cur = _cam([10.0, 20.0, 5.0], 0.6, -0.2)
    prev = _cam([10.6, 19.7, 4.2], 0.55, -0.17)
    vf.set_prev_camera(cur, prev)
    pos = np.zeros((H_, W_, 4), np.float32)
    # every pixel: a point along its own view ray at a random depth, some pixels sky
    ys, xs = np.mgrid[0:H_, 0:W_]
    rd = np.zeros((H_, W_, 3), np.float64)
    f3 = lambda v: (ctypes.c_float * 3)(*v)
    for y in range(H_):
        for x in range(W_):
            o = np.zeros(3, np.float32)
            O.lib().vfo_primary_dir(f3(cur.cam_right), f3(cur.cam_up), f3(cur.cam_fwd), cur.tan_half_fov, W_, H_, x,
                                    y, (ctypes.c_float * 2)(0, 0), o.ctypes.data)
            rd[y, x] = o
    t = rng.uniform(5, 200, (H_, W_))
    P = np.asarray(cur.cam_pos, np.float64) + rd * t[..., None]
    pos[..., :3] = P
    pos[..., 3] = t
    sky = rng.random((H_, W_)) < 0.15
    pos[sky, 3] = -1.0
    mot, dl = O.gbuffer_motion(W_, H_, cur, pos)
    m = np.stack([(mot & 0xffff).astype(np.uint16), (mot >> 16).astype(np.uint16)], -1).view(np.float16).astype(np.float64)
    uv, z = _project(cur, pos[..., :3].astype(np.float64), W_, H_)
    want = uv - np.stack([(xs + 0.5) / W_, (ys + 0.5) / H_], -1)
    ok = ~sky & (z > 0)
    assert np.abs(m[ok] - want[ok]).max() < 2e-3 * (1 + np.abs(want[ok]).max())
    assert np.all(dl == 0)
    # sky: the direction at infinity seen from the previous camera
    uvs, zs = _project(cur, np.asarray(cur.prev_cam_pos, np.float64) + rd * 1e6, W_, H_)
    wsky = uvs - np.stack([(xs + 0.5) / W_, (ys + 0.5) / H_], -1)
    assert np.abs(m[sky] - wsky[sky]).max() < 2e-3
    # still camera: ~0
    vf.set_prev_camera(cur, cur)
    mot, _ = O.gbuffer_motion(W_, H_, cur, pos)
    m0 = np.stack([(mot & 0xffff).astype(np.uint16), (mot >> 16).astype(np.uint16)], -1).view(np.float16)
    assert np.abs(m0.astype(np.float64)).max() < 1e-5
    # a point behind the previous camera: 0
    back = _cam([10.0, 20.0, 5.0], 0.6 + np.pi, 0.2)
    vf.set_prev_camera(cur, back)
    mot, _ = O.gbuffer_motion(W_, H_, cur, pos)
    assert np.all(mot[~sky] == 0)


def test_motion_of_a_moving_body():
    """A body hit carries its voxel by the previous pose: delta = P - Pprev = the body's displacement,
    .a = body + 1, and the motion follows Pprev."""
    import vftrace as vf
    from vftrace import bodies as Bd
    W_, H_ = 32, 16
    cur = _cam([0.0, 0.0, 0.0], 0.0, 0.0)
    r = Bd.rot_y(0.4)
    h = np.zeros(2, O.OBJ_HEADER_DTYPE)
    h[0] = Bd.header(r, np.array([5.0, 1.0, 30.0], np.float32), (8, 8, 8), 0)
    h[1] = Bd.header(r, np.array([-5.0, 1.0, 30.0], np.float32), (8, 8, 8), 1)
    prev_t = np.array([4.0, 1.5, 29.0], np.float32)        # body 0 was 1 voxel left, 0.5 down, 1 nearer
    ph = Bd.header(r, prev_t, (8, 8, 8), 0)
    h[0]["prev_obj_to_world"] = ph["obj_to_world"]
    h[1]["prev_obj_to_world"] = h[1]["obj_to_world"]       # body 1 did not move
    pos = np.zeros((H_, W_, 4), np.float32)
    rng = np.random.default_rng(3)
    pos[..., :3] = rng.uniform([-10, -5, 20], [10, 5, 40], (H_, W_, 3))
    pos[..., 3] = 30.0
    body = np.full((H_, W_), -1, np.int32)
    body[:, :10] = 0
    body[:, 10:20] = 1
    mot, dl = O.gbuffer_motion(W_, H_, cur, pos, body, h)
    d = f16_vals(dl)
    assert np.allclose(d[:, :10, :3], [1.0, -0.5, 1.0], atol=2e-3) and np.all(d[:, :10, 3] == 1.0)
    assert np.allclose(d[:, 10:20, :3], 0.0, atol=1e-5) and np.all(d[:, 10:20, 3] == 2.0)
    assert np.all(dl[:, 20:] == 0)
    m = np.stack([(mot & 0xffff).astype(np.uint16), (mot >> 16).astype(np.uint16)], -1).view(np.float16).astype(np.float64)
    uv, _ = _project(cur, pos[..., :3].astype(np.float64) - np.array([1.0, -0.5, 1.0]), W_, H_)
    ys, xs = np.mgrid[0:H_, 0:W_]
    want = uv - np.stack([(xs + 0.5) / W_, (ys + 0.5) / H_], -1)
    assert np.abs(m[:, :10] - want[:, :10]).max() < 2e-3
