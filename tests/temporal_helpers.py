"""Synthetic G-buffers and histories for the temporal-pass tests (gi.comp accumulation, lighting.comp
history, atrous.comp).  Layouts as include/vf_temporal.h: gb_pos f32 [h, w, 4], gb_normal u32 [h, w]
(RGBA8), gb_motion u32 [h, w] (RG16F bits), gb_delta u16 [h, w, 4] (RGBA16F bits), GI history u16
[hh, hw, 4] (RGBA16F bits), moments f32 [hh, hw, 4]."""
import numpy as np

UP = (128, 255, 128)      # RGBA8 bytes of a +y normal (decodes to (0.0039, 1, 0.0039))
DOWN = (128, 0, 128)
EAST = (255, 128, 128)


def f16_bits(a):
    return np.asarray(a, np.float32).astype(np.float16).view(np.uint16)


def f16_vals(bits):
    return np.asarray(bits, np.uint16).view(np.float16).astype(np.float32)


def pack_normal(b, alpha=255):
    return np.uint32(b[0] | (b[1] << 8) | (b[2] << 16) | (alpha << 24))


def plane_gbuffer(w, h, y=0.0, t=10.0, normal=UP, scale=0.25):
    """Every pixel on the plane at height y: P = (x*scale, y, row*scale), hit distance t."""
    pos = np.zeros((h, w, 4), np.float32)
    xs = np.arange(w, dtype=np.float32) * np.float32(scale)
    zs = np.arange(h, dtype=np.float32) * np.float32(scale)
    pos[:, :, 0] = xs[None, :]
    pos[:, :, 1] = y
    pos[:, :, 2] = zs[:, None]
    pos[:, :, 3] = t
    nrm = np.full((h, w), pack_normal(normal), np.uint32)
    return pos, nrm


def motion_bits(w, h, mx, my):
    m = np.zeros((h, w), np.uint32)
    bx = int(f16_bits(mx))
    by = int(f16_bits(my))
    m[:] = np.uint32(bx | (by << 16))
    return m


def zero_delta(w, h):
    return np.zeros((h, w, 4), np.uint16)


def gi_hist(hw, hh, rgb, count):
    """GI history (RGBA16F bits) with the given colour field [hh, hw, 3] (or a colour) and length."""
    v = np.zeros((hh, hw, 4), np.float32)
    v[..., :3] = rgb
    v[..., 3] = count
    return f16_bits(v)


def gi_mom(hw, hh, m1, m2, fast):
    m = np.zeros((hh, hw, 4), np.float32)
    m[..., 0], m[..., 1], m[..., 2] = m1, m2, fast
    return m


def random_frame(w, h, seed, sky_frac=0.1, max_count=60.0):
    """A random but well-formed pair of frames: positions on a few planes with jitter, RGBA8 normals,
    sky holes, motion vectors of up to ~2 texels (a quarter of them 0), small deltas, random histories."""
    rng = np.random.default_rng(seed)
    hw, hh = (w + 1) // 2, (h + 1) // 2
    out = {}
    for key in ("cur", "prev"):
        pos = np.zeros((h, w, 4), np.float32)
        pos[..., 0] = rng.uniform(0, 64, (h, w))
        pos[..., 1] = rng.choice([0.0, 2.0, 8.0], (h, w)) + rng.normal(0, 0.2, (h, w))
        pos[..., 2] = rng.uniform(0, 64, (h, w))
        pos[..., 3] = rng.uniform(0.5, 300, (h, w))
        sky = rng.random((h, w)) < sky_frac
        pos[sky, 3] = -1.0
        if key == "prev":
            pos[rng.random((h, w)) < 0.02, 3] = 0.0     # w == 0: valid now, rejected as history
        nb = np.where(rng.random((h, w)) < 0.8, 255, rng.integers(0, 256, (h, w)))
        n = (rng.integers(96, 160, (h, w)) | (nb << 8) | (rng.integers(96, 160, (h, w)) << 16) |
             (rng.integers(0, 256, (h, w)) << 24)).astype(np.uint32)
        n[sky] = 0
        out[key] = (pos, n)
    cur, prev = out["cur"], out["prev"]
    # make history taps plausible: previous positions near the current ones for most pixels
    near = rng.random((h, w)) < 0.7
    prev[0][near, :3] = cur[0][near, :3] + rng.normal(0, 0.05, (int(near.sum()), 3)).astype(np.float32)
    prev[1][near] = cur[1][near]
    mot = rng.normal(0, 1.5, (h, w, 2)) / np.array([w, h])
    mot[rng.random((h, w)) < 0.25] = 0.0
    mot[rng.random((h, w)) < 0.02] = rng.uniform(-2, 2, 2)        # far off screen
    mb = f16_bits(mot).astype(np.uint32)
    motion = (mb[..., 0] | (mb[..., 1] << 16)).astype(np.uint32)
    dl = np.zeros((h, w, 4), np.float32)
    moving = rng.random((h, w)) < 0.1
    dl[moving, :3] = rng.normal(0, 0.3, (int(moving.sum()), 3))
    delta = f16_bits(dl)
    hist = np.zeros((hh, hw, 4), np.float32)
    hist[..., :3] = rng.gamma(1.0, 0.5, (hh, hw, 3))
    hist[..., 3] = np.floor(rng.uniform(0, max_count, (hh, hw)))
    m1 = (hist[..., :3] @ np.array([0.2126, 0.7152, 0.0722])).astype(np.float32)
    mom = np.zeros((hh, hw, 4), np.float32)
    mom[..., 0] = m1
    mom[..., 1] = m1 * m1 + rng.gamma(1.0, 0.05, (hh, hw))
    mom[..., 2] = m1 * rng.uniform(0.5, 1.5, (hh, hw))
    indirect = np.zeros((hh, hw, 4), np.float32)
    indirect[..., :3] = rng.gamma(0.7, 0.6, (hh, hw, 3))
    indirect[rng.random((hh, hw)) < 0.01, :3] = 40.0                 # fireflies
    indirect[..., 3] = 1.0
    shadow = (rng.random((h, w)) < 0.6).astype(np.float32)
    shadow[rng.random((h, w)) < 0.1] = 0.5
    lhist = np.zeros((h, w, 4), np.float32)
    lhist[..., 0] = rng.random((h, w))
    lhist[..., 1] = lhist[..., 0] ** 2 + rng.gamma(1.0, 0.05, (h, w))
    lhist[..., 2] = np.floor(rng.uniform(0, max_count, (h, w)))
    lhist[..., 3] = rng.random((h, w))
    shift = rng.normal(0, 0.01, 3).astype(np.float32)
    return dict(gb_pos=cur[0], gb_normal=cur[1], gb_pos_prev=prev[0], gb_normal_prev=prev[1], gb_motion=motion,
                gb_delta=delta, hist_prev=f16_bits(hist), mom_prev=mom, indirect=indirect, shadow=shadow,
                light_hist_prev=lhist, shift=shift)
