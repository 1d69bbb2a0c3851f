"""Shared helpers of the parity tests: worlds, cameras and ray batches of the benchmark shapes."""
import numpy as np

import vftrace as vf
import oracle as O

_WORLDS = {}


def world(kind=1, **kw):
    """Cached generated world + its packed arrays + the oracle's copy."""
    key = (kind, tuple(sorted(kw.items())))
    if key not in _WORLDS:
        w = vf.GenWorld.generate(kind, **kw)
        a = w.arrays()
        _WORLDS[key] = (w, a, O.OracleWorld(a))
    return _WORLDS[key]


def dense_world(vox, world_win=(0, 0, 0, 0), palette=None):
    w = vf.GenWorld.from_dense(vox, world_win, palette)
    a = w.arrays()
    return w, a, O.OracleWorld(a)


def camera(pos, yaw, pitch, fov=70.0, frame=0):
    pos, right, up, fwd = vf.fly_camera(pos, yaw, pitch)
    pu = vf.make_pass(pos, right, up, fwd, fov, frame)
    cam = {"pos": pos, "right": right, "up": up, "fwd": fwd, "tan_half_fov": pu.tan_half_fov}
    return pu, cam


def config1_camera(frame=0):
    return camera([37.3, 96.4, 41.7], 0.8, -0.35, frame=frame)


def random_rays(n, lo, hi, seed=0, tiny_frac=0.1):
    """Rays with origins uniform in the box [lo, hi) and random directions; a share of them gets
    components below the 1e-5 substitution threshold or exactly zero (axis-parallel rays)."""
    rng = np.random.default_rng(seed)
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    o = lo + rng.random((n, 3)) * (hi - lo)
    d = rng.normal(size=(n, 3))
    k = int(n * tiny_frac)
    if k:
        idx = rng.choice(n, k, replace=False)
        comp = rng.integers(0, 3, k)
        d[idx, comp] = rng.choice([0.0, 3e-6, -3e-6, 1e-5, -1e-5], k)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], 1).astype(np.float32)


def hit_diff(a, b):
    """Indices where two hit-record arrays differ in any byte."""
    ab = a.view(np.uint8).reshape(-1, 20)
    bb = b.view(np.uint8).reshape(-1, 20)
    return np.nonzero((ab != bb).any(1))[0]


def random_bodies(rng, n, lo, hi, max_size=20):
    """n small bodies (crates, spheres, pillars, and blocks of random voxels) with random poses whose
    centres lie in the box [lo, hi)."""
    from vftrace import bodies as B
    vols, poses = [], []
    for k in range(n):
        kind = k % 4
        if kind == 0:
            v = B.crate(int(rng.integers(6, max_size)))
        elif kind == 1:
            v = B.sphere(int(rng.integers(3, max_size // 2)), int(rng.choice([17, 19, 21])))
        elif kind == 2:
            v = B.pillar(int(rng.integers(3, 10)), int(rng.integers(8, max_size + 8)), 8)
        else:
            s = int(rng.integers(1, 3)) * 8
            v = np.where(rng.random((s, s, s)) < 0.3, rng.integers(1, 30, (s, s, s)), 0).astype(np.uint8)
            v[v == 5] = 6   # no water: bodies trace with water clear
        c = lo + rng.random(3) * (np.asarray(hi) - np.asarray(lo))
        poses.append(B.pose(c, v.shape, float(rng.uniform(0, 6.283)), float(rng.uniform(-1.0, 1.0))))
        vols.append(v)
    return vols, poses


def body_world(vox, vols, poses, palette=None, **oracle_kw):
    """A dense world with bodies: (packed arrays incl. the body slots, headers, body grid, oracle world)."""
    from vftrace import bodies as B
    w = vf.GenWorld.from_dense(vox, (0, 0, 0, 0), palette)
    a = w.arrays()
    a2, hd, g = B.attach(a, vols, poses)
    ow = O.OracleWorld(a2, **oracle_kw)
    ow.set_bodies(hd, g)
    return a2, hd, g, ow
