"""Synthetic voxel bodies for the traversal workloads, packed the way GpuWorld::upload packs bodies.

The reference's demo scene has seven kinematic bodies around the world centre (src/objects.rs:275-304:
two crates, three spheres, two pillars, spinning about y and tilted about x).  This module builds
bodies of those shapes as dense voxel volumes, poses them (rotY(spin) * rotX(tilt) about the volume's
centre, src/objects.rs:92-98), and packs them: body bricks numbered into the world's atlas after its
own slots (gpu_world.rs:274-293), body grids concatenated (gpu_world.rs:276-297), 240-byte headers
(gpu_world.rs:513-539).
"""
import numpy as np

from .abi import OBJ_HEADER_DTYPE


def _ceil8(n):
    return ((int(n) + 7) // 8) * 8


def crate(size, wood=6, dark=2):
    """A size^3 wood box with plank stripes (y % 8 == 0, or x % 8 == 0 and z % 8 == 0 dark)."""
    n = _ceil8(size)
    v = np.zeros((n, n, n), np.uint8)
    z, y, x = np.meshgrid(np.arange(size), np.arange(size), np.arange(size), indexing="ij")
    stripe = (y % 8 == 0) | ((x % 8 == 0) & (z % 8 == 0))
    v[:size, :size, :size] = np.where(stripe, dark, wood)
    return v


def sphere(radius, mat):
    """A ball of radius r centred in a volume of ceil8(2r + 2) voxels (voxel centres within r)."""
    n = ((radius * 2 + 2 + 7) // 8) * 8
    c = np.float32(n * 0.5)
    r = np.float32(radius)
    idx = np.arange(n, dtype=np.float32) + np.float32(0.5) - c
    dz, dy, dx = np.meshgrid(idx, idx, idx, indexing="ij")
    rem = r * r - dz * dz - dy * dy
    inside = (rem >= 0) & (dx * dx <= rem)
    return np.where(inside, np.uint8(mat), np.uint8(0)).astype(np.uint8)


def pillar(w, h, mat):
    """A w x h x w column."""
    v = np.zeros((_ceil8(w), _ceil8(h), _ceil8(w)), np.uint8)
    v[:w, :h, :w] = mat
    return v


def rot_y(a):
    c, s = np.float32(np.cos(a)), np.float32(np.sin(a))
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]], np.float32)


def rot_x(a):
    c, s = np.float32(np.cos(a)), np.float32(np.sin(a))
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]], np.float32)


def pose(center, vol_shape, spin, tilt):
    """(r, t): local voxel coordinates -> world, rotating about the volume's centre placed at `center`."""
    r = (rot_y(spin) @ rot_x(tilt)).astype(np.float32)
    nz, ny, nx = vol_shape
    lc = np.array([nx, ny, nz], np.float32) * np.float32(0.5)
    t = np.asarray(center, np.float32) - r @ lc
    return r, t.astype(np.float32)


def header(r, t, vol_shape, grid_offset):
    """ObjHeader: worldToObj / objToWorld (column-major), dims in bricks + grid offset, world box."""
    nz, ny, nx = vol_shape
    h = np.zeros((), OBJ_HEADER_DTYPE)
    o2w = np.eye(4, dtype=np.float32)
    o2w[:3, :3] = r
    o2w[:3, 3] = t
    rt = r.T.astype(np.float32)
    w2o = np.eye(4, dtype=np.float32)
    w2o[:3, :3] = rt
    w2o[:3, 3] = -(rt @ t)
    h["obj_to_world"] = o2w.T.reshape(-1)   # column-major
    h["world_to_obj"] = w2o.T.reshape(-1)
    h["prev_obj_to_world"] = h["obj_to_world"]
    h["dims_bricks"] = [nx // 8, ny // 8, nz // 8, grid_offset]
    corners = np.array([[(i & 1) * nx, ((i >> 1) & 1) * ny, ((i >> 2) & 1) * nz] for i in range(8)], np.float32)
    w = corners @ r.T + t
    h["aabb_min"][:3] = w.min(axis=0)
    h["aabb_max"][:3] = w.max(axis=0)
    h["aabb_min"][3] = 1.0   # glow integrity (not read by the traversal)
    return h


def pack(vols, poses, slot_base):
    """Bodies -> (headers, grid, slot bricks [k, 512]).  Bricks: all air -> 0, one material -> that value,
    else 256 + slot with slots numbered from slot_base in body, then grid order."""
    headers = np.zeros(len(vols), OBJ_HEADER_DTYPE)
    grid, bricks = [], []
    for i, (v, (r, t)) in enumerate(zip(vols, poses)):
        nz, ny, nx = v.shape
        headers[i] = header(r, t, v.shape, len(grid))
        b = v.reshape(nz // 8, 8, ny // 8, 8, nx // 8, 8).transpose(0, 2, 4, 1, 3, 5).reshape(-1, 512)
        for brick in b:
            if not brick.any():
                grid.append(0)
            elif (brick == brick[0]).all():
                grid.append(int(brick[0]))
            else:
                grid.append(256 + slot_base + len(bricks))
                bricks.append(brick)
    slots = np.stack(bricks).astype(np.uint8) if bricks else np.zeros((0, 512), np.uint8)
    return headers, np.asarray(grid, np.uint32), slots


def attach(arrays, vols, poses):
    """A world's arrays with the bodies' bricks appended to its atlas (the masks are left to the library)
    -> (arrays, headers, grid)."""
    base = int(np.asarray(arrays["atlas"]).size // 512)
    headers, grid, slots = pack(vols, poses, base)
    out = dict(arrays)
    out["atlas"] = np.concatenate([np.asarray(arrays["atlas"], np.uint8).reshape(-1), slots.reshape(-1)])
    out["n_slots"] = out["atlas"].size // 512
    out["masks"] = None
    return out, headers, grid


def demo_bodies(center, ground, t=0.0):
    """The seven-body layout of the reference's demo scene around world point `center` (x, z): the shapes,
    offsets, hover heights, spin speeds and tilts of src/objects.rs:292-298, posed at time t.
    ground(x, z) -> terrain height there.  Returns (vols, poses)."""
    specs = [  # (volume, dx, dz, hover, spin, tilt, bob amplitude)
        (crate(24), -165.0, -97.0, 40.0, 0.6, 0.0, 4.0),
        (crate(36), 180.0, 18.0, 55.0, -0.35, 0.4, 6.0),
        (sphere(20, 17), -97.0, 180.0, 60.0, 1.2, 0.0, 5.0),
        (sphere(28, 21), 76.0, 225.0, 100.0, 0.8, 0.3, 8.0),
        (sphere(36, 19), 225.0, -165.0, 130.0, -0.5, 0.2, 10.0),
        (pillar(24, 112, 3), -188.0, 76.0, 90.0, 0.9, 0.5, 5.0),
        (pillar(16, 60, 8), 18.0, -188.0, 70.0, -1.4, 0.9, 7.0),
    ]
    vols, poses = [], []
    for k, (v, dx, dz, hover, spin, tilt, bob) in enumerate(specs):
        x, z = center[0] + dx, center[1] + dz
        y = ground(x, z) + hover + bob * np.sin(t * 0.8 + k * 1.3)
        vols.append(v)
        poses.append(pose((x, y, z), v.shape, spin * t, tilt))
    return vols, poses
