"""The reference's world save format (.vfworld, src/serialize.rs) for the synthetic worlds.

The reference's own loader (`--load`, src/app/mod.rs:34-45) reads these files, so the synthetic worlds
of the benchmark can be rendered by the reference itself.  Format (serialize.rs:1-13), all
little-endian:

    [magic u32 0x56464653 "VFFS"][version u32 = 6]
    [info size u32][info: name (u32 len + bytes), saved_at i64, thumb w u16, h u16,
                    (thumb: u32 len + zlib(RGBA8)) ]
    [payload raw size u64][chunks u32]{[raw u32][comp u32][zlib(<= 4 MiB of payload)]}...
    payload: [256 u32][256 materials x 12 f32]
             [kind u32: 0 bounded volume | 1 endless recipe]
             volume: [nx ny nz i32][grid RLE: total u32, (count u32, value u32)..., 0xFFFFFFFF]
                     [n u32][n bricks x 512 B]   (detailed grid entries 256 + i in grid order)
             [bodies u32]{volume, rotation 9 f32 (columns), t, vel, ang_vel 3 f32 each, glow f32}
             [has_camera u32][cam pos 3 f32][yaw f32][pitch f32][cvars u32 {str, str}]
             [next import slot u32][water cells u32 {x y z i32, volume f32}]

write_vfworld follows write_payload / write_volume / save_scene (serialize.rs:319-347, 441-591);
read_vfworld follows load_scene's parsing and checks (serialize.rs:612-872) for versions 5 and 6 (the
palette migration of older saves is not restated: such files are refused).
"""
import struct
import zlib

import numpy as np

MAGIC = 0x56464653
VERSION = 6
CHUNK = 4 << 20
MAX_BODY_DIM = 1024
MAX_DEFLATE_RATIO = 1100


class VfWorldError(ValueError):
    """A file that is not a readable save (kind: 'NotASave', 'Newer', 'WorldSize', 'Corrupt')."""

    def __init__(self, kind, msg=""):
        super().__init__(f"{kind}: {msg}" if msg else kind)
        self.kind = kind


# ---- writing --------------------------------------------------------------------------------------
def _str(s):
    b = s.encode()
    return struct.pack("<I", len(b)) + b


def _grid_rle(g):
    """(count, value) runs of the grid words, a 0xFFFFFFFF terminator (write_grid_rle)."""
    g = np.ascontiguousarray(g, dtype=np.uint32).reshape(-1)
    out = [struct.pack("<I", g.size)]
    if g.size:
        starts = np.concatenate([[0], np.nonzero(np.diff(g.astype(np.int64)))[0] + 1])
        counts = np.diff(np.concatenate([starts, [g.size]]))
        pairs = np.empty((starts.size, 2), np.uint32)
        pairs[:, 0] = counts
        pairs[:, 1] = g[starts]
        out.append(pairs.astype("<u4").tobytes())
    out.append(struct.pack("<I", 0xFFFFFFFF))
    return b"".join(out)


def volume_bytes(dims, grid, atlas):
    """write_volume: dims (nx, ny, nz) voxels, grid u32 (0, 1-255, 256 + slot), atlas slots of 512 B.
    Detailed entries are renumbered 256 + i in grid order and their bricks follow in that order."""
    g = np.ascontiguousarray(grid, dtype=np.uint32).reshape(-1)
    atlas = np.ascontiguousarray(atlas, dtype=np.uint8).reshape(-1, 512)
    det = g >= 256
    slots = (g[det] - 256).astype(np.int64)
    if slots.size and slots.max() >= atlas.shape[0]:
        raise ValueError("grid names a slot beyond the atlas")
    g2 = g.copy()
    g2[det] = 256 + np.arange(slots.size, dtype=np.uint32)
    head = struct.pack("<iii", *[int(d) for d in dims])
    return head + _grid_rle(g2) + struct.pack("<I", slots.size) + atlas[slots].tobytes()


def _material_bytes(palette):
    p = np.ascontiguousarray(palette, dtype=np.float32).reshape(256, 12)
    return struct.pack("<I", 256) + p.astype("<f4").tobytes()


def _chunks(payload, level):
    """The payload in CHUNK-byte pieces (the last one shorter), each deflated on its own."""
    n = len(payload)
    pieces = [payload[i:i + CHUNK] for i in range(0, n, CHUNK)] or [b""]
    out = [struct.pack("<QI", n, len(pieces))]
    for p in pieces:
        c = zlib.compress(p, level)
        out.append(struct.pack("<II", len(p), len(c)))
        out.append(c)
    return b"".join(out)


def write_vfworld(path, arrays, bodies=(), name="", saved_at=0, thumb=None, camera=None, cvars=(), water=(),
                  import_slot=64, level=2):
    """A bounded world save.
    arrays: world_bricks (bx, by, bz) with bx, bz multiples of 16 and by of 1 (world voxel dims multiples of
      128 / 8, at least 256 x 128 x 256), grid, atlas, palette (256 x 12 float32).
    bodies: dicts {dims, grid, atlas, r (3x3, body -> world rotation), t, vel, ang_vel, glow}.
    thumb: RGBA8 uint8 array [h, w, 4] or None.  camera: (pos[3], yaw, pitch) or None.
    cvars: (name, value) pairs.  water: (x, y, z, volume) cells."""
    wb = [int(x) for x in arrays["world_bricks"]]
    dims = (wb[0] * 8, wb[1] * 8, wb[2] * 8)
    if dims[0] % 128 or dims[2] % 128 or dims[1] % 8 or dims[0] < 256 or dims[2] < 256 or dims[1] < 128 or \
            dims[0] > 16384 or dims[2] > 16384 or dims[1] > 2048:
        raise ValueError(f"world size {dims} is not one the reference can create")
    pay = [_material_bytes(arrays["palette"]), struct.pack("<I", 0),
           volume_bytes(dims, arrays["grid"], arrays["atlas"]), struct.pack("<I", len(bodies))]
    for b in bodies:
        bd = tuple(int(x) for x in b["dims"])
        if any(d <= 0 or d > MAX_BODY_DIM for d in bd):
            raise ValueError("body dims out of range")
        r = np.asarray(b["r"], np.float32).reshape(3, 3)
        pay.append(volume_bytes(bd, b["grid"], b["atlas"]))
        pay.append(r.T.astype("<f4").tobytes())   # columns (Mat3::to_cols_array)
        for k in ("t", "vel", "ang_vel"):
            pay.append(np.asarray(b.get(k, (0.0, 0.0, 0.0)), "<f4").reshape(3).tobytes())
        pay.append(struct.pack("<f", float(b.get("glow", 1.0))))
    pos, yaw, pitch = camera if camera is not None else ((0.0, 0.0, 0.0), 0.0, 0.0)
    pay.append(struct.pack("<I3fff", int(camera is not None), *[float(x) for x in pos], float(yaw), float(pitch)))
    pay.append(struct.pack("<I", len(cvars)) + b"".join(_str(k) + _str(v) for k, v in cvars))
    pay.append(struct.pack("<II", int(import_slot), len(water)))
    pay.append(b"".join(struct.pack("<iiif", int(x), int(y), int(z), float(v)) for x, y, z, v in water))
    payload = b"".join(pay)

    info = [_str(name), struct.pack("<q", int(saved_at))]
    if thumb is not None and np.asarray(thumb).size:
        t = np.ascontiguousarray(thumb, dtype=np.uint8)
        h, w = t.shape[:2]
        c = zlib.compress(t.tobytes(), 6)
        info.append(struct.pack("<HHI", w, h, len(c)) + c)
    else:
        info.append(struct.pack("<HH", 0, 0))
    info = b"".join(info)
    data = struct.pack("<III", MAGIC, VERSION, len(info)) + info + _chunks(payload, level)
    with open(path, "wb") as f:
        f.write(data)
    return len(data)


# ---- reading --------------------------------------------------------------------------------------
class _Reader:
    """Reads past the end raise Corrupt."""

    def __init__(self, b):
        self.b = memoryview(b)
        self.at = 0

    def take(self, n):
        if n < 0 or self.at + n > len(self.b):
            raise VfWorldError("Corrupt", "truncated")
        v = self.b[self.at:self.at + n]
        self.at += n
        return v

    def unpack(self, fmt):
        return struct.unpack(fmt, self.take(struct.calcsize(fmt)))

    def u32(self):
        return self.unpack("<I")[0]

    def str(self, max_len):
        n = self.u32()
        if n > max_len:
            raise VfWorldError("Corrupt", "string too long")
        return bytes(self.take(n)).decode(errors="replace")


def _inflate_exact(src, n):
    d = zlib.decompressobj()
    try:
        out = d.decompress(bytes(src), n)
    except zlib.error as e:
        raise VfWorldError("Corrupt", str(e))
    if len(out) != n or d.unconsumed_tail or not d.eof:
        raise VfWorldError("Corrupt", "deflate stream size")
    return out


def _read_volume(r, world):
    nx, ny, nz = r.unpack("<iii")
    if nx <= 0 or ny <= 0 or nz <= 0:
        raise VfWorldError("Corrupt", "volume dims")
    if world:
        if nx % 128 or nz % 128 or ny % 8 or nx < 256 or nz < 256 or ny < 128 or nx > 16384 or nz > 16384 or ny > 2048:
            raise VfWorldError("WorldSize", f"{nx}x{ny}x{nz}")
    elif nx > MAX_BODY_DIM or ny > MAX_BODY_DIM or nz > MAX_BODY_DIM:
        raise VfWorldError("Corrupt", "body dims")
    bx, by, bz = (nx + 7) // 8, (ny + 7) // 8, (nz + 7) // 8
    total = r.u32()
    if total > 1 << 30 or total != bx * by * bz:
        raise VfWorldError("Corrupt", "grid size")
    runs = []
    filled = 0
    while True:
        count = r.u32()
        if count == 0xFFFFFFFF:
            break
        v = r.u32()
        if filled + count > total:
            raise VfWorldError("Corrupt", "grid runs")
        runs.append((count, v))
        filled += count
    if filled != total:
        raise VfWorldError("Corrupt", "grid runs")
    grid = np.repeat(np.array([v for _, v in runs], np.uint32), np.array([c for c, _ in runs], np.int64)) \
        if runs else np.zeros(0, np.uint32)
    n = r.u32()
    if n > 1 << 24:
        raise VfWorldError("Corrupt", "brick count")
    bricks = np.frombuffer(bytes(r.take(n * 512)), np.uint8).reshape(n, 512)
    if np.any((grid >= 256) & (grid - 256 >= n)):
        raise VfWorldError("Corrupt", "grid names a missing brick")
    return {"dims": (nx, ny, nz), "grid": grid, "bricks": bricks}


def read_vfworld(path):
    """load_scene for versions 5 and 6: {version, info, palette, volume (or None), endless, bodies, state,
    import_slot}; raises VfWorldError like the reference's SceneError."""
    with open(path, "rb") as f:
        file = f.read()
    if len(file) < 8:
        raise VfWorldError("NotASave")
    magic, ver = struct.unpack("<II", file[:8])
    if magic != MAGIC:
        raise VfWorldError("NotASave")
    if ver > VERSION:
        raise VfWorldError("Newer")
    if ver < 5:
        raise VfWorldError("Corrupt", f"version {ver}: saves before v5 need the reference's palette migration")
    r = _Reader(file)
    r.take(8)
    n = r.u32()
    ir = _Reader(r.take(n))
    info = {"name": ir.str(256), "saved_at": ir.unpack("<q")[0]}
    tw, th = ir.unpack("<HH")
    info["thumb"] = None
    if tw and th:
        c = ir.u32()
        src = ir.take(c)
        if tw <= 1024 and th <= 1024 and c:
            try:
                info["thumb"] = np.frombuffer(_inflate_exact(src, tw * th * 4), np.uint8).reshape(th, tw, 4)
            except VfWorldError:   # a broken thumbnail is not a broken save
                info["thumb"] = None
    raw, nc = r.unpack("<QI")
    rest = len(file) - r.at
    if raw > 1 << 34 or raw > rest * MAX_DEFLATE_RATIO or nc == 0 or nc > raw // CHUNK + 2:
        raise VfWorldError("Corrupt", "payload header")
    parts, got = [], 0
    for _ in range(nc):
        rn, cn = r.unpack("<II")
        src = r.take(cn)
        if rn > CHUNK or got + rn > raw:
            raise VfWorldError("Corrupt", "chunk size")
        parts.append(_inflate_exact(src, rn))
        got += rn
    if got != raw:
        raise VfWorldError("Corrupt", "payload size")
    p = _Reader(b"".join(parts))
    if p.u32() != 256:
        raise VfWorldError("Corrupt", "palette")
    palette = np.frombuffer(bytes(p.take(256 * 48)), "<f4").reshape(256, 12).astype(np.float32)
    kind = p.u32()
    if kind > 1:
        raise VfWorldError("Corrupt", "world kind")
    endless, volume = {"endless": False}, None
    if kind == 1:
        seed, flat, height, water_h, nl = p.unpack("<IIiiI")
        if nl > 256 or not 1 <= height <= 4096 or not 0 <= water_h <= 4096:
            raise VfWorldError("Corrupt", "flat world")
        layers = [p.unpack("<Ii") for _ in range(nl)]
        gen, ny = p.unpack("<Ii")
        scale = p.u32() if ver >= 6 else 1
        endless = {"endless": True, "seed": seed, "flat": bool(flat), "height": height, "water": water_h,
                   "layers": layers, "generator": gen, "ny": ny, "scale": scale}
    else:
        volume = _read_volume(p, True)
    nb = p.u32()
    if nb > 100000:
        raise VfWorldError("Corrupt", "body count")
    bodies = []
    for _ in range(nb):
        vol = _read_volume(p, False)
        cols = np.array(p.unpack("<9f"), np.float32).reshape(3, 3)
        rmat = cols.T
        t, vel, ang = (np.array(p.unpack("<3f"), np.float32) for _ in range(3))
        glow = p.unpack("<f")[0]
        finite = np.isfinite(rmat).all() and np.isfinite(t).all() and np.isfinite(vel).all() and np.isfinite(ang).all()
        if not finite or not np.isfinite(glow) or np.abs(rmat.T @ rmat - np.eye(3)).max() > 1e-3:
            raise VfWorldError("Corrupt", "body pose")
        bodies.append({**vol, "r": rmat, "t": t, "vel": vel, "ang_vel": ang, "glow": float(min(max(glow, 0.0), 1.0))})
    has_cam = p.u32() != 0
    cam = np.array(p.unpack("<3f"), np.float32)
    yaw, pitch = p.unpack("<ff")
    ncv = p.u32()
    if ncv > 4096:
        raise VfWorldError("Corrupt", "cvars")
    cvars = [(p.str(128), p.str(1024)) for _ in range(ncv)]
    import_slot = p.u32()
    nw = p.u32()
    if nw > 64 << 20:
        raise VfWorldError("Corrupt", "water cells")
    water = [p.unpack("<iiif") for _ in range(nw)]
    state = {"has_camera": has_cam and bool(np.isfinite(cam).all()), "cam_pos": cam, "cam_yaw": yaw,
             "cam_pitch": pitch, "cvars": cvars, "water": water}
    return {"version": ver, "info": info, "palette": palette, "volume": volume, "endless": endless, "bodies": bodies,
            "state": state, "import_slot": import_slot}


def arrays_from_save(save):
    """A bounded save's world as vf_world_upload arrays (grid entries 256 + i name brick i of the atlas)."""
    v = save["volume"]
    if v is None:
        raise ValueError("an endless save holds no voxels")
    nx, ny, nz = v["dims"]
    return {"world_bricks": ((nx + 7) // 8, (ny + 7) // 8, (nz + 7) // 8), "world_win": (0, 0, 0, 0),
            "grid": v["grid"], "atlas": v["bricks"].reshape(-1), "n_slots": v["bricks"].shape[0], "masks": None,
            "coarse1": None, "coarse2": None, "palette": save["palette"]}


def ring_to_bounded(arrays, chunks):
    """An endless-shape world (a ring grid of 16 x by x 16-brick chunks, `chunks` x `chunks` of them loaded
    around chunk (0, 0)) as a bounded world of the loaded chunks: (arrays, voxel offset added to absolute
    x and z).  The reference saves endless worlds as a recipe plus region files of its own generator's
    chunks; a synthetic world goes out as the bounded world it is."""
    bx, by, bz = (int(x) for x in arrays["world_bricks"])
    half = chunks // 2
    lo = -half * 16
    n = chunks * 16
    g = np.asarray(arrays["grid"], np.uint32).reshape(bz, by, bx)
    xs = (np.arange(n) + lo) & (bx - 1)
    zs = (np.arange(n) + lo) & (bz - 1)
    out = dict(arrays)
    out["grid"] = np.ascontiguousarray(g[np.ix_(zs, np.arange(by), xs)]).reshape(-1)
    out["world_bricks"] = (n, by, n)
    out["world_win"] = (0, 0, 0, 0)
    out["coarse1"] = out["coarse2"] = out["masks"] = None
    return out, -lo * 8


def main(argv=None):
    """python -m vftrace.vfworld OUT.vfworld [--kind 1|2] [--seed S] [--chunks C] [--trees T]: a benchmark
    world as a save the reference can load."""
    import argparse
    from . import GenWorld
    ap = argparse.ArgumentParser()
    ap.add_argument("out")
    ap.add_argument("--kind", type=int, default=1)
    ap.add_argument("--seed", type=int, default=1337)
    ap.add_argument("--chunks", type=int, default=13)
    ap.add_argument("--trees", type=int, default=400)
    a = ap.parse_args(argv)
    if a.kind == 1:
        arrays = GenWorld.generate(1, seed=a.seed).arrays()
        off = 0
    else:
        arrays, off = ring_to_bounded(GenWorld.generate(2, seed=a.seed, chunks=a.chunks, trees_per_chunk=a.trees).arrays(),
                                      a.chunks)
    n = write_vfworld(a.out, arrays, name=f"synthetic kind {a.kind} seed {a.seed}")
    print(f"{a.out}: {n} bytes; world {tuple(arrays['world_bricks'])} bricks, absolute x/z shifted by {off} voxels")


if __name__ == "__main__":
    main()
