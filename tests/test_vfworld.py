"""The .vfworld writer and reader (vftrace/vfworld.py) against the reference's save format
(src/serialize.rs): the reference's own round-trip case (serialize.rs:953-1006), byte-level layout,
payload chunking, damage detection (serialize.rs:1055-1076), and generated benchmark worlds that trace
identically after a save and load."""
import struct
import zlib

import numpy as np
import pytest

import oracle as O
import vftrace as vf
from vftrace import bodies as B
from vftrace import vfworld as V
from helpers import random_rays


def _dense(save_or_arrays):
    a = save_or_arrays
    bx, by, bz = a["world_bricks"]
    g = np.asarray(a["grid"], np.uint32).reshape(bz, by, bx)
    atlas = np.asarray(a["atlas"], np.uint8).reshape(-1, 512)
    out = np.zeros((bz * 8, by * 8, bx * 8), np.uint8)
    for z in range(bz):
        for y in range(by):
            for x in range(bx):
                v = g[z, y, x]
                if v == 0:
                    continue
                blk = atlas[v - 256].reshape(8, 8, 8) if v >= 256 else np.full((8, 8, 8), v, np.uint8)
                out[z * 8:z * 8 + 8, y * 8:y * 8 + 8, x * 8:x * 8 + 8] = blk
    return out


def _sample(tmp_path):
    """serialize.rs:953-966: a 256x128x256 flat world with a sphere and a custom colour, one rotated crate."""
    vox = np.zeros((256, 128, 256), np.uint8)
    vox[:, :20, :] = 2                                    # stone up to y 20
    z, y, x = np.ogrid[:256, :128, :256]
    vox[((x + 0.5 - 100) ** 2 + (y + 0.5 - 30) ** 2 + (z + 0.5 - 100) ** 2) <= 81] = 11
    w = vf.GenWorld.from_dense(vox)
    a = w.arrays()
    a["palette"] = a["palette"].copy()
    a["palette"][40] = [0.5, 0.25, 0.125, 0.0, 1.0, 0.1, 0.0, 0.0, 0.9, 0.0, 0.04, 0.0]
    crate = B.crate(16)
    r = B.rot_y(0.7)
    nz_, ny_, nx_ = crate.shape
    _, g, slots = B.pack([crate], [(r, np.zeros(3, np.float32))], 0)
    body = {"dims": (nx_, ny_, nz_), "grid": g, "atlas": slots, "r": r, "t": (50.0, 40.0, 60.0), "vel": (1.0, 2.0, 3.0),
            "ang_vel": (0.0, 0.0, 0.0), "glow": 0.5}
    return vox, a, body, crate


def test_reference_round_trip_case(tmp_path):
    vox, a, body, crate = _sample(tmp_path)
    path = tmp_path / "w.vfworld"
    thumb = np.array([[[1, 2, 3, 255], [4, 5, 6, 255]]], np.uint8)
    V.write_vfworld(path, a, [body], name="Test", saved_at=1234, thumb=thumb, camera=((1.0, 2.0, 3.0), 0.5, -0.25),
                    cvars=[("r.time.hour", "13.5")], water=[(5, 20, 7, 0.4)], import_slot=41)
    s = V.read_vfworld(path)
    assert s["version"] == 6 and s["volume"]["dims"] == (256, 128, 256)
    back = V.arrays_from_save(s)
    d = _dense(back)
    for p in [(100, 30, 100), (0, 0, 0), (10, 21, 10), (200, 100, 200)]:
        assert d[p[2], p[1], p[0]] == vox[p[2], p[1], p[0]]
    assert np.array_equal(d, vox)
    assert np.array_equal(s["palette"][40], a["palette"][40])
    assert s["info"]["name"] == "Test" and s["info"]["saved_at"] == 1234
    assert np.array_equal(s["info"]["thumb"], thumb)
    o = s["bodies"][0]
    assert np.abs(o["t"] - np.array([50, 40, 60])).max() < 1e-6 and np.abs(o["vel"] - np.array([1, 2, 3])).max() < 1e-6
    assert np.array_equal(o["r"], body["r"]) and o["glow"] == 0.5
    ob = _dense({"world_bricks": tuple(x // 8 for x in o["dims"]), "grid": o["grid"], "atlas": o["bricks"]})
    assert np.array_equal(ob, crate)
    st = s["state"]
    assert st["has_camera"] and np.allclose(st["cam_pos"], [1, 2, 3]) and st["cam_yaw"] == 0.5 and st["cam_pitch"] == -0.25
    assert st["cvars"] == [("r.time.hour", "13.5")]
    assert st["water"][0][:3] == (5, 20, 7) and abs(st["water"][0][3] - 0.4) < 1e-7
    assert s["import_slot"] == 41


def test_byte_layout(tmp_path):
    """Header, info block, chunk table and the payload's first sections at their serialize.rs offsets."""
    _, a, _, _ = _sample(tmp_path)
    path = tmp_path / "w.vfworld"
    V.write_vfworld(path, a, name="ab", saved_at=-5)
    b = path.read_bytes()
    assert b[:4] == b"SFFV" and struct.unpack("<I", b[4:8])[0] == 6
    (n,) = struct.unpack("<I", b[8:12])
    assert n == 4 + 2 + 8 + 4 and b[12:12 + n] == struct.pack("<I", 2) + b"ab" + struct.pack("<qHH", -5, 0, 0)
    at = 12 + n
    raw, nc = struct.unpack("<QI", b[at:at + 12])
    rn, cn = struct.unpack("<II", b[at + 12:at + 20])
    assert nc == 1 and rn == raw
    payload = zlib.decompress(b[at + 20:at + 20 + cn])
    assert len(payload) == raw
    assert struct.unpack("<I", payload[:4])[0] == 256
    assert payload[4:4 + 256 * 48] == np.asarray(a["palette"], "<f4").tobytes()
    off = 4 + 256 * 48
    assert struct.unpack("<Iiii", payload[off:off + 16]) == (0, 256, 128, 256)   # bounded, dims
    assert struct.unpack("<I", payload[off + 16:off + 20])[0] == 32 * 16 * 32      # grid words


def test_grid_rle_and_brick_renumbering():
    """write_grid_rle / write_volume on a hand-checked grid: runs of equal words, detailed entries
    renumbered 256 + i in grid order, their bricks in that order."""
    atlas = np.zeros((3, 512), np.uint8)
    atlas[0, 0], atlas[1, 1], atlas[2, 2] = 7, 8, 9
    grid = np.array([0, 0, 5, 258, 256, 256, 5, 0], np.uint32)   # slot 2, slot 0 twice
    blob = V.volume_bytes((16, 8, 32), grid, atlas)
    want = struct.pack("<iii", 16, 8, 32) + struct.pack("<I", 8) + struct.pack(
        "<12I", 2, 0, 1, 5, 1, 256, 1, 257, 1, 258, 1, 5) + struct.pack("<3I", 1, 0, 0xFFFFFFFF) + struct.pack("<I", 3)
    assert blob[:len(want)] == want
    bricks = np.frombuffer(blob[len(want):], np.uint8).reshape(3, 512)
    assert bricks[0, 2] == 9 and bricks[1, 0] == 7 and bricks[2, 0] == 7


def test_payload_chunks_of_4_mib(tmp_path):
    rng = np.random.default_rng(0)
    vox = np.where(rng.random((256, 128, 384)) < 0.02, 3, 0).astype(np.uint8)   # ~24k detailed bricks
    a = vf.GenWorld.from_dense(vox).arrays()
    path = tmp_path / "big.vfworld"
    V.write_vfworld(path, a)
    b = path.read_bytes()
    (n,) = struct.unpack("<I", b[8:12])
    at = 12 + n
    raw, nc = struct.unpack("<QI", b[at:at + 12])
    assert raw > 2 * V.CHUNK and nc == (raw + V.CHUNK - 1) // V.CHUNK
    at += 12
    sizes = []
    for _ in range(nc):
        rn, cn = struct.unpack("<II", b[at:at + 8])
        sizes.append(rn)
        at += 8 + cn
    assert sizes[:-1] == [V.CHUNK] * (nc - 1) and sum(sizes) == raw and at == len(b)
    assert np.array_equal(_dense(V.arrays_from_save(V.read_vfworld(path))), vox)


def test_damage_is_detected(tmp_path):
    _, a, body, _ = _sample(tmp_path)
    path = tmp_path / "w.vfworld"
    V.write_vfworld(path, a, [body])
    b = bytearray(path.read_bytes())
    n = len(b)
    b[n - 40] ^= 0x5A
    path.write_bytes(bytes(b))
    with pytest.raises(V.VfWorldError) as e:
        V.read_vfworld(path)
    assert e.value.kind == "Corrupt"
    path.write_bytes(bytes(b[:n // 2]))
    with pytest.raises(V.VfWorldError) as e:
        V.read_vfworld(path)
    assert e.value.kind == "Corrupt"
    path.write_bytes(b"not a save")
    with pytest.raises(V.VfWorldError) as e:
        V.read_vfworld(path)
    assert e.value.kind == "NotASave"
    path.write_bytes(struct.pack("<II", V.MAGIC, 7) + bytes(16))
    with pytest.raises(V.VfWorldError) as e:
        V.read_vfworld(path)
    assert e.value.kind == "Newer"


def test_generated_world_traces_identically_after_save_and_load(tmp_path):
    w = vf.GenWorld.generate(1)
    a = w.arrays()
    path = tmp_path / "cfg1.vfworld"
    V.write_vfworld(path, a)
    back = V.arrays_from_save(V.read_vfworld(path))
    assert np.array_equal(_dense(back), _dense(a))
    rays = random_rays(3000, (0, 0, 0), (256, 128, 256), seed=4)
    h0, _ = O.OracleWorld(a).trace_rays(rays)
    h1, _ = O.OracleWorld(back).trace_rays(rays)
    assert np.array_equal(h0.view(np.uint8), h1.view(np.uint8))
    assert h0["hit"].mean() > 0.2


def test_endless_shape_as_a_bounded_world():
    w = vf.GenWorld.generate(2, chunks=3, trees_per_chunk=20, world_bricks=(64, 32, 64))
    a = w.arrays()
    b, off = V.ring_to_bounded(a, 3)
    assert b["world_bricks"] == (48, 32, 48) and off == 16 * 8
    rng = np.random.default_rng(1)
    ob = O.OracleWorld(b, build_fields=False, build_skytop=False)
    for _ in range(2000):
        x, y, z = int(rng.integers(-24, 24) * 8 + rng.integers(0, 8)), int(rng.integers(0, 256)), int(rng.integers(-24, 24) * 8 + rng.integers(0, 8))
        assert ob.voxel(x + off, y, z + off) == w.voxel(x, y, z)
