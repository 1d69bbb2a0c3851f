"""C ABI checks that need no GPU: the library loads and exports every declared entry point, the shared
structs keep the reference's byte layouts, and host-side errors come back as codes with messages."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import vftrace as vf
from vftrace import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    inc = os.path.join(ROOT, "include")
    for fn in sorted(os.listdir(inc)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^[A-Za-z_][\w \*]*?\b(vf_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 30
    L = C.CDLL(vf.library_path())
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python host binds each one
    assert set(names) <= set(vf._SIGS)


def test_struct_sizes_match_reference_layouts():
    s = [C.c_size_t() for _ in range(5)]
    assert vf.lib().vf_abi_sizes(*[C.byref(x) for x in s]) == 0
    assert [x.value for x in s] == [1536, 288, 48, 24, 20]
    assert C.sizeof(abi.SceneUbo) == 1536 and C.sizeof(abi.PassUbo) == 288 and C.sizeof(abi.Material) == 48
    assert vf.lib().vf_abi_version() == 1


def test_context_without_device_is_an_error_not_a_crash():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(vf.VfError) as e:
        vf.Context(0)
    assert e.value.code == -2 and "device" in str(e.value)


def test_null_and_invalid_arguments():
    L = vf.lib()
    assert L.vf_sync(None) == -1
    assert b"null" in L.vf_last_error(None)
    assert L.vf_trace_rays(None, None, 0, 1e30, 0, None) == -1
    p = vf.GenWorld.defaults(1)
    p.octaves = 0
    h = C.c_void_p()
    assert L.vf_gen_world_create(C.byref(p), C.byref(h)) == -1
    assert b"octaves" in L.vf_last_error(None)
    p = vf.GenWorld.defaults(2)
    p.world_bricks[0] = 96          # not a power of two: the ring grid needs x & (x-1) == 0
    assert L.vf_gen_world_create(C.byref(p), C.byref(h)) == -1
    assert L.vf_gen_defaults(3, C.byref(p)) == -1


def test_dense_round_trip_and_packing():
    rng = np.random.default_rng(0)
    vox = np.zeros((32, 24, 40), np.uint8)
    vox[:, :5, :] = 3                                # uniform bricks
    vox[rng.random(vox.shape) < 0.05] = 7            # detailed bricks
    vox[16:24, 8:16, 0:8] = 254                      # torch flame: air on the GPU side
    w = vf.GenWorld.from_dense(vox)
    a = w.arrays()
    assert a["world_bricks"] == (5, 3, 4)
    exp = np.where(vox == 254, 0, vox)
    for _ in range(500):
        z, y, x = rng.integers(0, 32), rng.integers(0, 24), rng.integers(0, 40)
        assert w.voxel(x, y, z) == exp[z, y, x]
    g = a["grid"]
    slots = g[g >= 256] - 256
    assert np.array_equal(np.sort(slots), np.arange(a["n_slots"]))       # every slot used once
    assert np.all(np.diff(slots) > 0)                                    # slots in grid-index order


def test_generator_is_deterministic():
    a = vf.GenWorld.generate(1, world_bricks=(16, 16, 16)).arrays()
    b = vf.GenWorld.generate(1, world_bricks=(16, 16, 16)).arrays()
    c = vf.GenWorld.generate(1, world_bricks=(16, 16, 16), seed=7).arrays()
    assert np.array_equal(a["grid"], b["grid"]) and np.array_equal(a["atlas"], b["atlas"])
    assert not np.array_equal(a["atlas"], c["atlas"])


def test_endless_shape_generator_small():
    w = vf.GenWorld.generate(2, world_bricks=(64, 32, 64), chunks=3, trees_per_chunk=40)
    a = w.arrays()
    assert a["world_win"] == (-32, 0, -32, 1)
    st = w.stats()
    assert st["detailed_bricks"] > 1000
    # unloaded ring slots (the window's outer chunk) stay empty
    g = a["grid"].reshape(32, 64, 64)
    loaded_cols = (g != 0).any(1)
    assert loaded_cols.sum() <= 3 * 16 * 3 * 16
    # trees: wood and leaves exist above the ground
    assert (a["atlas"] == 6).any() and (a["atlas"] == 7).any()


def test_brick_arguments_packed_and_as_records():
    """The world calls' brick arguments: packed arrays (512-byte stride, gpu_slot copied) or brick records
    whose vox / gpu_slot fields are passed in place at the record stride (how World::bricks is handed over)."""
    packed = np.arange(3 * 512, dtype=np.uint8).reshape(3, 512)
    bp, n, bs, sp, ss, hold, slot = vf.Context._brick_args(packed, [5, -1, 7])
    assert (n, bs, ss) == (3, 512, 4) and bp == hold.ctypes.data and slot.tolist() == [5, -1, 7]
    assert sp == slot.ctypes.data
    rec_t = np.dtype([("vox", np.uint8, 512), ("gpu_slot", np.int32), ("mask_ok", np.uint8), ("mask", np.uint32, 16)],
                     align=True)
    rec = np.zeros(4, rec_t)
    bp, n, bs, sp, ss, hold, slot = vf.Context._brick_args(rec, None)
    assert (n, bs, ss) == (4, rec_t.itemsize, rec_t.itemsize) and hold is rec
    assert bp == rec.ctypes.data + rec_t.fields["vox"][1] and sp == rec.ctypes.data + rec_t.fields["gpu_slot"][1]
    slot[2] = 9                                      # a view of the records' own field
    assert rec["gpu_slot"][2] == 9
    with pytest.raises(ValueError):
        vf.Context._brick_args(rec, [0, 0, 0, 0])    # records carry their own slots
    with pytest.raises(ValueError):
        vf.Context._brick_args(packed, [0, 0])       # one slot per brick


def test_buffer_checks_refuse_host_and_undersized_tensors():
    """The Python host checks device buffers before a call: a host tensor, a non-contiguous one or one
    smaller than the pass writes is a ValueError, never a device write out of bounds."""
    import torch
    from vftrace import _dev
    assert _dev(None, 100, "x") is None
    with pytest.raises(ValueError, match="device"):
        _dev(torch.zeros(64, dtype=torch.uint8), 16, "hits")
    with pytest.raises(ValueError, match="device"):
        _dev(np.zeros(64, np.uint8), 16, "hits")
