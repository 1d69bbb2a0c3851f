"""Host-side mirror of VoxelForge's traced-pass interface over libvftrace.so (MI355X / gfx950).

The reference records its traced passes in Rust (src/gfx/renderer/record.rs `Renderer::go` per
pass, src/gfx/gpu_world.rs `GpuWorld::upload` / `sync` / `record_distance_field`).  Rust is not part
of this build, so this module is the host side above the C ABI (include/vf_trace.h): the same calls,
argument meanings and error behaviour (a failing call raises VfError with the library's message;
nothing falls back to another implementation).  Device buffers are torch tensors on the HIP device
(only their data pointers cross the ABI).
"""
import ctypes as C
import math
import os

import numpy as np

from .abi import (BVH_NODE_DTYPE, BvhNode, GenParams, GI_RECORD_DTYPE, GICACHE_INPUTS, GICACHE_PARAMS, GiCacheParams,
                  GiCacheStats, Hit, HIT_DTYPE, Material, OBJ_HEADER_DTYPE, ObjHeader, PassUbo, QueueCfg, Ray, RAY_DTYPE,
                  Request, SceneUbo, TuneBatchArgs, TuneGiArgs, TuneResult, VF_MAX_REQUESTS, VF_REQ_PROBES, VF_REQ_VIEW, VF_TUNE_BATCH,
                  VF_TUNE_GI_BATCH, WorldDesc, WorldSrc, Chunk)

__all__ = ["VfError", "lib", "Context", "GenWorld", "SceneUbo", "PassUbo", "Material", "Ray", "Hit", "HIT_DTYPE",
           "RAY_DTYPE", "GI_RECORD_DTYPE", "OBJ_HEADER_DTYPE", "BVH_NODE_DTYPE", "GiCache", "fly_camera", "sun_dir",
           "make_pass", "default_scene", "library_path"]

_HERE = os.path.dirname(os.path.abspath(__file__))


def library_path():
    return os.path.join(_HERE, "libvftrace.so")


class VfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"vftrace error {code}: {msg}")
        self.code = code


_LIB = None

_P = C.c_void_p
_SIGS = {
    "vf_abi_version": (C.c_int, []),
    "vf_abi_sizes": (C.c_int, [C.POINTER(C.c_size_t)] * 5),
    "vf_last_error": (C.c_char_p, [_P]),
    "vf_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "vf_ctx_destroy": (None, [_P]),
    "vf_set_stream": (C.c_int, [_P, _P]),
    "vf_sync": (C.c_int, [_P]),
    "vf_device_alloc": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "vf_device_free": (C.c_int, [_P, _P]),
    "vf_copy_to_device": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "vf_copy_to_host": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "vf_memset_device": (C.c_int, [_P, _P, C.c_int, C.c_size_t]),
    "vf_world_upload": (C.c_int, [_P, C.POINTER(WorldDesc)]),
    "vf_world_pack": (C.c_int, [_P, C.POINTER(WorldSrc), _P]),
    "vf_world_read": (C.c_int, [_P, _P, _P, C.c_uint32, _P]),
    "vf_world_sync_chunks": (C.c_int, [_P, C.POINTER(Chunk), C.c_int32, _P, C.c_uint32, C.c_uint32, _P, C.c_uint32, _P,
                                       C.c_int32]),
    "vf_world_atlas_state": (C.c_int, [_P, _P, _P, _P]),
    "vf_world_sync_edits": (C.c_int, [_P, _P, _P, C.c_uint32, _P, C.c_uint32, C.c_uint32, _P, C.c_uint32]),
    "vf_read_coarse": (C.c_int, [_P, _P, _P]),
    "vf_world_update": (C.c_int, [_P, _P, _P, C.c_uint32, _P, _P, C.c_uint32]),
    "vf_build_brickdist": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_int32)]),
    "vf_read_brickdist": (C.c_int, [_P, _P, _P]),
    "vf_read_skytop": (C.c_int, [_P, _P]),
    "vf_set_scene": (C.c_int, [_P, C.POINTER(SceneUbo)]),
    "vf_trace_primary": (C.c_int, [_P, C.POINTER(PassUbo), C.c_int32, C.c_int32, _P, _P, _P]),
    "vf_trace_gbuffer": (C.c_int, [_P, C.POINTER(PassUbo), C.c_int32, C.c_int32, _P, _P, _P, _P, _P, _P]),
    "vf_objects_upload": (C.c_int, [_P, _P, C.c_int32, _P, C.c_int64]),
    "vf_objects_set_headers": (C.c_int, [_P, _P, C.c_int32]),
    "vf_objects_set_bvh": (C.c_int, [_P, _P, C.c_int32, _P, C.c_int32]),
    "vf_objects_update_grid": (C.c_int, [_P, _P, _P, C.c_int32]),
    "vf_objects_read_bvh": (C.c_int, [_P, _P, C.c_int32, C.POINTER(C.c_int32), _P, C.c_int32, C.POINTER(C.c_int32)]),
    "vf_trace_shadow": (C.c_int, [_P, C.POINTER(PassUbo), C.c_int32, C.c_int32, _P, _P, _P, _P]),
    "vf_trace_sun_cone": (C.c_int, [_P, C.POINTER(PassUbo), C.c_int32, C.c_int32, _P, _P, _P, _P]),
    "vf_trace_primary_shadow": (C.c_int, [_P, C.POINTER(PassUbo), C.c_int32, C.c_int32, _P, _P, _P, _P, _P]),
    "vf_trace_gi": (C.c_int, [_P, C.POINTER(PassUbo), C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_int32, _P, _P,
                              _P, _P]),
    "vf_trace_rays": (C.c_int, [_P, _P, C.c_int64, C.c_float, C.c_uint32, _P]),
    "vf_trace_probes": (C.c_int, [_P, C.POINTER(C.c_float), C.c_float, C.POINTER(C.c_int32), C.c_int32, C.c_uint32,
                                  C.c_float, _P]),
    "vf_queue_defaults": (C.c_int, [C.POINTER(QueueCfg)]),
    "vf_trace_gi_batch": (C.c_int, [_P, C.POINTER(PassUbo), C.c_int32, C.c_int32, C.c_float, _P, _P, _P, _P,
                                    C.POINTER(QueueCfg)]),
    "vf_trace_batch": (C.c_int, [_P, C.POINTER(Request), C.c_int32, C.POINTER(QueueCfg)]),
    "vf_trace_batch_noalloc": (C.c_int, [_P, C.POINTER(Request), C.c_int32, C.POINTER(QueueCfg)]),
    "vf_read_batch_times": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_double)]),
    "vf_autotune": (C.c_int, [_P, C.c_int32, _P, C.POINTER(QueueCfg), C.POINTER(QueueCfg), C.POINTER(TuneResult),
                              C.c_int32, C.POINTER(C.c_int32)]),
    "vf_graph_begin": (C.c_int, [_P]),
    "vf_graph_end": (C.c_int, [_P, C.POINTER(_P)]),
    "vf_graph_launch": (C.c_int, [_P, _P]),
    "vf_graph_destroy": (None, [_P]),
    "vf_set_counting": (C.c_int, [_P, C.c_int32]),
    "vf_set_timing": (C.c_int, [_P, C.c_int32]),
    "vf_set_traversal": (C.c_int, [_P, C.c_int32]),
    "vf_read_counters": (C.c_int, [_P, C.POINTER(C.c_uint64), C.c_int32]),
    "vf_last_kernel_ms": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "vf_gicache_defaults": (C.c_int, [C.POINTER(GiCacheParams)]),
    "vf_gicache_create": (C.c_int, [_P, C.POINTER(GiCacheParams), C.POINTER(_P)]),
    "vf_gicache_destroy": (None, [_P]),
    "vf_gicache_reset": (C.c_int, [_P, _P]),
    "vf_trace_gi_path": (C.c_int, [_P, C.POINTER(PassUbo), C.c_int32, C.c_int32, C.c_float, C.c_int32, _P, _P,
                                   C.c_uint32, C.c_int32, C.c_uint32, _P]),
    "vf_trace_gi_path_n": (C.c_int, [_P, C.POINTER(PassUbo), C.c_int32, C.c_int32, C.c_float, C.c_int32, _P, _P,
                                     C.c_uint32, C.c_int32, C.c_uint32, C.c_int32, _P]),
    "vf_gicache_frame": (C.c_int, [_P, _P, C.POINTER(PassUbo), C.c_int32, C.c_int32, _P, _P, C.c_uint32, _P]),
    "vf_gicache_read_stats": (C.c_int, [_P, _P, C.POINTER(GiCacheStats)]),
    "vf_gicache_samples": (C.c_int, [_P, _P, C.POINTER(PassUbo), C.c_int32, C.c_int32, _P, _P, C.c_uint32, _P]),
    "vf_gicache_get_params": (C.c_int, [_P, _P, _P]),
    "vf_gicache_get_render_params": (C.c_int, [_P, _P, _P]),
    "vf_gicache_set_params": (C.c_int, [_P, _P, _P]),
    "vf_gicache_encode": (C.c_int, [_P, _P, _P, C.c_int64, _P]),
    "vf_gicache_infer": (C.c_int, [_P, _P, _P, C.c_int64, _P]),
    "vf_gicache_grad": (C.c_int, [_P, _P, _P, C.c_int64, _P, _P]),
    "vf_gicache_adam": (C.c_int, [_P, _P, _P]),
    "vf_gicache_train": (C.c_int, [_P, _P, _P, C.c_int64]),
    "vf_gbuffer_motion": (C.c_int, [_P, C.POINTER(PassUbo), C.c_int32, C.c_int32, _P, _P, _P, _P]),
    "vf_gi_accumulate": (C.c_int, [_P, C.POINTER(PassUbo), C.c_int32, C.c_int32, C.c_int32] + [_P] * 12),
    "vf_light_accumulate": (C.c_int, [_P, C.POINTER(PassUbo), C.c_int32, C.c_int32] + [_P] * 9),
    "vf_gi_atrous": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, C.c_int32] + [_P] * 7),
    "vf_gen_defaults": (C.c_int, [C.c_int32, C.POINTER(GenParams)]),
    "vf_gen_world_create": (C.c_int, [C.POINTER(GenParams), C.POINTER(_P)]),
    "vf_gen_world_from_dense": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _P, C.POINTER(_P)]),
    "vf_gen_world_destroy": (None, [_P]),
    "vf_gen_world_desc": (C.c_int, [_P, C.POINTER(WorldDesc)]),
    "vf_gen_world_stats": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "vf_gen_world_voxel": (C.c_uint32, [_P, C.c_int32, C.c_int32, C.c_int32]),
    "vf_gen_world_ground": (C.c_int32, [_P, C.c_int32, C.c_int32]),
}


def lib():
    """Loads libvftrace.so (built in-tree by __graft_entry__.build()); raises if it is missing."""
    global _LIB
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise RuntimeError(f"libvftrace.so not built: {path} is missing (run __graft_entry__.build())")
        L = C.CDLL(path)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def _check(rc, ctx=None):
    if rc != 0:
        msg = lib().vf_last_error(ctx).decode(errors="replace")
        raise VfError(rc, msg)


def _ptr(t):
    """Device pointer of a torch tensor (or None)."""
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


def _dev(t, nbytes, what):
    """Device pointer of a torch tensor that is on the GPU, contiguous and holds at least nbytes (None
    passes through): an undersized or host buffer is a ValueError here, not an out-of-bounds device
    write in the kernel."""
    if t is None:
        return None
    if not hasattr(t, "data_ptr") or not getattr(t, "is_cuda", False):
        raise ValueError(f"{what}: expected a device (cuda) tensor")
    if not t.is_contiguous():
        raise ValueError(f"{what}: tensor is not contiguous")
    have = t.numel() * t.element_size()
    if have < nbytes:
        raise ValueError(f"{what}: {have} bytes, {int(nbytes)} needed")
    return C.c_void_p(t.data_ptr())


def _half(n):
    return (int(n) + 1) // 2


def _i3(v):
    return (C.c_int32 * 3)(*[int(x) for x in v])


# --- uniforms ---------------------------------------------------------------------------------
def fly_camera(pos, yaw, pitch):
    """FlyCamera basis (src/camera.rs:34-42): forward = (cos p sin y, sin p, cos p cos y),
    right = normalize(forward x Y), up = right x forward; float32."""
    f32 = np.float32
    y, p = f32(yaw), f32(pitch)
    fwd = np.array([np.cos(p) * np.sin(y), np.sin(p), np.cos(p) * np.cos(y)], dtype=np.float32)
    r = np.cross(fwd, np.array([0, 1, 0], dtype=np.float32)).astype(np.float32)
    n = np.float32(np.sqrt(np.float32(r @ r)))
    right = (r / n).astype(np.float32) if n > 1e-6 else np.array([0, 1, 0], dtype=np.float32)
    up = np.cross(right, fwd).astype(np.float32)
    return np.asarray(pos, dtype=np.float32), right, up, fwd


def sun_dir(az_deg, el_deg):
    """Renderer::set_sun (src/gfx/renderer/frame.rs:44-48)."""
    az, el = math.radians(az_deg), math.radians(el_deg)
    s = np.array([math.cos(el) * math.cos(az), math.sin(el), math.cos(el) * math.sin(az)], dtype=np.float32)
    return (s / np.float32(np.linalg.norm(s))).astype(np.float32)


def make_pass(pos, right, up, fwd, fov_deg=70.0, frame_idx=0, jitter=(0.0, 0.0), leaf_light=0.6):
    """The camera part of PassUbo (set_frame_uniforms, src/gfx/renderer/frame.rs:549-566) and look_p[3] =
    r.foliage.light (frame.rs:606; default 0.6, cvars.rs:143): the sunlight each leaf lets through."""
    p = PassUbo()
    p.look_p[3] = float(leaf_light)
    thf = np.float32(math.tan(math.radians(fov_deg * 0.5)))
    for i in range(3):
        p.cam_pos[i] = float(pos[i]); p.cam_right[i] = float(right[i])
        p.cam_up[i] = float(up[i]); p.cam_fwd[i] = float(fwd[i])
        p.prev_cam_pos[i] = float(pos[i]); p.prev_cam_right[i] = float(right[i])
        p.prev_cam_up[i] = float(up[i]); p.prev_cam_fwd[i] = float(fwd[i])
    p.tan_half_fov = float(thf)
    p.prev_tan_half_fov = float(thf)
    p.frame_idx = int(frame_idx)
    p.jitter[0], p.jitter[1] = float(jitter[0]), float(jitter[1])
    return p


def set_prev_camera(p, prev):
    """Copies last frame's camera (pass `prev`'s cam_*, tan_half_fov) into p's prev_cam_* fields, as
    set_frame_uniforms does each frame (frame.rs:549-566); returns p."""
    for i in range(3):
        p.prev_cam_pos[i] = prev.cam_pos[i]; p.prev_cam_right[i] = prev.cam_right[i]
        p.prev_cam_up[i] = prev.cam_up[i]; p.prev_cam_fwd[i] = prev.cam_fwd[i]
    p.prev_tan_half_fov = prev.tan_half_fov
    return p


_DEFAULT_PASS = PassUbo()
_DEFAULT_PASS.look_p[3] = 0.6


def default_scene(world_bricks, world_win=(0, 0, 0, 0), sun=(32.0, 46.0), sky_model=0):
    """SceneUbo with the reference defaults the traced passes read (cvars.rs:152-160, 209-210):
    sun azimuth 32, elevation 46, classic sky, sky and GI intensity 1."""
    s = SceneUbo()
    for i in range(3):
        s.world_bricks[i] = int(world_bricks[i])
    for i in range(4):
        s.world_win[i] = int(world_win[i])
    d = sun_dir(*sun)
    for i in range(3):
        s.sun_dir[i] = float(d[i])
    s.sun_rgb[0], s.sun_rgb[1], s.sun_rgb[2] = 3.0, 2.85, 2.6
    s.sky_zenith[0], s.sky_zenith[1], s.sky_zenith[2] = 0.25, 0.45, 0.85
    s.sky_horizon[0], s.sky_horizon[1], s.sky_horizon[2] = 0.75, 0.82, 0.90
    s.sunset_rgb[0], s.sunset_rgb[1], s.sunset_rgb[2] = 0.3, 0.15, 0.05
    s.sky_scale = 1.0
    s.sun_angle = 0.028   # 0.028 * r.sun.softness (1) * (1 + 6 cloudy^2), clear sky (frame.rs:230)
    s.gi_scale = 1.0
    s.sky_model = int(sky_model)
    s.num_objects = 0
    return s


# --- worlds -----------------------------------------------------------------------------------
class GenWorld:
    """A synthetic world packed like GpuWorld::upload (include/vf_worldgen.h)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    @classmethod
    def defaults(cls, kind):
        p = GenParams()
        _check(lib().vf_gen_defaults(int(kind), C.byref(p)))
        return p

    @classmethod
    def generate(cls, kind=1, **overrides):
        p = cls.defaults(kind)
        for k, v in overrides.items():
            if k == "world_bricks":
                for i in range(3):
                    p.world_bricks[i] = int(v[i])
            else:
                setattr(p, k, v)
        h = C.c_void_p()
        _check(lib().vf_gen_world_create(C.byref(p), C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_dense(cls, vox, world_win=(0, 0, 0, 0), palette=None):
        """vox: uint8 array [z, y, x] of shape world_bricks*8 (z, y, x order)."""
        vox = np.ascontiguousarray(vox, dtype=np.uint8)
        nz, ny, nx = vox.shape
        if nx % 8 or ny % 8 or nz % 8:
            raise ValueError("dense world dims must be multiples of 8")
        wb = (C.c_int32 * 3)(nx // 8, ny // 8, nz // 8)
        win = (C.c_int32 * 4)(*[int(x) for x in world_win])
        pal = None
        if palette is not None:
            pal = (Material * 256)()
            arr = np.ascontiguousarray(palette, dtype=np.float32).reshape(256, 12)
            C.memmove(pal, arr.ctypes.data, 256 * 48)
        h = C.c_void_p()
        _check(lib().vf_gen_world_from_dense(vox.ctypes.data_as(C.c_void_p), wb, win,
                                             C.cast(pal, C.c_void_p) if pal is not None else None, C.byref(h)))
        return cls(h.value)

    def close(self):
        if self._h and self._h.value:
            lib().vf_gen_world_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def desc(self):
        d = WorldDesc()
        _check(lib().vf_gen_world_desc(self._h, C.byref(d)))
        return d

    def stats(self):
        out = (C.c_uint64 * 4)()
        _check(lib().vf_gen_world_stats(self._h, out))
        return {"nonempty_bricks": out[0], "detailed_bricks": out[1], "uniform_bricks": out[2],
                "atlas_solid_voxels": out[3]}

    def voxel(self, x, y, z):
        return int(lib().vf_gen_world_voxel(self._h, int(x), int(y), int(z)))

    def ground(self, x, z):
        return int(lib().vf_gen_world_ground(self._h, int(x), int(z)))

    def arrays(self):
        """Copies of the packed arrays: grid, atlas, masks, coarse1, coarse2, palette, dims."""
        d = self.desc()
        wb = tuple(d.world_bricks)
        n = wb[0] * wb[1] * wb[2]
        c1 = tuple((x + 3) // 4 for x in wb)
        c2 = tuple((x + 15) // 16 for x in wb)
        ns = d.n_slots

        def arr(ptr, count, dtype):
            if count == 0 or not ptr:
                return np.zeros(0, dtype=dtype)
            return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True)

        pal = np.ctypeslib.as_array(C.cast(d.palette, C.POINTER(C.c_float)), shape=(256 * 12,)).copy()
        return {
            "world_bricks": wb, "world_win": tuple(d.world_win), "n_slots": ns,
            "grid": arr(d.grid, n, np.uint32),
            "atlas": arr(d.atlas, ns * 512, np.uint8),
            "masks": arr(d.masks, ns * 16, np.uint32),
            "coarse1": arr(d.coarse1, c1[0] * c1[1] * c1[2], np.uint8),
            "coarse2": arr(d.coarse2, c2[0] * c2[1] * c2[2], np.uint8),
            "palette": pal.reshape(256, 12),
        }


# --- context ----------------------------------------------------------------------------------
class Context:
    """One vf_ctx: owns the world mirror on the device and a HIP stream."""

    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib().vf_ctx_create(int(device), C.byref(h)))
        self._h = h
        self.world_bricks = None

    def close(self):
        if self._h and self._h.value:
            lib().vf_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc):
        if rc != 0:
            raise VfError(rc, lib().vf_last_error(self._h).decode(errors="replace"))

    # state
    def set_stream(self, stream=None):
        """Use a torch.cuda.Stream (or None for the context's own stream)."""
        self._ck(lib().vf_set_stream(self._h, C.c_void_p(stream.cuda_stream) if stream is not None else None))

    def sync(self):
        self._ck(lib().vf_sync(self._h))

    def upload(self, world):
        d = world.desc() if isinstance(world, GenWorld) else world
        self._ck(lib().vf_world_upload(self._h, C.byref(d)))
        self.world_bricks = tuple(d.world_bricks)

    def update(self, cell_idx=None, cell_val=None, slot_idx=None, slot_vox=None):
        ci = np.ascontiguousarray(cell_idx if cell_idx is not None else [], dtype=np.uint32)
        cv = np.ascontiguousarray(cell_val if cell_val is not None else [], dtype=np.uint32)
        si = np.ascontiguousarray(slot_idx if slot_idx is not None else [], dtype=np.uint32)
        sv = np.ascontiguousarray(slot_vox if slot_vox is not None else np.zeros((0, 512)), dtype=np.uint8)
        if ci.size != cv.size or sv.size != si.size * 512:
            raise ValueError("update: mismatched array sizes")
        self._ck(lib().vf_world_update(self._h, ci.ctypes.data if ci.size else None,
                                       cv.ctypes.data if cv.size else None, ci.size,
                                       si.ctypes.data if si.size else None, sv.ctypes.data if sv.size else None,
                                       si.size))

    def build_brickdist(self, lo, hi, dirty_lo=None, dirty_hi=None):
        dlo = lo if dirty_lo is None else dirty_lo
        dhi = hi if dirty_hi is None else dirty_hi
        self._ck(lib().vf_build_brickdist(self._h, _i3(lo), _i3(hi), _i3(dlo), _i3(dhi)))

    def read_brickdist(self):
        n = int(np.prod(self.world_bricks))
        d = np.zeros(n, np.uint8)
        r = np.zeros(n, np.uint8)
        self._ck(lib().vf_read_brickdist(self._h, d.ctypes.data, r.ctypes.data))
        return d, r

    def read_skytop(self):
        wb = self.world_bricks
        s = np.zeros(wb[0] * 8 * wb[2] * 8, np.uint16)
        self._ck(lib().vf_read_skytop(self._h, s.ctypes.data))
        return s

    def set_scene(self, scene):
        self._ck(lib().vf_set_scene(self._h, C.byref(scene)))

    def upload_arrays(self, arrays):
        """vf_world_upload from numpy arrays (GenWorld.arrays() layout; masks / coarse pyramid optional)."""
        wb = tuple(int(x) for x in arrays["world_bricks"])
        grid = np.ascontiguousarray(arrays["grid"], dtype=np.uint32)
        atlas = np.ascontiguousarray(arrays["atlas"], dtype=np.uint8)
        pal = np.ascontiguousarray(arrays["palette"], dtype=np.float32).reshape(256, 12)
        d = WorldDesc()
        d.world_bricks[:] = wb
        d.world_win[:] = tuple(int(x) for x in arrays["world_win"])
        d.grid = grid.ctypes.data_as(C.POINTER(C.c_uint32))
        d.atlas = atlas.ctypes.data_as(C.POINTER(C.c_uint8)) if atlas.size else None
        d.n_slots = atlas.size // 512
        keep = [grid, atlas, pal]
        for k, typ in (("masks", C.c_uint32), ("coarse1", C.c_uint8), ("coarse2", C.c_uint8)):
            a = arrays.get(k)
            if a is not None and np.asarray(a).size:
                a = np.ascontiguousarray(a)
                keep.append(a)
                setattr(d, k, a.ctypes.data_as(C.POINTER(typ)))
        d.palette = pal.ctypes.data_as(C.POINTER(Material))
        self._ck(lib().vf_world_upload(self._h, C.byref(d)))
        self.world_bricks = wb
        del keep

    def pack_world(self, world_bricks, world_win, grid, bricks, palette, atlas_slots=0):
        """vf_world_pack: the world from the CPU side's grid (0 / 1-255 / 256 + brick index) and bricks
        ([n, 512] u8, or brick records as `_brick_args` takes them, whose gpu_slot fields are written),
        packed on the device.  Returns each brick's slot (-1: unnamed or did not fit)."""
        wb = tuple(int(x) for x in world_bricks)
        g = np.ascontiguousarray(grid, dtype=np.uint32)
        rec = getattr(bricks, "dtype", None) is not None and bricks.dtype.names is not None
        bp, nb, bs, sp, ss, hold, slot = self._brick_args(bricks, None if rec else
                                                          np.zeros(np.asarray(bricks).size // 512, np.int32))
        pal = np.ascontiguousarray(palette, dtype=np.float32).reshape(256, 12)
        d = WorldSrc()
        d.world_bricks[:] = wb
        d.world_win[:] = tuple(int(x) for x in world_win)
        d.grid = g.ctypes.data_as(C.POINTER(C.c_uint32))
        d.bricks = C.cast(bp, C.POINTER(C.c_uint8)) if bp else None
        d.n_bricks = nb
        d.atlas_slots = int(atlas_slots)
        d.palette = pal.ctypes.data_as(C.POINTER(Material))
        d.brick_stride, d.slot_stride = bs, ss
        self._ck(lib().vf_world_pack(self._h, C.byref(d), sp))
        self.world_bricks = wb
        del hold
        return np.array(slot, np.int32)

    def world_read(self, n_slots):
        """The device world's encoded grid, its first n_slots atlas slots and their brick masks."""
        wb = self.world_bricks
        grid = np.zeros(wb[0] * wb[1] * wb[2], np.uint32)
        atlas = np.zeros((max(n_slots, 1), 512), np.uint8)
        masks = np.zeros((max(n_slots, 1), 16), np.uint32)
        self._ck(lib().vf_world_read(self._h, grid.ctypes.data, atlas.ctypes.data, int(n_slots), masks.ctypes.data))
        return grid, atlas[:n_slots], masks[:n_slots]

    @staticmethod
    def _brick_args(bricks, gpu_slot):
        """(voxel pointer, count, stride, slot pointer, slot stride, keep-alive, slots) for the world calls:
        packed bricks [n, 512] u8 with gpu_slot [n] i32 (copied), or an array of brick records (a structured
        array with fields `vox` (512 u8) and `gpu_slot` (i32), like World::bricks), read and written in place."""
        if getattr(bricks, "dtype", None) is not None and bricks.dtype.names and "vox" in bricks.dtype.names:
            if gpu_slot is not None:
                raise ValueError("brick records carry their own gpu_slot")
            if not bricks.flags.c_contiguous:
                raise ValueError("brick records must be contiguous")
            base = bricks.ctypes.data
            vo, so = bricks.dtype.fields["vox"][1], bricks.dtype.fields["gpu_slot"][1]
            n = bricks.shape[0]
            return (base + vo if n else None, n, bricks.dtype.itemsize, base + so if n else None, bricks.dtype.itemsize,
                    bricks, bricks["gpu_slot"])
        b = np.ascontiguousarray(bricks, dtype=np.uint8).reshape(-1, 512)
        slot = np.array(gpu_slot, dtype=np.int32).reshape(-1)
        if slot.size != b.shape[0]:
            raise ValueError(f"gpu_slot has {slot.size} entries for {b.shape[0]} bricks")
        return (b.ctypes.data if b.size else None, b.shape[0], 512, slot.ctypes.data if slot.size else None, 4, b, slot)

    def sync_chunks(self, chunks, bricks, gpu_slot, released=()):
        """vf_world_sync_chunks (GpuWorld::sync_chunks): chunks = [(cx, cz, cells or None)], cells [16*wby*16] u32
        in (z, y, x) order (None: evicted and not refilled); bricks [n, 512] u8; gpu_slot [n] i32 (-1: none).
        Returns the updated gpu_slot."""
        keep = []
        arr = (Chunk * max(len(chunks), 1))()
        for i, (cx, cz, cells) in enumerate(chunks):
            arr[i].cx, arr[i].cz = int(cx), int(cz)
            if cells is None:
                arr[i].loaded, arr[i].cells = 0, None
            else:
                c = np.ascontiguousarray(cells, dtype=np.uint32).reshape(-1)
                keep.append(c)
                arr[i].loaded, arr[i].cells = 1, c.ctypes.data_as(C.POINTER(C.c_uint32))
        bp, nb, bs, sp, ss, hold, slot = self._brick_args(bricks, gpu_slot)
        rel = np.ascontiguousarray(released, dtype=np.int32).reshape(-1)
        self._ck(lib().vf_world_sync_chunks(self._h, arr, len(chunks), bp, nb, bs, sp, ss,
                                            rel.ctypes.data if rel.size else None, rel.size))
        del keep, hold
        return slot

    def sync_edits(self, cells, values, bricks, gpu_slot):
        """vf_world_sync_edits (GpuWorld::sync): the world's edited cells (grid indices) and their values
        (0 / 1-255 / 256 + brick index); bricks [n, 512] u8; gpu_slot [n] i32.  Returns (values after the
        compaction of bricks that became uniform, gpu_slot)."""
        c = np.ascontiguousarray(cells, dtype=np.uint32).reshape(-1)
        v = np.array(values, dtype=np.uint32).reshape(-1)
        if v.size != c.size:
            raise ValueError(f"{v.size} values for {c.size} cells")
        bp, nb, bs, sp, ss, hold, slot = self._brick_args(bricks, gpu_slot)
        self._ck(lib().vf_world_sync_edits(self._h, c.ctypes.data if c.size else None, v.ctypes.data if v.size else None,
                                           c.size, bp, nb, bs, sp, ss))
        del hold
        return v, slot

    def atlas_state(self):
        """(slots handed out, free-list length, capacity) of the device world's atlas."""
        v = (C.c_uint32 * 3)()
        self._ck(lib().vf_world_atlas_state(self._h, C.byref(v, 0), C.byref(v, 4), C.byref(v, 8)))
        return int(v[0]), int(v[1]), int(v[2])

    def read_coarse(self):
        """The device world's coarse occupancy pyramid: (coarse1, coarse2), flat u8."""
        wb = self.world_bricks
        c1 = np.zeros(int(np.prod([(x + 3) // 4 for x in wb])), np.uint8)
        c2 = np.zeros(int(np.prod([(x + 15) // 16 for x in wb])), np.uint8)
        self._ck(lib().vf_read_coarse(self._h, c1.ctypes.data, c2.ctypes.data))
        return c1, c2

    # bodies (GpuWorld::init_objects / update_objects / sync_objects)
    def objects_upload(self, headers, grid):
        h = np.ascontiguousarray(headers, dtype=OBJ_HEADER_DTYPE)
        g = np.ascontiguousarray(grid, dtype=np.uint32)
        self._ck(lib().vf_objects_upload(self._h, h.ctypes.data if h.size else None, int(h.size),
                                         g.ctypes.data if g.size else None, int(g.size)))

    def objects_set_headers(self, headers):
        h = np.ascontiguousarray(headers, dtype=OBJ_HEADER_DTYPE)
        self._ck(lib().vf_objects_set_headers(self._h, h.ctypes.data if h.size else None, int(h.size)))

    def objects_set_bvh(self, nodes, prims):
        nd = np.ascontiguousarray(nodes, dtype=BVH_NODE_DTYPE)
        pr = np.ascontiguousarray(prims, dtype=np.uint32)
        self._ck(lib().vf_objects_set_bvh(self._h, nd.ctypes.data if nd.size else None, int(nd.size),
                                          pr.ctypes.data if pr.size else None, int(pr.size)))

    def objects_update_grid(self, idx, val):
        i = np.ascontiguousarray(idx, dtype=np.uint32)
        v = np.ascontiguousarray(val, dtype=np.uint32)
        if i.size != v.size:
            raise ValueError("objects_update_grid: mismatched arrays")
        self._ck(lib().vf_objects_update_grid(self._h, i.ctypes.data if i.size else None,
                                              v.ctypes.data if v.size else None, int(i.size)))

    def objects_read_bvh(self):
        nn, npr = C.c_int32(), C.c_int32()
        self._ck(lib().vf_objects_read_bvh(self._h, None, 0, C.byref(nn), None, 0, C.byref(npr)))
        nodes = np.zeros(nn.value, BVH_NODE_DTYPE)
        prims = np.zeros(npr.value, np.uint32)
        self._ck(lib().vf_objects_read_bvh(self._h, nodes.ctypes.data if nodes.size else None, nn.value, C.byref(nn),
                                           prims.ctypes.data if prims.size else None, npr.value, C.byref(npr)))
        return nodes, prims

    # passes (device buffers: torch tensors)
    def trace_primary(self, pass_ubo, width, height, hits, gb_pos=None, gb_normal=None):
        n = int(width) * int(height)
        self._ck(lib().vf_trace_primary(self._h, C.byref(pass_ubo), int(width), int(height), _dev(hits, 20 * n, "hits"),
                                        _dev(gb_pos, 16 * n, "gb_pos"), _dev(gb_normal, 4 * n, "gb_normal")))

    def trace_gbuffer(self, pass_ubo, width, height, hits, gb_pos=None, gb_normal=None, body=None, shadow_raw=None,
                      shadow_hits=None):
        """P::Gbuf over the scene (+ P::ShTrace when shadow_raw is given) with the body index per pixel."""
        n = int(width) * int(height)
        self._ck(lib().vf_trace_gbuffer(self._h, C.byref(pass_ubo), int(width), int(height), _dev(hits, 20 * n, "hits"),
                                        _dev(gb_pos, 16 * n, "gb_pos"), _dev(gb_normal, 4 * n, "gb_normal"),
                                        _dev(body, 4 * n, "body"), _dev(shadow_raw, 4 * n, "shadow_raw"),
                                        _dev(shadow_hits, 20 * n, "shadow_hits")))

    def trace_shadow(self, width, height, gb_pos, gb_normal, shadow_raw, hits=None, pass_ubo=None):
        """shadow_trace.comp; the pass supplies look_p[3] (r.foliage.light, the reference default 0.6)."""
        p = pass_ubo if pass_ubo is not None else _DEFAULT_PASS
        n = int(width) * int(height)
        self._ck(lib().vf_trace_shadow(self._h, C.byref(p), int(width), int(height), _dev(gb_pos, 16 * n, "gb_pos"),
                                       _dev(gb_normal, 4 * n, "gb_normal"), _dev(shadow_raw, 4 * n, "shadow_raw"),
                                       _dev(hits, 20 * n, "hits")))

    def trace_sun_cone(self, pass_ubo, width, height, gb_pos, gb_normal, shadow, hits=None):
        """lighting.comp's cone-jittered sun ray (float per pixel)."""
        n = int(width) * int(height)
        self._ck(lib().vf_trace_sun_cone(self._h, C.byref(pass_ubo), int(width), int(height),
                                         _dev(gb_pos, 16 * n, "gb_pos"), _dev(gb_normal, 4 * n, "gb_normal"),
                                         _dev(shadow, 4 * n, "shadow"), _dev(hits, 20 * n, "hits")))

    def trace_primary_shadow(self, pass_ubo, width, height, hits, gb_pos, gb_normal, shadow_raw, shadow_hits=None):
        n = int(width) * int(height)
        self._ck(lib().vf_trace_primary_shadow(self._h, C.byref(pass_ubo), int(width), int(height),
                                               _dev(hits, 20 * n, "hits"), _dev(gb_pos, 16 * n, "gb_pos"),
                                               _dev(gb_normal, 4 * n, "gb_normal"),
                                               _dev(shadow_raw, 4 * n, "shadow_raw"),
                                               _dev(shadow_hits, 20 * n, "shadow_hits")))

    def trace_gi(self, pass_ubo, width, height, gb_pos, gb_normal, radiance, bounce=None, half_res=True,
                 gi_reach=48.0, lights_nee=True):
        n = int(width) * int(height)
        no = _half(width) * _half(height) if half_res else n
        self._ck(lib().vf_trace_gi(self._h, C.byref(pass_ubo), int(width), int(height), int(bool(half_res)),
                                   float(gi_reach), int(bool(lights_nee)), _dev(gb_pos, 16 * n, "gb_pos"),
                                   _dev(gb_normal, 4 * n, "gb_normal"), _dev(radiance, 16 * no, "radiance"),
                                   _dev(bounce, 20 * no, "bounce")))

    # temporal consumers (include/vf_temporal.h)
    def gbuffer_motion(self, pass_ubo, width, height, gb_pos, gb_motion, gb_delta, body=None):
        """gbuffer.comp's gbMotion (RG16F bits per pixel, int32 tensor) and gbDelta (RGBA16F bits)."""
        n = int(width) * int(height)
        self._ck(lib().vf_gbuffer_motion(self._h, C.byref(pass_ubo), int(width), int(height),
                                         _dev(gb_pos, 16 * n, "gb_pos"), _dev(body, 4 * n, "body"),
                                         _dev(gb_motion, 4 * n, "gb_motion"), _dev(gb_delta, 8 * n, "gb_delta")))

    def gi_accumulate(self, pass_ubo, width, height, gb_pos, gb_normal, gb_pos_prev, gb_normal_prev, gb_motion,
                      gb_delta, indirect, hist_prev, mom_prev, hist_cur, mom_cur, guide, bicubic=True):
        """gi.comp's reprojected accumulation of the traced half-res GI (RGBA16F history as int16/uint16
        tensors of half-float bits, RGBA32F moments, RGBA32UI guide)."""
        n, nh = int(width) * int(height), _half(width) * _half(height)
        sizes = (("gb_pos", 16 * n), ("gb_normal", 4 * n), ("gb_pos_prev", 16 * n), ("gb_normal_prev", 4 * n),
                 ("gb_motion", 4 * n), ("gb_delta", 8 * n), ("indirect", 16 * nh), ("hist_prev", 8 * nh),
                 ("mom_prev", 16 * nh), ("hist_cur", 8 * nh), ("mom_cur", 16 * nh), ("guide", 16 * nh))
        ts = (gb_pos, gb_normal, gb_pos_prev, gb_normal_prev, gb_motion, gb_delta, indirect, hist_prev, mom_prev,
              hist_cur, mom_cur, guide)
        self._ck(lib().vf_gi_accumulate(self._h, C.byref(pass_ubo), int(width), int(height), int(bool(bicubic)),
                                        *[_dev(t, b, w) for t, (w, b) in zip(ts, sizes)]))

    def light_accumulate(self, pass_ubo, width, height, gb_pos, gb_normal, gb_pos_prev, gb_normal_prev, gb_motion,
                         gb_delta, shadow, hist_prev, hist_cur):
        """lighting.comp's sun-shadow history (RGBA32F per pixel)."""
        n = int(width) * int(height)
        sizes = (("gb_pos", 16 * n), ("gb_normal", 4 * n), ("gb_pos_prev", 16 * n), ("gb_normal_prev", 4 * n),
                 ("gb_motion", 4 * n), ("gb_delta", 8 * n), ("shadow", 4 * n), ("hist_prev", 16 * n),
                 ("hist_cur", 16 * n))
        ts = (gb_pos, gb_normal, gb_pos_prev, gb_normal_prev, gb_motion, gb_delta, shadow, hist_prev, hist_cur)
        self._ck(lib().vf_light_accumulate(self._h, C.byref(pass_ubo), int(width), int(height),
                                           *[_dev(t, b, w) for t, (w, b) in zip(ts, sizes)]))

    def gi_atrous(self, width, height, levels, gb_pos, gb_normal, gi_mom, gi_hist, guide, filt0, filt1,
                  histfix=True):
        """`levels` a-trous levels; level i writes filt[i & 1] (odd counts end in filt0)."""
        n, nh = int(width) * int(height), _half(width) * _half(height)
        sizes = (("gb_pos", 16 * n), ("gb_normal", 4 * n), ("gi_mom", 16 * nh), ("gi_hist", 8 * nh),
                 ("guide", 16 * nh), ("filt0", 8 * nh), ("filt1", 8 * nh))
        ts = (gb_pos, gb_normal, gi_mom, gi_hist, guide, filt0, filt1)
        self._ck(lib().vf_gi_atrous(self._h, int(width), int(height), int(levels), int(bool(histfix)),
                                    *[_dev(t, b, w) for t, (w, b) in zip(ts, sizes)]))

    def trace_rays(self, rays, n, hits, t_cap=1e30, skip_mat=0):
        self._ck(lib().vf_trace_rays(self._h, _dev(rays, 24 * int(n), "rays"), int(n), float(t_cap), int(skip_mat),
                                     _dev(hits, 20 * int(n), "hits")))

    def trace_probes(self, origin, spacing, dims, n_rays, frame, reach, hits):
        o = (C.c_float * 3)(*[float(x) for x in origin])
        nh = int(np.prod([int(x) for x in dims])) * int(n_rays)
        self._ck(lib().vf_trace_probes(self._h, o, float(spacing), _i3(dims), int(n_rays), int(frame), float(reach),
                                       _dev(hits, 20 * max(nh, 0), "hits")))

    # persistent-queue batches
    @staticmethod
    def queue_cfg(tile=(8, 8), chunk=256, sort=True, groups_per_cu=0, refill=32, traversal=0):
        c = QueueCfg()
        c.tile_w, c.tile_h = int(tile[0]), int(tile[1])
        c.chunk, c.sort, c.groups_per_cu, c.refill = int(chunk), int(bool(sort)), int(groups_per_cu), int(refill)
        c.traversal = int(traversal)
        return c

    def trace_gi_batch(self, pass_ubo, width, height, reach, gb_pos, gb_normal, gi_hits, sun_hits, cfg=None):
        n = int(width) * int(height)
        self._ck(lib().vf_trace_gi_batch(self._h, C.byref(pass_ubo), int(width), int(height), float(reach),
                                         _dev(gb_pos, 16 * n, "gb_pos"), _dev(gb_normal, 4 * n, "gb_normal"),
                                         _dev(gi_hits, 20 * n, "gi_hits"), _dev(sun_hits, 20 * n, "sun_hits"),
                                         C.byref(cfg) if cfg is not None else None))

    @staticmethod
    def view_request(pass_ubo, width, height, hits):
        r = Request()
        r.kind = VF_REQ_VIEW
        r.width, r.height = int(width), int(height)
        for i in range(3):
            r.cam_pos[i], r.cam_right[i] = pass_ubo.cam_pos[i], pass_ubo.cam_right[i]
            r.cam_up[i], r.cam_fwd[i] = pass_ubo.cam_up[i], pass_ubo.cam_fwd[i]
        r.tan_half_fov = pass_ubo.tan_half_fov
        r.jitter[0], r.jitter[1] = pass_ubo.jitter[0], pass_ubo.jitter[1]
        r.d_hits = _dev(hits, 20 * int(width) * int(height), "hits").value
        return r

    @staticmethod
    def probe_request(origin, spacing, dims, n_rays, frame, reach, hits):
        r = Request()
        r.kind = VF_REQ_PROBES
        for i in range(3):
            r.origin[i] = float(origin[i])
            r.dims[i] = int(dims[i])
        r.spacing, r.n_rays, r.frame, r.reach = float(spacing), int(n_rays), int(frame), float(reach)
        r.d_hits = _dev(hits, 20 * int(np.prod([int(x) for x in dims])) * int(n_rays), "hits").value
        return r

    def trace_batch(self, requests, cfg=None):
        arr = (Request * len(requests))(*requests)
        self._ck(lib().vf_trace_batch(self._h, arr, len(requests), C.byref(cfg) if cfg is not None else None))

    def trace_batch_noalloc(self, requests, cfg=None):
        """vf_trace_batch_noalloc: the same batch with its counters in the library's per-GPU scratch (no
        device allocation)."""
        arr = (Request * len(requests))(*requests)
        self._ck(lib().vf_trace_batch_noalloc(self._h, arr, len(requests), C.byref(cfg) if cfg is not None else None))

    def batch_times(self, n):
        out = (C.c_double * n)()
        self._ck(lib().vf_read_batch_times(self._h, int(n), out))
        return list(out)

    def autotune_gi_batch(self, pass_ubo, width, height, reach, gb_pos, gb_normal, gi_hits, sun_hits):
        a = TuneGiArgs()
        a.pass_ = C.cast(C.pointer(pass_ubo), C.c_void_p)
        a.width, a.height, a.reach = int(width), int(height), float(reach)
        n = int(width) * int(height)
        a.d_gb_pos, a.d_gb_normal = _dev(gb_pos, 16 * n, "gb_pos").value, _dev(gb_normal, 4 * n, "gb_normal").value
        a.d_gi_hits, a.d_sun_hits = _dev(gi_hits, 20 * n, "gi_hits").value, _dev(sun_hits, 20 * n, "sun_hits").value
        return self._autotune(VF_TUNE_GI_BATCH, a)

    def autotune_batch(self, requests):
        arr = (Request * len(requests))(*requests)
        a = TuneBatchArgs()
        a.reqs = C.cast(arr, C.c_void_p)
        a.n_req = len(requests)
        self._keep = arr
        return self._autotune(VF_TUNE_BATCH, a)

    def _autotune(self, workload, args, base=None):
        best = QueueCfg()
        res = (TuneResult * 64)()
        n = C.c_int32()
        self._ck(lib().vf_autotune(self._h, workload, C.cast(C.pointer(args), C.c_void_p),
                                   C.byref(base) if base is not None else None, C.byref(best), res, 64, C.byref(n)))
        table = [((r.cfg.tile_w, r.cfg.tile_h), r.cfg.chunk, bool(r.cfg.sort), r.ms) for r in res[:n.value]]
        return best, table

    def trace_gi_path(self, pass_ubo, width, height, gb_pos, gb_normal, out_rgba, frame=0, spp=1, salt=0x68e31da4,
                      gi_reach=48.0, lights_nee=True, bounces=2):
        """Plain traced GI, two bounces (the reference's giPath) or three (mean of spp paths per pixel) into
        out_rgba (float32, 4 per pixel)."""
        n = int(width) * int(height)
        args = (self._h, C.byref(pass_ubo), int(width), int(height), float(gi_reach), int(bool(lights_nee)),
                _dev(gb_pos, 16 * n, "gb_pos"), _dev(gb_normal, 4 * n, "gb_normal"), int(frame) & 0xFFFFFFFF, int(spp),
                int(salt) & 0xFFFFFFFF)
        out = _dev(out_rgba, 16 * n, "out_rgba")
        if bounces == 2:
            self._ck(lib().vf_trace_gi_path(*args, out))
        else:
            self._ck(lib().vf_trace_gi_path_n(*args, int(bounces), out))

    # frame graphs
    def graph_begin(self):
        self._ck(lib().vf_graph_begin(self._h))

    def graph_end(self):
        g = C.c_void_p()
        self._ck(lib().vf_graph_end(self._h, C.byref(g)))
        return Graph(g.value)

    def graph_launch(self, graph):
        self._ck(lib().vf_graph_launch(self._h, graph._h))

    # measurement
    def set_counting(self, on):
        self._ck(lib().vf_set_counting(self._h, int(bool(on))))

    def counters(self, reset=True):
        out = (C.c_uint64 * 4)()
        self._ck(lib().vf_read_counters(self._h, out, int(bool(reset))))
        return {"rays": out[0], "outer": out[1], "voxel": out[2], "bytes": out[3]}

    def set_traversal(self, mode):
        """0: by launch size (default), 1: the plain traversal, 2: the small-frame traversal (same hits)."""
        self._ck(lib().vf_set_traversal(self._h, int(mode)))

    def set_timing(self, on):
        """Record HIP events around each trace call (for last_kernel_ms); off by default."""
        self._ck(lib().vf_set_timing(self._h, int(bool(on))))

    def last_kernel_ms(self):
        ms = C.c_float()
        self._ck(lib().vf_last_kernel_ms(self._h, C.byref(ms)))
        return ms.value


class GiCache:
    """The learned GI radiance cache (include/vf_gicache.h) on a Context."""

    N_PARAMS = GICACHE_PARAMS
    N_INPUTS = GICACHE_INPUTS

    @staticmethod
    def defaults():
        p = GiCacheParams()
        _check(lib().vf_gicache_defaults(C.byref(p)))
        return p

    def __init__(self, ctx, params=None, **overrides):
        p = params if params is not None else GiCache.defaults()
        for k, v in overrides.items():
            setattr(p, k, v)
        self.params = p
        self.ctx = ctx
        h = C.c_void_p()
        ctx._ck(lib().vf_gicache_create(ctx._h, C.byref(p), C.byref(h)))
        self._h = h

    def close(self):
        if self._h and self._h.value:
            lib().vf_gicache_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc):
        self.ctx._ck(rc)

    def reset(self):
        self._ck(lib().vf_gicache_reset(self.ctx._h, self._h))

    def frame(self, pass_ubo, width, height, gb_pos, gb_normal, out_rgba, frame):
        n = int(width) * int(height)
        self._ck(lib().vf_gicache_frame(self.ctx._h, self._h, C.byref(pass_ubo), int(width), int(height),
                                        _dev(gb_pos, 16 * n, "gb_pos"), _dev(gb_normal, 4 * n, "gb_normal"),
                                        int(frame) & 0xFFFFFFFF, _dev(out_rgba, 16 * n, "out_rgba")))

    def stats(self):
        s = GiCacheStats()
        self._ck(lib().vf_gicache_read_stats(self.ctx._h, self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in GiCacheStats._fields_}

    def samples(self, pass_ubo, width, height, gb_pos, gb_normal, frame, out):
        self._ck(lib().vf_gicache_samples(self.ctx._h, self._h, C.byref(pass_ubo), int(width), int(height),
                                          _ptr(gb_pos), _ptr(gb_normal), int(frame) & 0xFFFFFFFF, _ptr(out)))

    def get_params(self):
        out = np.zeros(GICACHE_PARAMS, np.float32)
        self._ck(lib().vf_gicache_get_params(self.ctx._h, self._h, out.ctypes.data))
        return out

    def get_render_params(self):
        """The weights the render uses (the exponential mean of the trained ones with weight_ema > 0)."""
        out = np.zeros(GICACHE_PARAMS, np.float32)
        self._ck(lib().vf_gicache_get_render_params(self.ctx._h, self._h, out.ctypes.data))
        return out

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == GICACHE_PARAMS
        self._ck(lib().vf_gicache_set_params(self.ctx._h, self._h, w.ctypes.data))

    def encode(self, recs, n, out):
        self._ck(lib().vf_gicache_encode(self.ctx._h, self._h, _ptr(recs), int(n), _ptr(out)))

    def infer(self, recs, n, out_h):
        self._ck(lib().vf_gicache_infer(self.ctx._h, self._h, _ptr(recs), int(n), _ptr(out_h)))

    def grad(self, samples, n, grad, loss):
        self._ck(lib().vf_gicache_grad(self.ctx._h, self._h, _ptr(samples), int(n), _ptr(grad), _ptr(loss)))

    def adam(self, grad):
        self._ck(lib().vf_gicache_adam(self.ctx._h, self._h, _ptr(grad)))

    def train(self, samples, n):
        self._ck(lib().vf_gicache_train(self.ctx._h, self._h, _ptr(samples), int(n)))


class Graph:
    """A recorded frame (vf_graph); replay with Context.graph_launch."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)

    def close(self):
        if self._h and self._h.value:
            lib().vf_graph_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def hits_to_numpy(t):
    """A torch uint8/int32 tensor holding vf_hit records -> structured numpy array."""
    raw = t.detach().cpu().contiguous().view(-1).numpy().view(np.uint8)
    return raw.view(HIT_DTYPE)
