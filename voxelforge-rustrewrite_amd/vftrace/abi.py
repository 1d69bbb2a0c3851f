"""ctypes mirrors of the C ABI structs (include/vf_trace.h, include/vf_worldgen.h, include/vf_gicache.h).

The field order and sizes follow the reference's #[repr(C)] / std140 structs: SceneUbo and PassUbo
(src/gfx/ubo.rs:10-131, offsets asserted at :76-86 and :132-139) and Material
(src/voxel/world.rs:21-31).  The asserts at the bottom of this module are the same ones.
"""
import ctypes as C

import numpy as np

f3 = C.c_float * 3
f4 = C.c_float * 4
i3 = C.c_int32 * 3
i4 = C.c_int32 * 4


class Material(C.Structure):
    _fields_ = [("r", C.c_float), ("g", C.c_float), ("b", C.c_float), ("emissive", C.c_float),
                ("props", f4), ("surf", f4)]


class SceneUbo(C.Structure):
    _fields_ = [
        ("atlas_bricks", C.c_uint32 * 3), ("night", C.c_float),
        ("world_bricks", i3), ("num_objects", C.c_int32),
        ("sun_dir", f3), ("grass_dist", C.c_float),
        ("sun_rgb", f3), ("sun_angle", C.c_float),
        ("fog_sun_rgb", f3), ("fog_density", C.c_float),
        ("fog_ambient", f3), ("fog_g", C.c_float),
        ("sky_zenith", f3), ("sky_scale", C.c_float),
        ("sky_horizon", f3), ("gi_scale", C.c_float),
        ("sunset_rgb", f3), ("sky_model", C.c_int32),
        ("fog_height", C.c_float), ("fog_falloff", C.c_float), ("fog_noise", C.c_float), ("fog_steps", C.c_int32),
        ("fog_shadow_max", C.c_float), ("tonemap_op", C.c_int32), ("saturation", C.c_float), ("contrast", C.c_float),
        ("white_balance", f3), ("grass_shadows", C.c_int32),
        ("sh_right", f3), ("sh_texel0", C.c_float),
        ("sh_up", f3), ("sh_texel1", C.c_float),
        ("sh_lz_top", C.c_float), ("sh_texel2", C.c_float), ("sh_n", C.c_int32), ("std_lighting", C.c_int32),
        ("sh_win_a", i4), ("sh_win_b", i4),
        ("probe_origin", i3), ("probe_spacing", C.c_float),
        ("probe_dims", i3), ("probe_update_offset", C.c_int32),
        ("probe_hyst", C.c_float), ("probe_normal_bias", C.c_float), ("std_soft_taps", C.c_float),
        ("std_cascade_blend", C.c_float),
        ("std_cascade_bias", C.c_int32), ("std_view_bias", C.c_float), ("sh_texel3", C.c_float),
        ("std_soft_min", C.c_float),
        ("probe_origin_c", i3), ("probe_spacing_c", C.c_float),
        ("world_win", i4), ("hor_a", i4), ("hor_b", i4),
        ("chunk_fade", C.c_uint32 * 256),
        ("gi_cache_cam", f3), ("gi_cache_lod", C.c_float),
        ("gi_cache_p", i4), ("gi_cache_q", f4), ("hor_c", i4),
        ("cloud_a", f4), ("cloud_b", f4), ("rain_a", f4),
    ]


class PassUbo(C.Structure):
    _fields_ = [
        ("cam_pos", f3), ("lock_active", C.c_float),
        ("cam_right", f3), ("fog_start", C.c_float),
        ("cam_up", f3), ("haze_k", C.c_float),
        ("cam_fwd", f3), ("tan_half_fov", C.c_float),
        ("prev_cam_pos", f3), ("prev_tan_half_fov", C.c_float),
        ("prev_cam_right", f3), ("haze_sea", C.c_float),
        ("prev_cam_up", f3), ("_pf", C.c_float),
        ("prev_cam_fwd", f3), ("frame_idx", C.c_uint32),
        ("exposure", C.c_float), ("u_step", C.c_int32), ("u_iteration", C.c_int32), ("debug_mode", C.c_int32),
        ("time", C.c_float), ("prev_time", C.c_float), ("jitter", C.c_float * 2),
        ("water_p", f4), ("look_p", f4), ("fog_p", f4), ("under_p", f4),
        ("opt_p", f4), ("wave_p", f4), ("shift_p", f4), ("facet_p", f4),
    ]


class Ray(C.Structure):
    _fields_ = [("o", f3), ("d", f3)]


class Hit(C.Structure):
    _fields_ = [("cell", i3), ("t", C.c_float), ("mat", C.c_uint8), ("face", C.c_uint8),
                ("hit", C.c_uint8), ("body", C.c_uint8)]


class ObjHeader(C.Structure):
    """std430 ObjHeader, 240 B (src/gfx/gpu_world.rs:34-47)."""
    _fields_ = [("world_to_obj", C.c_float * 16), ("obj_to_world", C.c_float * 16),
                ("prev_obj_to_world", C.c_float * 16), ("dims_bricks", i4), ("aabb_min", f4), ("aabb_max", f4)]


class BvhNode(C.Structure):
    """BvhNode, 32 B (shaders/common.glsl:77-81)."""
    _fields_ = [("min", f3), ("first", C.c_uint32), ("max", f3), ("count", C.c_uint32)]


class WorldDesc(C.Structure):
    _fields_ = [
        ("world_bricks", i3), ("world_win", i4),
        ("grid", C.POINTER(C.c_uint32)), ("atlas", C.POINTER(C.c_uint8)), ("n_slots", C.c_uint32),
        ("masks", C.POINTER(C.c_uint32)), ("coarse1", C.POINTER(C.c_uint8)), ("coarse2", C.POINTER(C.c_uint8)),
        ("palette", C.POINTER(Material)),
    ]


class WorldSrc(C.Structure):
    """vf_world_src: the world as the reference's CPU side holds it (grid with brick references, bricks)."""
    _fields_ = [
        ("world_bricks", i3), ("world_win", i4),
        ("grid", C.POINTER(C.c_uint32)), ("bricks", C.POINTER(C.c_uint8)), ("n_bricks", C.c_uint32),
        ("atlas_slots", C.c_uint32), ("palette", C.POINTER(Material)),
        ("brick_stride", C.c_uint32), ("slot_stride", C.c_uint32),
    ]


class Chunk(C.Structure):
    """vf_chunk: one 16 x wby x 16-brick chunk of a ring world for vf_world_sync_chunks."""
    _fields_ = [("cx", C.c_int32), ("cz", C.c_int32), ("loaded", C.c_int32), ("cells", C.POINTER(C.c_uint32))]


class GenParams(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("seed", C.c_uint32), ("world_bricks", i3), ("chunks", C.c_int32),
        ("octaves", C.c_int32), ("surface_mean", C.c_float), ("amplitude", C.c_float), ("period", C.c_float),
        ("water_level", C.c_int32), ("trees_per_chunk", C.c_int32), ("threads", C.c_int32),
    ]


class QueueCfg(C.Structure):
    _fields_ = [("tile_w", C.c_int32), ("tile_h", C.c_int32), ("chunk", C.c_int32), ("sort", C.c_int32),
                ("groups_per_cu", C.c_int32), ("refill", C.c_int32), ("traversal", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("kind", C.c_int32), ("width", C.c_int32), ("height", C.c_int32),
                ("cam_pos", f3), ("cam_right", f3), ("cam_up", f3), ("cam_fwd", f3), ("tan_half_fov", C.c_float),
                ("jitter", C.c_float * 2), ("origin", f3), ("spacing", C.c_float), ("dims", i3),
                ("n_rays", C.c_int32), ("frame", C.c_uint32), ("reach", C.c_float), ("d_hits", C.c_void_p)]


class TuneResult(C.Structure):
    _fields_ = [("cfg", QueueCfg), ("ms", C.c_float)]


class TuneGiArgs(C.Structure):
    _fields_ = [("pass_", C.c_void_p), ("width", C.c_int32), ("height", C.c_int32), ("reach", C.c_float),
                ("d_gb_pos", C.c_void_p), ("d_gb_normal", C.c_void_p), ("d_gi_hits", C.c_void_p),
                ("d_sun_hits", C.c_void_p)]


class TuneBatchArgs(C.Structure):
    _fields_ = [("reqs", C.c_void_p), ("n_req", C.c_int32)]


class GiCacheParams(C.Structure):
    _fields_ = [("terminate", C.c_int32), ("train_samples", C.c_int32), ("train_steps", C.c_int32), ("spp", C.c_int32),
                ("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float),
                ("loss_eps", C.c_float), ("fallback_ratio", C.c_float), ("warmup_frames", C.c_int32),
                ("enc_shift", C.c_int32), ("seed", C.c_uint32), ("salt", C.c_uint32), ("gi_reach", C.c_float),
                ("lights_nee", C.c_int32), ("bounces", C.c_int32), ("weight_ema", C.c_float)]


class GiCacheStats(C.Structure):
    _fields_ = [("frames", C.c_int64), ("steps", C.c_int64), ("last_loss", C.c_float), ("running_error", C.c_float),
                ("baseline", C.c_float), ("fallback", C.c_int32), ("n_queries", C.c_int32),
                ("n_valid_samples", C.c_int32)]


GICACHE_PARAMS = 23043
GICACHE_INPUTS = 96

VF_REQ_VIEW, VF_REQ_PROBES = 0, 1
VF_MAX_REQUESTS = 24   # requests per uneven batch (include/vf_trace.h)
VF_TUNE_GI_BATCH, VF_TUNE_BATCH = 0, 1

# numpy views of the 24-byte ray and 20-byte hit records
RAY_DTYPE = np.dtype([("o", "<f4", (3,)), ("d", "<f4", (3,))])
HIT_DTYPE = np.dtype([("cell", "<i4", (3,)), ("t", "<f4"), ("mat", "u1"), ("face", "u1"), ("hit", "u1"),
                      ("body", "u1")])
# 240-byte body header and 32-byte BVH node
OBJ_HEADER_DTYPE = np.dtype([("world_to_obj", "<f4", (16,)), ("obj_to_world", "<f4", (16,)),
                             ("prev_obj_to_world", "<f4", (16,)), ("dims_bricks", "<i4", (4,)),
                             ("aabb_min", "<f4", (4,)), ("aabb_max", "<f4", (4,))])
BVH_NODE_DTYPE = np.dtype([("min", "<f4", (3,)), ("first", "<u4"), ("max", "<f4", (3,)), ("count", "<u4")])
# 64-byte query / training record (vf_gi_record)
GI_RECORD_DTYPE = np.dtype([("pos", "<f4", (3,)), ("tag", "<u4"), ("n", "<f4", (3,)), ("pad0", "<f4"),
                            ("view", "<f4", (3,)), ("pad1", "<f4"), ("rgb", "<f4", (3,)), ("pad2", "<f4")])

# Layout asserts (src/gfx/ubo.rs:76-86, 132-139; src/voxel/world.rs:31).
assert C.sizeof(SceneUbo) == 1536
assert SceneUbo.fog_height.offset == 144
assert SceneUbo.sh_win_a.offset == 240
assert SceneUbo.probe_origin.offset == 272
assert SceneUbo.world_win.offset == 352
assert SceneUbo.chunk_fade.offset == 400
assert SceneUbo.gi_cache_cam.offset == 1424
assert SceneUbo.hor_c.offset == 1472
assert SceneUbo.rain_a.offset == 1520
assert C.sizeof(PassUbo) == 288
assert PassUbo.prev_tan_half_fov.offset == 76
assert PassUbo.prev_cam_fwd.offset == 112
assert PassUbo.frame_idx.offset == 124
assert PassUbo.jitter.offset == 152
assert PassUbo.facet_p.offset == 272
assert C.sizeof(Material) == 48
assert C.sizeof(Ray) == 24 and RAY_DTYPE.itemsize == 24
assert C.sizeof(Hit) == 20 and HIT_DTYPE.itemsize == 20
assert GI_RECORD_DTYPE.itemsize == 64
assert C.sizeof(GiCacheParams) == 72 and C.sizeof(GiCacheStats) == 40

assert C.sizeof(ObjHeader) == 240 and OBJ_HEADER_DTYPE.itemsize == 240
assert C.sizeof(BvhNode) == 32 and BVH_NODE_DTYPE.itemsize == 32
