"""ctypes wrapper of the CPU oracle (liboracle.so) — TEST INFRASTRUCTURE ONLY.

Used by tests/, the smoke check and the CPU-baseline leg of bench.py as the checker; the product
library (libvftrace.so) never imports, links or runs anything here.  See vf_oracle.h for what each
function restates (reference file:line) and how it is pinned.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

HIT_DTYPE = np.dtype([("cell", "<i4", (3,)), ("t", "<f4"), ("mat", "u1"), ("face", "u1"), ("hit", "u1"),
                      ("body", "u1")])


class VfoWorld(C.Structure):
    _fields_ = [("wb", C.c_int32 * 3), ("win", C.c_int32 * 4), ("grid", C.c_void_p), ("atlas", C.c_void_p),
                ("n_slots", C.c_uint32), ("dist", C.c_void_p), ("dry", C.c_void_p), ("coarse1", C.c_void_p),
                ("coarse2", C.c_void_p), ("c1", C.c_int32 * 3), ("c2", C.c_int32 * 3), ("use_dist", C.c_int32),
                ("coarse_skip", C.c_int32), ("n_bodies", C.c_int32), ("bodies", C.c_void_p), ("obj_grid", C.c_void_p),
                ("bvh", C.c_void_p), ("n_nodes", C.c_int32), ("bvh_prims", C.c_void_p), ("use_bvh", C.c_int32),
                ("palette", C.c_void_p)]


class VfoHit(C.Structure):
    _fields_ = [("hit", C.c_int32), ("t", C.c_float), ("n", C.c_int32 * 3), ("mat", C.c_uint32),
                ("cell", C.c_int32 * 3), ("body", C.c_int32)]


class VfoCount(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("outer", C.c_uint64), ("voxel", C.c_uint64), ("bytes", C.c_uint64)]

    def as_dict(self):
        return {"rays": self.rays, "outer": self.outer, "voxel": self.voxel, "bytes": self.bytes}


class VfoCam(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("right", C.c_float * 3), ("up", C.c_float * 3), ("fwd", C.c_float * 3),
                ("tan_half_fov", C.c_float), ("jitter", C.c_float * 2), ("frame_idx", C.c_uint32)]


class VfoTail(C.Structure):
    _fields_ = [("valid", C.c_int32), ("p", C.c_float * 3), ("n", C.c_float * 3), ("dir", C.c_float * 3),
                ("thr", C.c_float * 3), ("body", C.c_int32)]


# 240-byte body header (ObjHeader, common.glsl:65-72; the C ABI's vf_obj_header)
OBJ_HEADER_DTYPE = np.dtype([("world_to_obj", "<f4", (16,)), ("obj_to_world", "<f4", (16,)),
                             ("prev_obj_to_world", "<f4", (16,)), ("dims_bricks", "<i4", (4,)),
                             ("aabb_min", "<f4", (4,)), ("aabb_max", "<f4", (4,))])

# 64-byte query / training record (vfo_record = the C ABI's vf_gi_record)
RECORD_DTYPE = np.dtype([("pos", "<f4", (3,)), ("tag", "<u4"), ("n", "<f4", (3,)), ("pad0", "<f4"),
                         ("view", "<f4", (3,)), ("pad1", "<f4"), ("rgb", "<f4", (3,)), ("pad2", "<f4")])


class VfoSky(C.Structure):
    _fields_ = [("sun_dir", C.c_float * 3), ("sun_rgb", C.c_float * 3), ("sky_zenith", C.c_float * 3),
                ("sky_horizon", C.c_float * 3), ("sunset_rgb", C.c_float * 3), ("sky_scale", C.c_float),
                ("gi_scale", C.c_float), ("night", C.c_float), ("sky_model", C.c_int32), ("gi_reach", C.c_float),
                ("lights_nee", C.c_int32), ("palette", C.c_void_p), ("skytop", C.c_void_p)]


_L = None


def build():
    """Compiles liboracle.so with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        sig = {
            "vfo_trace": (None, [P, P, P, C.c_float, C.c_uint32, P, P]),
            "vfo_trace_batch": (None, [P, P, C.c_int64, C.c_float, C.c_uint32, P, P, C.c_int]),
            "vfo_trace_box_batch": (None, [P, P, C.c_int64, C.c_float, C.c_uint32, C.c_float, C.c_int, P, P, C.c_int]),
            "vfo_march_f64": (None, [P, P, P, C.c_double, C.c_uint32, P, P, P, P, P]),
            "vfo_raycast_cpu": (None, [P, P, P, C.c_float, C.c_uint32, P, P, P, P, P]),
            "vfo_voxel": (C.c_uint32, [P, C.c_int32, C.c_int32, C.c_int32]),
            "vfo_brick_mask": (None, [P, P]),
            "vfo_coarse": (None, [P, P, P, P]),
            "vfo_brickdist_pass": (None, [P, P, P, P, P, P, P, C.c_int32, C.c_int32]),
            "vfo_record_distance_field": (None, [P, P, P, C.c_int32, P, P, P, P, P, P, P, P]),
            "vfo_brickdist_brute": (None, [P, P, P, C.c_int32, C.c_int32, P]),
            "vfo_skytop": (None, [P, P, P]),
            "vfo_pcg": (C.c_uint32, [C.c_uint32]),
            "vfo_pixel_seed": (C.c_uint32, [C.c_int32, C.c_int32, C.c_uint32, C.c_uint32]),
            "vfo_ld_sample2": (None, [C.c_int32, C.c_int32, C.c_uint32, C.c_uint32, P]),
            "vfo_cosine_hemisphere_u": (None, [P, P, P]),
            "vfo_sin": (C.c_float, [C.c_float]),
            "vfo_cos": (C.c_float, [C.c_float]),
            "vfo_primary_dir": (None, [P, P, P, C.c_float, C.c_int32, C.c_int32, C.c_int32, C.c_int32, P, P]),
            "vfo_probe_dir": (None, [C.c_int32, C.c_int32, C.c_uint32, P]),
            "vfo_pass_primary": (None, [P, P, C.c_int32, C.c_int32, P, P, P, P, P, C.c_int]),
            "vfo_build_bvh": (None, [P, C.c_int32, P, P, P]),
            "vfo_trace_objects": (None, [P, P, P, P, C.c_float, C.c_int32, P]),
            "vfo_trace_scene": (None, [P, P, P, C.c_float, C.c_uint32, P, P]),
            "vfo_trace_scene_batch": (None, [P, P, C.c_int64, C.c_float, C.c_uint32, P, P, C.c_int]),
            "vfo_hit_normal": (None, [P, P, P]),
            "vfo_pass_primary_rows": (None, [P, P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, P, P, P, P,
                                              C.c_int]),
            "vfo_pass_shadow": (None, [P, P, P, C.c_float, C.c_int32, C.c_int32, P, P, P, P, P, C.c_int]),
            "vfo_sun_visibility_t": (C.c_float, [P, P, P, P, C.c_float, P, C.c_int32, P, P, P]),
            "vfo_sample_cone": (None, [P, C.c_float, P, P]),
            "vfo_pass_sun_cone": (None, [P, P, P, C.c_float, C.c_float, C.c_uint32, C.c_int32, C.c_int32, P, P, P, P,
                                         C.c_int]),
            "vfo_pass_gi": (None, [P, P, C.c_uint32, C.c_int32, C.c_int32, C.c_int32, P, P, P, P, P, C.c_int]),
            "vfo_gi_path": (None, [P, P, P, P, P, P, P, P]),
            "vfo_sky_open": (C.c_float, [P, P, P]),
            "vfo_gi_rays": (None, [C.c_int32, C.c_int32, C.c_uint32, P, P, P, P]),
            "vfo_cosine_hemisphere": (None, [P, P, P]),
            "vfo_voxel_mat_at": (C.c_uint32, [P, P]),
            "vfo_exit_solid": (C.c_float, [P, P, P, C.c_int32]),
            "vfo_gi_path_ex": (None, [P, P, P, P, P, P, C.c_int32, C.c_int32, P, P, P, P]),
            "vfo_gi_path3": (None, [P, P, P, P, P, P, P, P]),
            "vfo_pass_gi_path": (None, [P, P, C.c_int32, C.c_int32, P, P, C.c_uint32, C.c_int32, C.c_uint32,
                                        C.c_int32, P, C.c_int]),
            "vfo_gicache_prepass": (None, [P, P, C.c_int32, C.c_int32, P, P, C.c_uint32, C.c_int32, C.c_uint32,
                                           C.c_int32, P, P, C.c_int]),
            "vfo_gicache_samples": (None, [P, P, P, C.c_int32, C.c_int32, P, P, C.c_uint32, C.c_int32, C.c_int32,
                                           C.c_uint32, C.c_int32, P, C.c_int]),
            "vfo_exp": (C.c_float, [C.c_float]),
            "vfo_pow16": (C.c_float, [C.c_float]),
            "vfo_pow64": (C.c_float, [C.c_float]),
            "vfo_f16_to_f32": (C.c_float, [C.c_uint16]),
            "vfo_f32_to_f16": (C.c_uint16, [C.c_float]),
            "vfo_gbuffer_motion": (None, [C.c_int32, C.c_int32, P, P, P, C.c_float, P, P, P, P, P, C.c_float, P, P, P,
                                          C.c_int32, P, P]),
            "vfo_gi_accumulate": (None, [C.c_int32, C.c_int32, C.c_uint32, P, C.c_int32] + [P] * 12),
            "vfo_light_accumulate": (None, [C.c_int32, C.c_int32, P] + [P] * 9),
            "vfo_gi_atrous_level": (None, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32] + [P] * 7),
            "vfo_gi_atrous": (None, [C.c_int32, C.c_int32, C.c_int32, C.c_int32] + [P] * 7),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _L = L
    return _L


def _p(a):
    return a.ctypes.data if a is not None and a.size else None


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class OracleWorld:
    """The oracle's own copy of a world: its grid and atlas plus the derived fields it builds itself
    (coarse pyramid, both distance fields by record_distance_field, sky-top map)."""

    def __init__(self, arrays, use_dist=True, coarse_skip=True, build_fields=True, build_skytop=True):
        self.wb = tuple(int(x) for x in arrays["world_bricks"])
        self.win = tuple(int(x) for x in arrays["world_win"])
        self.grid = np.ascontiguousarray(arrays["grid"], dtype=np.uint32)
        self.atlas = np.ascontiguousarray(arrays["atlas"], dtype=np.uint8)
        if self.atlas.size == 0:
            self.atlas = np.zeros(512, np.uint8)
        self.atlas = np.where(self.atlas == 254, 0, self.atlas).astype(np.uint8)   # gpu_world.rs:58-67
        self.palette = np.ascontiguousarray(arrays["palette"], dtype=np.float32).reshape(256, 12)
        n = self.wb[0] * self.wb[1] * self.wb[2]
        self.c1d = tuple((x + 3) // 4 for x in self.wb)
        self.c2d = tuple((x + 15) // 16 for x in self.wb)
        self.c1 = np.zeros(int(np.prod(self.c1d)), np.uint8)
        self.c2 = np.zeros(int(np.prod(self.c2d)), np.uint8)
        wb = (C.c_int32 * 3)(*self.wb)
        lib().vfo_coarse(_p(self.grid), wb, _p(self.c1), _p(self.c2))
        self.d = [np.zeros(n, np.uint8) for _ in range(4)]
        if build_fields:
            self.rebuild_fields()
        self.s = VfoWorld()
        self.s.wb[:] = self.wb
        self.s.win[:] = self.win
        self.s.grid = self.grid.ctypes.data
        self.s.atlas = self.atlas.ctypes.data
        self.s.n_slots = self.atlas.size // 512
        self.s.dist = self.d[0].ctypes.data
        self.s.dry = self.d[2].ctypes.data
        self.s.coarse1 = self.c1.ctypes.data
        self.s.coarse2 = self.c2.ctypes.data
        self.s.c1[:] = self.c1d
        self.s.c2[:] = self.c2d
        self.s.use_dist = int(bool(use_dist))
        self.s.coarse_skip = int(bool(coarse_skip))
        self.s.palette = self.palette.ctypes.data
        self.skytop = np.zeros(self.wb[0] * 8 * self.wb[2] * 8, np.uint16)
        if build_skytop:
            kinds = np.ascontiguousarray(self.palette[:, 6])
            lib().vfo_skytop(C.byref(self.s), _p(kinds), _p(self.skytop))

    def ring(self):
        return self.win[3] != 0

    # bodies (common.glsl:65-82, 664-793)
    def set_bodies(self, headers, grid, use_bvh=True, bvh=None):
        """headers: OBJ_HEADER_DTYPE records (or float32 [n, 60]); grid: the body grids (u32).  The BVH is
        built by the oracle's build_bvh unless bvh = (nodes float32 [m, 8], prims u32) is given."""
        h = np.ascontiguousarray(headers)
        self.bodies = np.ascontiguousarray(h.view(np.float32).reshape(-1, 60))
        n = self.bodies.shape[0]
        self.obj_grid = np.ascontiguousarray(grid, dtype=np.uint32)
        if self.obj_grid.size == 0:
            self.obj_grid = np.zeros(1, np.uint32)
        if bvh is None:
            nodes = np.zeros((max(2 * n - 1, 1), 8), np.float32)
            prims = np.zeros(max(n, 1), np.uint32)
            nn = C.c_int32()
            lib().vfo_build_bvh(_p(self.bodies), int(n), _p(nodes), C.byref(nn), _p(prims))
            self.bvh_nodes, self.bvh_prims = nodes[:nn.value].copy(), prims[:n].copy()
        else:
            self.bvh_nodes = np.ascontiguousarray(bvh[0], dtype=np.float32).reshape(-1, 8)
            self.bvh_prims = np.ascontiguousarray(bvh[1], dtype=np.uint32)
        self.s.n_bodies = int(n)
        self.s.bodies = self.bodies.ctypes.data
        self.s.obj_grid = self.obj_grid.ctypes.data
        self.s.bvh = self.bvh_nodes.ctypes.data if self.bvh_nodes.size else None
        self.s.n_nodes = int(self.bvh_nodes.shape[0])
        self.s.bvh_prims = self.bvh_prims.ctypes.data if self.bvh_prims.size else None
        self.s.use_bvh = int(bool(use_bvh))

    def use_bvh(self, on):
        self.s.use_bvh = int(bool(on))

    def trace_scene_one(self, ro, rd, t_cap=1e30, skip_mat=0):
        """traceSceneS for one ray -> VfoHit (body = index + 1, 0 for the world)."""
        h = VfoHit()
        ro32, rd32 = f32(ro), f32(rd)
        lib().vfo_trace_scene(C.byref(self.s), _p(ro32), _p(rd32), float(t_cap), int(skip_mat), C.byref(h), None)
        return h

    def trace_scene_rays(self, rays, t_cap=1e30, skip_mat=0, threads=0):
        """traceSceneS per ray: (hits, body index per ray, -1 = world or miss)."""
        r = f32(rays).reshape(-1, 6)
        out = np.zeros(r.shape[0], HIT_DTYPE)
        body = np.zeros(r.shape[0], np.int32)
        lib().vfo_trace_scene_batch(C.byref(self.s), _p(r), r.shape[0], float(t_cap), int(skip_mat), _p(out), _p(body),
                                    int(threads))
        return out, body

    def objects_only(self, ro, rd, limit, any_hit=False):
        """traceObjectsA from an empty hit: VfoHit."""
        h = VfoHit()
        h.t = -1.0
        ro32, rd32 = f32(ro), f32(rd)
        lib().vfo_trace_objects(C.byref(self.s), _p(ro32), _p(rd32), C.byref(h), float(limit), int(bool(any_hit)), None)
        return h

    def hit_normal(self, h):
        out = (C.c_float * 3)()
        lib().vfo_hit_normal(C.byref(self.s), C.byref(h), out)
        return np.array(out[:], np.float32)

    def full_box(self):
        lo = [self.win[0] if self.ring() else 0, 0, self.win[2] if self.ring() else 0]
        hi = [lo[0] + self.wb[0] - 1, self.wb[1] - 1, lo[2] + self.wb[2] - 1]
        return lo, hi

    def rebuild_fields(self, lo=None, hi=None, dlo=None, dhi=None):
        flo, fhi = self.full_box()
        lo = flo if lo is None else lo
        hi = fhi if hi is None else hi
        dlo = lo if dlo is None else dlo
        dhi = hi if dhi is None else dhi
        i3 = lambda v: (C.c_int32 * 3)(*[int(x) for x in v])
        lib().vfo_record_distance_field(_p(self.grid), _p(self.atlas), i3(self.wb), int(self.ring()), i3(lo), i3(hi),
                                        i3(dlo), i3(dhi), _p(self.d[0]), _p(self.d[1]), _p(self.d[2]),
                                        _p(self.d[3]))

    def voxel(self, x, y, z):
        return int(lib().vfo_voxel(C.byref(self.s), int(x), int(y), int(z)))

    def trace_rays(self, rays, t_cap=1e30, skip_mat=0, threads=0):
        """rays: float32 [N, 6] (origin, direction) -> (hits, counters)."""
        r = f32(rays).reshape(-1, 6)
        out = np.zeros(r.shape[0], HIT_DTYPE)
        cnt = VfoCount()
        lib().vfo_trace_batch(C.byref(self.s), _p(r), r.shape[0], float(t_cap), int(skip_mat), _p(out), C.byref(cnt),
                              int(threads))
        return out, cnt.as_dict()

    def trace_rays_box(self, rays, t_cap=1e30, skip_mat=0, margin=0.05, mode=0, threads=0):
        """trace_rays with the library's small-frame mask test before each walk -> (hits, walks skipped,
        walks run); mode 1 plants a mistake in the test (the entry voxel left out of the box)."""
        r = f32(rays).reshape(-1, 6)
        out = np.zeros(r.shape[0], HIT_DTYPE)
        st = np.zeros(2, np.int64)
        lib().vfo_trace_box_batch(C.byref(self.s), _p(r), r.shape[0], float(t_cap), int(skip_mat), float(margin),
                                  int(mode), _p(out), _p(st), int(threads))
        return out, int(st[0]), int(st[1])

    def trace_one(self, ro, rd, t_cap=1e30, skip_mat=0):
        h = VfoHit()
        cnt = VfoCount()
        ro32, rd32 = f32(ro), f32(rd)
        lib().vfo_trace(C.byref(self.s), _p(ro32), _p(rd32), float(t_cap), int(skip_mat), C.byref(h), C.byref(cnt))
        return h, cnt.as_dict()

    def march_f64(self, ro, rd, t_cap=1e30, skip_mat=0):
        hit, t, mat = C.c_int32(), C.c_double(), C.c_uint32()
        cell, n = (C.c_int32 * 3)(), (C.c_int32 * 3)()
        ro64 = np.ascontiguousarray(ro, dtype=np.float64)
        rd64 = np.ascontiguousarray(rd, dtype=np.float64)
        lib().vfo_march_f64(C.byref(self.s), _p(ro64), _p(rd64), float(t_cap), int(skip_mat), C.byref(hit),
                            C.byref(t), cell, n, C.byref(mat))
        return bool(hit.value), t.value, tuple(cell), tuple(n), mat.value

    def raycast_cpu(self, ro, rd, max_t, skip_mat=0):
        hit, t, mat = C.c_int32(), C.c_float(), C.c_uint32()
        cell, n = (C.c_int32 * 3)(), (C.c_int32 * 3)()
        ro32, rd32 = f32(ro), f32(rd)
        lib().vfo_raycast_cpu(C.byref(self.s), _p(ro32), _p(rd32), float(max_t), int(skip_mat), C.byref(hit),
                              C.byref(t), cell, n, C.byref(mat))
        return bool(hit.value), t.value, tuple(cell), tuple(n), mat.value

    def pass_primary(self, cam, width, height, threads=0):
        c = VfoCam()
        c.pos[:] = [float(x) for x in cam["pos"]]
        c.right[:] = [float(x) for x in cam["right"]]
        c.up[:] = [float(x) for x in cam["up"]]
        c.fwd[:] = [float(x) for x in cam["fwd"]]
        c.tan_half_fov = float(cam["tan_half_fov"])
        c.jitter[:] = [float(x) for x in cam.get("jitter", (0.0, 0.0))]
        n = width * height
        hits = np.zeros(n, HIT_DTYPE)
        pos = np.zeros((n, 4), np.float32)
        nrm = np.zeros(n, np.uint32)
        cnt = VfoCount()
        body = np.zeros(n, np.int32)
        lib().vfo_pass_primary(C.byref(self.s), C.byref(c), int(width), int(height), _p(hits), _p(pos), _p(nrm),
                               _p(body), C.byref(cnt), int(threads))
        self.last_body = body
        return hits, pos, nrm, cnt.as_dict()

    def pass_primary_rows(self, cam, width, height, y0, ystep, nrows, threads=0):
        c = VfoCam()
        c.pos[:] = [float(x) for x in cam["pos"]]
        c.right[:] = [float(x) for x in cam["right"]]
        c.up[:] = [float(x) for x in cam["up"]]
        c.fwd[:] = [float(x) for x in cam["fwd"]]
        c.tan_half_fov = float(cam["tan_half_fov"])
        n = width * nrows
        hits = np.zeros(n, HIT_DTYPE)
        pos = np.zeros((n, 4), np.float32)
        nrm = np.zeros(n, np.uint32)
        cnt = VfoCount()
        lib().vfo_pass_primary_rows(C.byref(self.s), C.byref(c), int(width), int(height), int(y0), int(ystep),
                                    int(nrows), _p(hits), _p(pos), _p(nrm), C.byref(cnt), int(threads))
        return hits, pos, nrm, cnt.as_dict()

    def pass_shadow(self, sun, width, height, gb_pos, gb_normal, threads=0, leaf_light=0.6):
        """shadow_trace.comp with FOLIAGE_LIGHT (leaf_light = lookP.w, the reference default 0.6)."""
        n = width * height
        sh = np.zeros((n, 2), np.float32)
        hits = np.zeros(n, HIT_DTYPE)
        cnt = VfoCount()
        sun32, pos32, nrm = f32(sun), f32(gb_pos), np.ascontiguousarray(gb_normal, dtype=np.uint32)
        lib().vfo_pass_shadow(C.byref(self.s), _p(self.palette), _p(sun32), float(leaf_light), int(width), int(height),
                              _p(pos32), _p(nrm), _p(sh), _p(hits), C.byref(cnt), int(threads))
        return sh, hits, cnt.as_dict()

    def pass_sun_cone(self, sun, sun_angle, leaf_light, frame, width, height, gb_pos, gb_normal, threads=0):
        """lighting.comp's stochastic cone-jittered sun sample per pixel, and the first segments' hits."""
        n = width * height
        out = np.zeros(n, np.float32)
        hits = np.zeros(n, HIT_DTYPE)
        sun32, pos32, nrm = f32(sun), f32(gb_pos), np.ascontiguousarray(gb_normal, dtype=np.uint32)
        lib().vfo_pass_sun_cone(C.byref(self.s), _p(self.palette), _p(sun32), float(sun_angle), float(leaf_light),
                                int(frame) & 0xFFFFFFFF, int(width), int(height), _p(pos32), _p(nrm), _p(out), _p(hits),
                                int(threads))
        return out, hits

    def sun_visibility(self, o, d, leaf_light, seed=0, stochastic=False):
        """sunVisibilityT: (visibility, tFirst, seed after)."""
        o32, d32 = f32(o), f32(d)
        sd = C.c_uint32(int(seed) & 0xFFFFFFFF)
        tf = C.c_float()
        v = lib().vfo_sun_visibility_t(C.byref(self.s), _p(self.palette), _p(o32), _p(d32), float(leaf_light),
                                       C.byref(sd), int(bool(stochastic)), C.byref(tf), None, None)
        return float(v), float(tf.value), sd.value

    def sky(self, scene, gi_reach=48.0, lights_nee=True):
        s = VfoSky()
        for k in ("sun_dir", "sun_rgb", "sky_zenith", "sky_horizon", "sunset_rgb"):
            getattr(s, k)[:] = [float(x) for x in getattr(scene, k)]
        s.sky_scale, s.gi_scale, s.night = scene.sky_scale, scene.gi_scale, scene.night
        s.sky_model = scene.sky_model
        s.gi_reach = float(gi_reach)
        s.lights_nee = int(bool(lights_nee))
        s.palette = self.palette.ctypes.data
        s.skytop = self.skytop.ctypes.data
        self._sky_keep = s
        return s

    def pass_gi(self, scene, frame, width, height, gb_pos, gb_normal, half_res=True, gi_reach=48.0, lights_nee=True,
                threads=0):
        s = self.sky(scene, gi_reach, lights_nee)
        ow, oh = ((width + 1) // 2, (height + 1) // 2) if half_res else (width, height)
        rad = np.zeros((ow * oh, 4), np.float32)
        bounce = np.zeros(ow * oh, HIT_DTYPE)
        cnt = VfoCount()
        pos32, nrm = f32(gb_pos), np.ascontiguousarray(gb_normal, dtype=np.uint32)
        lib().vfo_pass_gi(C.byref(self.s), C.byref(s), int(frame), int(width), int(height), int(bool(half_res)),
                          _p(pos32), _p(nrm), _p(rad), _p(bounce), C.byref(cnt), int(threads))
        return rad, bounce, cnt.as_dict()


    def gi_path(self, scene, P, N, d, seed, last_stage=3, trace_on=True, tail=False, gi_reach=48.0, lights_nee=True):
        """One giPathImpl path; returns (L, seed after, tail dict or None)."""
        s = self.sky(scene, gi_reach, lights_nee)
        P32, N32, d32 = f32(P), f32(N), f32(d)
        L = np.zeros(3, np.float32)
        sd = C.c_uint32(int(seed) & 0xFFFFFFFF)
        t = VfoTail()
        lib().vfo_gi_path_ex(C.byref(self.s), C.byref(s), _p(P32), _p(N32), _p(d32), C.byref(sd), int(last_stage),
                             int(bool(trace_on)), _p(L), None, C.byref(t) if tail else None, None)
        tl = None
        if tail:
            tl = {"valid": bool(t.valid), "p": np.array(t.p[:], np.float32), "n": np.array(t.n[:], np.float32),
                  "dir": np.array(t.dir[:], np.float32), "thr": np.array(t.thr[:], np.float32)}
        return L, sd.value, tl

    def gi_path3(self, scene, P, N, d, seed, gi_reach=48.0, lights_nee=True):
        """One three-bounce path (vfo_gi_path3); returns (L, seed after)."""
        s = self.sky(scene, gi_reach, lights_nee)
        P32, N32, d32 = f32(P), f32(N), f32(d)
        L = np.zeros(3, np.float32)
        sd = C.c_uint32(int(seed) & 0xFFFFFFFF)
        lib().vfo_gi_path3(C.byref(self.s), C.byref(s), _p(P32), _p(N32), _p(d32), C.byref(sd), _p(L), None)
        return L, sd.value

    def pass_gi_path(self, scene, width, height, gb_pos, gb_normal, frame, spp, salt, gi_reach=48.0, lights_nee=True,
                     threads=0, bounces=2):
        """Plain traced GI, two bounces (giPath) or three: RGBA per pixel (mean of spp paths)."""
        s = self.sky(scene, gi_reach, lights_nee)
        out = np.zeros((width * height, 4), np.float32)
        pos32, nrm = f32(gb_pos), np.ascontiguousarray(gb_normal, dtype=np.uint32)
        lib().vfo_pass_gi_path(C.byref(self.s), C.byref(s), int(width), int(height), _p(pos32), _p(nrm),
                               int(frame) & 0xFFFFFFFF, int(spp), int(salt) & 0xFFFFFFFF, int(bounces), _p(out),
                               int(threads))
        return out

    def gicache_prepass(self, scene, width, height, gb_pos, gb_normal, frame, spp, salt, gi_reach=48.0,
                        lights_nee=True, threads=0, bounces=2):
        """The cached render's traced part: (pre RGBA per pixel, query record per pixel*spp slot)."""
        s = self.sky(scene, gi_reach, lights_nee)
        pre = np.zeros((width * height, 4), np.float32)
        q = np.zeros(width * height * spp, RECORD_DTYPE)
        pos32, nrm = f32(gb_pos), np.ascontiguousarray(gb_normal, dtype=np.uint32)
        lib().vfo_gicache_prepass(C.byref(self.s), C.byref(s), int(width), int(height), _p(pos32), _p(nrm),
                                  int(frame) & 0xFFFFFFFF, int(spp), int(salt) & 0xFFFFFFFF, int(bounces), _p(pre),
                                  _p(q), int(threads))
        return pre, q

    def gicache_samples(self, scene, cam_pos, width, height, gb_pos, gb_normal, frame, n, terminate, salt,
                        gi_reach=48.0, lights_nee=True, threads=0, bounces=2):
        """The training records of one frame."""
        s = self.sky(scene, gi_reach, lights_nee)
        out = np.zeros(n, RECORD_DTYPE)
        cam = f32(cam_pos)
        pos32, nrm = f32(gb_pos), np.ascontiguousarray(gb_normal, dtype=np.uint32)
        lib().vfo_gicache_samples(C.byref(self.s), C.byref(s), _p(cam), int(width), int(height), _p(pos32), _p(nrm),
                                  int(frame) & 0xFFFFFFFF, int(n), int(terminate), int(salt) & 0xFFFFFFFF,
                                  int(bounces), _p(out), int(threads))
        return out


    def voxel_mat_at(self, p):
        p32 = f32(p)
        return int(lib().vfo_voxel_mat_at(C.byref(self.s), _p(p32)))

    def exit_solid(self, ro, rd, max_steps=2048):
        """exitSolid (common.glsl:990-1016)."""
        ro32, rd32 = f32(ro), f32(rd)
        return float(lib().vfo_exit_solid(C.byref(self.s), _p(ro32), _p(rd32), int(max_steps)))


def face_code(n):
    """Hit face code of an integer normal (vf_hit.face): 1 +x, 2 -x, 3 +y, 4 -y, 5 +z, 6 -z, 0 none."""
    for a in range(3):
        if n[a]:
            return 2 * a + (1 if n[a] > 0 else 2)
    return 0


def pcg(v):
    return int(lib().vfo_pcg(int(v) & 0xFFFFFFFF))


def cosine_hemisphere(n, seed):
    """cosineHemisphere(n, seed) (common.glsl:1271-1274): (direction, seed after)."""
    n32 = f32(n)
    out = np.zeros(3, np.float32)
    sd = C.c_uint32(int(seed) & 0xFFFFFFFF)
    lib().vfo_cosine_hemisphere(_p(n32), C.byref(sd), _p(out))
    return out, sd.value


def brick_mask(vox):
    v = np.ascontiguousarray(vox, dtype=np.uint8).reshape(512)
    out = np.zeros(16, np.uint32)
    lib().vfo_brick_mask(_p(v), _p(out))
    return out


def brickdist_brute(grid, atlas, wb, ring, dry):
    out = np.zeros(int(np.prod(wb)), np.uint8)
    g = np.ascontiguousarray(grid, dtype=np.uint32)
    a = np.ascontiguousarray(atlas, dtype=np.uint8)
    if a.size == 0:
        a = np.zeros(512, np.uint8)
    lib().vfo_brickdist_brute(_p(g), _p(a), (C.c_int32 * 3)(*wb), int(bool(ring)), int(bool(dry)), _p(out))
    return out


def primary_dirs(cam, width, height):
    out = np.zeros((height * width, 3), np.float32)
    jit = np.zeros(2, np.float32)
    L = lib()
    R, U, F = f32(cam["right"]), f32(cam["up"]), f32(cam["fwd"])
    tmp = np.zeros(3, np.float32)
    for py in range(height):
        for px in range(width):
            L.vfo_primary_dir(_p(R), _p(U), _p(F), float(cam["tan_half_fov"]), width, height, px, py, _p(jit), _p(tmp))
            out[py * width + px] = tmp
    return out


def probe_dirs(n_rays, frame):
    out = np.zeros((n_rays, 3), np.float32)
    tmp = np.zeros(3, np.float32)
    for i in range(n_rays):
        lib().vfo_probe_dir(i, n_rays, int(frame), _p(tmp))
        out[i] = tmp
    return out


def gi_rays(width, height, frame, gb_pos, gb_normal):
    """GI batch rays (float32 [N, 6]) and the non-sky mask of a full-resolution G-buffer."""
    rays = np.zeros((width * height, 6), np.float32)
    valid = np.zeros(width * height, np.uint8)
    pos32 = f32(gb_pos)
    nrm = np.ascontiguousarray(gb_normal, dtype=np.uint32)
    lib().vfo_gi_rays(int(width), int(height), int(frame), _p(pos32), _p(nrm), _p(rays), _p(valid))
    return rays, valid.astype(bool)


def _body_record(b):
    return ((b.cell[0], b.cell[1], b.cell[2]), b.t, b.mat, face_code(b.n), 1, 1)


def gi_batch(ow, width, height, frame, gb_pos, gb_normal, sun, reach, return_counts=False, bodies=False):
    """The GI batch on the oracle: GI rays (t_cap = reach, skipMat 0), then for each hit that is not
    water and faces the sun, its sun ray from hit + N*0.03 (giPathImpl stage 1).  Per-pixel hits.
    bodies (GI_OBJECTS, common.glsl:780, 2267-2271): a GI ray also meets the bodies nearer than min(its
    world hit or reach, 16), its sun origin then uses the body's rotated normal, and a sun ray that missed
    the world meets any body (counters then cover the world walks only)."""
    rays, valid = gi_rays(width, height, frame, gb_pos, gb_normal)
    gi = np.zeros(width * height, HIT_DTYPE)
    gi["t"] = -1.0
    idx = np.nonzero(valid)[0]
    h, cnt = ow.trace_rays(rays[idx], reach, 0)
    sun = np.asarray(sun, np.float32)
    face = h["face"].astype(np.int64)
    hn = np.zeros((len(h), 3), np.float32)
    ax = (face - 1) // 2
    sgn = np.where(face % 2 == 1, 1.0, -1.0).astype(np.float32)
    has = face > 0
    hn[np.nonzero(has)[0], ax[has]] = sgn[has]
    if bodies:
        for k in range(len(h)):
            lim = min(float(h["t"][k]) if h["hit"][k] else float(reach), 16.0)
            b = ow.objects_only(rays[idx[k], :3], rays[idx[k], 3:], lim)
            if b.hit:
                h[k] = _body_record(b)
                hn[k] = ow.hit_normal(b)
    gi[idx] = h
    with np.errstate(invalid="ignore"):
        ndl = (hn[:, 0] * sun[0] + hn[:, 1] * sun[1]) + hn[:, 2] * sun[2]
        emit = (h["hit"] == 1) & (h["mat"] != 5) & (np.maximum(ndl, np.float32(0.0)) > 0)
    o = rays[idx, :3]
    d = rays[idx, 3:]
    t = h["t"].astype(np.float32)[:, None]
    hp = o + d * t
    so = hp + hn * np.float32(3e-2)
    sun_rays = np.concatenate([so, np.broadcast_to(sun, so.shape)], 1)[emit].astype(np.float32)
    sh = np.zeros(width * height, HIT_DTYPE)
    sh["t"] = -1.0
    if len(sun_rays):
        s_hits, c2 = ow.trace_rays(sun_rays, 1e30, 0)
        if bodies:
            for k in np.nonzero(s_hits["hit"] == 0)[0]:
                b = ow.objects_only(sun_rays[k, :3], sun_rays[k, 3:], 1e30, any_hit=True)
                if b.hit:
                    s_hits[k] = _body_record(b)
        sh[idx[emit]] = s_hits
        cnt = {k: cnt[k] + c2[k] for k in cnt}
    if return_counts:
        return gi, sh, cnt
    return gi, sh


# ---- temporal consumers (vfo_temporal.c) -------------------------------------------------------------
def half_dims(w, h):
    """The half-resolution GI targets' size: w.div_ceil(2) x h.div_ceil(2) (renderer/mod.rs:525)."""
    return (w + 1) // 2, (h + 1) // 2


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def gi_accumulate(w, h, frame_idx, shift, gb_pos, gb_normal, gb_pos_prev, gb_normal_prev, gb_motion, gb_delta,
                  indirect, hist_prev, mom_prev, bicubic=True):
    """gi.comp's accumulation (gi.comp:37-63, 93-184).  Inputs as numpy arrays: gb_pos f32 [h, w, 4],
    gb_normal u32 [h, w], gb_motion u32 [h, w] (RG16F bits), gb_delta u16 [h, w, 4] (RGBA16F bits),
    indirect f32 [hh, hw, 4], hist_prev u16 [hh, hw, 4] (RGBA16F bits), mom_prev f32 [hh, hw, 4].
    Returns (hist_cur u16, mom_cur f32, guide u32 [hh, hw, 4])."""
    hw, hh = half_dims(w, h)
    ins = [_c(gb_pos, np.float32), _c(gb_normal, np.uint32), _c(gb_pos_prev, np.float32),
           _c(gb_normal_prev, np.uint32), _c(gb_motion, np.uint32), _c(gb_delta, np.uint16), _c(indirect, np.float32),
           _c(hist_prev, np.uint16), _c(mom_prev, np.float32)]
    hist = np.zeros((hh, hw, 4), np.uint16)
    mom = np.zeros((hh, hw, 4), np.float32)
    guide = np.zeros((hh, hw, 4), np.uint32)
    sh = (C.c_float * 3)(*[float(x) for x in shift])
    lib().vfo_gi_accumulate(w, h, frame_idx, sh, int(bool(bicubic)), *[a.ctypes.data for a in ins],
                            hist.ctypes.data, mom.ctypes.data, guide.ctypes.data)
    return hist, mom, guide


def light_accumulate(w, h, shift, gb_pos, gb_normal, gb_pos_prev, gb_normal_prev, gb_motion, gb_delta, shadow,
                     hist_prev):
    """lighting.comp's shadow history (lighting.comp:36-42, 92-139).  shadow f32 [h, w]; hist_prev f32
    [h, w, 4].  Returns hist_cur f32 [h, w, 4] = (shadow mean, 2nd moment, count, fast)."""
    ins = [_c(gb_pos, np.float32), _c(gb_normal, np.uint32), _c(gb_pos_prev, np.float32),
           _c(gb_normal_prev, np.uint32), _c(gb_motion, np.uint32), _c(gb_delta, np.uint16), _c(shadow, np.float32),
           _c(hist_prev, np.float32)]
    out = np.zeros((h, w, 4), np.float32)
    sh = (C.c_float * 3)(*[float(x) for x in shift])
    lib().vfo_light_accumulate(w, h, sh, *[a.ctypes.data for a in ins], out.ctypes.data)
    return out


def gi_atrous_level(w, h, it, step, gb_pos, gb_normal, gi_mom, gi_hist, guide, in_cv=None, histfix=True):
    """One a-trous level (atrous.comp); returns outCV u16 [hh, hw, 4] (RGBA16F bits)."""
    hw, hh = half_dims(w, h)
    ins = [_c(gb_pos, np.float32), _c(gb_normal, np.uint32), _c(gi_mom, np.float32), _c(gi_hist, np.uint16),
           _c(guide, np.uint32)]
    icv = _c(in_cv, np.uint16) if in_cv is not None else np.zeros((hh, hw, 4), np.uint16)
    out = np.zeros((hh, hw, 4), np.uint16)
    lib().vfo_gi_atrous_level(w, h, it, step, int(bool(histfix)), *[a.ctypes.data for a in ins], icv.ctypes.data,
                              out.ctypes.data)
    return out


def gi_atrous(w, h, levels, gb_pos, gb_normal, gi_mom, gi_hist, guide, histfix=True):
    """`levels` a-trous levels with the reference's ping-pong (level i writes filt[i & 1]); returns
    (filt0, filt1)."""
    hw, hh = half_dims(w, h)
    ins = [_c(gb_pos, np.float32), _c(gb_normal, np.uint32), _c(gi_mom, np.float32), _c(gi_hist, np.uint16),
           _c(guide, np.uint32)]
    f0 = np.zeros((hh, hw, 4), np.uint16)
    f1 = np.zeros((hh, hw, 4), np.uint16)
    lib().vfo_gi_atrous(w, h, levels, int(bool(histfix)), *[a.ctypes.data for a in ins], f0.ctypes.data,
                        f1.ctypes.data)
    return f0, f1


def gbuffer_motion(w, h, pass_ubo, gb_pos, body=None, bodies=None):
    """gbuffer.comp's gbMotion (u32 [h, w], RG16F bits) and gbDelta (u16 [h, w, 4], RGBA16F bits) from the
    written G-buffer, the pass's current and previous camera, and optional per-pixel body indices with
    the bodies' headers (OBJ_HEADER_DTYPE records)."""
    pu = pass_ubo
    f3 = lambda v: (C.c_float * 3)(*[float(x) for x in v])
    pos = _c(gb_pos, np.float32)
    bd = _c(body, np.int32) if body is not None else None
    hb = np.ascontiguousarray(np.asarray(bodies).view(np.float32).reshape(-1, 60)) if bodies is not None else None
    nb = 0 if hb is None else hb.shape[0]
    mot = np.zeros((h, w), np.uint32)
    dl = np.zeros((h, w, 4), np.uint16)
    lib().vfo_gbuffer_motion(w, h, f3(pu.cam_right), f3(pu.cam_up), f3(pu.cam_fwd), float(pu.tan_half_fov),
                             (C.c_float * 2)(*[float(x) for x in pu.jitter]), f3(pu.prev_cam_pos), f3(pu.prev_cam_right),
                             f3(pu.prev_cam_up), f3(pu.prev_cam_fwd), float(pu.prev_tan_half_fov), pos.ctypes.data,
                             None if bd is None else bd.ctypes.data, None if hb is None else hb.ctypes.data, nb,
                             mot.ctypes.data, dl.ctypes.data)
    return mot, dl


# ---- world pack (GpuWorld::upload's CPU pack) ----------------------------------------------------------
TORCH_FLAME = 254


def majority_material(vox):
    """encode_brick's full-atlas value (gpu_world.rs:202-213): the most frequent of materials 1..255 other
    than TORCH_FLAME (strictly more to replace, starting from 1), if it fills half the brick, else 0."""
    count = np.bincount(np.asarray(vox, np.uint8).reshape(-1), minlength=256)
    best = 1
    for m in range(2, 256):
        if count[m] > count[best] and m != TORCH_FLAME:
            best = m
    return best if count[best] * 2 >= 512 else 0


def pack_world_loop(grid_src, bricks, atlas_slots=0):
    """The reference's loop (gpu_world.rs:245-263 + encode_brick :191-214) transcribed for small inputs:
    cells in order; a detailed cell's brick takes the next slot unless it has one; none left: majority."""
    grid_src = np.asarray(grid_src, np.uint32)
    bricks = np.asarray(bricks, np.uint8).reshape(-1, 512)
    need = len({int(v) - 256 for v in grid_src if v >= 256})
    cap = atlas_slots or need
    gpu_slot = [-1] * bricks.shape[0]
    n_used = 0
    out = np.zeros_like(grid_src)
    for gi, v in enumerate(grid_src):
        if v < 256:
            out[gi] = v
            continue
        b = int(v) - 256
        if gpu_slot[b] < 0 and n_used < cap:
            gpu_slot[b] = n_used
            n_used += 1
        out[gi] = 256 + gpu_slot[b] if gpu_slot[b] >= 0 else majority_material(bricks[b])
    return out, np.asarray(gpu_slot, np.int64), cap


def pack_world(grid_src, bricks, atlas_slots=0):
    """GpuWorld::upload's pack, vectorised: (encoded grid u32, atlas [cap, 512] with TORCH_FLAME as air,
    brick masks [cap, 16] of the bricks' own voxels, slot per brick (-1 none))."""
    grid_src = np.asarray(grid_src, np.uint32)
    bricks = np.asarray(bricks, np.uint8).reshape(-1, 512)
    det = np.nonzero(grid_src >= 256)[0]
    ref = grid_src[det].astype(np.int64) - 256
    uniq, first = np.unique(ref, return_index=True)
    order = np.argsort(first, kind="stable")          # bricks by their first cell in grid order
    named = uniq[order]
    cap = atlas_slots or len(named)
    slot = np.full(bricks.shape[0], -1, np.int64)
    fit = named[:cap]
    slot[fit] = np.arange(len(fit))
    out = grid_src.copy()
    s = slot[ref]
    enc = np.where(s >= 0, 256 + s, 0).astype(np.uint32)
    for k in np.nonzero(s < 0)[0]:
        enc[k] = majority_material(bricks[ref[k]])
    out[det] = enc
    atlas = np.zeros((cap, 512), np.uint8)
    masks = np.zeros((cap, 16), np.uint32)
    if len(fit):
        atlas[:len(fit)] = np.where(bricks[fit] == TORCH_FLAME, 0, bricks[fit])
        for i, b in enumerate(fit):
            masks[i] = brick_mask(bricks[b])
    return out, atlas, masks, slot



# ---- chunk sync (GpuWorld::sync_chunks) ------------------------------------------------------------------
RING_CHUNK_B = 16


class GpuWorldState:
    """GpuWorld's state that sync and sync_chunks read and write (gpu_world.rs:140-170): the World's
    source grid (0 / 1-255 / 256 + brick index), the encoded grid, the atlas with its brick masks, the
    coarse pyramid, the free list, the slots handed out and which slots hold world bricks."""

    def __init__(self, wb, src, bricks, capacity, win=(0, 0, 0, 0)):
        """The state right after GpuWorld::upload packs `src` (pack_world) into an atlas of `capacity`;
        win: the ring window (x, y, z origin in bricks, ring flag)."""
        self.wb = tuple(int(x) for x in wb)
        self.win = tuple(int(x) for x in win)
        self.src = np.array(src, np.uint32).reshape(-1)
        out, atlas, masks, slot = pack_world(self.src, bricks, capacity)
        self.cap = int(capacity)
        self.grid = out
        self.atlas = np.zeros((self.cap, 512), np.uint8)
        self.masks = np.zeros((self.cap, 16), np.uint32)
        self.used = int((slot >= 0).sum())
        self.atlas[:self.used], self.masks[:self.used] = atlas[:self.used], masks[:self.used]
        self.free = []
        self.world_slot = np.zeros(self.cap, np.uint8)
        self.world_slot[:self.used] = 1
        self.slot0 = slot.astype(np.int32)
        self.c1d = tuple((x + 3) // 4 for x in self.wb)
        self.c2d = tuple((x + 15) // 16 for x in self.wb)
        self.c1 = np.zeros(int(np.prod(self.c1d)), np.uint8)
        self.c2 = np.zeros(int(np.prod(self.c2d)), np.uint8)
        lib().vfo_coarse(_p(self.src), (C.c_int32 * 3)(*self.wb), _p(self.c1), _p(self.c2))

    def alloc_slot(self):
        """gpu_world.rs:178-186: the free list's back, else the next unused slot, else -1."""
        if self.free:
            return self.free.pop()
        if self.used < self.cap:
            self.used += 1
            return self.used - 1
        return -1

    def coarse_cell1(self, ci):
        """gpu_world.rs:944-957 over the source grid."""
        c = self.c1d
        wbx, wby, wbz = self.wb
        cx, cy, cz = ci % c[0], (ci // c[0]) % c[1], ci // (c[0] * c[1])
        g = self.src.reshape(wbz, wby, wbx)
        return np.uint8(g[cz * 4:min((cz + 1) * 4, wbz), cy * 4:min((cy + 1) * 4, wby),
                          cx * 4:min((cx + 1) * 4, wbx)].any())

    def coarse_cell2(self, ci):
        """gpu_world.rs:959-973 over coarse1."""
        c, d = self.c2d, self.c1d
        cx, cy, cz = ci % c[0], (ci // c[0]) % c[1], ci // (c[0] * c[1])
        g = self.c1.reshape(d[2], d[1], d[0])
        return np.uint8(g[cz * 4:min((cz + 1) * 4, d[2]), cy * 4:min((cy + 1) * 4, d[1]),
                          cx * 4:min((cx + 1) * 4, d[0])].any())


def sync_chunks(st, chunks, bricks, gpu_slot, released=()):
    """GpuWorld::sync_chunks (gpu_world.rs:1194-1376) on a GpuWorldState, transcribed: chunks = [(cx, cz, cells or
    None)] (None: evicted and not refilled), cells [16*wby*16] in (z, y, x) order.  Returns (gpu_slot, boxes):
    the bricks' slots after the sync and the brick boxes whose distance fields it rebuilds, in order."""
    wbx, wby, wbz = st.wb
    cb = RING_CHUNK_B
    bricks = np.asarray(bricks, np.uint8).reshape(-1, 512)
    gpu_slot = np.array(gpu_slot, np.int64).reshape(-1)
    for sl in sorted((int(x) for x in released), reverse=True):          # :1199-1207
        if 0 <= sl < st.world_slot.size and st.world_slot[sl]:
            st.world_slot[sl] = 0
            st.free.append(sl)
    ring_w = wbx // cb
    done = set()
    up, boxes = [], []
    grid3 = st.grid.reshape(wbz, wby, wbx)
    src3 = st.src.reshape(wbz, wby, wbx)
    c1_3 = st.c1.reshape(st.c1d[2], st.c1d[1], st.c1d[0])
    c2_3 = st.c2.reshape(st.c2d[2], st.c2d[1], st.c2d[0])
    for cx, cz, cells in chunks:
        slot = (cz & (ring_w - 1)) * ring_w + (cx & (ring_w - 1))
        if slot in done:
            continue
        done.add(slot)
        tx, tz = (cx * cb) & (wbx - 1), (cz * cb) & (wbz - 1)
        boxes.append(([cx * cb, 0, cz * cb], [cx * cb + cb - 1, wby - 1, cz * cb + cb - 1]))
        if cells is None:                                                # :1228-1253
            grid3[tz:tz + cb, :, tx:tx + cb] = 0
            src3[tz:tz + cb, :, tx:tx + cb] = 0
            c1_3[tz // 4:(tz + cb) // 4, :, tx // 4:(tx + cb) // 4] = 0
            c2_3[tz // 16:(tz + cb) // 16, :, tx // 16:(tx + cb) // 16] = 0
            continue
        cells = np.asarray(cells, np.uint32).reshape(cb, wby, cb)
        src3[tz:tz + cb, :, tx:tx + cb] = cells
        enc = cells.copy()
        flat = enc.reshape(-1)
        for k in np.nonzero(flat >= 256)[0]:                             # :1255-1273, encode_brick :191-214
            b = int(flat[k]) - 256
            if gpu_slot[b] < 0:
                gpu_slot[b] = st.alloc_slot()
            if gpu_slot[b] >= 0:
                st.world_slot[gpu_slot[b]] = 1
                flat[k] = 256 + gpu_slot[b]
                up.append(b)
            else:
                flat[k] = majority_material(bricks[b])
        grid3[tz:tz + cb, :, tx:tx + cb] = enc
        for z in range(tz // 4, (tz + cb) // 4):                          # :1274-1284
            for y in range(st.c1d[1]):
                for x in range(tx // 4, (tx + cb) // 4):
                    ci = (z * st.c1d[1] + y) * st.c1d[0] + x
                    st.c1[ci] = st.coarse_cell1(ci)
        for z in range(tz // 16, (tz + cb) // 16):                        # :1285-1295
            for y in range(st.c2d[1]):
                for x in range(tx // 16, (tx + cb) // 16):
                    ci = (z * st.c2d[1] + y) * st.c2d[0] + x
                    st.c2[ci] = st.coarse_cell2(ci)
    for b in up:                                                         # :1300-1333, gpu_voxels, masks_of
        st.atlas[gpu_slot[b]] = np.where(bricks[b] == TORCH_FLAME, 0, bricks[b])
        st.masks[gpu_slot[b]] = brick_mask(bricks[b])
    return gpu_slot.astype(np.int32), boxes


DIST_CAP = 16


def brick_coords(st, gi):
    """World::brick_coords (sparse_volume.rs:206-215): grid index -> brick coordinates, absolute in a ring
    world (the window's chunk in each ring slot)."""
    wbx, wby, wbz = st.wb
    b = [gi % wbx, (gi // wbx) % wby, gi // (wbx * wby)]
    if st.win[3]:
        b[0] = st.win[0] + (b[0] - st.win[0]) % wbx
        b[2] = st.win[2] + (b[2] - st.win[2]) % wbz
    return b


def edit_boxes(st, dirty):
    """GpuWorld::sync's clusters (gpu_world.rs:1103-1130): a dirty brick grows the first box within DIST_CAP
    of it, else starts one; then boxes within 2 DIST_CAP of each other merge, first pair first, until none."""
    boxes = []
    for gi in dirty:
        b = brick_coords(st, gi)
        for x in boxes:
            if all(x[0][k] - DIST_CAP <= b[k] <= x[1][k] + DIST_CAP for k in range(3)):
                x[0] = [min(x[0][k], b[k]) for k in range(3)]
                x[1] = [max(x[1][k], b[k]) for k in range(3)]
                break
        else:
            boxes.append([list(b), list(b)])
    merged = True
    while merged:
        merged = False
        for i in range(len(boxes)):
            for j in range(i + 1, len(boxes)):
                if all(boxes[j][0][k] <= boxes[i][1][k] + 2 * DIST_CAP and boxes[j][1][k] >= boxes[i][0][k] - 2 * DIST_CAP
                       for k in range(3)):
                    bj = boxes.pop(j)
                    boxes[i][0] = [min(boxes[i][0][k], bj[0][k]) for k in range(3)]
                    boxes[i][1] = [max(boxes[i][1][k], bj[1][k]) for k in range(3)]
                    merged = True
                    break
            if merged:
                break
    return [(x[0], x[1]) for x in boxes]


def sync_edits(st, cells, values, bricks, gpu_slot):
    """GpuWorld::sync (gpu_world.rs:985-1142) on a GpuWorldState, transcribed: the World's edited cells take
    their values (a cell listed twice: its last value), then in grid order, once per cell, compact_brick
    (sparse_volume.rs:354-366: a uniform brick becomes its material and frees its slot) and encode_brick;
    the touched coarse cells; the clusters.  Returns (values after compaction per input entry, gpu_slot,
    boxes)."""
    bricks = np.asarray(bricks, np.uint8).reshape(-1, 512)
    gpu_slot = np.array(gpu_slot, np.int64).reshape(-1)
    cells = [int(c) for c in cells]
    for gi, v in zip(cells, values):
        st.src[gi] = int(v)
    dirty = sorted(set(cells))
    up = []
    for gi in dirty:
        v = int(st.src[gi])
        if v >= 256:
            b = v - 256
            if (bricks[b] == bricks[b][0]).all():                       # compact_brick
                freed = int(gpu_slot[b])
                gpu_slot[b] = -1
                st.src[gi] = bricks[b][0]
                if freed >= 0:
                    st.world_slot[freed] = 0
                    st.free.append(freed)
        v = int(st.src[gi])
        if v < 256:
            st.grid[gi] = v
            continue
        b = v - 256
        if gpu_slot[b] < 0:
            gpu_slot[b] = st.alloc_slot()
        if gpu_slot[b] >= 0:
            st.world_slot[gpu_slot[b]] = 1
            st.grid[gi] = 256 + gpu_slot[b]
            up.append(b)
        else:
            st.grid[gi] = majority_material(bricks[b])
    for b in up:
        st.atlas[gpu_slot[b]] = np.where(bricks[b] == TORCH_FLAME, 0, bricks[b])
        st.masks[gpu_slot[b]] = brick_mask(bricks[b])
    wbx, wby, _ = st.wb
    c1 = sorted({((gi // (wbx * wby)) // 4 * st.c1d[1] + (gi // wbx) % wby // 4) * st.c1d[0] + gi % wbx // 4 for gi in dirty})
    c2 = sorted({((gi // (wbx * wby)) // 16 * st.c2d[1] + (gi // wbx) % wby // 16) * st.c2d[0] + gi % wbx // 16 for gi in dirty})
    for ci in c1:
        st.c1[ci] = st.coarse_cell1(ci)
    for ci in c2:
        st.c2[ci] = st.coarse_cell2(ci)
    return (np.array([st.src[gi] for gi in cells], np.uint32), gpu_slot.astype(np.int32), edit_boxes(st, dirty))
