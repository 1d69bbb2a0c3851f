/*
 * vf_trace.h — C ABI of the MI355X (gfx950) voxel traversal library (libvftrace.so).
 *
 * Drop-in boundary for VoxelForge's traced passes.  Each entry point replaces one GPU call of the
 * reference renderer; the structs keep the reference's #[repr(C)] / std140 byte layouts so the
 * Rust host can pass its own `bytemuck::Pod` values straight through.  Citations are relative to
 * the reference repository root.
 *
 * Conventions
 *   - Every function returns 0 on success or a negative VF_E_* code; vf_last_error() gives text.
 *     No exception or abort crosses the ABI; invalid input is an error, never a crash.
 *   - Pointers named `d_*` are DEVICE pointers (caller-owned, e.g. from vf_device_alloc); all
 *     other pointers are host pointers, read during the call only.
 *   - Work is enqueued on the context's stream (vf_set_stream) and is asynchronous; vf_sync() is
 *     the fence.  The world must not be updated while a trace on the same context is in flight.
 *   - One context per host thread, or external synchronisation.
 */
#ifndef VF_TRACE_H
#define VF_TRACE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VF_ABI_VERSION 1

enum {
    VF_OK = 0,
    VF_E_INVALID = -1,     /* bad argument (null pointer, dimension, slot out of range) */
    VF_E_DEVICE = -2,      /* HIP runtime error */
    VF_E_NO_WORLD = -3,    /* a trace before vf_world_upload */
    VF_E_OOM = -4,         /* device allocation failed */
    VF_E_UNSUPPORTED = -5  /* a feature outside this library (e.g. the allocation-free batch on a device it cannot reach) */
};

/* ---- layouts shared with the reference host ------------------------------------------------ */

/* Palette entry, 48 B (src/voxel/world.rs:21-31; common.glsl Material). */
typedef struct vf_material {
    float r, g, b, emissive;
    float props[4];   /* density, jitter, kind (1 water, 2 grass, 3 foliage), variant */
    float surf[4];    /* roughness, metalness, dielectric F0, detail */
} vf_material;

/* SceneUbo, std140, 1536 B (src/gfx/ubo.rs:10-86; offsets asserted there and in vftrace/abi.py). */
typedef struct vf_scene_ubo {
    uint32_t atlas_bricks[3]; float night;
    int32_t world_bricks[3];  int32_t num_objects;
    float sun_dir[3];         float grass_dist;
    float sun_rgb[3];         float sun_angle;
    float fog_sun_rgb[3];     float fog_density;
    float fog_ambient[3];     float fog_g;
    float sky_zenith[3];      float sky_scale;
    float sky_horizon[3];     float gi_scale;
    float sunset_rgb[3];      int32_t sky_model;
    float fog_height, fog_falloff, fog_noise; int32_t fog_steps;
    float fog_shadow_max; int32_t tonemap_op; float saturation, contrast;
    float white_balance[3];   int32_t grass_shadows;
    float sh_right[3];        float sh_texel0;
    float sh_up[3];           float sh_texel1;
    float sh_lz_top, sh_texel2; int32_t sh_n, std_lighting;
    int32_t sh_win_a[4];
    int32_t sh_win_b[4];
    int32_t probe_origin[3];  float probe_spacing;
    int32_t probe_dims[3];    int32_t probe_update_offset;
    float probe_hyst, probe_normal_bias, std_soft_taps, std_cascade_blend;
    int32_t std_cascade_bias; float std_view_bias, sh_texel3, std_soft_min;
    int32_t probe_origin_c[3]; float probe_spacing_c;
    int32_t world_win[4];     /* ring window min corner (bricks) in .x/.z; .w != 0 = ring mode */
    int32_t hor_a[4];
    int32_t hor_b[4];
    uint32_t chunk_fade[256];
    float gi_cache_cam[3];    float gi_cache_lod;
    int32_t gi_cache_p[4];
    float gi_cache_q[4];
    int32_t hor_c[4];
    float cloud_a[4];
    float cloud_b[4];
    float rain_a[4];
} vf_scene_ubo;

/* PassUbo, std140, 288 B (src/gfx/ubo.rs:91-131). */
typedef struct vf_pass_ubo {
    float cam_pos[3];      float lock_active;
    float cam_right[3];    float fog_start;
    float cam_up[3];       float haze_k;
    float cam_fwd[3];      float tan_half_fov;
    float prev_cam_pos[3]; float prev_tan_half_fov;
    float prev_cam_right[3]; float haze_sea;
    float prev_cam_up[3];  float _pf;
    float prev_cam_fwd[3]; uint32_t frame_idx;
    float exposure; int32_t u_step, u_iteration, debug_mode;
    float time, prev_time; float jitter[2];
    float water_p[4];
    float look_p[4];
    float fog_p[4];
    float under_p[4];
    float opt_p[4];
    float wave_p[4];
    float shift_p[4];
    float facet_p[4];
} vf_pass_ubo;

/* The layout asserts of the reference (src/gfx/ubo.rs:76-86, 132-139; src/voxel/world.rs:31), checked by
 * every C or C++ compiler that includes this header. */
#ifdef __cplusplus
#define VF_LAYOUT_ASSERT(c, m) static_assert(c, m)
#else
#define VF_LAYOUT_ASSERT(c, m) _Static_assert(c, m)
#endif
VF_LAYOUT_ASSERT(sizeof(vf_material) == 48, "Material is 48 B");
VF_LAYOUT_ASSERT(sizeof(vf_scene_ubo) == 1536, "SceneUbo is 1536 B");
VF_LAYOUT_ASSERT(offsetof(vf_scene_ubo, fog_height) == 144, "SceneUbo.fog_height @144");
VF_LAYOUT_ASSERT(offsetof(vf_scene_ubo, sh_win_a) == 240, "SceneUbo.sh_win_a @240");
VF_LAYOUT_ASSERT(offsetof(vf_scene_ubo, probe_origin) == 272, "SceneUbo.probe_origin @272");
VF_LAYOUT_ASSERT(offsetof(vf_scene_ubo, world_win) == 352, "SceneUbo.world_win @352");
VF_LAYOUT_ASSERT(offsetof(vf_scene_ubo, chunk_fade) == 400, "SceneUbo.chunk_fade @400");
VF_LAYOUT_ASSERT(offsetof(vf_scene_ubo, gi_cache_cam) == 1424, "SceneUbo.gi_cache_cam @1424");
VF_LAYOUT_ASSERT(offsetof(vf_scene_ubo, hor_c) == 1472, "SceneUbo.hor_c @1472");
VF_LAYOUT_ASSERT(offsetof(vf_scene_ubo, rain_a) == 1520, "SceneUbo.rain_a @1520");
VF_LAYOUT_ASSERT(sizeof(vf_pass_ubo) == 288, "PassUbo is 288 B");
VF_LAYOUT_ASSERT(offsetof(vf_pass_ubo, prev_tan_half_fov) == 76, "PassUbo.prev_tan_half_fov @76");
VF_LAYOUT_ASSERT(offsetof(vf_pass_ubo, prev_cam_fwd) == 112, "PassUbo.prev_cam_fwd @112");
VF_LAYOUT_ASSERT(offsetof(vf_pass_ubo, frame_idx) == 124, "PassUbo.frame_idx @124");
VF_LAYOUT_ASSERT(offsetof(vf_pass_ubo, jitter) == 152, "PassUbo.jitter @152");
VF_LAYOUT_ASSERT(offsetof(vf_pass_ubo, facet_p) == 272, "PassUbo.facet_p @272");

/* Explicit ray, 24 B: origin and direction (voxel units; direction need not be normalised). */
typedef struct vf_ray { float o[3]; float d[3]; } vf_ray;

/* Hit record, 20 B (common.glsl:454-461 Hit, without the body index).
 * face: 0 = none (the ray started inside the world box), 1 +x, 2 -x, 3 +y, 4 -y, 5 +z, 6 -z
 * (the outward normal of the face the ray entered through).  Miss: hit = 0, t = -1, rest 0.
 * body = 1: the hit is on a body (Hit.obj >= 0); cell and face are then in the body's own voxel
 * space, and the body index comes with vf_trace_gbuffer's d_body. */
typedef struct vf_hit {
    int32_t cell[3];   /* absolute voxel coordinates (world), or body voxel coordinates */
    float t;
    uint8_t mat, face, hit, body;
} vf_hit;

/* The world as GpuWorld::upload (src/gfx/gpu_world.rs:228-437) packs it. */
typedef struct vf_world_desc {
    int32_t world_bricks[3];     /* grid dims; ring mode: x and z powers of two */
    int32_t world_win[4];        /* SceneUbo.world_win */
    const uint32_t *grid;        /* [wbx*wby*wbz], (z*wby + y)*wbx + x: 0 empty, 1-255 uniform, 256+slot */
    const uint8_t *atlas;        /* [n_slots*512], per slot voxel lz*64 + ly*8 + lx (Brick::vox order) */
    uint32_t n_slots;
    const uint32_t *masks;       /* [n_slots*16] brick_mask layout (sparse_volume.rs:48-63), or NULL */
    const uint8_t *coarse1;      /* [ceil(wb/4)^3-shaped] or NULL (then built here) */
    const uint8_t *coarse2;      /* [ceil(wb/16)] or NULL (then built here) */
    const vf_material *palette;  /* [256] */
} vf_world_desc;

typedef struct vf_ctx vf_ctx;

/* ---- context ------------------------------------------------------------------------------- */
/* Creates a context on HIP device `device` with its own default-priority stream.
 * Replaces the Vulkan device/queue the reference passes run on (src/gfx/context.rs:67). */
int vf_ctx_create(int device, vf_ctx **out);
void vf_ctx_destroy(vf_ctx *ctx);
/* Last error text of ctx (or of the calling thread when ctx is NULL); never NULL. */
const char *vf_last_error(const vf_ctx *ctx);
int vf_abi_version(void);
/* Sizes of the shared structs, for the host's layout asserts (ubo.rs:76-86, 132-139). */
int vf_abi_sizes(size_t *scene_ubo, size_t *pass_ubo, size_t *material, size_t *ray, size_t *hit);
/* Run on a caller stream (a hipStream_t passed as void*; NULL = the context's own stream). */
int vf_set_stream(vf_ctx *ctx, void *hip_stream);
/* Fence: waits for all work enqueued on the context's stream (end_one_shot_and_wait). */
int vf_sync(vf_ctx *ctx);

/* ---- device memory helpers (the Rust host's Buffer equivalents) ------------------------------ */
int vf_device_alloc(vf_ctx *ctx, size_t bytes, void **d_out);
int vf_device_free(vf_ctx *ctx, void *d_ptr);
int vf_copy_to_device(vf_ctx *ctx, void *d_dst, const void *src, size_t bytes);
int vf_copy_to_host(vf_ctx *ctx, void *dst, const void *d_src, size_t bytes);
int vf_memset_device(vf_ctx *ctx, void *d_dst, int value, size_t bytes);

/* ---- world mirror -------------------------------------------------------------------------- */
/* GpuWorld::upload (gpu_world.rs:228): grid, atlas, masks, coarse pyramid and palette, then the full
 * brick distance field (gpu_world.rs:422) and the sky-top map (grasstop.comp skyTop). */
int vf_world_upload(vf_ctx *ctx, const vf_world_desc *desc);
/* The world as the reference's CPU side holds it (World::grid and SparseVolume::bricks): grid cells
 * 0 empty, 1-255 uniform, 256 + i = detailed brick i of `bricks` (512 voxels each, Brick::vox order). */
typedef struct vf_world_src {
    int32_t world_bricks[3];
    int32_t world_win[4];
    const uint32_t *grid;        /* [wbx*wby*wbz], (z*wby + y)*wbx + x */
    const uint8_t *bricks;       /* [n_bricks*512] */
    uint32_t n_bricks;
    uint32_t atlas_slots;        /* atlas capacity; 0 = as many slots as the grid names distinct bricks */
    const vf_material *palette;  /* [256] */
    uint32_t brick_stride;       /* bytes from one brick's voxels to the next's; 0 = 512 (packed) */
    uint32_t slot_stride;        /* bytes between slot_of_brick entries (multiple of 4); 0 = 4 */
} vf_world_src;
/* GpuWorld::upload's pack (gpu_world.rs:228-330 with encode_brick :191-214 and gpu_voxels :58-67) on the
 * device, then what vf_world_upload builds: each detailed brick gets the next atlas slot at its first
 * cell in grid order (so a brick named by several cells shares one slot), its voxels go to the slot with
 * TORCH_FLAME as air, `masks` of the slot = brick_mask of its voxels (sparse_volume.rs:48-63); past the
 * capacity a brick's cells take its majority material (gpu_world.rs:202-213).  The coarse pyramid
 * comes from the source grid.  slot_of_brick (nullable, [n_bricks]): each brick's slot (the
 * reference's Brick::gpu_slot), -1 for a brick the grid does not name or that did not fit.  With the
 * strides, `bricks` and `slot_of_brick` may point at the `vox` and `gpu_slot` fields of an array of
 * the caller's brick records (World::bricks) and are read and written in place. */
int vf_world_pack(vf_ctx *ctx, const vf_world_src *src, int32_t *slot_of_brick);
/* One chunk of GpuWorld::sync_chunks (gpu_world.rs:1194-1376): 16 x wby x 16 bricks of a ring world. */
typedef struct vf_chunk {
    int32_t cx, cz;              /* chunk coordinates (16 bricks a chunk, absolute) */
    int32_t loaded;              /* World::chunk_loaded: 0 = evicted and not refilled (its cells are cleared) */
    const uint32_t *cells;       /* loaded: [16*wby*16], (z*wby + y)*16 + x: 0 / 1-255 / 256 + i = brick i of `bricks` */
} vf_chunk;
/* GpuWorld::sync_chunks on the device (ring worlds).  `released` (World::released_gpu_slots) slots that
 * hold world bricks go back to the free list, highest first, so pops hand out ascending slots.  Chunks
 * are taken in order, once per ring slot.  A loaded chunk's cells are encoded in (z, y, x) order: a brick
 * with gpu_slot[i] < 0 takes a slot from the free list, else the next unused one, else its cells take
 * its majority material (encode_brick); every brick it names is written to its slot (TORCH_FLAME as air)
 * with its brick masks; its coarse cells come from its cells.  An unloaded chunk's cells and coarse cells
 * are cleared.  Then the distance field over each chunk's box +-16 bricks, the cell words and the
 * sky-top map.  gpu_slot: [n_bricks], in and out (Brick::gpu_slot).  brick_stride / slot_stride: bytes
 * between consecutive bricks' voxels / slots (0 = packed: 512 / 4), so both may point into the caller's
 * brick records. */
int vf_world_sync_chunks(vf_ctx *ctx, const vf_chunk *chunks, int32_t n_chunks, const uint8_t *bricks, uint32_t n_bricks,
                         uint32_t brick_stride, int32_t *gpu_slot, uint32_t slot_stride, const int32_t *released,
                         int32_t n_released);
/* GpuWorld::sync (gpu_world.rs:985-1142) on the device: the world's edited cells.  cells[i] is a grid index,
 * values[i] the world's value there (0 / 1-255 / 256 + i = brick i of `bricks`); a cell listed twice takes
 * its last value.  In grid order, once per cell: a brick that became uniform turns back into a uniform cell
 * and frees its slot (SparseVolume::compact_brick; values[] returns the new value, gpu_slot -1), then the
 * cell is encoded as vf_world_sync_chunks encodes.  The touched coarse cells are rebuilt from the world's
 * values; the dirty bricks are clustered (boxes within 16 bricks grow, boxes within 32 merge) and each
 * cluster gets its sky-top columns and its distance field over the box +-16 bricks.  gpu_slot: [n_bricks],
 * in and out; the strides as for vf_world_sync_chunks. */
int vf_world_sync_edits(vf_ctx *ctx, const uint32_t *cells, uint32_t *values, uint32_t n, const uint8_t *bricks,
                        uint32_t n_bricks, uint32_t brick_stride, int32_t *gpu_slot, uint32_t slot_stride);
/* The atlas bookkeeping: slots handed out, the free list's length, capacity. */
int vf_world_atlas_state(vf_ctx *ctx, uint32_t *used, uint32_t *n_free, uint32_t *capacity);
/* Copies the device world's encoded grid ([wbx*wby*wbz]) and the first n_slots atlas slots back. */
int vf_world_read(vf_ctx *ctx, uint32_t *grid, uint8_t *atlas, uint32_t n_slots, uint32_t *masks);
/* GpuWorld::sync (gpu_world.rs:985) minus the Vulkan staging: set grid cells and (re)write atlas
 * slots.  cells: grid indices + new values; slots: slot ids + 512 B each.  The caller then rebuilds
 * the distance field over the dirty box with vf_build_brickdist, as the reference does. */
int vf_world_update(vf_ctx *ctx, const uint32_t *cell_idx, const uint32_t *cell_val, uint32_t n_cells,
                    const uint32_t *slot_idx, const uint8_t *slot_vox, uint32_t n_slot_updates);
/* GpuWorld::record_distance_field (gpu_world.rs:1376-1428): the separable L-inf brick distance
 * field (brickdist.comp) over bricks [lo, hi] and the dry occupancy over [dirty_lo, dirty_hi],
 * inclusive, absolute brick coordinates.  Then refreshes the cell words the traversal reads.  vf_last_kernel_ms
 * afterwards gives the distance-field passes' time. */
int vf_build_brickdist(vf_ctx *ctx, const int32_t lo[3], const int32_t hi[3], const int32_t dirty_lo[3],
                       const int32_t dirty_hi[3]);
/* Read back the two distance fields the traversal uses (brickDistTex, brickDistDryTex), host. */
int vf_read_brickdist(vf_ctx *ctx, uint8_t *dist, uint8_t *dry);
/* Read back the sky-top map (R16UI, (wbx*8) x (wbz*8)), host. */
int vf_read_skytop(vf_ctx *ctx, uint16_t *skytop);
/* Read back the coarse occupancy pyramid (coarse1: 4^3 bricks a cell, coarse2: 16^3), host. */
int vf_read_coarse(vf_ctx *ctx, uint8_t *coarse1, uint8_t *coarse2);

/* ---- bodies (dynamic voxel objects) -------------------------------------------------------- */
/* Body header, std430, 240 B (gpu_world.rs:34-47 ObjHeader; common.glsl:65-72).  Matrices are
 * column-major (GLSL mat4; math.rs:30 Transform::to_mat4).  prev_obj_to_world and the .w of the boxes
 * (glow integrity) are not read by the traversal. */
typedef struct vf_obj_header {
    float world_to_obj[16];
    float obj_to_world[16];
    float prev_obj_to_world[16];
    int32_t dims_bricks[4];   /* bricks x, y, z; offset of the body's grid in the body grid array */
    float aabb_min[4];        /* the body's world box */
    float aabb_max[4];
} vf_obj_header;

/* BVH node, 32 B (common.glsl:77-81 BvhNode, as VkGpuWorld::writeBvh / gpu_world.rs:556-558 writes it):
 * count 0 = interior node with children first and first + 1, else a leaf of prims [first, first+count). */
typedef struct vf_bvh_node {
    float min[3]; uint32_t first;
    float max[3]; uint32_t count;
} vf_bvh_node;

/* GpuWorld::init_objects (gpu_world.rs:661-707): n bodies with their grids and headers, replacing any
 * before.  grid: every body's brick grid, body i's brick (x, y, z) at dims_bricks[3] + (z*by + y)*bx + x,
 * encoded like the world grid (0 empty, 1-255 uniform, 256 + slot); body bricks live in the world's
 * atlas (slots uploaded with vf_world_upload / rewritten with vf_world_update, as the reference numbers
 * them into the same atlas, gpu_world.rs:274-293).  The BVH over the bodies' world boxes is built here
 * (build_bvh, gpu_world.rs:550-600).  n = 0 removes the bodies.  The traces include the bodies when
 * the scene's num_objects is non-zero (it must then equal n). */
int vf_objects_upload(vf_ctx *ctx, const vf_obj_header *headers, int32_t n, const uint32_t *grid, int64_t grid_len);
/* GpuWorld::update_objects (gpu_world.rs:650-656): new headers (poses, boxes) for the same n bodies and
 * a rebuilt BVH; written in place, so recorded frame graphs see them.  Not during a graph capture. */
int vf_objects_set_headers(vf_ctx *ctx, const vf_obj_header *headers, int32_t n);
/* The caller's own BVH (the layout above) instead of the one built here: checked to be a tree over
 * the nodes with prim ranges and indices in range and at most 24 pending subtrees on the reference's
 * traversal stack (common.glsl:697).  Replaced again by the next upload / set_headers. */
int vf_objects_set_bvh(vf_ctx *ctx, const vf_bvh_node *nodes, int32_t n_nodes, const uint32_t *prims, int32_t n_prims);
/* GpuWorld::sync_objects (gpu_world.rs:710-745): body grid entries idx[k] <- val[k]. */
int vf_objects_update_grid(vf_ctx *ctx, const uint32_t *idx, const uint32_t *val, int32_t n);
/* The BVH in use, host copies (counts always written; arrays when large enough, else VF_E_INVALID). */
int vf_objects_read_bvh(vf_ctx *ctx, vf_bvh_node *nodes, int32_t max_nodes, int32_t *n_nodes, uint32_t *prims,
                        int32_t max_prims, int32_t *n_prims);

/* ---- per-frame uniforms -------------------------------------------------------------------- */
/* Renderer::set_frame_uniforms (frame.rs:549): the scene block for the following traces. */
int vf_set_scene(vf_ctx *ctx, const vf_scene_ubo *scene);

/* ---- traced passes ------------------------------------------------------------------------- */
/* P::Gbuf (record.rs:327; gbuffer.comp:189-266, 470-480): one pinhole primary ray per pixel over the
 * scene (the world, and the bodies when the scene's num_objects is set).  A camera inside solid terrain starts each ray where it leaves the solid (exitSolid,
 * common.glsl:990-1016; t counts from the camera, rays that only leave the world are misses); a
 * camera in water traces with skipMat 5.  Writes hits [w*h], and when non-NULL gbPos (RGBA32F: P, t;
 * sky (0,0,0,-1)) and gbNormal (RGBA8 unorm: n*0.5+0.5; a = matFlag, gbuffer.comp:524-526: 0.8 (204) on a
 * foliage material (palette kind 3), 1.0 (255) otherwise; sky 0).  Row 0 is the image bottom.
 * A frame graph records the camera case with the pass: record it again after an edit of the
 * camera's voxel. */
int vf_trace_primary(vf_ctx *ctx, const vf_pass_ubo *pass, int32_t width, int32_t height, vf_hit *d_hits,
                     float *d_gb_pos, uint32_t *d_gb_normal);
/* P::Gbuf (+ P::ShTrace) over the scene (traceScenePrimary = traceSceneS, common.glsl:787-793, 955): the
 * world, then the bodies closer than its hit, with every per-pixel output: as vf_trace_primary_shadow,
 * plus d_body (nullable): the hit body's index per pixel (gbuffer.comp h.obj; -1 = world or sky).
 * d_shadow_raw NULL: no sun ray (vf_trace_primary's outputs).  A body hit's gbNormal is
 * normalize(mat3(objToWorld) * n) (common.glsl:678) in RGBA8 (NaN -> 0). */
int vf_trace_gbuffer(vf_ctx *ctx, const vf_pass_ubo *pass, int32_t width, int32_t height, vf_hit *d_hits,
                     float *d_gb_pos, uint32_t *d_gb_normal, int32_t *d_body, uint32_t *d_shadow_raw,
                     vf_hit *d_shadow_hits);
/* P::ShTrace (record.rs:463; shadow_trace.comp:17-105 without the hybrid map, grass and under-water
 * refraction): one sun ray per G-buffer hit from P + N*0.03 through the dry field (water clear), with
 * FOLIAGE_LIGHT: sunVisibilityT (common.glsl:1964-1981, deterministic) lets pass->look_p[3]
 * (r.foliage.light) of the light through each leaf voxel, up to 6 (0 = a hard shadow ray).  A texel
 * facing away from the sun gets no ray, unless it is foliage (gbNormal.a in (0.62, 0.88)) with the sun
 * above the horizon: FOLIAGE_SSS (shadow_trace.comp:35-46, the reference default) traces sunVisibility
 * from foliageSunOrigin (P - N*0.5 + sun*0.9; a blade, a in (0.62, 0.75): P + sun*0.05) and stores no
 * blocker distance.  shadowRaw is RG16F (visibility, min(tFirstBlocker, 60000)); d_hits (nullable) gets
 * the first segment's hit. */
int vf_trace_shadow(vf_ctx *ctx, const vf_pass_ubo *pass, int32_t width, int32_t height, const float *d_gb_pos,
                    const uint32_t *d_gb_normal, uint32_t *d_shadow_raw, vf_hit *d_hits);
/* The SVGF lighting pass's sun ray (lighting.comp:44-84): per lit G-buffer texel, or back-lit foliage
 * texel with the sun above the horizon (FOLIAGE_SSS, from foliageSunOrigin as in vf_trace_shadow), one ray
 * along sampleCone(sunDir, scene sun_angle) (common.glsl:1259-1265) seeded pixelSeed(pix, frame_idx, 0),
 * through leaves with probability look_p[3] (sunVisibility, stochastic).  d_shadow: float per pixel,
 * the raw sample (1 sky, 0 facing away / blocked, 1 lit); d_hits (nullable): the first segment. */
int vf_trace_sun_cone(vf_ctx *ctx, const vf_pass_ubo *pass, int32_t width, int32_t height, const float *d_gb_pos,
                      const uint32_t *d_gb_normal, float *d_shadow, vf_hit *d_hits);
/* P::Gbuf followed by P::ShTrace (record.rs:436, 463: record_standard's order) in ONE kernel: the same
 * outputs as vf_trace_primary + vf_trace_shadow (the sun ray starts from the written, quantised
 * G-buffer values), without the second launch and the G-buffer read-back.  gb_pos / gb_normal may be
 * NULL when only hits and shadowRaw are wanted; shadow_hits is nullable. */
int vf_trace_primary_shadow(vf_ctx *ctx, const vf_pass_ubo *pass, int32_t width, int32_t height, vf_hit *d_hits,
                            float *d_gb_pos, uint32_t *d_gb_normal, uint32_t *d_shadow_raw, vf_hit *d_shadow_hits);
/* P::Gi (record.rs:350/361; gi.comp:34-84 + giPath 1 bounce): cosine-hemisphere GI ray from the
 * quantised G-buffer normal (R2 sequence, GI_LDS), its sun ray, sky escape via the sky-top map.
 * half_res != 0: the reference's half-resolution dispatch ((w+1)/2 x (h+1)/2 outputs); else one
 * ray per full-res pixel.  radiance: RGBA32F (L.rgb, 1; sky pixel 0).  d_bounce (nullable): the
 * bounce ray's hit.  gi_reach: GI_REACH in voxels (r.gi.reach, default 48). */
int vf_trace_gi(vf_ctx *ctx, const vf_pass_ubo *pass, int32_t width, int32_t height, int32_t half_res,
                float gi_reach, int32_t lights_nee, const float *d_gb_pos, const uint32_t *d_gb_normal,
                float *d_radiance, vf_hit *d_bounce);
/* traceVolumeS over the world for an explicit batch (common.glsl:541-658): skip_mat 0 (water is a
 * hit) or 5 (water is clear, dry field).  t_cap is tCap (1e30 = none). */
int vf_trace_rays(vf_ctx *ctx, const vf_ray *d_rays, int64_t n, float t_cap, uint32_t skip_mat, vf_hit *d_hits);
/* probe_trace.comp ray source (probes.glsl:155-170) for a probe grid: n_rays spherical-Fibonacci
 * directions per probe, rotated per frame; probe p's origin = origin + spacing*(p.x,p.y,p.z),
 * p = (i % dims.x, i / dims.x % dims.y, i / dims.x / dims.y).  Hits [n_probes * n_rays]. */
int vf_trace_probes(vf_ctx *ctx, const float origin[3], float spacing, const int32_t dims[3], int32_t n_rays,
                    uint32_t frame, float reach, vf_hit *d_hits);

/* ---- persistent work-queue batches ------------------------------------------------------- */
/* Queue parameters: the pixel tile that consecutive ray ids cover (8x8, 16x8, 16x16, 8x16, 32x8), rays
 * a wavefront claims per queue fetch (64..4096, multiple of 64), whether GI rays are grouped by
 * direction octant within each 256-pixel block, an optional cap on workgroups per CU (0 = as many as
 * fit) and how many lanes must be idle before a wavefront fetches new rays (1..64; 0 = chunk mode: a
 * wavefront traces its claimed chunk 64 rays at a time, each to completion).  Results never depend
 * on these. */
/* refill -1: no queue.  vf_trace_gi_batch: one lane per pixel traces its GI ray and then its sun ray.
 * vf_trace_batch: one launch with a wavefront per 64 consecutive ray ids of the requests (in order, tile
 * order within a view), each traced to completion; completion stamps as in the queue modes.  The autotuner
 * measures it beside the queue configurations. */
/* traversal (vf_trace_batch with refill -1 only): 0 the plain traversal, 1 the small-frame one (see
 * vf_set_traversal); the same hits either way. */
typedef struct vf_queue_cfg {
    int32_t tile_w, tile_h, chunk, sort, groups_per_cu, refill;
    int32_t traversal;
} vf_queue_cfg;
int vf_queue_defaults(vf_queue_cfg *cfg);

/* Config-3 style batch: one GI ray per full-resolution G-buffer pixel (gi.comp:55-79 with fpix = pix:
 * cosine-hemisphere direction from the quantised normal and the R2 sequence, origin P + N*0.03,
 * t_cap = reach) and one sun ray (giPathImpl's, common.glsl:2318-2330: from hit + N*0.03 toward
 * sunDir, no cap) per GI hit that is not water and faces the sun.  Both go through the persistent
 * queue kernel.  gi_hits / sun_hits: one record per pixel (miss where no ray).  Not a reference
 * entry point: it batches what P::Gi traces, for incoherent-ray throughput. */
/* With bodies in the scene (GI_OBJECTS, common.glsl:780, 2267-2271) a GI ray also meets the bodies nearer than
 * min(its world hit or reach, 16); a body hit's sun ray leaves along the body's rotated normal; a sun ray that
 * missed the world meets any body.  Body hits carry the body flag. */
int vf_trace_gi_batch(vf_ctx *ctx, const vf_pass_ubo *pass, int32_t width, int32_t height, float reach,
                      const float *d_gb_pos, const uint32_t *d_gb_normal, vf_hit *d_gi_hits, vf_hit *d_sun_hits,
                      const vf_queue_cfg *cfg);

#define VF_MAX_REQUESTS 24
enum { VF_REQ_VIEW = 0, VF_REQ_PROBES = 1 };
/* One request of an uneven batch: a camera view (primary rays as vf_trace_primary, hits [w*h]) or a
 * DDGI probe grid (as vf_trace_probes, hits [n_probes*n_rays]).  With bodies in the scene both meet them
 * as those entry points do (traceSceneS for a view, GI_ADD_OBJECTS for a probe ray). */
typedef struct vf_request {
    int32_t kind;
    int32_t width, height;
    float cam_pos[3], cam_right[3], cam_up[3], cam_fwd[3];
    float tan_half_fov;
    float jitter[2];
    float origin[3];
    float spacing;
    int32_t dims[3];
    int32_t n_rays;
    uint32_t frame;
    float reach;
    vf_hit *d_hits;
} vf_request;
/* All requests in ONE persistent launch; each request's completion time is stamped on the device.
 * Request order is the service order (put the latency-critical ones first).  Recordable into a graph. */
int vf_trace_batch(vf_ctx *ctx, const vf_request *reqs, int32_t n_req, const vf_queue_cfg *cfg);
/* vf_trace_batch without any allocation: no reference counterpart.  For callers that cannot allocate device
 * memory (code under HIP graph capture before the batch ever ran eagerly, serving loops and harnesses that
 * forbid hipMalloc).  Same requests, outputs and limits as vf_trace_batch (1..VF_MAX_REQUESTS requests, else
 * VF_E_INVALID); the batch's counters and completion stamps live in a 512-byte array that the library keeps
 * on each GPU until the process ends, instead of the context's queue buffers.  Every call on one GPU shares
 * that array (the batch clears it first): calls that overlap on one GPU, from any context or stream, race.
 * vf_read_batch_times reads the stamps of whichever batch ran last on the context. */
int vf_trace_batch_noalloc(vf_ctx *ctx, const vf_request *reqs, int32_t n_req, const vf_queue_cfg *cfg);
/* Per-request completion times of the last batch, ms after the launch started (host, synchronises). */
int vf_read_batch_times(vf_ctx *ctx, int32_t n_req, double *ms);

/* Autotuner: times every refill mode {chunk mode, refill 32} x tile shape {8x8, 16x8, 16x16} x chunk
 * {64..1024} (x sort off/on for the GI batch), plus refill -1 (no queue: the GI batch at 16x16, the batch
 * at each tile shape with either traversal), on the given batch (median of 3 runs) and returns the fastest
 * configuration; base (nullable) supplies groups_per_cu. */
enum { VF_TUNE_GI_BATCH = 0, VF_TUNE_BATCH = 1 };
typedef struct vf_tune_result { vf_queue_cfg cfg; float ms; } vf_tune_result;
typedef struct vf_tune_gi_args {
    const vf_pass_ubo *pass;
    int32_t width, height;
    float reach;
    const float *d_gb_pos;
    const uint32_t *d_gb_normal;
    vf_hit *d_gi_hits, *d_sun_hits;
} vf_tune_gi_args;
typedef struct vf_tune_batch_args { const vf_request *reqs; int32_t n_req; } vf_tune_batch_args;
int vf_autotune(vf_ctx *ctx, int32_t workload, const void *args, const vf_queue_cfg *base, vf_queue_cfg *best,
                vf_tune_result *results, int32_t max_results, int32_t *n_results);

/* ---- frame graphs ------------------------------------------------------------------------- */
/* A frame's trace calls recorded once and replayed (the reference records a command buffer per
 * frame, record.rs:320).  Between vf_graph_begin and vf_graph_end the trace entry points are
 * recorded, not run; calls that synchronise or copy (uploads, reads, vf_sync) are refused.  The
 * recorded calls keep their arguments (buffers, uniforms) as they were at capture time.  The
 * context's stream must not be the null stream. */
typedef struct vf_graph vf_graph;
int vf_graph_begin(vf_ctx *ctx);
int vf_graph_end(vf_ctx *ctx, vf_graph **out);
int vf_graph_launch(vf_ctx *ctx, vf_graph *graph);
void vf_graph_destroy(vf_graph *graph);

/* ---- measurement --------------------------------------------------------------------------- */
/* When on, the traces launch their counting variants, which add the reference visit sequence's
 * fetch bytes (grid 4 B per brick step, 1 B per distance/dry/coarse/atlas fetch) to counters. */
int vf_set_counting(vf_ctx *ctx, int32_t on);
/* counters: {rays, brick steps, voxel steps, fetch bytes}; reset = 1 zeroes them after reading. */
int vf_read_counters(vf_ctx *ctx, uint64_t out[4], int32_t reset);
/* Which traversal the primary and primary+sun passes (vf_trace_primary, vf_trace_primary_shadow; camera in
 * air) and vf_trace_rays launch; both return the same hits.  0 (default): by launch size.  Launches of at
 * most 16 wavefronts per SIMD (small frames, whose time is that of their slowest rays) take the
 * small-frame traversal: it tests each detailed brick's occupancy mask over the rows and columns its walk
 * can visit and skips walks that cannot hit, walks two voxels per loop trip, and raises the issue priority
 * of long-running waves.  Larger launches (issue-bound) take the plain one.  1: always the plain one.
 * 2: always the small-frame one.  The counting variants (vf_set_counting) always take the plain one.  No
 * reference counterpart: a kernel-selection knob of this library. */
int vf_set_traversal(vf_ctx *ctx, int32_t mode);
/* When on, every trace entry point records a HIP event before and after its launches, for
 * vf_last_kernel_ms.  Off by default: the two event records add GPU time to each call. */
int vf_set_timing(vf_ctx *ctx, int32_t on);
/* Kernel time of the last trace entry point on the context's stream, from HIP events (ms); needs
 * vf_set_timing(ctx, 1) before that call. */
int vf_last_kernel_ms(vf_ctx *ctx, float *ms);

#ifdef __cplusplus
}
#endif
#endif
