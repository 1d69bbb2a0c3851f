/*
 * CPU oracle for the brick-DDA voxel traversal — TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference's traversal (shaders/common.glsl
 * traceVolumeS / brickVoxels), its distance-field build (shaders/brickdist.comp), its
 * ray generators (gbuffer.comp, gi.comp, probes.glsl) and its CPU-side mirror builders
 * (gpu_world.rs, sparse_volume.rs).  It exists only to check the HIP kernels: tests,
 * the smoke check and the CPU-baseline leg of the benchmark may call it; the product
 * library never links or loads it.
 *
 * Numerics: FP32 in the reference's operation order, compiled with -ffp-contract=off and
 * IEEE division / square root, so that GPU results can be compared bit for bit.  GLSL's
 * min/max/clamp are restated with the GLSL definitions (min(x,y) = y < x ? y : x), which
 * matters for signed zeros.
 *
 * Pinning (see the tests under tests/):
 *   - traversal: the plain brick-by-brick DDA (vfo_world.use_dist = 0, coarse_skip = 0) and an
 *     independent FP64 plain voxel march (vfo_march_f64); analytic scenes; the reference's
 *     raycast_hits_floor case (sparse_volume.rs:669-679) through vfo_raycast_cpu.
 *   - brick masks: the reference's mask_bits vector (sparse_volume.rs:660-668).
 *   - distance field: the closed form min(16, L-inf distance) by brute force, incl. ring wrap.
 *   - coarse pyramid and sky-top map: closed forms on small grids.
 *   - RNG: pcgHash is a bijection check plus the published RXS-M-XS structure; the exact
 *     values are "parity unpinned" beyond GPU-vs-oracle agreement (no reference vectors).
 *   - GI radiance: estimator structure only (statistical), "parity unpinned" bitwise.
 *   - two-bounce paths: a path stopped at its tail and continued from there by hand equals the whole
 *     path; one-bounce = the pass the GI kernels were first pinned with.
 */
#ifndef VF_ORACLE_H
#define VF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One world as the traversal sees it (the GPU world mirror's textures, common.glsl:11-44). */
typedef struct vfo_world {
    int32_t wb[3];            /* worldBricks (SceneUbo@16) */
    int32_t win[4];           /* worldWin (SceneUbo@352); .w != 0 -> ring (toroidal x/z) */
    const uint32_t *grid;     /* gridTex R32UI, index (z*wby + y)*wbx + x (wrapped coords) */
    const uint8_t *atlas;     /* slot-major: slot*512 + lz*64 + ly*8 + lx */
    uint32_t n_slots;
    const uint8_t *dist;      /* brickDistTex R8UI, grid layout */
    const uint8_t *dry;       /* brickDistDryTex R8UI, grid layout */
    const uint8_t *coarse1;   /* coarse1Tex R8UI, ceil(wb/4) */
    const uint8_t *coarse2;   /* coarse2Tex R8UI, ceil(wb/16) */
    int32_t c1[3], c2[3];
    int32_t use_dist;         /* USE_DIST_FIELD (default on, cvars.rs:110) */
    int32_t coarse_skip;      /* kCoarseSkip (common.glsl:20); 0 = plain brick-by-brick DDA */
    /* bodies (common.glsl:65-82): numObjects, the ObjHeaders (240 B = 60 floats each: worldToObj,
     * objToWorld, prevObjToWorld column-major, dimsBricks as int bits, aabbMin, aabbMax), objGrid, and the
     * BVH (8 floats per node: min.xyz, first bits, max.xyz, count bits) with its prims. */
    int32_t n_bodies;
    const float *bodies;
    const uint32_t *obj_grid;
    const float *bvh;
    int32_t n_nodes;
    const uint32_t *bvh_prims;
    int32_t use_bvh;          /* USE_OBJ_BVH (default on, renderer/mod.rs:96); 0 = the loop over every body */
    const float *palette;     /* materials[256] (12 floats each, kind at [6]): gbNormal.a; NULL = no foliage */
} vfo_world;

typedef struct vfo_hit {
    int32_t hit;
    float t;
    int32_t n[3];             /* face normal, one axis +-1, or all 0 (ray started inside the box) */
    uint32_t mat;
    int32_t cell[3];          /* absolute voxel coordinates (a body's: its own voxel coordinates) */
    int32_t body;             /* 0 = the world, else the body index + 1 (Hit.obj + 1) */
} vfo_hit;

/* Fetch counters of the algorithmic byte model: grid 4 B per outer iteration, 1 B per
 * distance / dry / coarse fetch, 1 B per atlas voxel fetch. */
typedef struct vfo_count {
    uint64_t rays, outer, voxel, bytes;
} vfo_count;

/* Explicit ray record (24 B) and hit record (20 B), the same as the C-ABI's vf_ray / vf_hit. */
typedef struct vfo_ray { float o[3]; float d[3]; } vfo_ray;
typedef struct vfo_hitrec { int32_t cell[3]; float t; uint8_t mat, face, hit, body; } vfo_hitrec;

/* --- traversal (common.glsl:474-658) --- */
void vfo_trace(const vfo_world *w, const float ro[3], const float rd[3], float tcap, uint32_t skip_mat,
               vfo_hit *out, vfo_count *cnt);
void vfo_trace_batch(const vfo_world *w, const vfo_ray *rays, int64_t n, float tcap, uint32_t skip_mat,
                     vfo_hitrec *out, vfo_count *cnt, int threads);
/* vfo_trace_batch with the library's small-frame mask test before each walk (margin 0.05 and mode 0 as the
 * library; mode 1 a planted mistake); stats = {walks skipped, walks run}.  Checks that the test never
 * skips a walk that hits. */
void vfo_trace_box_batch(const vfo_world *w, const vfo_ray *rays, int64_t n, float tcap, uint32_t skip_mat, float margin,
                         int mode, vfo_hitrec *out, int64_t stats[2], int threads);
/* --- bodies (common.glsl:664-720, 787-793; gpu_world.rs:550-600) --- */
/* build_bvh over the bodies' world boxes: nodes (8 floats each, at most 2n - 1) and prims (n). */
void vfo_build_bvh(const float *bodies, int32_t n, float *nodes, int32_t *n_nodes, uint32_t *prims);
/* traceObjectsA: the nearest body hit closer than limit merged into *best (any_hit: the first found). */
void vfo_trace_objects(const vfo_world *w, const float ro[3], const float rd[3], vfo_hit *best, float limit,
                       int32_t any_hit, vfo_count *cnt);
/* traceSceneS: the world, then the bodies closer than its hit. */
void vfo_trace_scene(const vfo_world *w, const float ro[3], const float rd[3], float tcap, uint32_t skip_mat,
                     vfo_hit *out, vfo_count *cnt);
void vfo_trace_scene_batch(const vfo_world *w, const vfo_ray *rays, int64_t n, float tcap, uint32_t skip_mat,
                           vfo_hitrec *out, int32_t *body, int threads);
/* The hit's world normal: the face normal, or normalize(mat3(objToWorld) * n) for a body. */
void vfo_hit_normal(const vfo_world *w, const vfo_hit *h, float out[3]);
/* FP64 plain voxel march over the same voxels (independent pin for t). */
void vfo_march_f64(const vfo_world *w, const double ro[3], const double rd[3], double tcap, uint32_t skip_mat,
                   int32_t *hit, double *t, int32_t cell[3], int32_t n[3], uint32_t *mat);
/* The reference's CPU raycast (sparse_volume.rs:573-637) over a world's voxels. */
void vfo_raycast_cpu(const vfo_world *w, const float ro[3], const float rd[3], float max_t, uint32_t skip_mat,
                     int32_t *hit, float *t, int32_t cell[3], int32_t n[3], uint32_t *mat);
uint32_t vfo_voxel(const vfo_world *w, int32_t x, int32_t y, int32_t z);
int32_t vfo_face_code(const int32_t n[3]);

/* --- mirror builders (gpu_world.rs:300-313, sparse_volume.rs:48-63, brickdist.comp) --- */
void vfo_brick_mask(const uint8_t vox[512], uint32_t out[16]);
void vfo_coarse(const uint32_t *grid, const int32_t wb[3], uint8_t *c1, uint8_t *c2);
/* One brickdist.comp dispatch (exact restatement).  src: grid (u32) for first_pass 1/2,
 * else a u8 field.  dst: u8 field.  atlas: slot-major. */
void vfo_brickdist_pass(const uint32_t *grid, const uint8_t *src_u8, uint8_t *dst, const uint8_t *atlas,
                        const int32_t wb[3], const int32_t box_min[3], const int32_t box_max[3],
                        int32_t axis, int32_t first_pass);
/* record_distance_field (gpu_world.rs:1376-1428): 7 passes over 4 fields (d0 = normal, d2 = dry). */
void vfo_record_distance_field(const uint32_t *grid, const uint8_t *atlas, const int32_t wb[3], int32_t ring,
                               const int32_t lo[3], const int32_t hi[3], const int32_t dlo[3], const int32_t dhi[3],
                               uint8_t *d0, uint8_t *d1, uint8_t *d2, uint8_t *d3);
/* Closed form: min(16, L-inf distance to the nearest occupied brick), brute force. dry: water-only = empty. */
void vfo_brickdist_brute(const uint32_t *grid, const uint8_t *atlas, const int32_t wb[3], int32_t ring, int32_t dry,
                         uint8_t *out);
/* Sky-top map (grasstop.comp:121-146): per voxel column, 1 + the highest voxel whose kind is not
 * water (1) or foliage (3); 0 if none.  kinds[256] = palette props.z. */
void vfo_skytop(const vfo_world *w, const float *kinds, uint16_t *out);

/* --- RNG and ray generators (common.glsl:1233-1288, gbuffer.comp:194-202, probes.glsl:155-170) --- */
uint32_t vfo_pcg(uint32_t v);
uint32_t vfo_pixel_seed(int32_t px, int32_t py, uint32_t frame, uint32_t salt);
void vfo_ld_sample2(int32_t px, int32_t py, uint32_t k, uint32_t salt, float out[2]);
void vfo_cosine_hemisphere_u(const float n[3], const float u[2], float out[3]);
float vfo_sin(float x);
float vfo_cos(float x);
void vfo_primary_dir(const float cam_right[3], const float cam_up[3], const float cam_fwd[3], float tan_half_fov,
                     int32_t w, int32_t h, int32_t px, int32_t py, const float jitter[2], float out[3]);
void vfo_probe_dir(int32_t i, int32_t n_rays, uint32_t frame, float out[3]);

/* --- camera inside solid terrain (gbuffer.comp:233-266) --- */
uint32_t vfo_voxel_mat_at(const vfo_world *w, const float p[3]);
/* exitSolid (common.glsl:990-1016): distance to the first non-solid voxel; -2 leaves the world, -1 too long. */
float vfo_exit_solid(const vfo_world *w, const float ro[3], const float rd[3], int32_t max_steps);

/* --- passes --- */
typedef struct vfo_cam {
    float pos[3], right[3], up[3], fwd[3];
    float tan_half_fov;
    float jitter[2];
    uint32_t frame_idx;
} vfo_cam;

typedef struct vfo_sky {
    float sun_dir[3], sun_rgb[3];
    float sky_zenith[3], sky_horizon[3], sunset_rgb[3];
    float sky_scale, gi_scale, night;
    int32_t sky_model;
    float gi_reach;
    int32_t lights_nee;
    const float *palette;     /* 256 x 12 floats (Material, 48 B) */
    const uint16_t *skytop;   /* (wbx*8) x (wbz*8), index z*(wbx*8) + x */
} vfo_sky;

/* gbuffer.comp primary pass over the scene (traceScenePrimary; a camera inside solid terrain or water as
 * gbuffer.comp:233-266): hits + gbPos (xyz, t) + gbNormal (RGBA8) + the body index (-1 world / sky). */
void vfo_pass_primary(const vfo_world *w, const vfo_cam *c, int32_t width, int32_t height,
                      vfo_hitrec *hits, float *gb_pos, uint32_t *gb_normal, int32_t *body, vfo_count *cnt, int threads);
void vfo_pass_primary_rows(const vfo_world *w, const vfo_cam *c, int32_t width, int32_t height, int32_t y0,
                           int32_t ystep, int32_t nrows, vfo_hitrec *hits, float *gb_pos, uint32_t *gb_normal,
                           vfo_count *cnt, int threads);
/* shadow_trace.comp sun ray per pixel with FOLIAGE_LIGHT (leaf_light = lookP.w): shadowRaw as two floats
 * (visibility, min(tFirst, 60000)); hits = the first segment. palette: 256 x 12 floats. */
void vfo_pass_shadow(const vfo_world *w, const float *palette, const float sun_dir[3], float leaf_light, int32_t width,
                     int32_t height, const float *gb_pos, const uint32_t *gb_normal, float *shadow, vfo_hitrec *hits,
                     vfo_count *cnt, int threads);
/* sunVisibilityT (common.glsl:1964-1981); stochastic as SVGF's lighting pass, else as shadow_trace.comp. */
float vfo_sun_visibility_t(const vfo_world *w, const float *palette, const float o[3], const float d[3],
                           float leaf_light, uint32_t *seed, int32_t stochastic, float *tfirst, vfo_hit *first,
                           vfo_count *cnt);
void vfo_sample_cone(const float dir[3], float ang, uint32_t *state, float out[3]);
/* lighting.comp's cone-jittered stochastic sun sample per pixel (1 sky, 0 / 1 otherwise). */
void vfo_pass_sun_cone(const vfo_world *w, const float *palette, const float sun_dir[3], float sun_angle,
                       float leaf_light, uint32_t frame, int32_t width, int32_t height, const float *gb_pos,
                       const uint32_t *gb_normal, float *out, vfo_hitrec *hits, int threads);
/* gi.comp sampling + giPath (1 bounce): radiance rgb + the bounce ray's hit. half_res: the reference's
 * half-resolution dispatch (gi.comp:34-84); else one ray per full-res pixel. */
void vfo_pass_gi(const vfo_world *w, const vfo_sky *s, uint32_t frame_idx, int32_t width, int32_t height,
                 int32_t half_res, const float *gb_pos, const uint32_t *gb_normal, float *radiance,
                 vfo_hitrec *bounce, vfo_count *cnt, int threads);
/* GI path from an explicit point/normal/direction (giPathImpl with lastStage 1). */
void vfo_gi_path(const vfo_world *w, const vfo_sky *s, const float P[3], const float N[3], const float dir[3],
                 float L[3], vfo_hit *bounce, vfo_count *cnt);
/* The point where a two-bounce path's bounce 2 would start: hit point, its normal, the bounce-1 ray's
 * direction and the throughput after `thr *= hitAlb * giScale`. */
typedef struct vfo_tail { int32_t valid; float p[3], n[3], dir[3], thr[3]; int32_t body; } vfo_tail;
/* Query / training record, the same 64 B as the C-ABI's vf_gi_record. */
typedef struct vfo_record {
    float pos[3]; uint32_t tag; float n[3]; float pad0; float view[3]; float pad1; float rgb[3]; float pad2;
} vfo_record;
void vfo_cosine_hemisphere(const float n[3], uint32_t *state, float out[3]);
/* giPathImpl (common.glsl:2250-2366) with lastStage 1 or 3, traceOn, and an optional stop at the tail. */
void vfo_gi_path_ex(const vfo_world *w, const vfo_sky *s, const float P[3], const float N[3], const float dir[3],
                    uint32_t *seed, int32_t last_stage, int32_t trace_on, float L[3], vfo_hit *bounce, vfo_tail *tail,
                    vfo_count *cnt);
/* A three-bounce path (the library's, not the reference's: its giPath stops after two): the two-bounce path
 * stopped where bounce 2 would start (a body hit included), then a two-bounce path from there with the next
 * cosine draw of the seed and no trace-on, weighted by the tail's throughput. */
void vfo_gi_path3(const vfo_world *w, const vfo_sky *s, const float P[3], const float N[3], const float dir[3],
                  uint32_t *seed, float L[3], vfo_count *cnt);
/* Plain traced GI (bounces 2: giPath; 3: vfo_gi_path3) with spp samples per pixel (RGBA32F out; a = 0 for sky). */
void vfo_pass_gi_path(const vfo_world *w, const vfo_sky *s, int32_t width, int32_t height, const float *gb_pos,
                      const uint32_t *gb_normal, uint32_t frame, int32_t spp, uint32_t salt, int32_t bounces, float *out,
                      int threads);
/* Cached render, traced part (tail at the bounce-1 hit; bounces 3: a body's tail is traced on instead). */
void vfo_gicache_prepass(const vfo_world *w, const vfo_sky *s, int32_t width, int32_t height, const float *gb_pos,
                         const uint32_t *gb_normal, uint32_t frame, int32_t spp, uint32_t salt, int32_t bounces,
                         float *pre, vfo_record *queries, int threads);
/* Training samples of one frame (terminate 0: at the G-buffer point; 1: at the bounce-1 hit). */
void vfo_gicache_samples(const vfo_world *w, const vfo_sky *s, const float cam_pos[3], int32_t width, int32_t height,
                         const float *gb_pos, const uint32_t *gb_normal, uint32_t frame, int32_t n,
                         int32_t terminate, uint32_t salt, int32_t bounces, vfo_record *out, int threads);
float vfo_sky_open(const vfo_world *w, const vfo_sky *s, const float p[3]);
/* GI rays of a full-resolution G-buffer (the GI batch: one per non-sky pixel). */
void vfo_gi_rays(int32_t width, int32_t height, uint32_t frame, const float *gb_pos, const uint32_t *gb_normal,
                 vfo_ray *rays, uint8_t *valid);

/* ---- temporal consumers (vfo_temporal.c): gi.comp accumulation, lighting.comp history, a-trous ----
 * Layouts: gb_pos RGBA32F and gb_normal RGBA8 per full-res pixel; gb_motion RG16F (x in the low
 * half); gb_delta RGBA16F; GI history RGBA16F, GI moments RGBA32F, guide RGBA32UI per half-res
 * texel ((w+1)/2 x (h+1)/2); indirect RGBA32F per half-res texel (vf_trace_gi's radiance); shadow
 * float and shadow history RGBA32F per full-res pixel. */
float vfo_exp(float x);
float vfo_pow16(float x);
float vfo_pow64(float x);
float vfo_f16_to_f32(uint16_t h);
uint16_t vfo_f32_to_f16(float f);
/* gbuffer.comp's gbMotion / gbDelta from the written G-buffer (body: index per pixel or -1, nullable;
 * bodies: 60 floats per body as vf_obj_header). */
void vfo_gbuffer_motion(int32_t w, int32_t h, const float cam_right[3], const float cam_up[3], const float cam_fwd[3],
                        float thf, const float jit[2], const float prev_pos[3], const float prev_right[3],
                        const float prev_up[3], const float prev_fwd[3], float prev_thf, const float *gb_pos,
                        const int32_t *body, const float *bodies, int32_t n_bodies, uint32_t *motion, uint16_t *delta);
void vfo_gi_accumulate(int32_t w, int32_t h, uint32_t frame_idx, const float shift[3], int32_t bicubic,
                       const float *gb_pos, const uint32_t *gb_normal, const float *gb_pos_prev,
                       const uint32_t *gb_normal_prev, const uint32_t *gb_motion, const uint16_t *gb_delta,
                       const float *indirect, const uint16_t *hist_prev, const float *mom_prev, uint16_t *hist_cur,
                       float *mom_cur, uint32_t *guide);
void vfo_light_accumulate(int32_t w, int32_t h, const float shift[3], const float *gb_pos, const uint32_t *gb_normal,
                          const float *gb_pos_prev, const uint32_t *gb_normal_prev, const uint32_t *gb_motion,
                          const uint16_t *gb_delta, const float *shadow, const float *hist_prev, float *hist_cur);
void vfo_gi_atrous_level(int32_t w, int32_t h, int32_t it, int32_t step, int32_t histfix, const float *gb_pos,
                         const uint32_t *gb_normal, const float *gi_mom, const uint16_t *gi_hist, const uint32_t *guide,
                         const uint16_t *in_cv, uint16_t *out_cv);
void vfo_gi_atrous(int32_t w, int32_t h, int32_t levels, int32_t histfix, const float *gb_pos, const uint32_t *gb_normal,
                   const float *gi_mom, const uint16_t *gi_hist, const uint32_t *guide, uint16_t *filt0, uint16_t *filt1);

#ifdef __cplusplus
}
#endif
#endif
