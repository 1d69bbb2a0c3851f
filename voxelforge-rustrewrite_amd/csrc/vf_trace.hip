// MI355X (gfx950) brick-DDA voxel traversal: HIP kernels and the C ABI of include/vf_trace.h.
//
// The traversal restates traceVolumeS / brickVoxels (reference shaders/common.glsl:474-658) with
// the same FP32 operation order (built with -ffp-contract=off, IEEE division and square root),
// so hits match the reference algorithm bit for bit.  What changes is how the data reaches the
// arithmetic:
//   * one 8-byte "cell word" per brick packs the grid value (gridTex), both distance fields
//     (brickDistTex, brickDistDryTex) and the two coarse-pyramid flags of the brick's cells
//     (coarse1Tex, coarse2Tex), so an outer DDA step is ONE load instead of the reference's
//     dependent chain grid -> dist -> coarse2 -> coarse1.  Cell words are stored in 4x4x4-brick
//     tiles (512 B) so a ray's neighbouring bricks share cache lines.
//   * each detailed brick has two 512-bit occupancy masks (m != 0, and m not in {0, water}),
//     fetched as four 16-byte loads when the ray enters the brick; the voxel DDA then tests bits
//     in registers and touches the 512-byte atlas brick once, for the hit's material.
//   * 256-thread workgroups trace 16x16 pixel tiles (one 8x8 tile per wavefront), and workgroup
//     ids are remapped so that each of the 8 XCDs (each with its own L2) gets a contiguous band
//     of tiles.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/vf_trace.h"
#include "vf_ctx.h"
#include "vf_device.h"
#include "vf_gi.h"
#include "vf_internal.h"

// ------------------------------------------------------------------------------------------------
// pass kernels
// ------------------------------------------------------------------------------------------------
struct PrimArgs {
    float pos[3], right[3], up[3], fwd[3];
    float thf;
    float jit[2];
    int32_t w, h, tiles_x, ntiles;
    int32_t cam_air;   // the host's grid copy shows the camera in an empty brick (or outside the world)
    float leaf;        // lookP.w = r.foliage.light: sunlight passing through each leaf (FOLIAGE_LIGHT)
    float aspect;      // (float)w / (float)h, the IEEE quotient, computed once on the host
};

// The camera voxel decides which primary kernel variant's results stand (gbuffer.comp:233-241).  A camera
// in air (the common case): every ray starts at the camera (a uniform origin, kept in scalar registers)
// with skipMat 0; the camera voxel is loaded at the start but only checked before the results are
// written, so its latency overlaps the trace.  A camera inside solid terrain (each ray starts where it
// leaves the solid: a per-lane origin) or in water (skipMat 5): the SPECIAL variant, a small persistent
// grid that returns at once for a camera in air.  Both are launched.
__device__ __forceinline__ bool camera_special(uint32_t cam_mat) { return cam_mat != 0u; }

// Pixel (px, py) of a 16x16-tile pass: its primary ray direction (gbuffer.comp:194-202).
__device__ __forceinline__ void primary_dir(const PrimArgs &A, int32_t px, int32_t py, float rd[3]) {
    const float sx = ((float)px + 0.5f) + A.jit[0];
    const float sy = ((float)py + 0.5f) + A.jit[1];
    const float nx = (sx / (float)A.w) * 2.0f - 1.0f;
    const float ny = (sy / (float)A.h) * 2.0f - 1.0f;
    const float a = (nx * A.aspect) * A.thf;
    const float b = ny * A.thf;
#pragma unroll
    for (int i = 0; i < 3; ++i) rd[i] = (A.fwd[i] + A.right[i] * a) + A.up[i] * b;
    normalize3(rd);
}

template <bool COUNT, bool SPECIAL, int LS = 256, bool BOX = false>
__device__ __forceinline__ void primary_hit(const DevWorld &W, const PrimArgs &A, uint32_t cam_mat, const float rd[3],
                                            Hit &h, Cnt &c) {
    if (SPECIAL) primary_trace<COUNT, true>(W, A.pos, cam_mat, rd, h, c);
    else trace_world<0, COUNT, true, LS, false, BOX>(W, A.pos, rd, 1e30f, h, c);
}

// Tile L's pixel for this thread (8x8 per wave, 16x16 per workgroup).
__device__ __forceinline__ void tile_pixel_of(int32_t L, int32_t tiles_x, int32_t &px, int32_t &py) {
    const int32_t tx = L % tiles_x, ty = L / tiles_x;
    const int32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    px = tx * 16 + (wv & 1) * 8 + (lane & 7);
    py = ty * 16 + (wv >> 1) * 8 + (lane >> 3);
}

// gbuffer.comp primary ray (gbuffer.comp:194-202, 233-266, 470-480) for a 16x16 tile per workgroup.
template <bool COUNT, bool SPECIAL, int LS = 256, bool BOX = false>
__device__ __forceinline__ void primary_pixel(const DevWorld &W, const PrimArgs &A, uint32_t cam_mat, int32_t px, int32_t py,
                                              vf_hit *hits, float4 *gbpos, uint32_t *gbn, Cnt &c) {
    float rd[3];
    primary_dir(A, px, py, rd);
    Hit h;
    primary_hit<COUNT, SPECIAL, LS, BOX>(W, A, cam_mat, rd, h, c);
    if (camera_special(cam_mat) != SPECIAL) return;
    const size_t i = (size_t)py * (size_t)A.w + (size_t)px;
    store_hit(hits + i, h);
    if (gbpos) {
        float4 p;
        if (h.hit) p = make_float4(A.pos[0] + rd[0] * h.t, A.pos[1] + rd[1] * h.t, A.pos[2] + rd[2] * h.t, h.t);
        else p = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
        gbpos[i] = p;
    }
    if (gbn) {
        uint32_t v = 0u;
        if (h.hit) v = face_normal_bytes(h.face) | (mat_flag_byte(W, h.mat) << 24);
        gbn[i] = v;
    }
}

// SPECIAL = false: one workgroup per tile (XCD-remapped); SPECIAL = true: a small persistent grid over
// the tiles.
template <bool COUNT, bool SPECIAL>
__global__ void __launch_bounds__(256) k_primary(DevWorld W, PrimArgs A, vf_hit *hits, float4 *gbpos, uint32_t *gbn,
                                                 unsigned long long *cnt) {
    const uint32_t cam_mat = A.cam_air ? 0u : voxel_mat_at(W, A.pos);
    if (SPECIAL && !camera_special(cam_mat)) return;
    Cnt c = {0, 0, 0, 0};
    if (!SPECIAL) {
        int32_t px, py;
        tile_pixel_of(xcd_remap((int32_t)blockIdx.x, (int32_t)gridDim.x), A.tiles_x, px, py);
        if (px < A.w && py < A.h) primary_pixel<COUNT, false>(W, A, cam_mat, px, py, hits, gbpos, gbn, c);
    } else {
        for (int32_t L = (int32_t)blockIdx.x; L < A.ntiles; L += (int32_t)gridDim.x) {
            int32_t px, py;
            tile_pixel_of(L, A.tiles_x, px, py);
            if (px < A.w && py < A.h) primary_pixel<COUNT, true>(W, A, cam_mat, px, py, hits, gbpos, gbn, c);
        }
    }
    if (camera_special(cam_mat) == SPECIAL) flush_count<COUNT>(c, cnt);   // only the variant whose results stand
}

struct TileArgs {
    int32_t w, h, tiles_x, ntiles;
    float leaf;        // lookP.w (FOLIAGE_LIGHT)
};

__device__ __forceinline__ void store_miss(vf_hit *out) {
    Hit m;
    m.hit = 0; m.t = -1.0f; m.face = 0; m.mat = 0; m.cell[0] = m.cell[1] = m.cell[2] = 0;
    store_hit(out, m);
}

// shadow_trace.comp's result for one G-buffer texel (shadow_trace.comp:21-49, 87, 96-105): sky -> lit;
// facing away from the sun -> shadowed without a ray, unless the texel is foliage (gbNormal.a) and the sun
// is above the horizon: then FOLIAGE_SSS traces sunVisibility from foliageSunOrigin (the light through the
// leaf, no blocker distance); else sunVisibilityT (FOLIAGE_LIGHT, deterministic) from P + N*0.03:
// visibility (leaves let leaf of the light through) and the distance to the first blocker.
// NDEC: N holds the decoded normal already (only nq's alpha is read, which is then 204 or 255).
template <bool COUNT, bool BODIES = false, bool LDSM = false, int LS = 256, bool PRE = false, bool NDEC = false,
          bool BOX = false>
__device__ __forceinline__ uint32_t shadow_texel(const DevWorld &W, float4 p, uint32_t nq, const float sun[3], float leaf,
                                                 vf_hit *hits, size_t i, Cnt &c, const DevBodies *B = nullptr,
                                                 const DirPre *pre = nullptr, const float *Ndec = nullptr) {
    float vis = 0.0f, tb = 0.0f;
    if (p.w < 0.0f) {
        vis = 1.0f;
        if (hits) store_miss(hits + i);
    } else {
        const float N[3] = {NDEC ? Ndec[0] : unorm8_dec(nq & 255u), NDEC ? Ndec[1] : unorm8_dec((nq >> 8) & 255u),
                            NDEC ? Ndec[2] : unorm8_dec((nq >> 16) & 255u)};
        // the ray's origin: P + N*0.03 for a texel facing the sun; FOLIAGE_SSS (shadow_trace.comp:35-46):
        // foliageSunOrigin for back-lit foliage with the sun above the horizon, which stores no blocker
        // distance; one call site of the traversal for both
        const bool lit = dot3(N, sun) > 0.0f;
        bool trace = lit;
        float so[3] = {p.x + N[0] * 3e-2f, p.y + N[1] * 3e-2f, p.z + N[2] * 3e-2f};
        if (!lit) {
            const float a = NDEC ? ((nq >> 24) == kFoliageAlpha ? 0.8f : 1.0f) : gbn_alpha(nq);
            if (foliage_flag(a) && sun[1] > 0.0f) {
                const float P[3] = {p.x, p.y, p.z};
                foliage_sun_origin(P, N, sun, blade_flag(a), so);
                trace = true;
            }
        }
        if (trace) {
            uint32_t seed = 0u;
            vis = sun_visibility_t<false, COUNT, BODIES, LDSM, LS, PRE, BOX>(W, so, sun, leaf, seed, tb, hits, i, c, B, pre);
            if (!lit) tb = 0.0f;
        } else if (hits) {
            store_miss(hits + i);
        }
    }
    const uint32_t hv = (uint32_t)__half_as_ushort(__float2half_rn(vis));
    const uint32_t ht = (uint32_t)__half_as_ushort(__float2half_rn(gmin(tb, 60000.0f)));
    return hv | (ht << 16);
}

// shadow_trace.comp (shadow_trace.comp:17-34, 87, 96-105) for a 16x16 tile per workgroup.
template <bool COUNT, bool BODIES>
__global__ void __launch_bounds__(256) k_shadow(DevWorld W, DevBodies B, TileArgs A, float sx, float sy, float sz,
                                                const float4 *gbpos, const uint32_t *gbn, uint32_t *shadow, vf_hit *hits,
                                                unsigned long long *cnt, DirPre sp) {
    const int32_t L = xcd_remap((int32_t)blockIdx.x, (int32_t)gridDim.x);
    const int32_t tx = L % A.tiles_x, ty = L / A.tiles_x;
    const int32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int32_t px = tx * 16 + (wv & 1) * 8 + (lane & 7);
    const int32_t py = ty * 16 + (wv >> 1) * 8 + (lane >> 3);
    if (px >= A.w || py >= A.h) return;
    const size_t i = (size_t)py * (size_t)A.w + (size_t)px;
    const float sun[3] = {sx, sy, sz};
    Cnt c = {0, 0, 0, 0};
    shadow[i] = shadow_texel<COUNT, BODIES, true, 256, true>(W, gbpos[i], gbn[i], sun, A.leaf, hits, i, c, &B, &sp);
    flush_count<COUNT>(c, cnt);
}

// gbuffer.comp then shadow_trace.comp for the same pixel in one kernel: the sun ray starts from the
// G-buffer values this lane has just written (position, RGBA8-quantised normal), exactly what the
// separate shadow pass would read back, so the results are those of the two passes.
template <bool COUNT, bool SPECIAL, int LS = 256, bool PRE = false, bool BOX = false>
__device__ __forceinline__ void primary_shadow_pixel(const DevWorld &W, const PrimArgs &A, uint32_t cam_mat, int32_t px,
                                                     int32_t py, const float sun[3], vf_hit *hits, float4 *gbpos,
                                                     uint32_t *gbn, uint32_t *shadow, vf_hit *shits, Cnt &c,
                                                     const DirPre *pre = nullptr) {
    float rd[3];
    primary_dir(A, px, py, rd);
    Hit h;
    primary_hit<COUNT, SPECIAL, LS, BOX>(W, A, cam_mat, rd, h, c);
    if (camera_special(cam_mat) != SPECIAL) return;
    const size_t i = (size_t)py * (size_t)A.w + (size_t)px;
    store_hit(hits + i, h);
    float4 p;
    uint32_t nq = 0u;
    // the quantised face normal as the shadow pass decodes it: unorm8 bytes 0 / 128 / 255 give exactly
    // -1, (128 / 255) * 2 - 1 and 1 (the middle value folded at compile time, IEEE like the division)
    constexpr float kDec128 = (128.0f / 255.0f) * 2.0f - 1.0f;
    float Nd[3] = {kDec128, kDec128, kDec128};
    if (h.hit) {
        p = make_float4(A.pos[0] + rd[0] * h.t, A.pos[1] + rd[1] * h.t, A.pos[2] + rd[2] * h.t, h.t);
        float n[3];
        face_normal(h.face, n);
        nq = face_normal_bytes(h.face) | (mat_flag_byte(W, h.mat) << 24);
#pragma unroll
        for (int a = 0; a < 3; ++a) Nd[a] = n[a] > 0.0f ? 1.0f : (n[a] < 0.0f ? -1.0f : kDec128);
    } else {
        p = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    }
    if (gbpos) gbpos[i] = p;
    if (gbn) gbn[i] = nq;
    // shadow_trace.comp on (p, nq)
    shadow[i] = shadow_texel<COUNT, false, true, LS, PRE, true, BOX>(W, p, nq, sun, A.leaf, shits, i, c, nullptr, pre, Nd);
}

template <bool COUNT, bool SPECIAL>
__global__ void __launch_bounds__(256) k_primary_shadow(DevWorld W, PrimArgs A, float sx, float sy, float sz, vf_hit *hits,
                                                        float4 *gbpos, uint32_t *gbn, uint32_t *shadow, vf_hit *shits,
                                                        unsigned long long *cnt) {
    const uint32_t cam_mat = A.cam_air ? 0u : voxel_mat_at(W, A.pos);
    if (SPECIAL && !camera_special(cam_mat)) return;
    const float sun[3] = {sx, sy, sz};
    Cnt c = {0, 0, 0, 0};
    if (!SPECIAL) {
        int32_t px, py;
        tile_pixel_of(xcd_remap((int32_t)blockIdx.x, (int32_t)gridDim.x), A.tiles_x, px, py);
        if (px < A.w && py < A.h)
            primary_shadow_pixel<COUNT, false>(W, A, cam_mat, px, py, sun, hits, gbpos, gbn, shadow, shits, c);
    } else {
        for (int32_t L = (int32_t)blockIdx.x; L < A.ntiles; L += (int32_t)gridDim.x) {
            int32_t px, py;
            tile_pixel_of(L, A.tiles_x, px, py);
            if (px < A.w && py < A.h)
                primary_shadow_pixel<COUNT, true>(W, A, cam_mat, px, py, sun, hits, gbpos, gbn, shadow, shits, c);
        }
    }
    if (camera_special(cam_mat) == SPECIAL) flush_count<COUNT>(c, cnt);   // only the variant whose results stand
}

// The fused kernel for a camera in air with one 8x8-pixel wave per workgroup: a wave that finishes
// frees its slot at once instead of waiting for the other three of a 16x16 tile.  Block L is quadrant
// L & 3 of 16x16 tile L >> 2, so the pixels of a wave are those of the 256-thread kernel.
template <bool COUNT>
__global__ void __launch_bounds__(64, 7) k_primary_shadow64(DevWorld W, PrimArgs A, float sx, float sy, float sz, vf_hit *hits,
                                                         float4 *gbpos, uint32_t *gbn, uint32_t *shadow, vf_hit *shits,
                                                         unsigned long long *cnt, DirPre sp) {
    const float sun[3] = {sx, sy, sz};
    Cnt c = {0, 0, 0, 0};
    const int32_t L = xcd_remap64((int32_t)blockIdx.x, (int32_t)gridDim.x);
    const int32_t T = L >> 2, q = L & 3, lane = (int32_t)threadIdx.x;
    const int32_t px = (T % A.tiles_x) * 16 + (q & 1) * 8 + (lane & 7);
    const int32_t py = (T / A.tiles_x) * 16 + (q >> 1) * 8 + (lane >> 3);
    if (px < A.w && py < A.h)
        primary_shadow_pixel<COUNT, false, 64, true>(W, A, 0u, px, py, sun, hits, gbpos, gbn, shadow, shits, c, &sp);
    flush_count<COUNT>(c, cnt);
}

// gbuffer.comp's primary pass with the small-frame traversal (trace_world's BOX) for a camera in air, one
// 8x8-pixel wave per workgroup, for launches of at most 16 waves per SIMD.
__global__ void __launch_bounds__(64, 5) k_primary_small(DevWorld W, PrimArgs A, vf_hit *hits, float4 *gbpos, uint32_t *gbn) {
    Cnt c = {0, 0, 0, 0};
    const int32_t L = xcd_remap64((int32_t)blockIdx.x, (int32_t)gridDim.x);
    const int32_t T = L >> 2, q = L & 3, lane = (int32_t)threadIdx.x;
    const int32_t px = (T % A.tiles_x) * 16 + (q & 1) * 8 + (lane & 7);
    const int32_t py = (T / A.tiles_x) * 16 + (q >> 1) * 8 + (lane >> 3);
    if (px < A.w && py < A.h) primary_pixel<false, false, 64, true>(W, A, 0u, px, py, hits, gbpos, gbn, c);
}

// The same pass with the small-frame traversal (trace_world's BOX), for launches of at most 16 waves per
// SIMD: bounded at 5 waves per SIMD (it takes 78 VGPRs; bounds of 6 and 7 measured no faster on these
// launches, which hold at most 16 waves per SIMD in all).
__global__ void __launch_bounds__(64, 5) k_primary_shadow_small(DevWorld W, PrimArgs A, float sx, float sy, float sz,
                                                               vf_hit *hits, float4 *gbpos, uint32_t *gbn, uint32_t *shadow,
                                                               vf_hit *shits, DirPre sp) {
    const float sun[3] = {sx, sy, sz};
    Cnt c = {0, 0, 0, 0};
    const int32_t L = xcd_remap64((int32_t)blockIdx.x, (int32_t)gridDim.x);
    const int32_t T = L >> 2, q = L & 3, lane = (int32_t)threadIdx.x;
    const int32_t px = (T % A.tiles_x) * 16 + (q & 1) * 8 + (lane & 7);
    const int32_t py = (T / A.tiles_x) * 16 + (q >> 1) * 8 + (lane >> 3);
    if (px < A.w && py < A.h)
        primary_shadow_pixel<false, false, 64, true, true>(W, A, 0u, px, py, sun, hits, gbpos, gbn, shadow, shits, c, &sp);
}

// gbuffer.comp over the scene (traceScenePrimary = traceSceneS, common.glsl:787-793: the world, then the
// bodies closer than its hit; a camera in solid terrain or water as gbuffer.comp:233-266), optionally
// followed by shadow_trace.comp on the values just written (SHADOW), for a 16x16 tile per workgroup.
// body (nullable): h.obj per pixel, -1 for the world and the sky.
template <bool COUNT, bool SHADOW>
__global__ void __launch_bounds__(256) k_scene(DevWorld W, DevBodies B, PrimArgs A, float sx, float sy, float sz, vf_hit *hits,
                                               float4 *gbpos, uint32_t *gbn, int32_t *body, uint32_t *shadow, vf_hit *shits,
                                               unsigned long long *cnt) {
    const uint32_t cam_mat = A.cam_air ? 0u : voxel_mat_at(W, A.pos);
    int32_t px, py;
    tile_pixel_of(xcd_remap((int32_t)blockIdx.x, (int32_t)gridDim.x), A.tiles_x, px, py);
    if (px >= A.w || py >= A.h) return;
    Cnt c = {0, 0, 0, 0};
    float rd[3];
    primary_dir(A, px, py, rd);
    Hit h;
    scene_primary_trace<COUNT, true>(W, B, A.pos, cam_mat, rd, h, c);
    const size_t i = (size_t)py * (size_t)A.w + (size_t)px;
    if (hits) store_hit(hits + i, h, h.hit && h.obj >= 0);
    float4 p;
    uint32_t nq = 0u;
    if (h.hit) {
        p = make_float4(A.pos[0] + rd[0] * h.t, A.pos[1] + rd[1] * h.t, A.pos[2] + rd[2] * h.t, h.t);
        float n[3];
        hit_normal<true>(&B, h, n);
        nq = pack_normal(n, mat_flag_byte(W, h.mat));
    } else {
        p = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    }
    if (gbpos) gbpos[i] = p;
    if (gbn) gbn[i] = nq;
    if (body) body[i] = h.hit ? h.obj : -1;
    if (SHADOW) {
        const float sun[3] = {sx, sy, sz};
        shadow[i] = shadow_texel<COUNT, true, true>(W, p, nq, sun, A.leaf, shits, i, c, &B);
    }
    flush_count<COUNT>(c, cnt);
}

struct GiArgs {
    int32_t w, h, ow, oh, tiles_x, ntiles, half_res, lights_nee;
    uint32_t frame;
    float reach;
};

// gi.comp (gi.comp:34-84) + giPathImpl with one bounce (common.glsl:2250-2366).
template <bool COUNT, bool BODIES>
__global__ void __launch_bounds__(256) k_gi(DevWorld W, SkyParams S, DevBodies B, GiArgs A, const float4 *gbpos,
                                            const uint32_t *gbn, float4 *radiance, vf_hit *bounce, unsigned long long *cnt) {
    const int32_t L = xcd_remap((int32_t)blockIdx.x, (int32_t)gridDim.x);
    const int32_t tx = L % A.tiles_x, ty = L / A.tiles_x;
    const int32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int32_t px = tx * 16 + (wv & 1) * 8 + (lane & 7);
    const int32_t py = ty * 16 + (wv >> 1) * 8 + (lane >> 3);
    if (px >= A.ow || py >= A.oh) return;
    const size_t oi = (size_t)py * (size_t)A.ow + (size_t)px;
    int32_t fx = px, fy = py;
    bool sky = false;
    if (A.half_res) {
        const int32_t ox = (int32_t)(A.frame & 1u), oy = (int32_t)((A.frame >> 1u) & 1u);
        fx = imin(px * 2 + ox, A.w - 1);
        fy = imin(py * 2 + oy, A.h - 1);
        int32_t gx = imin(px * 2, A.w - 1), gy = imin(py * 2, A.h - 1);
        for (int k = 0; k < 4; ++k) {
            const int32_t qx = imin(px * 2 + (k & 1), A.w - 1), qy = imin(py * 2 + (k >> 1), A.h - 1);
            if (gbpos[(size_t)qy * A.w + qx].w >= 0.0f) { gx = qx; gy = qy; break; }
        }
        if (gbpos[(size_t)gy * A.w + gx].w < 0.0f) sky = true;
        else if (gbpos[(size_t)fy * A.w + fx].w < 0.0f) { fx = gx; fy = gy; }
    } else if (gbpos[(size_t)fy * A.w + fx].w < 0.0f) {
        sky = true;
    }
    Cnt c = {0, 0, 0, 0};
    Hit bh;
    bh.hit = 0; bh.t = -1.0f; bh.face = 0; bh.mat = 0; bh.cell[0] = bh.cell[1] = bh.cell[2] = 0; bh.obj = -1;
    float Lr[3] = {0.0f, 0.0f, 0.0f};
    if (!sky) {
        const size_t fi = (size_t)fy * A.w + fx;
        const float4 P4 = gbpos[fi];
        const uint32_t nb = gbn[fi];
        const float N[3] = {unorm8_dec(nb & 255u), unorm8_dec((nb >> 8) & 255u), unorm8_dec((nb >> 16) & 255u)};
        float u0, u1;
        ld_sample2(fx, fy, A.frame >> 2u, 0x68e31da4u, u0, u1);   // gi.comp:70-72
        float rdir[3];
        cosine_hemisphere_u(N, u0, u1, rdir);
        const float P[3] = {P4.x, P4.y, P4.z};
        const GiPathCfg G = {A.reach, A.lights_nee};
        uint32_t seed = 0u;
        gi_path<1, false, COUNT, BODIES, true>(W, S, G, P, N, rdir, seed, true, Lr, &bh, nullptr, c, &B);
    }
    radiance[oi] = make_float4(Lr[0], Lr[1], Lr[2], sky ? 0.0f : 1.0f);
    if (bounce) store_hit(bounce + oi, bh, BODIES && bh.hit && bh.obj >= 0);
    flush_count<COUNT>(c, cnt);
}

// lighting.comp's sun ray (SVGF, lighting.comp:44-84): per pixel seed pixelSeed(pix, frameIdx, 0), a
// lit texel (N.L > 0) traces sunVisibility (FOLIAGE_LIGHT, stochastic) from P + N*0.03 along
// sampleCone(sunDir, sunAngle), a back-lit foliage texel (FOLIAGE_SSS, sun above the horizon) from
// foliageSunOrigin; the raw sample is 1 for sky, 0 for another texel facing away, else 0 or 1.
template <bool COUNT, bool BODIES>
__global__ void __launch_bounds__(256) k_sun_cone(DevWorld W, DevBodies B, TileArgs A, float sx, float sy, float sz, float ang,
                                                  uint32_t frame, const float4 *gbpos, const uint32_t *gbn, float *out,
                                                  vf_hit *hits, unsigned long long *cnt) {
    const int32_t L = xcd_remap((int32_t)blockIdx.x, (int32_t)gridDim.x);
    const int32_t tx = L % A.tiles_x, ty = L / A.tiles_x;
    const int32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int32_t px = tx * 16 + (wv & 1) * 8 + (lane & 7);
    const int32_t py = ty * 16 + (wv >> 1) * 8 + (lane >> 3);
    if (px >= A.w || py >= A.h) return;
    const size_t i = (size_t)py * (size_t)A.w + (size_t)px;
    const float4 p = gbpos[i];
    bool stored = false;
    Cnt c = {0, 0, 0, 0};
    float v = 1.0f;
    if (p.w >= 0.0f) {
        const uint32_t nq = gbn[i];
        const float N[3] = {unorm8_dec(nq & 255u), unorm8_dec((nq >> 8) & 255u), unorm8_dec((nq >> 16) & 255u)};
        const float sun[3] = {sx, sy, sz};
        uint32_t seed = pixel_seed(px, py, frame, 0u);
        v = 0.0f;
        const bool lit = gmax(dot3(N, sun), 0.0f) > 0.0f;
        bool ray = lit;
        float so[3] = {p.x + N[0] * 3e-2f, p.y + N[1] * 3e-2f, p.z + N[2] * 3e-2f};
        if (!lit) {   // FOLIAGE_SSS (lighting.comp:66-71): back-lit foliage traces from foliageSunOrigin
            const float a = gbn_alpha(nq);
            if (foliage_flag(a) && sun[1] > 0.0f) {
                const float P[3] = {p.x, p.y, p.z};
                foliage_sun_origin(P, N, sun, blade_flag(a), so);
                ray = true;
            }
        }
        if (ray) {
            float sd[3], tf;
            sample_cone(sun, ang, seed, sd);
            v = sun_visibility_t<true, COUNT, BODIES, true>(W, so, sd, A.leaf, seed, tf, hits, i, c, &B);
            stored = true;
        }
    }
    if (hits && !stored) store_miss(hits + i);
    out[i] = v;
    flush_count<COUNT>(c, cnt);
}

// Explicit ray batch: traceVolumeS per ray (1-D, XCD-remapped 256-ray blocks); BOXT: the small-frame
// traversal (not in the counting variant).
template <int SKIP, bool COUNT, bool BOXT = false>
__global__ void __launch_bounds__(256) k_rays(DevWorld W, const vf_ray *rays, long long n, float tcap, vf_hit *hits,
                                              unsigned long long *cnt) {
    constexpr bool BOX = BOXT && !COUNT;
    const int32_t L = xcd_remap((int32_t)blockIdx.x, (int32_t)gridDim.x);
    const long long i = (long long)L * 256 + threadIdx.x;
    if (i >= n) return;
    const float *r = reinterpret_cast<const float *>(rays + i);
    const float ro[3] = {r[0], r[1], r[2]}, rd[3] = {r[3], r[4], r[5]};
    Cnt c = {0, 0, 0, 0};
    Hit h;
    trace_world<SKIP, COUNT, true, 256, false, BOX>(W, ro, rd, tcap, h, c);
    store_hit(hits + i, h);
    flush_count<COUNT>(c, cnt);
}

struct ProbeArgs {
    float origin[3];
    float spacing;
    int32_t dims[3];
    int32_t n_rays;
    uint32_t frame;
    float reach;
    long long n;
};

// probe_trace.comp ray source (probes.glsl:155-170) traced with skipMat 5 (kSkipMat).
// BODIES (GI_OBJECTS, probe_trace.comp:47): bodies nearer than min(the world hit or the reach, 16).
template <bool COUNT, bool BODIES>
__global__ void __launch_bounds__(256) k_probes(DevWorld W, DevBodies B, ProbeArgs A, vf_hit *hits, unsigned long long *cnt) {
    const int32_t L = xcd_remap((int32_t)blockIdx.x, (int32_t)gridDim.x);
    const long long idx = (long long)L * 256 + threadIdx.x;
    if (idx >= A.n) return;
    const int32_t i = (int32_t)(idx % A.n_rays);
    const long long p = idx / A.n_rays;
    const int32_t pxi = (int32_t)(p % A.dims[0]), pyi = (int32_t)((p / A.dims[0]) % A.dims[1]),
                  pzi = (int32_t)(p / ((long long)A.dims[0] * A.dims[1]));
    const float ro[3] = {A.origin[0] + A.spacing * (float)pxi, A.origin[1] + A.spacing * (float)pyi,
                         A.origin[2] + A.spacing * (float)pzi};
    const float golden = 2.39996323f;
    const float z = 1.0f - (2.0f * (float)i + 1.0f) / (float)A.n_rays;
    const float rr = sqrtf(gmax(0.0f, 1.0f - z * z));
    const float ph = (float)i * golden;
    float s, c;
    sincos_red(ph, s, c);
    const float d[3] = {c * rr, z, s * rr};
    const uint32_t h0 = pcg(A.frame * 747796405u + 1u);
    const float u1 = (float)(h0 & 0xFFFFu) / 65535.0f, u2 = (float)(h0 >> 16) / 65535.0f;
    const uint32_t h2 = pcg(h0);
    const float u3 = (float)(h2 & 0xFFFFu) / 65535.0f;
    float s2, c2, s3, c3;
    sincos_red(6.2831853f * u2, s2, c2);
    sincos_red(6.2831853f * u3, s3, c3);
    const float qa = sqrtf(1.0f - u1), qb = sqrtf(u1);
    const float q[3] = {qa * s2, qa * c2, qb * s3};
    const float qw = qb * c3;
    float c1v[3], tmp[3], c2v[3], rd[3];
    cross3(q, d, c1v);
#pragma unroll
    for (int k = 0; k < 3; ++k) tmp[k] = c1v[k] + qw * d[k];
    cross3(q, tmp, c2v);
#pragma unroll
    for (int k = 0; k < 3; ++k) rd[k] = d[k] + 2.0f * c2v[k];
    Cnt cc = {0, 0, 0, 0};
    Hit h;
    trace_world<5, COUNT, true>(W, ro, rd, A.reach, h, cc);
    if (BODIES) {
        h.obj = -1;
        trace_objects<COUNT>(W, B, ro, rd, h, gmin(h.hit ? h.t : A.reach, 16.0f), false, cc);
    }
    store_hit(hits + idx, h, BODIES && h.hit && h.obj >= 0);
    flush_count<COUNT>(cc, cnt);
}

// ------------------------------------------------------------------------------------------------
// world build kernels: occupancy masks, coarse pyramid, brick distance field, cell words, sky-top
// ------------------------------------------------------------------------------------------------
// One thread per (slot, 32-bit word): 32 voxels = 4 rows of 8.
__global__ void __launch_bounds__(256) k_masks(const uint8_t *atlas, const uint32_t *slots, uint32_t n_list, uint32_t n_slots,
                                               uint32_t *occ, uint32_t *occdry, uint8_t *dry_any) {
    const uint64_t gid = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    const uint64_t k = gid >> 4;
    const uint32_t wd = (uint32_t)(gid & 15u);
    const bool live = k < (slots ? n_list : n_slots);
    const uint32_t slot = live ? (slots ? slots[k] : (uint32_t)k) : 0u;
    const bool ok = live && slot < n_slots;
    uint32_t m0 = 0u, m5 = 0u;
    if (ok) {
        const uint8_t *b = atlas + (size_t)slot * 512u + (size_t)wd * 32u;
        for (int i = 0; i < 32; ++i) {
            const uint32_t v = b[i];
            m0 |= (uint32_t)(v != 0u) << i;
            m5 |= (uint32_t)(v != 0u && v != 5u) << i;
        }
        occ[(size_t)slot * 16u + wd] = m0;
        occdry[(size_t)slot * 16u + wd] = m5;
    }
    // the slot's 16 words are 16 consecutive lanes: whether any of them has a dry voxel
    const uint64_t b = __ballot(m5 != 0u);
    if (ok && wd == 0u) dry_any[slot] = ((b >> (__lane_id() & 48u)) & 0xFFFFull) != 0ull ? 1u : 0u;
}

// coarse1: one thread per 4^3-brick cell (gpu_world.rs:300-313).
__global__ void __launch_bounds__(256) k_coarse1(const uint32_t *grid, int32_t wbx, int32_t wby, int32_t wbz, int32_t c1x,
                                                 int32_t c1y, int32_t c1z, uint8_t *c1) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)c1x * c1y * c1z) return;
    const int32_t x = (int32_t)(i % c1x), y = (int32_t)((i / c1x) % c1y), z = (int32_t)(i / ((int64_t)c1x * c1y));
    uint8_t any = 0;
    for (int32_t dz = 0; dz < 4 && !any; ++dz)
        for (int32_t dy = 0; dy < 4 && !any; ++dy)
            for (int32_t dx = 0; dx < 4 && !any; ++dx) {
                const int32_t bx = x * 4 + dx, by = y * 4 + dy, bz = z * 4 + dz;
                if (bx >= wbx || by >= wby || bz >= wbz) continue;
                any = grid[((size_t)bz * wby + by) * wbx + bx] != 0u;
            }
    c1[i] = any;
}
// coarse2 from coarse1: a 16^3-brick cell is 4^3 coarse1 cells.
__global__ void __launch_bounds__(256) k_coarse2(const uint8_t *c1, int32_t c1x, int32_t c1y, int32_t c1z, int32_t c2x,
                                                 int32_t c2y, int32_t c2z, uint8_t *c2) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)c2x * c2y * c2z) return;
    const int32_t x = (int32_t)(i % c2x), y = (int32_t)((i / c2x) % c2y), z = (int32_t)(i / ((int64_t)c2x * c2y));
    uint8_t any = 0;
    for (int32_t dz = 0; dz < 4 && !any; ++dz)
        for (int32_t dy = 0; dy < 4 && !any; ++dy)
            for (int32_t dx = 0; dx < 4 && !any; ++dx) {
                const int32_t qx = x * 4 + dx, qy = y * 4 + dy, qz = z * 4 + dz;
                if (qx >= c1x || qy >= c1y || qz >= c1z) continue;
                any = c1[((size_t)qz * c1y + qy) * c1x + qx];
            }
    c2[i] = any;
}

struct DistArgs {
    int32_t bmin[3], bmax[3];
    int32_t dims[3];
    int32_t axis, fp, ring;
};

// One brickdist.comp dispatch (brickdist.comp:31-68): fp 1 = grid occupancy + pass, 0 = pass over
// a u8 field, 2 = dry occupancy (from the dry occupancy mask: any bit set = a solid non-water voxel).
__global__ void __launch_bounds__(256) k_dist(DistArgs A, const uint32_t *grid, const uint8_t *src, uint8_t *dst,
                                              const uint4 *occdry) {
    const int32_t ex = A.bmax[0] - A.bmin[0] + 1, ey = A.bmax[1] - A.bmin[1] + 1, ez = A.bmax[2] - A.bmin[2] + 1;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)ex * ey * ez) return;
    const int32_t c[3] = {A.bmin[0] + (int32_t)(i % ex), A.bmin[1] + (int32_t)((i / ex) % ey),
                          A.bmin[2] + (int32_t)(i / ((int64_t)ex * ey))};
    const int32_t cw[3] = {A.ring ? (c[0] & (A.dims[0] - 1)) : c[0], c[1], A.ring ? (c[2] & (A.dims[2] - 1)) : c[2]};
    const size_t ci = ((size_t)cw[2] * A.dims[1] + cw[1]) * A.dims[0] + cw[0];
    const uint32_t CAP = 16u;
    if (A.fp == 2) {
        const uint32_t g = grid[ci];
        bool solid = g != 0u && g != 5u;
        if (g >= 256u) {
            const uint4 *m = occdry + (size_t)(g - 256u) * 4u;
            const uint4 a = m[0], b = m[1], cc = m[2], d = m[3];
            solid = ((a.x | a.y | a.z | a.w) | (b.x | b.y | b.z | b.w) | (cc.x | cc.y | cc.z | cc.w) |
                     (d.x | d.y | d.z | d.w)) != 0u;
        }
        dst[ci] = solid ? 0u : (uint8_t)CAP;
        return;
    }
    uint32_t best = CAP;
    for (int32_t k = -16; k <= 16; ++k) {
        int32_t q[3] = {c[0], c[1], c[2]};
        q[A.axis] += k;
        if (A.ring) {
            if (q[1] < 0 || q[1] >= A.dims[1]) continue;
            q[0] &= A.dims[0] - 1;
            q[2] &= A.dims[2] - 1;
        } else if (q[0] < 0 || q[1] < 0 || q[2] < 0 || q[0] >= A.dims[0] || q[1] >= A.dims[1] || q[2] >= A.dims[2]) {
            continue;
        }
        const size_t qi = ((size_t)q[2] * A.dims[1] + q[1]) * A.dims[0] + q[0];
        const uint32_t f = A.fp ? (grid[qi] != 0u ? 0u : CAP) : (uint32_t)src[qi];
        const uint32_t ak = (uint32_t)(k < 0 ? -k : k);
        const uint32_t m = ak > f ? ak : f;
        best = m < best ? m : best;
    }
    dst[ci] = (uint8_t)best;
}

// The separable passes staged in LDS.  A workgroup takes a tile of kDX consecutive x cells; a tap outside the
// world reads the sentinel kOut, larger than any max(|k|, f), so it never wins (the shader's `continue`).
// A tap at |k| = 16 gives max(16, f) >= the cap 16, so only |k| <= 15 can lower a result: the taps run in
// ascending |k| and a wavefront stops once every lane's best is <= |k| (no later tap can lower it).
constexpr int kDX = 32;          // x cells per tile; the x pass's window is kDX + 32 = 64 cells (one 64-bit mask)
constexpr int kDMaxY = 64;       // box y extent the fused x+y kernel holds in LDS
constexpr uint32_t kOut = 255u;

typedef unsigned short vf_us2 __attribute__((ext_vector_type(2)));

// Tiles start at a multiple of 4 cells in x (x0 & 3 == 0, wrapped or not), so with dims.x % 4 == 0 (VEC) every
// 4-cell quad of a row is one aligned 32-bit word: loaded and stored whole.  Cells of a tile outside the box
// are read (inside the world) but never written.
__device__ __forceinline__ int32_t dist_tile_x0(const DistArgs &A, int32_t tile, int32_t width) {
    return (A.bmin[0] & ~3) + tile * width;
}

// 4 cells of a row (x .. x + 3, x % 4 == 0) as bytes of one word, each outside the world = kOut.
template <bool VEC>
__device__ __forceinline__ uint32_t load_quad(const DistArgs &A, const uint8_t *src, int32_t zl, int32_t y, int32_t x,
                                              bool zin) {
    const size_t row = ((size_t)zl * A.dims[1] + y) * A.dims[0];
    if (VEC) {
        int32_t xq = x;
        if (A.ring) xq &= A.dims[0] - 1;
        const bool in = zin && (A.ring || xq < A.dims[0]);   // dims.x % 4 == 0: a quad is all in or all out
        const uint32_t w = reinterpret_cast<const uint32_t *>(src + row)[(in ? xq : 0) >> 2];
        return in ? w : 0xFFFFFFFFu;
    }
    uint32_t w = 0u;
#pragma unroll
    for (int32_t e = 0; e < 4; ++e) {
        int32_t xe = x + e;
        if (A.ring) xe &= A.dims[0] - 1;
        const bool in = zin && (A.ring || xe < A.dims[0]);
        const uint32_t b = src[row + (in ? xe : 0)];
        w |= (in ? b : kOut) << (8 * e);
    }
    return w;
}

// Writes the cells of quad (x .. x + 3) that lie in the box [bmin.x, bmax.x] (absolute x).
template <bool VEC>
__device__ __forceinline__ void store_quad(const DistArgs &A, uint8_t *dst, int32_t zw, int32_t y, int32_t x, uint32_t w) {
    const size_t row = ((size_t)zw * A.dims[1] + y) * A.dims[0];
    if (VEC && x >= A.bmin[0] && x + 3 <= A.bmax[0]) {
        const int32_t xq = A.ring ? (x & (A.dims[0] - 1)) : x;
        reinterpret_cast<uint32_t *>(dst + row)[xq >> 2] = w;
        return;
    }
#pragma unroll
    for (int32_t e = 0; e < 4; ++e) {
        const int32_t xe = x + e;
        if (xe < A.bmin[0] || xe > A.bmax[0]) continue;
        dst[row + (A.ring ? (xe & (A.dims[0] - 1)) : xe)] = (uint8_t)(w >> (8 * e));
    }
}


// x pass of one row from its window (window cell j = x0 - 16 + j; output col = window cell col + 16).  A binary
// row (every value 0, 16 or outside): min(16, distance to the nearest 0) from the row's occupancy mask.
__device__ __forceinline__ uint32_t dist_x(const uint8_t *F, uint64_t occ, bool binary, int32_t col) {
    const int32_t p = col + 16;
    if (binary) {
        const uint64_t r = occ >> p, l = occ << (63 - p);
        const uint32_t dr = r ? (uint32_t)__builtin_ctzll(r) : 64u, dl = l ? (uint32_t)__builtin_clzll(l) : 64u;
        return min(16u, min(dl, dr));
    }
    uint32_t best = 16u;
#pragma unroll
    for (int32_t k = -15; k <= 15; ++k) best = min(best, max((uint32_t)(k < 0 ? -k : k), (uint32_t)F[p + k]));
    return best;
}

// brickdist.comp's x pass and y pass of one field in one kernel (record_distance_field passes 0-1 or 4-5)
// for a tile of kDX x cells x the box's y extent at one z.  x pass: f from the grid (FIRST: occupied = g != 0)
// or from the dry occupancy field `src`, over x +-16 (any x of the world: the source is complete).  y pass:
// a row inside the box reads the x pass's value (held in LDS: the x pass's field is overwritten by the z pass
// over the same box, so it is never stored); a row outside it reads `stale`, the field that pass reads in the
// reference (d0 / d2 as they stand).  Writes the y pass's field `dst` (d1) over the box.
template <bool FIRST>
__global__ void __launch_bounds__(256) k_dist_xy(DistArgs A, const uint32_t *grid, const uint8_t *src, const uint8_t *stale,
                                                 uint8_t *dst) {
    __shared__ uint8_t F[kDMaxY][kDX + 32];
    __shared__ uint8_t X[kDMaxY][kDX];
    __shared__ uint64_t M[kDMaxY];
    __shared__ int32_t nonbinary;
    const int32_t ey = A.bmax[1] - A.bmin[1] + 1;
    const int32_t x0 = A.bmin[0] + (int32_t)blockIdx.x * kDX;
    const int32_t z = A.bmin[2] + (int32_t)blockIdx.y;
    const int32_t zw = A.ring ? (z & (A.dims[2] - 1)) : z;
    const int32_t t = (int32_t)threadIdx.x, lane = t & 63, wv = t >> 6;
    if (t == 0) nonbinary = 0;
    __syncthreads();
    // window rows: wavefront w takes rows w, w + 4, ..., lane j = window cell j.  All its loads are issued
    // first (up to kDMaxY / 4 per lane), then each row's occupancy becomes a ballot.
    {
        int32_t xa = x0 - 16 + lane;
        if (A.ring) xa &= A.dims[0] - 1;
        const bool inside = A.ring || (xa >= 0 && xa < A.dims[0]);
        uint32_t v[kDMaxY / 4];
        const int32_t xl = inside ? xa : 0;
#pragma unroll
        for (int32_t j = 0; j < kDMaxY / 4; ++j) {   // unconditional loads from valid addresses, masked after
            const int32_t row = min(wv + 4 * j, ey - 1);
            const size_t qi = ((size_t)zw * A.dims[1] + (A.bmin[1] + row)) * A.dims[0] + xl;
            v[j] = FIRST ? grid[qi] : (uint32_t)src[qi];
        }
#pragma unroll
        for (int32_t j = 0; j < kDMaxY / 4; ++j)
            if (!(wv + 4 * j < ey && inside)) v[j] = kOut;
#pragma unroll
        for (int32_t j = 0; j < kDMaxY / 4; ++j) {
            const int32_t row = wv + 4 * j;
            if (row >= ey) break;   // wave-uniform
            const uint32_t f = FIRST ? (v[j] == kOut ? kOut : (v[j] != 0u ? 0u : 16u)) : v[j];
            F[row][lane] = (uint8_t)f;
            const uint64_t occ = __ballot(f == 0u);
            if (!FIRST && __any(f != 0u && f != 16u && f != kOut) && lane == 0) nonbinary = 1;
            if (lane == 0) M[row] = occ;
        }
    }
    __syncthreads();
    const bool binary = FIRST || nonbinary == 0;
    for (int32_t i = t; i < ey * kDX; i += 256) {
        const int32_t row = i / kDX, col = i - row * kDX;
        X[row][col] = (uint8_t)dist_x(F[row], M[row], binary, col);
    }
    __syncthreads();
    const int32_t nx = min(kDX, A.bmax[0] - x0 + 1);
    for (int32_t i0 = 0; i0 < ey * kDX; i0 += 256) {   // the same trip count in every lane (wavefront votes)
        const int32_t i = i0 + t;
        const int32_t row = i / kDX, col = i - row * kDX;
        const bool act = i < ey * kDX && col < nx;
        const int32_t y = A.bmin[1] + row;
        const int32_t xa = x0 + col, xw = A.ring ? (xa & (A.dims[0] - 1)) : xa;
        uint32_t best = act ? 16u : 0u;
        for (int32_t k = 0; k < 16; ++k) {
            if (!__any(best > (uint32_t)k)) break;
            if (!act) continue;
#pragma unroll
            for (int32_t sg = 0; sg < 2; ++sg) {
                if (sg == 1 && k == 0) break;
                const int32_t qy = sg ? y + k : y - k;
                if (qy < 0 || qy >= A.dims[1]) continue;
                const uint32_t f = (qy >= A.bmin[1] && qy <= A.bmax[1])
                                       ? (uint32_t)X[qy - A.bmin[1]][col]
                                       : (uint32_t)stale[((size_t)zw * A.dims[1] + qy) * A.dims[0] + xw];
                best = min(best, max((uint32_t)k, f));
            }
        }
        if (act) dst[((size_t)zw * A.dims[1] + y) * A.dims[0] + xw] = (uint8_t)best;
    }
}

// brickdist.comp's x and y passes of one field for a box spanning the whole world height (every y tap is inside
// the box: no stale rows), one wavefront per tile of kDX x cells at one z.  With a binary x-pass input
// (occupancy: the grid's g != 0, or the dry occupancy field holding 0 / 16) the x+y result is min(16, Chebyshev
// distance in the xy plane to the nearest occupied cell): it is <= r exactly when the (2r+1)^2 square around
// the cell holds an occupied cell.  So the result = the number of radii r in [0, 15] whose square dilation of
// the occupancy misses the cell.  Lane j holds window column x0 - 16 + j as a bit mask over the rows; each
// radius is one more dilation step (rows: a shift within the lane; columns: the neighbour lanes through DPP
// wavefront shifts), counted by a 5-plane bit-sliced counter per lane.  A lane's error from the window's edge
// moves one lane per step, so the 32 output lanes 16..47 are exact after 15 steps.  A non-binary dry input
// falls back to the passes' taps.  MaskT: 32 or 64 rows.
__device__ __forceinline__ uint32_t lane_from_below(uint32_t v) {   // lane i gets lane i - 1's value, lane 0 gets 0
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xf, 0xf, true);
}
__device__ __forceinline__ uint32_t lane_from_above(uint32_t v) {   // lane i gets lane i + 1's value, lane 63 gets 0
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x130, 0xf, 0xf, true);
}
__device__ __forceinline__ uint64_t lane_from_below(uint64_t v) {
    return (uint64_t)lane_from_below((uint32_t)v) | ((uint64_t)lane_from_below((uint32_t)(v >> 32)) << 32);
}
__device__ __forceinline__ uint64_t lane_from_above(uint64_t v) {
    return (uint64_t)lane_from_above((uint32_t)v) | ((uint64_t)lane_from_above((uint32_t)(v >> 32)) << 32);
}

// LDS written by one wavefront and read by its other lanes: ordered within the wavefront, and visible once its
// own LDS operations are done (each wavefront of these kernels has LDS of its own, so no workgroup barrier).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The square-dilation count of one tile (lane j = window column x0 - 16 + j, MaskT bits = rows) into 5 counter
// planes, then the result bytes through LDS as rows of 32-bit quads to `dst` over the box.
template <typename MaskT>
__device__ __forceinline__ void dist_xy_square(MaskT occ, int32_t ey, MaskT &c0, MaskT &c1, MaskT &c2, MaskT &c3, MaskT &c4) {
    constexpr int32_t kRows = (int32_t)sizeof(MaskT) * 8;
    const MaskT rows = ey >= kRows ? ~(MaskT)0 : (((MaskT)1 << ey) - (MaskT)1);
    MaskT S = occ;
    c0 = 0; c1 = 0; c2 = 0; c3 = 0; c4 = 0;
#pragma unroll
    for (int32_t r = 0; r < 16; ++r) {
        if (r > 0) {
            const MaskT t = (S | (S << 1) | (S >> 1)) & rows;   // rows r - 1 -> r (no row outside the world)
            S = t | lane_from_below(t) | lane_from_above(t);     // columns
        }
        MaskT carry = ~S, q;   // rows of this column not yet within r of an occupied cell
        q = c0 & carry; c0 ^= carry; carry = q;
        q = c1 & carry; c1 ^= carry; carry = q;
        q = c2 & carry; c2 ^= carry; carry = q;
        q = c3 & carry; c3 ^= carry; carry = q;
        c4 ^= carry;
    }
}

template <typename MaskT>
__device__ __forceinline__ void dist_xy_emit(const DistArgs &A, int32_t x0, int32_t zw, int32_t ey, MaskT c0, MaskT c1,
                                             MaskT c2, MaskT c3, MaskT c4, uint32_t (*D)[kDX / 4], uint8_t *dst) {
    constexpr int32_t kRows = (int32_t)sizeof(MaskT) * 8;
    const int32_t lane = (int32_t)(threadIdx.x & 63u), col = lane - 16;
    if (col >= 0 && col < kDX) {   // lanes 16..47 hold the tile's columns
        uint8_t *Db = reinterpret_cast<uint8_t *>(&D[0][0]);
#pragma unroll
        for (int32_t row = 0; row < kRows; ++row) {
            const uint32_t d = (uint32_t)((c0 >> row) & 1u) | ((uint32_t)((c1 >> row) & 1u) << 1) |
                               ((uint32_t)((c2 >> row) & 1u) << 2) | ((uint32_t)((c3 >> row) & 1u) << 3) |
                               ((uint32_t)((c4 >> row) & 1u) << 4);
            Db[row * kDX + col] = (uint8_t)d;
        }
    }
    wave_lds_sync();
    for (int32_t i = lane; i < ey * (kDX / 4); i += 64) {
        const int32_t row = i / (kDX / 4), q = i - row * (kDX / 4);
        if (A.dims[0] % 4 == 0) store_quad<true>(A, dst, zw, row, x0 + 4 * q, D[row][q]);
        else store_quad<false>(A, dst, zw, row, x0 + 4 * q, D[row][q]);
    }
}

// The passes' taps over LDS for a tile whose x-pass input is not 0 / 16 (F: the window values).
__device__ __forceinline__ void dist_xy_taps(const DistArgs &A, int32_t x0, int32_t zw, int32_t ey, const uint8_t (*F)[kDX + 32],
                                             uint8_t (*X)[kDX], uint8_t *dst) {
    const int32_t lane = (int32_t)(threadIdx.x & 63u);
    wave_lds_sync();
    for (int32_t i = lane; i < ey * kDX; i += 64) {
        const int32_t row = i / kDX, col = i - row * kDX;
        X[row][col] = (uint8_t)dist_x(F[row], 0ull, false, col);
    }
    wave_lds_sync();
    for (int32_t i = lane; i < ey * kDX; i += 64) {
        const int32_t row = i / kDX, col = i - row * kDX;
        const int32_t xc = x0 + col;
        if (xc < A.bmin[0] || xc > A.bmax[0]) continue;
        uint32_t best = 16u;
        for (int32_t k = -15; k <= 15; ++k) {
            const int32_t qy = row + k;
            if (qy < 0 || qy >= ey) continue;
            best = min(best, max((uint32_t)(k < 0 ? -k : k), (uint32_t)X[qy][col]));
        }
        const int32_t xw = A.ring ? (xc & (A.dims[0] - 1)) : xc;
        dst[((size_t)zw * A.dims[1] + row) * A.dims[0] + xw] = (uint8_t)best;
    }
}

template <bool FIRST, typename MaskT>
__global__ void __launch_bounds__(64) k_dist_xy_mask(DistArgs A, const uint32_t *grid, const uint8_t *src, uint8_t *dst) {
    constexpr int32_t kRows = (int32_t)sizeof(MaskT) * 8;
    __shared__ uint8_t F[kRows][kDX + 32];   // the window values (the non-binary fallback only)
    __shared__ uint8_t X[kRows][kDX];
    __shared__ uint32_t D[kRows][kDX / 4];   // the results, rows x quads (vector stores)
    const int32_t ey = A.dims[1];
    const int32_t x0 = dist_tile_x0(A, (int32_t)blockIdx.x, kDX);
    const int32_t z = A.bmin[2] + (int32_t)blockIdx.y;
    const int32_t zw = A.ring ? (z & (A.dims[2] - 1)) : z;
    const int32_t lane = (int32_t)threadIdx.x;
    int32_t xa = x0 - 16 + lane;
    if (A.ring) xa &= A.dims[0] - 1;
    const bool inside = A.ring || (xa >= 0 && xa < A.dims[0]);
    const int32_t xl = inside ? xa : 0;   // a valid address for every lane; outside lanes read the sentinel
    MaskT occ = 0;
    bool nonbinary = false;
    // every row's value loaded from a valid address with no branch around it (a load the compiler sinks into
    // a conditional block is waited for before the next one), then masked
    uint32_t raw[kRows];
    const size_t base = (size_t)zw * A.dims[1] * A.dims[0] + xl;
#pragma unroll
    for (int32_t j = 0; j < kRows; ++j) {
        const size_t qi = base + (size_t)(j < ey ? j : ey - 1) * A.dims[0];
        raw[j] = FIRST ? grid[qi] : (uint32_t)src[qi];
    }
#pragma unroll
    for (int32_t j = 0; j < kRows; ++j) {
        const uint32_t valid = (uint32_t)(inside && j < ey);
        if (FIRST) {
            occ |= (MaskT)((raw[j] != 0u ? 1u : 0u) & valid) << j;
        } else {
            const uint32_t f = valid ? raw[j] : kOut;
            occ |= (MaskT)((f == 0u ? 1u : 0u)) << j;
            nonbinary = nonbinary || (f != 0u && f != 16u && f != kOut);
            if (j < ey) F[j][lane] = (uint8_t)f;
        }
    }
    if (!FIRST && __any(nonbinary)) {
        dist_xy_taps(A, x0, zw, ey, F, X, dst);
        return;
    }
    MaskT c0, c1, c2, c3, c4;
    dist_xy_square<MaskT>(occ, ey, c0, c1, c2, c3, c4);
    dist_xy_emit<MaskT>(A, x0, zw, ey, c0, c1, c2, c3, c4, D, dst);
}

struct DirtyBox {
    int32_t lo[3], hi[3];
};

// Is absolute brick coordinate v of axis a inside [lo, hi] of that axis, on the ring (x / z of a ring world:
// modulo the grid; the box is at most one ring wide) or plainly?
__device__ __forceinline__ bool in_span(const DistArgs &A, int a, int32_t v, int32_t lo, int32_t hi) {
    if (A.ring && a != 1) return (uint32_t)((v - lo) & (A.dims[a] - 1)) <= (uint32_t)(hi - lo);
    return v >= lo && v <= hi;
}

// Both fields' x and y passes in one launch (blockIdx.z: 0 = the normal field, 1 = the dry field) for a box
// over the world's whole height that contains the dirty box.  Normal: as k_dist_xy_mask<true> into `tmp`.
// Dry: the x-pass input is the dry occupancy of record_distance_field's pass 3 inside the dirty box (computed
// here: a uniform brick other than air and water, or a detailed brick with a dry voxel, dry_any) and the
// dry occupancy field `d3` as it stands outside it; the pass-3 values of the tile's cells in the dirty box are
// written to d3; the x+y result goes to d1.
template <typename MaskT>
__device__ __forceinline__ void dist_xy2_dry(const DistArgs &A, const DirtyBox &Db, int32_t tx, int32_t tz,
                                             const uint32_t *grid, const uint8_t *dry_any, uint8_t *d3, uint8_t *d1,
                                             uint8_t (*F)[kDX + 32], uint8_t (*X)[kDX], uint32_t (*D)[kDX / 4]) {
    constexpr int32_t kRows = (int32_t)sizeof(MaskT) * 8;
    const int32_t ey = A.dims[1];
    const int32_t x0 = dist_tile_x0(A, tx, kDX);
    const int32_t z = A.bmin[2] + tz;
    const int32_t zw = A.ring ? (z & (A.dims[2] - 1)) : z;
    const int32_t lane = (int32_t)(threadIdx.x & 63u);
    const int32_t xabs = x0 - 16 + lane;
    int32_t xa = xabs;
    if (A.ring) xa &= A.dims[0] - 1;
    const bool inside = A.ring || (xa >= 0 && xa < A.dims[0]);
    const int32_t xl = inside ? xa : 0;
    const bool dirty_xz = in_span(A, 0, xabs, Db.lo[0], Db.hi[0]) && in_span(A, 2, z, Db.lo[2], Db.hi[2]);
    const size_t base = (size_t)zw * A.dims[1] * A.dims[0] + xl;
    uint32_t g[kRows], o[kRows];
#pragma unroll
    for (int32_t j = 0; j < kRows; ++j) {   // all loads first, then the slots' dry flags they name
        const size_t qi = base + (size_t)(j < ey ? j : ey - 1) * A.dims[0];
        g[j] = grid[qi];
        o[j] = d3[qi];
    }
    uint32_t a[kRows];
#pragma unroll
    for (int32_t j = 0; j < kRows; ++j) a[j] = dry_any[g[j] >= 256u ? g[j] - 256u : 0u];
    MaskT solid = 0;
#pragma unroll
    for (int32_t j = 0; j < kRows; ++j) {   // arithmetic on every loaded flag (a select would sink each load into a branch)
        const uint32_t det = (uint32_t)(g[j] >= 256u), uni = (uint32_t)(g[j] != 0u && g[j] != 5u);
        const uint32_t sj = (a[j] & det) | (uni & (det ^ 1u));
        solid |= (MaskT)(sj & 1u) << j;
    }
    const MaskT drows = Db.hi[1] >= kRows - 1 ? ~(MaskT)0 << Db.lo[1]
                                              : ((((MaskT)1 << (Db.hi[1] + 1)) - (MaskT)1) & (~(MaskT)0 << Db.lo[1]));
    const MaskT rows = ey >= kRows ? ~(MaskT)0 : (((MaskT)1 << ey) - (MaskT)1);
    // x-pass input: pass 3's value in the dirty box (0 where solid), d3 elsewhere; outside the world: none
    const MaskT indirty = dirty_xz ? drows : (MaskT)0;
    MaskT occ = 0;
    bool nonbinary = false;
#pragma unroll
    for (int32_t j = 0; j < kRows; ++j) {
        const bool outside_old = ((indirty >> j) & 1) == 0;
        occ |= (MaskT)(outside_old && o[j] == 0u ? 1u : 0u) << j;
        nonbinary = nonbinary || (outside_old && j < ey && o[j] != 0u && o[j] != 16u);
    }
    occ = inside ? ((occ & ~indirty) | (solid & indirty)) & rows : (MaskT)0;
    if (lane >= 16 && lane < 16 + kDX && dirty_xz) {   // pass 3's writes of this tile's columns
        for (int32_t j = Db.lo[1]; j <= Db.hi[1]; ++j) d3[base + (size_t)j * A.dims[0]] = ((solid >> j) & 1) ? 0u : 16u;
    }
    if (__any(nonbinary)) {   // a d3 value that is not 0 / 16: the passes' taps
#pragma unroll
        for (int32_t j = 0; j < kRows; ++j) {
            if (j >= ey) continue;
            const bool dj = ((indirty >> j) & 1) != 0;
            const uint32_t f = !inside ? kOut : (dj ? (((solid >> j) & 1) ? 0u : 16u) : o[j]);
            F[j][lane] = (uint8_t)f;
        }
        dist_xy_taps(A, x0, zw, ey, F, X, d1);
        return;
    }
    MaskT c0, c1, c2, c3, c4;
    dist_xy_square<MaskT>(occ, ey, c0, c1, c2, c3, c4);
    dist_xy_emit<MaskT>(A, x0, zw, ey, c0, c1, c2, c3, c4, D, d1);
}

template <typename MaskT>
__device__ __forceinline__ void dist_xy2_normal(const DistArgs &A, int32_t tx, int32_t tz, const uint32_t *grid, uint8_t *tmp,
                                                uint32_t (*D)[kDX / 4]) {
    constexpr int32_t kRows = (int32_t)sizeof(MaskT) * 8;
    const int32_t ey = A.dims[1];
    const int32_t x0 = dist_tile_x0(A, tx, kDX);
    const int32_t z = A.bmin[2] + tz;
    const int32_t zw = A.ring ? (z & (A.dims[2] - 1)) : z;
    const int32_t lane = (int32_t)(threadIdx.x & 63u);
    int32_t xa = x0 - 16 + lane;
    if (A.ring) xa &= A.dims[0] - 1;
    const bool inside = A.ring || (xa >= 0 && xa < A.dims[0]);
    const size_t base = (size_t)zw * A.dims[1] * A.dims[0] + (inside ? xa : 0);
    uint32_t g[kRows];
#pragma unroll
    for (int32_t j = 0; j < kRows; ++j) g[j] = grid[base + (size_t)(j < ey ? j : ey - 1) * A.dims[0]];
    MaskT occ = 0;
#pragma unroll
    for (int32_t j = 0; j < kRows; ++j) occ |= (MaskT)((g[j] != 0u ? 1u : 0u) & (uint32_t)(inside && j < ey)) << j;
    MaskT c0, c1, c2, c3, c4;
    dist_xy_square<MaskT>(occ, ey, c0, c1, c2, c3, c4);
    dist_xy_emit<MaskT>(A, x0, zw, ey, c0, c1, c2, c3, c4, D, tmp);
}

// One workgroup of two wavefronts per tile: wavefront 0 the normal field, 1 the dry field (the second one's grid
// loads hit the cache the first one filled).  Workgroups are dealt round-robin to the 8 XCDs; each XCD gets a
// contiguous run of tiles (whole z slices), so the windows that neighbouring tiles share stay in one L2.
template <typename MaskT>
__global__ void __launch_bounds__(128) k_dist_xy2(DistArgs A, DirtyBox Db, int32_t tiles_x, int32_t n_tiles,
                                                  const uint32_t *grid, const uint8_t *dry_any, uint8_t *d3, uint8_t *tmp,
                                                  uint8_t *d1) {
    constexpr int32_t kRows = (int32_t)sizeof(MaskT) * 8;
    __shared__ uint8_t F[kRows][kDX + 32];
    __shared__ uint8_t X[kRows][kDX];
    __shared__ uint32_t D[2][kRows][kDX / 4];
    const int32_t L = (int32_t)blockIdx.x, full = (n_tiles / 8) * 8;
    const int32_t T = L < full ? (L & 7) * (full / 8) + (L >> 3) : L;
    const int32_t tx = T % tiles_x, tz = T / tiles_x;
    if (threadIdx.x < 64) dist_xy2_normal<MaskT>(A, tx, tz, grid, tmp, D[0]);
    else dist_xy2_dry<MaskT>(A, Db, tx, tz, grid, dry_any, d3, d1, F, X, D[1]);
}

constexpr int kZX = 64;   // x cells per z-pass tile (16 quads)
constexpr int kZZ = 16;   // z cells per z-pass tile

// brickdist.comp's z pass (record_distance_field passes 2 and 6) over a tile of kZX x cells x kZZ z cells at one
// y: the source over z +-16 staged in LDS as 16-bit values, one quad (4 x cells) per lane with packed 16-bit
// max / min; written to `dst` over the box.  A window row inside the box's z span reads `src_in`, one outside
// it `src_out` (one field for the seven-pass build; the two-launch build keeps the normal field's x+y values
// apart from d1).  A tile whose window holds no value below 16 (open sky) is 16 throughout: every tap's
// max(|k|, f) is >= 16.
template <bool VEC>
__device__ __forceinline__ void dist_z_tile(const DistArgs &A, int32_t ztile, const uint8_t *src_in, const uint8_t *src_out,
                                            uint8_t *dst) {
    constexpr int32_t Q = kZX / 4, ROWS = kZZ + 32, PER = ROWS * Q / 256;
    __shared__ uint2 F[ROWS][Q];
    const int32_t x0 = dist_tile_x0(A, (int32_t)blockIdx.x, kZX);
    const int32_t y = A.bmin[1] + (int32_t)blockIdx.y;
    const int32_t z0 = A.bmin[2] + ztile * kZZ;
    const int32_t t = (int32_t)threadIdx.x;
    bool low = false;
    {
        uint32_t v[PER];
#pragma unroll
        for (int32_t j = 0; j < PER; ++j) {   // every load issued before the first is used
            const int32_t i = t + 256 * j, row = i / Q, q = i - row * Q;
            const int32_t zabs = z0 - 16 + row;
            int32_t za = zabs;
            if (A.ring) za &= A.dims[2] - 1;
            const bool zin = A.ring || (za >= 0 && za < A.dims[2]);
            const uint8_t *src = in_span(A, 2, zabs, A.bmin[2], A.bmax[2]) ? src_in : src_out;
            v[j] = load_quad<VEC>(A, src, zin ? za : 0, y, x0 + 4 * q, zin);
        }
#pragma unroll
        for (int32_t j = 0; j < PER; ++j) {
            const int32_t i = t + 256 * j, row = i / Q, q = i - row * Q;
            const uint32_t w = v[j];
            low = low || ((w & 0xFFu) < 16u) || (((w >> 8) & 0xFFu) < 16u) || (((w >> 16) & 0xFFu) < 16u) || ((w >> 24) < 16u);
            F[row][q] = make_uint2((w & 0xFFu) | ((w & 0xFF00u) << 8), ((w >> 16) & 0xFFu) | ((w >> 8) & 0xFF0000u));
        }
    }
    const bool any_low = __syncthreads_or(low) != 0;
    const int32_t row = t / Q, q = t - (t / Q) * Q;
    const int32_t za = z0 + row;
    if (za > A.bmax[2]) return;
    const int32_t zw = A.ring ? (za & (A.dims[2] - 1)) : za;
    if (!any_low) {
        store_quad<VEC>(A, dst, zw, y, x0 + 4 * q, 0x10101010u);
        return;
    }
    vf_us2 b01 = vf_us2{16, 16}, b23 = b01;
    for (int32_t k = 0; k < 16; ++k) {
        const vf_us2 mx = __builtin_elementwise_max(b01, b23);
        if (!__any((uint32_t)max(mx.x, mx.y) > (uint32_t)k)) break;
        const vf_us2 kk = {(unsigned short)k, (unsigned short)k};
        const uint2 lo = F[row + 16 - k][q], hi = F[row + 16 + k][q];
        b01 = __builtin_elementwise_min(b01, __builtin_elementwise_max(kk, __builtin_bit_cast(vf_us2, lo.x)));
        b23 = __builtin_elementwise_min(b23, __builtin_elementwise_max(kk, __builtin_bit_cast(vf_us2, lo.y)));
        b01 = __builtin_elementwise_min(b01, __builtin_elementwise_max(kk, __builtin_bit_cast(vf_us2, hi.x)));
        b23 = __builtin_elementwise_min(b23, __builtin_elementwise_max(kk, __builtin_bit_cast(vf_us2, hi.y)));
    }
    store_quad<VEC>(A, dst, zw, y, x0 + 4 * q,
                    (uint32_t)b01.x | ((uint32_t)b01.y << 8) | ((uint32_t)b23.x << 16) | ((uint32_t)b23.y << 24));
}

template <bool VEC>
__global__ void __launch_bounds__(256) k_dist_z(DistArgs A, const uint8_t *src, uint8_t *dst) {
    dist_z_tile<VEC>(A, (int32_t)blockIdx.z, src, src, dst);
}

// Both fields' z passes in one launch (blockIdx.z = 2 x z tile + field): the normal field reads its x+y values
// (`tmp`) inside the box and d1 outside it, as record_distance_field's pass 2 finds them, into d0; the dry
// field reads d1 into d2 (pass 6).
template <bool VEC>
__global__ void __launch_bounds__(256) k_dist_z2(DistArgs A, const uint8_t *tmp, const uint8_t *d1, uint8_t *d0, uint8_t *d2) {
    const bool dry = (blockIdx.z & 1u) != 0u;
    dist_z_tile<VEC>(A, (int32_t)(blockIdx.z >> 1), dry ? d1 : tmp, d1, dry ? d2 : d0);
}

// Cell words over the whole grid: grid value, both distance fields and the coarse-cell flags.
__global__ void __launch_bounds__(256) k_pack(const uint32_t *grid, const uint8_t *d0, const uint8_t *d2, const uint8_t *c1,
                                              const uint8_t *c2, int32_t wbx, int32_t wby, int32_t wbz, int32_t c1x,
                                              int32_t c1y, int32_t c2x, int32_t c2y, uint2 *cells) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)wbx * wby * wbz) return;
    const int32_t x = (int32_t)(i % wbx), y = (int32_t)((i / wbx) % wby), z = (int32_t)(i / ((int64_t)wbx * wby));
    const uint32_t e2 = c2[((size_t)(z >> 4) * c2y + (y >> 4)) * c2x + (x >> 4)] == 0u ? 1u : 0u;
    const uint32_t e1 = c1[((size_t)(z >> 2) * c1y + (y >> 2)) * c1x + (x >> 2)] == 0u ? 1u : 0u;
    const uint32_t meta = (uint32_t)d0[i] | ((uint32_t)d2[i] << 8) | (e2 << 16) | (e1 << 17);
    cells[i] = make_uint2(grid[i], meta);
}

// Sky-top map (grasstop.comp:121-146 skyTop): 1 + the highest voxel of the column whose material
// kind is neither water (1) nor foliage (3); 0 if none.  Columns in grid (wrapped) coordinates.
// One thread per column of the window [x0, x0 + nx) x [z0, z0 + nz) (voxels).
__device__ __forceinline__ void skytop_column(const uint32_t *grid, const uint8_t *atlas, const float4 *palette, int32_t wbx,
                                              int32_t wby, int32_t wbz, int32_t xa, int32_t za, uint16_t *skytop) {
    // grid (wrapped) column: a ring world's window may start at negative absolute voxels and cross its edge
    const int32_t sx = wbx * 8, sz = wbz * 8;
    const int32_t x = (xa % sx + sx) % sx, z = (za % sz + sz) % sz;
    const int64_t i = (int64_t)z * sx + x;
    const uint32_t *col = grid + (size_t)(z >> 3) * wby * wbx + (x >> 3);
    uint32_t sky = 0u;
    // the column's bricks 8 at a time, top down: their grid words are loaded together (one wait per group)
    for (int32_t top = wby - 1; top >= 0 && sky == 0u; top -= 8) {
        uint32_t g[8];
#pragma unroll
        for (int32_t k = 0; k < 8; ++k) g[k] = col[(size_t)max(top - k, 0) * wbx];
        for (int32_t k = 0; k < 8 && sky == 0u; ++k) {
            const int32_t by = top - k;
            if (by < 0 || g[k] == 0u) continue;
            for (int32_t ly = 7; ly >= 0; --ly) {
                const uint32_t m = g[k] < 256u ? g[k] : atlas[(size_t)(g[k] - 256u) * 512u + (uint32_t)((z & 7) * 64 + ly * 8 + (x & 7))];
                if (m == 0u) continue;
                const float kind = palette[m * 3u + 1u].z;
                if (kind != 1.0f && kind != 3.0f) { sky = (uint32_t)(by * 8 + ly + 1); break; }
            }
        }
    }
    skytop[i] = (uint16_t)sky;
}

__global__ void __launch_bounds__(256) k_skytop(const uint32_t *grid, const uint8_t *atlas, const float4 *palette,
                                                int32_t wbx, int32_t wby, int32_t wbz, int32_t x0, int32_t z0, int32_t nx,
                                                int32_t nz, uint16_t *skytop) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= (int64_t)nx * nz) return;
    skytop_column(grid, atlas, palette, wbx, wby, wbz, x0 + (int32_t)(j % nx), z0 + (int32_t)(j / nx), skytop);
}

// The same over several windows of equal size in one launch (blockIdx.y = window; the windows of one call must
// not share a column).
struct SkyWindows {
    int32_t n, nx, nz;
    int32_t x0[64], z0[64];
};
__global__ void __launch_bounds__(256) k_skytop_windows(const uint32_t *grid, const uint8_t *atlas, const float4 *palette,
                                                        int32_t wbx, int32_t wby, int32_t wbz, SkyWindows Wn,
                                                        uint16_t *skytop) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int32_t w = (int32_t)blockIdx.y;
    if (j >= (int64_t)Wn.nx * Wn.nz || w >= Wn.n) return;
    skytop_column(grid, atlas, palette, wbx, wby, wbz, Wn.x0[w] + (int32_t)(j % Wn.nx), Wn.z0[w] + (int32_t)(j / Wn.nx),
                  skytop);
}

__global__ void k_update_cells(uint32_t *grid, uint8_t *srcnz, const uint32_t *idx, const uint32_t *val, uint32_t n) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    grid[idx[i]] = val[i];
    srcnz[idx[i]] = val[i] != 0u ? 1u : 0u;
}
__global__ void k_nonzero(const uint32_t *grid, int64_t n, uint8_t *nz) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) nz[i] = grid[i] != 0u ? 1u : 0u;
}
__global__ void k_update_slots(uint8_t *atlas, const uint32_t *slots, const uint8_t *vox, uint32_t n) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;   // one thread per 4 bytes
    if (i >= n * 128u) return;
    const uint32_t k = i >> 7, w = i & 127u;
    const uint32_t *src = reinterpret_cast<const uint32_t *>(vox + (size_t)k * 512u);
    uint32_t v = src[w];
    // TORCH_FLAME (254) is air for every ray (gpu_world.rs:58-67)
    uint32_t o = 0u;
    for (int b = 0; b < 4; ++b) {
        uint32_t byte = (v >> (8 * b)) & 255u;
        if (byte == 254u) byte = 0u;
        o |= byte << (8 * b);
    }
    reinterpret_cast<uint32_t *>(atlas + (size_t)slots[k] * 512u)[w] = o;
}

// ------------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------------

namespace vf {
std::string &thread_error() {
    static thread_local std::string e;
    return e;
}
void set_ctx_error(vf_ctx *ctx, const std::string &msg) { ctx->err = msg; }
}  // namespace vf

namespace vf {

int set_device(vf_ctx *ctx) {
    VF_HIP(ctx, hipSetDevice(ctx->device));
    return VF_OK;
}

int launch_check(vf_ctx *ctx, const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return vf::fail(ctx, VF_E_DEVICE, "%s launch failed: %s", what, hipGetErrorString(e));
    return VF_OK;
}

int begin_timed(vf_ctx *ctx) {
    if (ctx->capturing || !ctx->timing) return VF_OK;
    VF_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    return VF_OK;
}

int end_timed(vf_ctx *ctx) {
    if (ctx->capturing || !ctx->timing) return VF_OK;
    VF_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    ctx->timed = true;
    return VF_OK;
}

// Calls that synchronise or upload cannot be recorded into a graph.
int no_capture(vf_ctx *ctx, const char *what) {
    if (ctx && ctx->capturing) return vf::fail(ctx, VF_E_INVALID, "%s: not allowed while a graph is being captured", what);
    return VF_OK;
}

int check_ready(vf_ctx *ctx, const char *what) {
    if (!ctx) return vf::fail(nullptr, VF_E_INVALID, "%s: null context", what);
    if (!ctx->has_world) return vf::fail(ctx, VF_E_NO_WORLD, "%s: no world uploaded (vf_world_upload)", what);
    return set_device(ctx);
}

}  // namespace vf

using vf::begin_timed;
using vf::blocks_for;
using vf::check_ready;
using vf::end_timed;
using vf::launch_check;
using vf::no_capture;
using vf::set_device;

namespace {

void free_world(vf_ctx *ctx) {
    void *ptrs[] = {ctx->grid, ctx->cells, ctx->atlas, ctx->occ, ctx->occdry, ctx->masks_ref, ctx->dist[0], ctx->dist[1],
                    ctx->dist[2], ctx->dist[3], ctx->c1, ctx->c2, ctx->skytop, ctx->palette, ctx->srcnz, ctx->dry_any,
                    ctx->dist_tmp};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    ctx->grid = nullptr; ctx->cells = nullptr; ctx->atlas = nullptr; ctx->occ = nullptr; ctx->occdry = nullptr;
    ctx->masks_ref = nullptr;
    ctx->dry_any = nullptr;
    ctx->dist_tmp = nullptr;
    for (auto &d : ctx->dist) d = nullptr;
    ctx->c1 = ctx->c2 = nullptr;
    ctx->srcnz = nullptr;
    ctx->skytop = nullptr;
    ctx->palette = nullptr;
    ctx->has_world = false;
}



int rebuild_coarse(vf_ctx *ctx) {
    const int32_t *wb = ctx->W.wb;
    const int64_t n1 = (int64_t)ctx->c1d[0] * ctx->c1d[1] * ctx->c1d[2];
    const int64_t n2 = (int64_t)ctx->c2d[0] * ctx->c2d[1] * ctx->c2d[2];
    hipLaunchKernelGGL(k_coarse1, dim3(blocks_for(n1)), dim3(256), 0, ctx->stream, ctx->grid, wb[0], wb[1], wb[2], ctx->c1d[0],
                       ctx->c1d[1], ctx->c1d[2], ctx->c1);
    hipLaunchKernelGGL(k_coarse2, dim3(blocks_for(n2)), dim3(256), 0, ctx->stream, ctx->c1, ctx->c1d[0], ctx->c1d[1], ctx->c1d[2],
                       ctx->c2d[0], ctx->c2d[1], ctx->c2d[2], ctx->c2);
    return launch_check(ctx, "coarse pyramid");
}

int repack(vf_ctx *ctx) {
    const int32_t *wb = ctx->W.wb;
    hipLaunchKernelGGL(k_pack, dim3(blocks_for((int64_t)ctx->n_cells)), dim3(256), 0, ctx->stream, ctx->grid, ctx->dist[0],
                       ctx->dist[2], ctx->c1, ctx->c2, wb[0], wb[1], wb[2], ctx->c1d[0], ctx->c1d[1], ctx->c2d[0],
                       ctx->c2d[1], ctx->cells);
    return launch_check(ctx, "cell words");
}

int skytop_window(vf_ctx *ctx, int32_t x0, int32_t z0, int32_t nx, int32_t nz) {
    const int32_t *wb = ctx->W.wb;
    const int64_t n = (int64_t)nx * nz;
    if (n <= 0) return VF_OK;
    hipLaunchKernelGGL(k_skytop, dim3(blocks_for(n)), dim3(256), 0, ctx->stream, ctx->grid, ctx->atlas, ctx->palette, wb[0],
                       wb[1], wb[2], x0, z0, nx, nz, ctx->skytop);
    return launch_check(ctx, "sky-top");
}

int rebuild_skytop(vf_ctx *ctx) { return skytop_window(ctx, 0, 0, ctx->W.wb[0] * 8, ctx->W.wb[2] * 8); }

int dist_pass(vf_ctx *ctx, const int32_t bmin[3], const int32_t bmax[3], int axis, int fp, const uint8_t *src, uint8_t *dst) {
    DistArgs A;
    for (int a = 0; a < 3; ++a) {
        A.bmin[a] = bmin[a];
        A.bmax[a] = bmax[a];
        A.dims[a] = ctx->W.wb[a];
        if (bmax[a] < bmin[a]) return VF_OK;
    }
    A.axis = axis;
    A.fp = fp;
    A.ring = ctx->W.ring;
    const int64_t n = (int64_t)(bmax[0] - bmin[0] + 1) * (bmax[1] - bmin[1] + 1) * (bmax[2] - bmin[2] + 1);
    hipLaunchKernelGGL(k_dist, dim3(blocks_for(n)), dim3(256), 0, ctx->stream, A, ctx->grid, src, dst,
                       reinterpret_cast<const uint4 *>(ctx->occdry));
    return launch_check(ctx, "brickdist pass");
}

// record_distance_field (gpu_world.rs:1376-1428): box clamping, then 7 passes over 4 fields.
int distance_field(vf_ctx *ctx, const int32_t lo_in[3], const int32_t hi_in[3], const int32_t dlo_in[3], const int32_t dhi_in[3]) {
    const int32_t *wb = ctx->W.wb;
    int32_t lo[3], hi[3], dlo[3], dhi[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = lo_in[k]; hi[k] = hi_in[k]; dlo[k] = dlo_in[k]; dhi[k] = dhi_in[k];
        if (ctx->W.ring && k != 1) {
            dhi[k] = std::min(dhi[k], dlo[k] + wb[k] - 1);
            hi[k] = std::min(hi[k], lo[k] + wb[k] - 1);
            continue;
        }
        lo[k] = std::clamp(lo[k], 0, wb[k] - 1);
        hi[k] = std::clamp(hi[k], 0, wb[k] - 1);
        dlo[k] = std::clamp(dlo[k], 0, wb[k] - 1);
        dhi[k] = std::clamp(dhi[k], 0, wb[k] - 1);
    }
    uint8_t **d = ctx->dist;
    int rc;
    if (hi[1] - lo[1] + 1 > kDMaxY) {   // a box taller than the fused kernel holds: the seven passes one by one
        if ((rc = dist_pass(ctx, lo, hi, 0, 1, nullptr, d[0]))) return rc;
        if ((rc = dist_pass(ctx, lo, hi, 1, 0, d[0], d[1]))) return rc;
        if ((rc = dist_pass(ctx, lo, hi, 2, 0, d[1], d[0]))) return rc;
        if ((rc = dist_pass(ctx, dlo, dhi, 0, 2, nullptr, d[3]))) return rc;
        if ((rc = dist_pass(ctx, lo, hi, 0, 0, d[3], d[2]))) return rc;
        if ((rc = dist_pass(ctx, lo, hi, 1, 0, d[2], d[1]))) return rc;
        if ((rc = dist_pass(ctx, lo, hi, 2, 0, d[1], d[2]))) return rc;
        return VF_OK;
    }
    for (int a = 0; a < 3; ++a)
        if (hi[a] < lo[a]) return VF_OK;
    DistArgs A;
    for (int a = 0; a < 3; ++a) {
        A.bmin[a] = lo[a];
        A.bmax[a] = hi[a];
        A.dims[a] = wb[a];
    }
    A.axis = 0;
    A.fp = 0;
    A.ring = ctx->W.ring;
    const int32_t ex = hi[0] - lo[0] + 1, ey = hi[1] - lo[1] + 1, ez = hi[2] - lo[2] + 1;
    const int32_t exa = hi[0] - (lo[0] & ~3) + 1;   // from the 4-aligned tile origin
    const dim3 gxy((unsigned)((ex + kDX - 1) / kDX), (unsigned)ez),
        gxym((unsigned)((exa + kDX - 1) / kDX), (unsigned)ez),
        gz((unsigned)((exa + kZX - 1) / kZX), (unsigned)ey, (unsigned)((ez + kZZ - 1) / kZZ));
    using ZKernel = void (*)(DistArgs, const uint8_t *, uint8_t *);
    const ZKernel kz = wb[0] % 4 == 0 ? k_dist_z<true> : k_dist_z<false>;
    // a box over the world's whole height reads no row outside itself: the square-dilation count
    const bool full_height = lo[1] == 0 && hi[1] == wb[1] - 1 && wb[1] <= 64;
    bool dirty_inside = true;
    for (int a = 0; a < 3; ++a) dirty_inside = dirty_inside && dlo[a] >= lo[a] && dhi[a] <= hi[a] && dlo[a] <= dhi[a];
    if (full_height && dirty_inside) {   // both fields in two launches: x+y, then z
        DirtyBox Db;
        for (int a = 0; a < 3; ++a) { Db.lo[a] = dlo[a]; Db.hi[a] = dhi[a]; }
        const int32_t tiles_x = (exa + kDX - 1) / kDX, n_tiles = tiles_x * ez;
        if (wb[1] <= 32)
            hipLaunchKernelGGL(k_dist_xy2<uint32_t>, dim3((unsigned)n_tiles), dim3(128), 0, ctx->stream, A, Db, tiles_x, n_tiles,
                               ctx->grid, ctx->dry_any, d[3], ctx->dist_tmp, d[1]);
        else
            hipLaunchKernelGGL(k_dist_xy2<uint64_t>, dim3((unsigned)n_tiles), dim3(128), 0, ctx->stream, A, Db, tiles_x, n_tiles,
                               ctx->grid, ctx->dry_any, d[3], ctx->dist_tmp, d[1]);
        if ((rc = launch_check(ctx, "brickdist x+y"))) return rc;
        const dim3 gz2(gz.x, gz.y, gz.z * 2u);
        if (wb[0] % 4 == 0)
            hipLaunchKernelGGL(k_dist_z2<true>, gz2, dim3(256), 0, ctx->stream, A, ctx->dist_tmp, d[1], d[0], d[2]);
        else
            hipLaunchKernelGGL(k_dist_z2<false>, gz2, dim3(256), 0, ctx->stream, A, ctx->dist_tmp, d[1], d[0], d[2]);
        return launch_check(ctx, "brickdist z");
    }
    using MaskKernel = void (*)(DistArgs, const uint32_t *, const uint8_t *, uint8_t *);
    const MaskKernel mask_first = wb[1] <= 32 ? k_dist_xy_mask<true, uint32_t> : k_dist_xy_mask<true, uint64_t>;
    const MaskKernel mask_dry = wb[1] <= 32 ? k_dist_xy_mask<false, uint32_t> : k_dist_xy_mask<false, uint64_t>;
    // passes 0-2: x and y (d0's x values live in LDS only), then z into d0
    if (full_height)
        hipLaunchKernelGGL(mask_first, gxym, dim3(64), 0, ctx->stream, A, ctx->grid, nullptr, d[1]);
    else
        hipLaunchKernelGGL(k_dist_xy<true>, gxy, dim3(256), 0, ctx->stream, A, ctx->grid, nullptr, d[0], d[1]);
    if ((rc = launch_check(ctx, "brickdist x+y"))) return rc;
    hipLaunchKernelGGL(kz, gz, dim3(256), 0, ctx->stream, A, d[1], d[0]);
    if ((rc = launch_check(ctx, "brickdist z"))) return rc;
    // pass 3: the dry occupancy over the dirty box; passes 4-6 as above on it, into d2
    if ((rc = dist_pass(ctx, dlo, dhi, 0, 2, nullptr, d[3]))) return rc;
    if (full_height)
        hipLaunchKernelGGL(mask_dry, gxym, dim3(64), 0, ctx->stream, A, ctx->grid, d[3], d[1]);
    else
        hipLaunchKernelGGL(k_dist_xy<false>, gxy, dim3(256), 0, ctx->stream, A, ctx->grid, d[3], d[2], d[1]);
    if ((rc = launch_check(ctx, "brickdist dry x+y"))) return rc;
    hipLaunchKernelGGL(kz, gz, dim3(256), 0, ctx->stream, A, d[1], d[2]);
    return launch_check(ctx, "brickdist dry z");
}




}  // namespace

namespace {
// The variants of a primary kernel: the tile grid for a camera in air, and a persistent grid of one
// workgroup per compute unit for a camera in solid terrain or water (it returns at once otherwise); the
// host's grid copy launches only the one that applies unless the camera is in a detailed brick.  A frame
// graph records that choice with the camera (record it again after an edit of the camera's brick).
}  // namespace

namespace vf {
// Occupancy bits of one 512-voxel slot: word lz, bit ly * 8 + lx = voxel != 0.
void slot_occupancy(const uint8_t *vox, uint64_t out[8]) {
    for (int w = 0; w < 8; ++w) {
        uint64_t m = 0ull;
        for (int b = 0; b < 64; ++b) m |= (uint64_t)(vox[w * 64 + b] != 0u) << b;
        out[w] = m;
    }
}
void host_occupancy(std::vector<uint64_t> &occ, const uint8_t *atlas, uint32_t n_slots) {
    occ.assign((size_t)n_slots * 8u, 0ull);
    if (!n_slots || !atlas) return;
    const unsigned nt = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            for (size_t sl = t; sl < n_slots; sl += nt) slot_occupancy(atlas + sl * 512u, &occ[sl * 8u]);
        });
    for (auto &x : th) x.join();
}
}  // namespace vf

namespace {
// The camera voxel from the host copies of the grid and of the slots' occupancy bits: 0 = air (or
// outside the loaded world), 1 = solid or water (the SPECIAL variant decides which on the device).
int camera_class(const vf_ctx *ctx, const float pos[3]) {
    const DevWorld &W = ctx->W;
    int32_t v[3];
    for (int a = 0; a < 3; ++a) {
        const float f = std::floor(pos[a]);
        v[a] = f != f ? 0 : (f >= 2147483648.0f ? INT32_MAX : (f <= -2147483648.0f ? INT32_MIN : (int32_t)f));
        if (v[a] < W.minb[a] * 8 || v[a] >= (W.minb[a] + W.wb[a]) * 8) return 0;
    }
    int32_t bx = v[0] >> 3, by = v[1] >> 3, bz = v[2] >> 3;
    if (W.ring) { bx &= W.wb[0] - 1; bz &= W.wb[2] - 1; }
    const uint32_t g = ctx->host_grid[((size_t)bz * W.wb[1] + by) * W.wb[0] + bx];
    if (g < 256u) return g == 0u ? 0 : 1;
    const uint32_t k = (uint32_t)(((v[2] & 7) << 6) | ((v[1] & 7) << 3) | (v[0] & 7));
    return ((ctx->host_occ[(size_t)(g - 256u) * 8u + (k >> 6)] >> (k & 63u)) & 1ull) ? 1 : 0;
}

}  // namespace
namespace vf {
bool camera_in_air(const vf_ctx *ctx, const float pos[3]) { return camera_class(ctx, pos) == 0; }
}  // namespace vf
namespace {

template <bool COUNT>
void launch_primary(vf_ctx *ctx, PrimArgs A, vf_hit *hits, float4 *gbpos, uint32_t *gbn) {
    const int cls = camera_class(ctx, A.pos);
    A.cam_air = cls == 0 ? 1 : 0;
    if (cls == 0 && !COUNT && vf::small_frame(ctx, (int64_t)A.ntiles * 4))
        hipLaunchKernelGGL(k_primary_small, dim3(A.ntiles * 4), dim3(64), 0, ctx->stream, ctx->W, A, hits, gbpos, gbn);
    else if (cls != 1)
        hipLaunchKernelGGL((k_primary<COUNT, false>), dim3(A.ntiles), dim3(256), 0, ctx->stream, ctx->W, A, hits, gbpos,
                           gbn, ctx->d_cnt);
    const int32_t g = A.ntiles < ctx->n_cu ? A.ntiles : ctx->n_cu;
    if (cls != 0)
        hipLaunchKernelGGL((k_primary<COUNT, true>), dim3(g), dim3(256), 0, ctx->stream, ctx->W, A, hits, gbpos, gbn,
                           ctx->d_cnt);
}
template <bool COUNT>
void launch_primary_shadow(vf_ctx *ctx, PrimArgs A, const float *s, vf_hit *hits, float4 *gbpos, uint32_t *gbn,
                           uint32_t *shadow, vf_hit *shits) {
    const int cls = camera_class(ctx, A.pos);
    A.cam_air = cls == 0 ? 1 : 0;
    const DirPre sp = dir_pre(s);   // the sun ray's set-up, as trace_world computes it
    if (cls == 0 && !COUNT && vf::small_frame(ctx, (int64_t)A.ntiles * 4))
        hipLaunchKernelGGL(k_primary_shadow_small, dim3(A.ntiles * 4), dim3(64), 0, ctx->stream, ctx->W, A, s[0], s[1], s[2],
                           hits, gbpos, gbn, shadow, shits, sp);
    else if (cls == 0)   // camera in air: one wave per workgroup (measured 0.5 % faster than 16x16 tiles)
        hipLaunchKernelGGL((k_primary_shadow64<COUNT>), dim3(A.ntiles * 4), dim3(64), 0, ctx->stream, ctx->W, A, s[0], s[1],
                           s[2], hits, gbpos, gbn, shadow, shits, ctx->d_cnt, sp);
    else if (cls != 1)
        hipLaunchKernelGGL((k_primary_shadow<COUNT, false>), dim3(A.ntiles), dim3(256), 0, ctx->stream, ctx->W, A, s[0],
                           s[1], s[2], hits, gbpos, gbn, shadow, shits, ctx->d_cnt);
    const int32_t g = A.ntiles < ctx->n_cu ? A.ntiles : ctx->n_cu;
    if (cls != 0)
        hipLaunchKernelGGL((k_primary_shadow<COUNT, true>), dim3(g), dim3(256), 0, ctx->stream, ctx->W, A, s[0], s[1], s[2],
                           hits, gbpos, gbn, shadow, shits, ctx->d_cnt);
}
// The scene variant (bodies present): one tile grid for every camera case.
template <bool COUNT, bool SHADOW>
void launch_scene(vf_ctx *ctx, PrimArgs A, const float *s, vf_hit *hits, float4 *gbpos, uint32_t *gbn, int32_t *body,
                  uint32_t *shadow, vf_hit *shits) {
    A.cam_air = camera_class(ctx, A.pos) == 0 ? 1 : 0;
    hipLaunchKernelGGL((k_scene<COUNT, SHADOW>), dim3(A.ntiles), dim3(256), 0, ctx->stream, ctx->W, vf::dev_bodies(ctx), A,
                       s[0], s[1], s[2], hits, gbpos, gbn, body, shadow, shits, ctx->d_cnt);
}

PrimArgs prim_args(const vf_pass_ubo *p, int32_t w, int32_t h) {
    PrimArgs A;
    for (int a = 0; a < 3; ++a) {
        A.pos[a] = p->cam_pos[a]; A.right[a] = p->cam_right[a]; A.up[a] = p->cam_up[a]; A.fwd[a] = p->cam_fwd[a];
    }
    A.thf = p->tan_half_fov;
    A.jit[0] = p->jitter[0];
    A.jit[1] = p->jitter[1];
    A.w = w; A.h = h;
    A.tiles_x = (w + 15) / 16;
    A.ntiles = A.tiles_x * ((h + 15) / 16);
    A.cam_air = 0;
    A.leaf = p->look_p[3];
    A.aspect = (float)w / (float)h;
    return A;
}

__global__ void k_fill_i32(int32_t *p, long long n, int32_t v) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

// P::Gbuf (+ P::ShTrace) with every output: the world-only kernels when there are no bodies, else the
// scene kernel.
int trace_gbuffer(vf_ctx *ctx, const char *what, const vf_pass_ubo *p, int32_t w, int32_t h, vf_hit *hits, float *gb_pos,
                  uint32_t *gb_normal, int32_t *body, uint32_t *shadow, vf_hit *shadow_hits) {
    int rc = check_ready(ctx, what);
    if (rc) return rc;
    if (shadow && !ctx->has_scene) return vf::fail(ctx, VF_E_INVALID, "%s: no scene (vf_set_scene)", what);
    if (!p) return vf::fail(ctx, VF_E_INVALID, "%s: null pass", what);
    if (w <= 0 || h <= 0 || (int64_t)w * h > (1ll << 30)) return vf::fail(ctx, VF_E_INVALID, "%s: bad size %dx%d", what, w, h);
    if ((rc = vf::check_bodies(ctx, what))) return rc;
    const PrimArgs A = prim_args(p, w, h);
    const float *s = ctx->S.sun_dir;
    float4 *gp = reinterpret_cast<float4 *>(gb_pos);
    if ((rc = begin_timed(ctx))) return rc;
    if (vf::dev_bodies(ctx).n > 0) {
        if (shadow) {
            if (ctx->counting) launch_scene<true, true>(ctx, A, s, hits, gp, gb_normal, body, shadow, shadow_hits);
            else launch_scene<false, true>(ctx, A, s, hits, gp, gb_normal, body, shadow, shadow_hits);
        } else {
            if (ctx->counting) launch_scene<true, false>(ctx, A, s, hits, gp, gb_normal, body, nullptr, nullptr);
            else launch_scene<false, false>(ctx, A, s, hits, gp, gb_normal, body, nullptr, nullptr);
        }
        if ((rc = launch_check(ctx, "k_scene"))) return rc;
        return end_timed(ctx);
    }
    if (shadow) {
        if (ctx->counting) launch_primary_shadow<true>(ctx, A, s, hits, gp, gb_normal, shadow, shadow_hits);
        else launch_primary_shadow<false>(ctx, A, s, hits, gp, gb_normal, shadow, shadow_hits);
        if ((rc = launch_check(ctx, "k_primary_shadow"))) return rc;
    } else {
        if (ctx->counting) launch_primary<true>(ctx, A, hits, gp, gb_normal);
        else launch_primary<false>(ctx, A, hits, gp, gb_normal);
        if ((rc = launch_check(ctx, "k_primary"))) return rc;
    }
    if (body) {
        const long long n = (long long)w * h;
        hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(n)), dim3(256), 0, ctx->stream, body, n, -1);
        if ((rc = launch_check(ctx, "body index fill"))) return rc;
    }
    return end_timed(ctx);
}
}  // namespace

namespace vf {

// The device world's shape and buffers for a world of wb bricks with n_slots atlas slots: checks, frees
// the previous world, fills the DevWorld description and allocates every buffer (vf_world_upload and
// vf_world_pack share it).
int world_prepare(vf_ctx *ctx, const int32_t wb_in[3], const int32_t win[4], uint32_t n_slots, const char *what) {
    for (int a = 0; a < 3; ++a)
        if (wb_in[a] <= 0 || wb_in[a] > 4096)
            return vf::fail(ctx, VF_E_INVALID, "%s: world_bricks[%d] = %d out of range", what, a, wb_in[a]);
    const int ring = win[3] != 0;
    if (ring) {
        const int32_t x = wb_in[0], z = wb_in[2];
        if ((x & (x - 1)) || (z & (z - 1))) return vf::fail(ctx, VF_E_INVALID, "%s: ring mode needs power-of-two x/z brick dims", what);
    }
    if (n_slots > (1u << 24)) return vf::fail(ctx, VF_E_INVALID, "%s: %u atlas slots > 2^24", what, n_slots);
    int rc = set_device(ctx);
    if (rc) return rc;
    VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    free_world(ctx);
    DevWorld &W = ctx->W;
    W = DevWorld{};
    for (int a = 0; a < 3; ++a) {
        W.wb[a] = wb_in[a];
        ctx->c1d[a] = (W.wb[a] + 3) / 4;
        ctx->c2d[a] = (W.wb[a] + 15) / 16;
    }
    W.ring = ring;
    W.minb[0] = ring ? win[0] : 0;
    W.minb[1] = 0;
    W.minb[2] = ring ? win[2] : 0;
    W.use_dist = 1;
    W.max_outer = 3 * (W.wb[0] + W.wb[1] + W.wb[2]) + 4;
    const size_t n = (size_t)W.wb[0] * W.wb[1] * W.wb[2];
    ctx->n_cells = n;
    ctx->n_slots = n_slots;
    const size_t n1 = (size_t)ctx->c1d[0] * ctx->c1d[1] * ctx->c1d[2], n2 = (size_t)ctx->c2d[0] * ctx->c2d[1] * ctx->c2d[2];
    const size_t slots = std::max<size_t>(n_slots, 1);
    const size_t nsky = (size_t)W.wb[0] * 8 * W.wb[2] * 8;
#define VF_ALLOC(ptr, bytes)                                                                                     \
    do {                                                                                                         \
        hipError_t e_ = hipMalloc((void **)&(ptr), (bytes));                                                     \
        if (e_ != hipSuccess) {                                                                                  \
            free_world(ctx);                                                                                     \
            return vf::fail(ctx, VF_E_OOM, "%s: hipMalloc(%zu): %s", what, (size_t)(bytes), hipGetErrorString(e_)); \
        }                                                                                                        \
    } while (0)
    VF_ALLOC(ctx->grid, n * 4);
    VF_ALLOC(ctx->cells, n * 8);
    VF_ALLOC(ctx->atlas, slots * 512);
    VF_ALLOC(ctx->occ, slots * 64);
    VF_ALLOC(ctx->occdry, slots * 64);
    VF_ALLOC(ctx->dry_any, slots);
    VF_ALLOC(ctx->dist_tmp, n);
    VF_ALLOC(ctx->masks_ref, slots * 64);
    for (auto &p : ctx->dist) VF_ALLOC(p, n);
    VF_ALLOC(ctx->c1, n1);
    VF_ALLOC(ctx->c2, n2);
    VF_ALLOC(ctx->srcnz, n);
    VF_ALLOC(ctx->skytop, nsky * 2);
    VF_ALLOC(ctx->palette, 256 * 48);
#undef VF_ALLOC
    VF_HIP(ctx, hipMemsetAsync(ctx->cells, 0, n * 8, ctx->stream));
    VF_HIP(ctx, hipMemsetAsync(ctx->atlas, 0, slots * 512, ctx->stream));
    VF_HIP(ctx, hipMemsetAsync(ctx->masks_ref, 0, slots * 64, ctx->stream));
    VF_HIP(ctx, hipMemsetAsync(ctx->occ, 0, slots * 64, ctx->stream));
    VF_HIP(ctx, hipMemsetAsync(ctx->occdry, 0, slots * 64, ctx->stream));
    VF_HIP(ctx, hipMemsetAsync(ctx->dry_any, 0, slots, ctx->stream));
    return VF_OK;
}

// Everything derived from the grid and the atlas once both are on the device: occupancy masks, the
// coarse pyramid (unless the caller's was copied), the distance fields, the cell words, the sky-top map.
int world_finish(vf_ctx *ctx, bool coarse_given, bool source_given) {
    int rc;
    DevWorld &W = ctx->W;
    if (!source_given) {   // the source grid is not known: its non-air cells are the encoded grid's
        hipLaunchKernelGGL(k_nonzero, dim3(blocks_for((int64_t)ctx->n_cells)), dim3(256), 0, ctx->stream,
                           (const uint32_t *)ctx->grid, (int64_t)ctx->n_cells, ctx->srcnz);
        if ((rc = launch_check(ctx, "source cells"))) return rc;
    }
    if (ctx->n_slots) {
        hipLaunchKernelGGL(k_masks, dim3(blocks_for((int64_t)ctx->n_slots * 16)), dim3(256), 0, ctx->stream, ctx->atlas, nullptr,
                           0u, ctx->n_slots, ctx->occ, ctx->occdry, ctx->dry_any);
        if ((rc = launch_check(ctx, "occupancy masks"))) return rc;
    }
    if (!coarse_given && (rc = rebuild_coarse(ctx))) return rc;
    W.cells = ctx->cells;
    W.atlas = ctx->atlas;
    W.occ = reinterpret_cast<const uint4 *>(ctx->occ);
    W.occdry = reinterpret_cast<const uint4 *>(ctx->occdry);
    W.skytop = ctx->skytop;
    W.palette = ctx->palette;
    const int32_t lo[3] = {W.minb[0], 0, W.minb[2]};
    const int32_t hi[3] = {W.minb[0] + W.wb[0] - 1, W.wb[1] - 1, W.minb[2] + W.wb[2] - 1};
    if ((rc = distance_field(ctx, lo, hi, lo, hi))) return rc;
    if ((rc = repack(ctx))) return rc;
    if ((rc = rebuild_skytop(ctx))) return rc;
    VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->has_world = true;
    return VF_OK;
}

// The coarse pyramid of a grid that is not (yet) the world's: coarse1 = any non-empty brick, coarse2 =
// any coarse1 cell (gpu_world.rs:300-313).
int build_coarse(vf_ctx *ctx, const uint32_t *d_grid) {
    const int32_t *wb = ctx->W.wb;
    hipLaunchKernelGGL(k_nonzero, dim3(blocks_for((int64_t)ctx->n_cells)), dim3(256), 0, ctx->stream, d_grid,
                       (int64_t)ctx->n_cells, ctx->srcnz);
    const int64_t n1 = (int64_t)ctx->c1d[0] * ctx->c1d[1] * ctx->c1d[2];
    const int64_t n2 = (int64_t)ctx->c2d[0] * ctx->c2d[1] * ctx->c2d[2];
    hipLaunchKernelGGL(k_coarse1, dim3(blocks_for(n1)), dim3(256), 0, ctx->stream, d_grid, wb[0], wb[1], wb[2], ctx->c1d[0],
                       ctx->c1d[1], ctx->c1d[2], ctx->c1);
    hipLaunchKernelGGL(k_coarse2, dim3(blocks_for(n2)), dim3(256), 0, ctx->stream, ctx->c1, ctx->c1d[0], ctx->c1d[1], ctx->c1d[2],
                       ctx->c2d[0], ctx->c2d[1], ctx->c2d[2], ctx->c2);
    return launch_check(ctx, "coarse pyramid");
}

// Occupancy masks of the listed slots (device list), after their atlas bricks changed.
int masks_for(vf_ctx *ctx, const uint32_t *d_slots, uint32_t n) {
    if (!n) return VF_OK;
    hipLaunchKernelGGL(k_masks, dim3(blocks_for((int64_t)n * 16)), dim3(256), 0, ctx->stream, ctx->atlas, d_slots, n, ctx->n_slots,
                       ctx->occ, ctx->occdry, ctx->dry_any);
    return launch_check(ctx, "occupancy masks");
}

// The brick distance field over one box (record_distance_field's box semantics).
int distance_box(vf_ctx *ctx, const int32_t lo[3], const int32_t hi[3], const int32_t dlo[3], const int32_t dhi[3]) {
    return distance_field(ctx, lo, hi, dlo, dhi);
}

// The sky-top map over the voxel columns [x0, x0 + nx) x [z0, z0 + nz) (grid coordinates).
int skytop_columns(vf_ctx *ctx, int32_t x0, int32_t z0, int32_t nx, int32_t nz) { return skytop_window(ctx, x0, z0, nx, nz); }

// The sky-top map over n windows of nx x nz columns at (x0[k], z0[k]) that share no column, 64 per launch.
int skytop_windows(vf_ctx *ctx, const int32_t *x0, const int32_t *z0, int32_t n, int32_t nx, int32_t nz) {
    const int32_t *wb = ctx->W.wb;
    for (int32_t k0 = 0; k0 < n; k0 += 64) {
        SkyWindows Wn;
        Wn.n = std::min(64, n - k0);
        Wn.nx = nx;
        Wn.nz = nz;
        for (int32_t k = 0; k < Wn.n; ++k) { Wn.x0[k] = x0[k0 + k]; Wn.z0[k] = z0[k0 + k]; }
        hipLaunchKernelGGL(k_skytop_windows, dim3(blocks_for((int64_t)nx * nz), (unsigned)Wn.n), dim3(256), 0, ctx->stream,
                           ctx->grid, ctx->atlas, ctx->palette, wb[0], wb[1], wb[2], Wn, ctx->skytop);
        if (int rc = launch_check(ctx, "sky-top")) return rc;
    }
    return VF_OK;
}

// The cell words (grid value, distance fields, coarse flags) of the whole grid.
int cell_words(vf_ctx *ctx) { return repack(ctx); }

// record_grass_top's window (gpu_world.rs:1431-1441): voxel columns [x0, x1] x [z0, z1], absolute in a ring
// world (at most one ring wide), clamped to the grid otherwise.
int skytop_box(vf_ctx *ctx, int32_t x0, int32_t z0, int32_t x1, int32_t z1) {
    const int32_t nx = ctx->W.wb[0] * 8, nz = ctx->W.wb[2] * 8;
    if (ctx->W.ring) {
        x1 = std::min(x1, x0 + nx - 1);
        z1 = std::min(z1, z0 + nz - 1);
    } else {
        x0 = std::clamp(x0, 0, nx - 1); x1 = std::clamp(x1, 0, nx - 1);
        z0 = std::clamp(z0, 0, nz - 1); z1 = std::clamp(z1, 0, nz - 1);
    }
    if (x1 < x0 || z1 < z0) return VF_OK;
    return skytop_window(ctx, x0, z0, x1 - x0 + 1, z1 - z0 + 1);
}

// Cell words and sky-top map after grid, field or atlas changes.
int refresh_world(vf_ctx *ctx) {
    int rc = repack(ctx);
    return rc ? rc : rebuild_skytop(ctx);
}

}  // namespace vf

extern "C" {

int vf_abi_version(void) { return VF_ABI_VERSION; }

int vf_abi_sizes(size_t *scene, size_t *pass, size_t *mat, size_t *ray, size_t *hit) {
    if (scene) *scene = sizeof(vf_scene_ubo);
    if (pass) *pass = sizeof(vf_pass_ubo);
    if (mat) *mat = sizeof(vf_material);
    if (ray) *ray = sizeof(vf_ray);
    if (hit) *hit = sizeof(vf_hit);
    return VF_OK;
}

const char *vf_last_error(const vf_ctx *ctx) {
    if (ctx) return ctx->err.c_str();
    return vf::thread_error().c_str();
}

int vf_ctx_create(int device, vf_ctx **out) {
    if (!out) return vf::fail(nullptr, VF_E_INVALID, "vf_ctx_create: null out");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        return vf::fail(nullptr, VF_E_DEVICE, "vf_ctx_create: no HIP device (%s)", e != hipSuccess ? hipGetErrorString(e) : "count 0");
    if (device < 0 || device >= n) return vf::fail(nullptr, VF_E_INVALID, "vf_ctx_create: device %d of %d", device, n);
    vf_ctx *ctx = new (std::nothrow) vf_ctx();
    if (!ctx) return vf::fail(nullptr, VF_E_OOM, "vf_ctx_create: out of host memory");
    ctx->device = device;
    int rc = VF_OK;
    if ((e = hipSetDevice(device)) != hipSuccess || (e = hipStreamCreate(&ctx->own_stream)) != hipSuccess ||
        (e = hipMalloc(&ctx->d_cnt, 4 * sizeof(unsigned long long))) != hipSuccess ||
        (e = hipMemset(ctx->d_cnt, 0, 4 * sizeof(unsigned long long))) != hipSuccess ||
        (e = hipEventCreate(&ctx->ev0)) != hipSuccess || (e = hipEventCreate(&ctx->ev1)) != hipSuccess) {
        rc = vf::fail(nullptr, VF_E_DEVICE, "vf_ctx_create: %s", hipGetErrorString(e));
        vf_ctx_destroy(ctx);
        return rc;
    }
    ctx->stream = ctx->own_stream;
    if ((rc = vf::queue_scratch_init(ctx))) {
        vf_ctx_destroy(ctx);
        return rc;
    }
    {
        int cu = 0, khz = 0;
        if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cu > 0) ctx->n_cu = cu;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0) ctx->wall_khz = khz;
    }
    *out = ctx;
    return VF_OK;
}

void vf_ctx_destroy(vf_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    free_world(ctx);
    vf::free_queue(ctx);
    vf::free_bodies(ctx);
    if (ctx->d_cnt) (void)hipFree(ctx->d_cnt);
    if (ctx->pin) (void)hipHostFree(ctx->pin);
    if (ctx->sync_ws) (void)hipFree(ctx->sync_ws);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
}

int vf_set_stream(vf_ctx *ctx, void *s) {
    if (!ctx) return vf::fail(nullptr, VF_E_INVALID, "vf_set_stream: null context");
    if (int rc = no_capture(ctx, "vf_set_stream")) return rc;
    ctx->stream = s ? reinterpret_cast<hipStream_t>(s) : ctx->own_stream;
    return VF_OK;
}

int vf_sync(vf_ctx *ctx) {
    if (!ctx) return vf::fail(nullptr, VF_E_INVALID, "vf_sync: null context");
    int rc = no_capture(ctx, "vf_sync");
    if (rc) return rc;
    rc = set_device(ctx);
    if (rc) return rc;
    VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VF_OK;
}

int vf_device_alloc(vf_ctx *ctx, size_t bytes, void **d_out) {
    if (!ctx || !d_out) return vf::fail(ctx, VF_E_INVALID, "vf_device_alloc: null argument");
    if (int rc = no_capture(ctx, "vf_device_alloc")) return rc;
    int rc = set_device(ctx);
    if (rc) return rc;
    *d_out = nullptr;
    VF_HIP(ctx, hipMalloc(d_out, bytes ? bytes : 1));
    return VF_OK;
}
int vf_device_free(vf_ctx *ctx, void *p) {
    if (!ctx) return vf::fail(nullptr, VF_E_INVALID, "vf_device_free: null context");
    if (int rc = no_capture(ctx, "vf_device_free")) return rc;
    if (!p) return VF_OK;
    int rc = set_device(ctx);
    if (rc) return rc;
    VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    VF_HIP(ctx, hipFree(p));
    return VF_OK;
}
int vf_copy_to_device(vf_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return vf::fail(ctx, VF_E_INVALID, "vf_copy_to_device: null argument");
    if (int rc = no_capture(ctx, "vf_copy_to_device")) return rc;
    int rc = set_device(ctx);
    if (rc) return rc;
    VF_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VF_OK;
}
int vf_copy_to_host(vf_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return vf::fail(ctx, VF_E_INVALID, "vf_copy_to_host: null argument");
    if (int rc = no_capture(ctx, "vf_copy_to_host")) return rc;
    int rc = set_device(ctx);
    if (rc) return rc;
    VF_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VF_OK;
}
int vf_memset_device(vf_ctx *ctx, void *dst, int value, size_t bytes) {
    if (!ctx || (!dst && bytes)) return vf::fail(ctx, VF_E_INVALID, "vf_memset_device: null argument");
    if (int rc = no_capture(ctx, "vf_memset_device")) return rc;
    int rc = set_device(ctx);
    if (rc) return rc;
    VF_HIP(ctx, hipMemsetAsync(dst, value, bytes, ctx->stream));
    return VF_OK;
}

int vf_world_upload(vf_ctx *ctx, const vf_world_desc *d) {
    if (!ctx || !d) return vf::fail(ctx, VF_E_INVALID, "vf_world_upload: null argument");
    if (int rc = no_capture(ctx, "vf_world_upload")) return rc;
    if (!d->grid || !d->palette) return vf::fail(ctx, VF_E_INVALID, "vf_world_upload: grid and palette are required");
    if (d->n_slots && !d->atlas) return vf::fail(ctx, VF_E_INVALID, "vf_world_upload: n_slots > 0 without an atlas");
    for (int a = 0; a < 3; ++a)
        if (d->world_bricks[a] <= 0 || d->world_bricks[a] > 4096)
            return vf::fail(ctx, VF_E_INVALID, "vf_world_upload: world_bricks[%d] = %d out of range", a, d->world_bricks[a]);
    const size_t n = (size_t)d->world_bricks[0] * d->world_bricks[1] * d->world_bricks[2];
    for (size_t i = 0; i < n; ++i)
        if (d->grid[i] >= 256u && d->grid[i] - 256u >= d->n_slots)
            return vf::fail(ctx, VF_E_INVALID, "vf_world_upload: grid[%zu] = %u names slot %u of %u", i, d->grid[i],
                            d->grid[i] - 256u, d->n_slots);
    int rc = vf::world_prepare(ctx, d->world_bricks, d->world_win, d->n_slots, "vf_world_upload");
    if (rc) return rc;
    ctx->host_grid.assign(d->grid, d->grid + n);
    ctx->atlas_used = d->n_slots;
    ctx->free_slots.clear();
    ctx->world_slot.assign(d->n_slots, 0);
    for (size_t i = 0; i < n; ++i)
        if (d->grid[i] >= 256u) ctx->world_slot[d->grid[i] - 256u] = 1;
    vf::host_occupancy(ctx->host_occ, d->atlas, d->n_slots);
    VF_HIP(ctx, hipMemcpyAsync(ctx->grid, d->grid, n * 4, hipMemcpyHostToDevice, ctx->stream));
    if (d->n_slots) {
        // TORCH_FLAME -> air as the GPU atlas sees it (gpu_world.rs:58-67)
        std::vector<uint8_t> atlas(d->atlas, d->atlas + (size_t)d->n_slots * 512);
        for (auto &v : atlas)
            if (v == 254) v = 0;
        VF_HIP(ctx, hipMemcpyAsync(ctx->atlas, atlas.data(), atlas.size(), hipMemcpyHostToDevice, ctx->stream));
        VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (d->masks)
            VF_HIP(ctx, hipMemcpyAsync(ctx->masks_ref, d->masks, (size_t)d->n_slots * 64, hipMemcpyHostToDevice, ctx->stream));
    }
    VF_HIP(ctx, hipMemcpyAsync(ctx->palette, d->palette, 256 * 48, hipMemcpyHostToDevice, ctx->stream));
    const bool coarse = d->coarse1 && d->coarse2;
    if (coarse) {
        const size_t n1 = (size_t)ctx->c1d[0] * ctx->c1d[1] * ctx->c1d[2], n2 = (size_t)ctx->c2d[0] * ctx->c2d[1] * ctx->c2d[2];
        VF_HIP(ctx, hipMemcpyAsync(ctx->c1, d->coarse1, n1, hipMemcpyHostToDevice, ctx->stream));
        VF_HIP(ctx, hipMemcpyAsync(ctx->c2, d->coarse2, n2, hipMemcpyHostToDevice, ctx->stream));
    }
    return vf::world_finish(ctx, coarse);
}

int vf_world_update(vf_ctx *ctx, const uint32_t *cell_idx, const uint32_t *cell_val, uint32_t n_cells, const uint32_t *slot_idx,
                    const uint8_t *slot_vox, uint32_t n_slot_updates) {
    int rc = check_ready(ctx, "vf_world_update");
    if (rc) return rc;
    if ((rc = no_capture(ctx, "vf_world_update"))) return rc;
    if ((n_cells && (!cell_idx || !cell_val)) || (n_slot_updates && (!slot_idx || !slot_vox)))
        return vf::fail(ctx, VF_E_INVALID, "vf_world_update: null arrays");
    for (uint32_t i = 0; i < n_cells; ++i) {
        if (cell_idx[i] >= ctx->n_cells) return vf::fail(ctx, VF_E_INVALID, "vf_world_update: cell %u out of range", cell_idx[i]);
        if (cell_val[i] >= 256u && cell_val[i] - 256u >= ctx->n_slots)
            return vf::fail(ctx, VF_E_INVALID, "vf_world_update: value %u names a slot beyond %u", cell_val[i], ctx->n_slots);
    }
    for (uint32_t i = 0; i < n_slot_updates; ++i)
        if (slot_idx[i] >= ctx->n_slots) return vf::fail(ctx, VF_E_INVALID, "vf_world_update: slot %u out of range", slot_idx[i]);
    VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // staging (the reference's staging ring): temporary device buffers for this update
    void *d_ci = nullptr, *d_cv = nullptr, *d_si = nullptr, *d_sv = nullptr;
    auto cleanup = [&]() {
        for (void *p : {d_ci, d_cv, d_si, d_sv})
            if (p) (void)hipFree(p);
    };
    hipError_t e = hipSuccess;
    if (n_cells) {
        if ((e = hipMalloc(&d_ci, (size_t)n_cells * 4)) != hipSuccess || (e = hipMalloc(&d_cv, (size_t)n_cells * 4)) != hipSuccess ||
            (e = hipMemcpy(d_ci, cell_idx, (size_t)n_cells * 4, hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(d_cv, cell_val, (size_t)n_cells * 4, hipMemcpyHostToDevice)) != hipSuccess) {
            cleanup();
            return vf::fail(ctx, VF_E_DEVICE, "vf_world_update: %s", hipGetErrorString(e));
        }
        hipLaunchKernelGGL(k_update_cells, dim3(blocks_for(n_cells)), dim3(256), 0, ctx->stream, ctx->grid, ctx->srcnz, (const uint32_t *)d_ci,
                           (const uint32_t *)d_cv, n_cells);
    }
    for (uint32_t i = 0; i < n_cells; ++i) ctx->host_grid[cell_idx[i]] = cell_val[i];
    for (uint32_t i = 0; i < n_slot_updates; ++i) vf::slot_occupancy(slot_vox + (size_t)i * 512u, &ctx->host_occ[(size_t)slot_idx[i] * 8u]);
    if (n_slot_updates) {
        if ((e = hipMalloc(&d_si, (size_t)n_slot_updates * 4)) != hipSuccess ||
            (e = hipMalloc(&d_sv, (size_t)n_slot_updates * 512)) != hipSuccess ||
            (e = hipMemcpy(d_si, slot_idx, (size_t)n_slot_updates * 4, hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(d_sv, slot_vox, (size_t)n_slot_updates * 512, hipMemcpyHostToDevice)) != hipSuccess) {
            cleanup();
            return vf::fail(ctx, VF_E_DEVICE, "vf_world_update: %s", hipGetErrorString(e));
        }
        hipLaunchKernelGGL(k_update_slots, dim3(blocks_for((int64_t)n_slot_updates * 128)), dim3(256), 0, ctx->stream, ctx->atlas,
                           (const uint32_t *)d_si, (const uint8_t *)d_sv, n_slot_updates);
        hipLaunchKernelGGL(k_masks, dim3(blocks_for((int64_t)n_slot_updates * 16)), dim3(256), 0, ctx->stream, ctx->atlas,
                           (const uint32_t *)d_si, n_slot_updates, ctx->n_slots, ctx->occ, ctx->occdry, ctx->dry_any);
    }
    if ((rc = launch_check(ctx, "world update"))) { cleanup(); return rc; }
    if ((rc = rebuild_coarse(ctx)) || (rc = repack(ctx)) || (rc = rebuild_skytop(ctx))) { cleanup(); return rc; }
    e = hipStreamSynchronize(ctx->stream);
    cleanup();
    if (e != hipSuccess) return vf::fail(ctx, VF_E_DEVICE, "vf_world_update: %s", hipGetErrorString(e));
    return VF_OK;
}

int vf_graph_begin(vf_ctx *ctx) {
    if (!ctx) return vf::fail(nullptr, VF_E_INVALID, "vf_graph_begin: null context");
    if (ctx->capturing) return vf::fail(ctx, VF_E_INVALID, "vf_graph_begin: already capturing");
    if (!ctx->stream) return vf::fail(ctx, VF_E_INVALID, "vf_graph_begin: the null stream cannot be captured");
    int rc = set_device(ctx);
    if (rc) return rc;
    VF_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    ctx->capturing = true;
    return VF_OK;
}

int vf_graph_end(vf_ctx *ctx, vf_graph **out) {
    if (!ctx || !out) return vf::fail(ctx, VF_E_INVALID, "vf_graph_end: null argument");
    *out = nullptr;
    if (!ctx->capturing) return vf::fail(ctx, VF_E_INVALID, "vf_graph_end: no capture in progress");
    ctx->capturing = false;
    hipGraph_t g = nullptr;
    VF_HIP(ctx, hipStreamEndCapture(ctx->stream, &g));
    vf_graph *G = new (std::nothrow) vf_graph();
    if (!G) {
        (void)hipGraphDestroy(g);
        return vf::fail(ctx, VF_E_OOM, "vf_graph_end: out of host memory");
    }
    G->device = ctx->device;
    G->graph = g;
    hipError_t e = hipGraphInstantiate(&G->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        delete G;
        return vf::fail(ctx, VF_E_DEVICE, "vf_graph_end: hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    *out = G;
    return VF_OK;
}

int vf_graph_launch(vf_ctx *ctx, vf_graph *g) {
    if (!ctx || !g) return vf::fail(ctx, VF_E_INVALID, "vf_graph_launch: null argument");
    if (ctx->capturing) return vf::fail(ctx, VF_E_INVALID, "vf_graph_launch: capture in progress");
    if (g->device != ctx->device) return vf::fail(ctx, VF_E_INVALID, "vf_graph_launch: graph of another device");
    int rc = set_device(ctx);
    if (rc) return rc;
    if ((rc = begin_timed(ctx))) return rc;
    VF_HIP(ctx, hipGraphLaunch(g->exec, ctx->stream));
    return end_timed(ctx);
}

void vf_graph_destroy(vf_graph *g) {
    if (!g) return;
    (void)hipSetDevice(g->device);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
}

int vf_build_brickdist(vf_ctx *ctx, const int32_t lo[3], const int32_t hi[3], const int32_t dlo[3], const int32_t dhi[3]) {
    int rc = check_ready(ctx, "vf_build_brickdist");
    if (rc) return rc;
    if (!lo || !hi || !dlo || !dhi) return vf::fail(ctx, VF_E_INVALID, "vf_build_brickdist: null box");
    for (int a = 0; a < 3; ++a)
        if (hi[a] - lo[a] > (1 << 20) || dhi[a] - dlo[a] > (1 << 20))
            return vf::fail(ctx, VF_E_INVALID, "vf_build_brickdist: box too large");
    if ((rc = begin_timed(ctx))) return rc;   // the timed region: the distance-field passes
    if ((rc = distance_field(ctx, lo, hi, dlo, dhi))) return rc;
    if ((rc = end_timed(ctx))) return rc;
    return repack(ctx);
}

int vf_read_brickdist(vf_ctx *ctx, uint8_t *dist, uint8_t *dry) {
    int rc = check_ready(ctx, "vf_read_brickdist");
    if (rc) return rc;
    if ((rc = no_capture(ctx, "vf_read_brickdist"))) return rc;
    if (dist) VF_HIP(ctx, hipMemcpyAsync(dist, ctx->dist[0], ctx->n_cells, hipMemcpyDeviceToHost, ctx->stream));
    if (dry) VF_HIP(ctx, hipMemcpyAsync(dry, ctx->dist[2], ctx->n_cells, hipMemcpyDeviceToHost, ctx->stream));
    VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VF_OK;
}

int vf_read_skytop(vf_ctx *ctx, uint16_t *sky) {
    int rc = check_ready(ctx, "vf_read_skytop");
    if (rc) return rc;
    if ((rc = no_capture(ctx, "vf_read_skytop"))) return rc;
    if (!sky) return vf::fail(ctx, VF_E_INVALID, "vf_read_skytop: null output");
    VF_HIP(ctx, hipMemcpyAsync(sky, ctx->skytop, (size_t)ctx->W.wb[0] * 8 * ctx->W.wb[2] * 8 * 2, hipMemcpyDeviceToHost, ctx->stream));
    VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VF_OK;
}

int vf_set_scene(vf_ctx *ctx, const vf_scene_ubo *s) {
    if (!ctx || !s) return vf::fail(ctx, VF_E_INVALID, "vf_set_scene: null argument");
    if (ctx->has_world) {
        for (int a = 0; a < 3; ++a)
            if (s->world_bricks[a] != ctx->W.wb[a])
                return vf::fail(ctx, VF_E_INVALID, "vf_set_scene: world_bricks differ from the uploaded world");
        if ((s->world_win[3] != 0) != (ctx->W.ring != 0))
            return vf::fail(ctx, VF_E_INVALID, "vf_set_scene: world_win.w (ring mode) differs from the uploaded world");
        if (ctx->W.ring) { ctx->W.minb[0] = s->world_win[0]; ctx->W.minb[2] = s->world_win[2]; }
    }
    if (s->num_objects < 0) return vf::fail(ctx, VF_E_INVALID, "vf_set_scene: num_objects = %d", s->num_objects);
    ctx->scene = *s;
    SkyParams &S = ctx->S;
    for (int a = 0; a < 3; ++a) {
        S.sun_dir[a] = s->sun_dir[a];
        S.sun_rgb[a] = s->sun_rgb[a];
        S.sky_zenith[a] = s->sky_zenith[a];
        S.sky_horizon[a] = s->sky_horizon[a];
        S.sunset_rgb[a] = s->sunset_rgb[a];
    }
    S.sky_scale = s->sky_scale;
    S.gi_scale = s->gi_scale;
    S.night = s->night;
    S.sky_model = s->sky_model;
    ctx->has_scene = true;
    return VF_OK;
}

int vf_set_counting(vf_ctx *ctx, int32_t on) {
    if (!ctx) return vf::fail(nullptr, VF_E_INVALID, "vf_set_counting: null context");
    ctx->counting = on != 0;
    return VF_OK;
}

int vf_read_counters(vf_ctx *ctx, uint64_t out[4], int32_t reset) {
    if (!ctx || !out) return vf::fail(ctx, VF_E_INVALID, "vf_read_counters: null argument");
    if (int rc = no_capture(ctx, "vf_read_counters")) return rc;
    int rc = set_device(ctx);
    if (rc) return rc;
    unsigned long long h[4];
    VF_HIP(ctx, hipMemcpyAsync(h, ctx->d_cnt, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int i = 0; i < 4; ++i) out[i] = h[i];
    if (reset) VF_HIP(ctx, hipMemsetAsync(ctx->d_cnt, 0, sizeof(h), ctx->stream));
    return VF_OK;
}

int vf_set_traversal(vf_ctx *ctx, int32_t mode) {
    if (!ctx) return vf::fail(ctx, VF_E_INVALID, "vf_set_traversal: null context");
    if (mode < 0 || mode > 2) return vf::fail(ctx, VF_E_INVALID, "vf_set_traversal: mode %d is not 0, 1 or 2", mode);
    ctx->traversal = mode;
    return VF_OK;
}

int vf_set_timing(vf_ctx *ctx, int32_t on) {
    if (!ctx) return vf::fail(ctx, VF_E_INVALID, "vf_set_timing: null context");
    ctx->timing = on != 0;
    if (!ctx->timing) ctx->timed = false;
    return VF_OK;
}

int vf_last_kernel_ms(vf_ctx *ctx, float *ms) {
    if (!ctx || !ms) return vf::fail(ctx, VF_E_INVALID, "vf_last_kernel_ms: null argument");
    if (int rc = no_capture(ctx, "vf_last_kernel_ms")) return rc;
    if (!ctx->timing) return vf::fail(ctx, VF_E_INVALID, "vf_last_kernel_ms: timing is off (vf_set_timing)");
    if (!ctx->timed) return vf::fail(ctx, VF_E_INVALID, "vf_last_kernel_ms: nothing timed yet");
    int rc = set_device(ctx);
    if (rc) return rc;
    VF_HIP(ctx, hipEventSynchronize(ctx->ev1));
    VF_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return VF_OK;
}

int vf_trace_primary(vf_ctx *ctx, const vf_pass_ubo *p, int32_t w, int32_t h, vf_hit *hits, float *gb_pos, uint32_t *gb_normal) {
    if (ctx && !hits) return vf::fail(ctx, VF_E_INVALID, "vf_trace_primary: null hits");
    return trace_gbuffer(ctx, "vf_trace_primary", p, w, h, hits, gb_pos, gb_normal, nullptr, nullptr, nullptr);
}

int vf_trace_gbuffer(vf_ctx *ctx, const vf_pass_ubo *p, int32_t w, int32_t h, vf_hit *hits, float *gb_pos, uint32_t *gb_normal,
                     int32_t *body, uint32_t *shadow, vf_hit *shadow_hits) {
    if (ctx && !hits) return vf::fail(ctx, VF_E_INVALID, "vf_trace_gbuffer: null hits");
    return trace_gbuffer(ctx, "vf_trace_gbuffer", p, w, h, hits, gb_pos, gb_normal, body, shadow, shadow_hits);
}

int vf_trace_shadow(vf_ctx *ctx, const vf_pass_ubo *p, int32_t w, int32_t h, const float *gb_pos, const uint32_t *gb_normal,
                    uint32_t *shadow, vf_hit *hits) {
    int rc = check_ready(ctx, "vf_trace_shadow");
    if (rc) return rc;
    if (!ctx->has_scene) return vf::fail(ctx, VF_E_INVALID, "vf_trace_shadow: no scene (vf_set_scene)");
    if (!p || !gb_pos || !gb_normal || !shadow) return vf::fail(ctx, VF_E_INVALID, "vf_trace_shadow: null pass or buffer");
    if (w <= 0 || h <= 0 || (int64_t)w * h > (1ll << 30)) return vf::fail(ctx, VF_E_INVALID, "vf_trace_shadow: bad size %dx%d", w, h);
    if ((rc = vf::check_bodies(ctx, "vf_trace_shadow"))) return rc;
    TileArgs A{w, h, (w + 15) / 16, 0, p->look_p[3]};
    A.ntiles = A.tiles_x * ((h + 15) / 16);
    const float *s = ctx->S.sun_dir;
    const DevBodies B = vf::dev_bodies(ctx);
    const float4 *gp = reinterpret_cast<const float4 *>(gb_pos);
    const DirPre sp = dir_pre(s);   // the sun ray's set-up, as trace_world computes it
    if ((rc = begin_timed(ctx))) return rc;
    if (B.n > 0) {
        if (ctx->counting)
            hipLaunchKernelGGL((k_shadow<true, true>), dim3(A.ntiles), dim3(256), 0, ctx->stream, ctx->W, B, A, s[0], s[1], s[2],
                               gp, gb_normal, shadow, hits, ctx->d_cnt, sp);
        else
            hipLaunchKernelGGL((k_shadow<false, true>), dim3(A.ntiles), dim3(256), 0, ctx->stream, ctx->W, B, A, s[0], s[1], s[2],
                               gp, gb_normal, shadow, hits, ctx->d_cnt, sp);
    } else if (ctx->counting) {
        hipLaunchKernelGGL((k_shadow<true, false>), dim3(A.ntiles), dim3(256), 0, ctx->stream, ctx->W, B, A, s[0], s[1], s[2], gp,
                           gb_normal, shadow, hits, ctx->d_cnt, sp);
    } else {
        hipLaunchKernelGGL((k_shadow<false, false>), dim3(A.ntiles), dim3(256), 0, ctx->stream, ctx->W, B, A, s[0], s[1], s[2], gp,
                           gb_normal, shadow, hits, ctx->d_cnt, sp);
    }
    if ((rc = launch_check(ctx, "k_shadow"))) return rc;
    return end_timed(ctx);
}

int vf_trace_primary_shadow(vf_ctx *ctx, const vf_pass_ubo *p, int32_t w, int32_t h, vf_hit *hits, float *gb_pos,
                            uint32_t *gb_normal, uint32_t *shadow, vf_hit *shadow_hits) {
    if (ctx && (!hits || !shadow)) return vf::fail(ctx, VF_E_INVALID, "vf_trace_primary_shadow: null hits or shadow");
    return trace_gbuffer(ctx, "vf_trace_primary_shadow", p, w, h, hits, gb_pos, gb_normal, nullptr, shadow, shadow_hits);
}

int vf_trace_sun_cone(vf_ctx *ctx, const vf_pass_ubo *p, int32_t w, int32_t h, const float *gb_pos, const uint32_t *gb_normal,
                      float *shadow, vf_hit *hits) {
    int rc = check_ready(ctx, "vf_trace_sun_cone");
    if (rc) return rc;
    if (!ctx->has_scene) return vf::fail(ctx, VF_E_INVALID, "vf_trace_sun_cone: no scene (vf_set_scene)");
    if (!p || !gb_pos || !gb_normal || !shadow) return vf::fail(ctx, VF_E_INVALID, "vf_trace_sun_cone: null pass or buffer");
    if (w <= 0 || h <= 0 || (int64_t)w * h > (1ll << 30)) return vf::fail(ctx, VF_E_INVALID, "vf_trace_sun_cone: bad size %dx%d", w, h);
    TileArgs A{w, h, (w + 15) / 16, 0, p->look_p[3]};
    A.ntiles = A.tiles_x * ((h + 15) / 16);
    const float *s = ctx->S.sun_dir;
    const float ang = ctx->scene.sun_angle;
    if ((rc = vf::check_bodies(ctx, "vf_trace_sun_cone"))) return rc;
    const DevBodies B = vf::dev_bodies(ctx);
    const float4 *gp = reinterpret_cast<const float4 *>(gb_pos);
    if ((rc = begin_timed(ctx))) return rc;
    if (B.n > 0) {
        if (ctx->counting)
            hipLaunchKernelGGL((k_sun_cone<true, true>), dim3(A.ntiles), dim3(256), 0, ctx->stream, ctx->W, B, A, s[0], s[1], s[2],
                               ang, p->frame_idx, gp, gb_normal, shadow, hits, ctx->d_cnt);
        else
            hipLaunchKernelGGL((k_sun_cone<false, true>), dim3(A.ntiles), dim3(256), 0, ctx->stream, ctx->W, B, A, s[0], s[1], s[2],
                               ang, p->frame_idx, gp, gb_normal, shadow, hits, ctx->d_cnt);
    } else if (ctx->counting) {
        hipLaunchKernelGGL((k_sun_cone<true, false>), dim3(A.ntiles), dim3(256), 0, ctx->stream, ctx->W, B, A, s[0], s[1], s[2], ang,
                           p->frame_idx, gp, gb_normal, shadow, hits, ctx->d_cnt);
    } else {
        hipLaunchKernelGGL((k_sun_cone<false, false>), dim3(A.ntiles), dim3(256), 0, ctx->stream, ctx->W, B, A, s[0], s[1], s[2],
                           ang, p->frame_idx, gp, gb_normal, shadow, hits, ctx->d_cnt);
    }
    if ((rc = launch_check(ctx, "k_sun_cone"))) return rc;
    return end_timed(ctx);
}

int vf_trace_gi(vf_ctx *ctx, const vf_pass_ubo *p, int32_t w, int32_t h, int32_t half_res, float gi_reach, int32_t lights_nee,
                const float *gb_pos, const uint32_t *gb_normal, float *radiance, vf_hit *bounce) {
    int rc = check_ready(ctx, "vf_trace_gi");
    if (rc) return rc;
    if (!ctx->has_scene) return vf::fail(ctx, VF_E_INVALID, "vf_trace_gi: no scene (vf_set_scene)");
    if (!p || !gb_pos || !gb_normal || !radiance) return vf::fail(ctx, VF_E_INVALID, "vf_trace_gi: null buffer");
    if (w <= 0 || h <= 0 || (int64_t)w * h > (1ll << 30)) return vf::fail(ctx, VF_E_INVALID, "vf_trace_gi: bad size %dx%d", w, h);
    if (!(gi_reach > 0.0f) || gi_reach > 1e6f) return vf::fail(ctx, VF_E_INVALID, "vf_trace_gi: gi_reach %g out of range", gi_reach);
    GiArgs A;
    A.w = w; A.h = h;
    A.half_res = half_res ? 1 : 0;
    A.ow = half_res ? (w + 1) / 2 : w;
    A.oh = half_res ? (h + 1) / 2 : h;
    A.tiles_x = (A.ow + 15) / 16;
    A.ntiles = A.tiles_x * ((A.oh + 15) / 16);
    A.lights_nee = lights_nee ? 1 : 0;
    A.frame = p->frame_idx;
    A.reach = gi_reach;
    if ((rc = vf::check_bodies(ctx, "vf_trace_gi"))) return rc;
    const DevBodies B = vf::dev_bodies(ctx);
    const float4 *gp = reinterpret_cast<const float4 *>(gb_pos);
    float4 *rad = reinterpret_cast<float4 *>(radiance);
    if ((rc = begin_timed(ctx))) return rc;
    if (B.n > 0) {
        if (ctx->counting)
            hipLaunchKernelGGL((k_gi<true, true>), dim3(A.ntiles), dim3(256), 0, ctx->stream, ctx->W, ctx->S, B, A, gp, gb_normal,
                               rad, bounce, ctx->d_cnt);
        else
            hipLaunchKernelGGL((k_gi<false, true>), dim3(A.ntiles), dim3(256), 0, ctx->stream, ctx->W, ctx->S, B, A, gp, gb_normal,
                               rad, bounce, ctx->d_cnt);
    } else if (ctx->counting) {
        hipLaunchKernelGGL((k_gi<true, false>), dim3(A.ntiles), dim3(256), 0, ctx->stream, ctx->W, ctx->S, B, A, gp, gb_normal, rad,
                           bounce, ctx->d_cnt);
    } else {
        hipLaunchKernelGGL((k_gi<false, false>), dim3(A.ntiles), dim3(256), 0, ctx->stream, ctx->W, ctx->S, B, A, gp, gb_normal,
                           rad, bounce, ctx->d_cnt);
    }
    if ((rc = launch_check(ctx, "k_gi"))) return rc;
    return end_timed(ctx);
}

int vf_trace_rays(vf_ctx *ctx, const vf_ray *rays, int64_t n, float tcap, uint32_t skip_mat, vf_hit *hits) {
    int rc = check_ready(ctx, "vf_trace_rays");
    if (rc) return rc;
    if (n < 0 || n > (1ll << 31) * 64) return vf::fail(ctx, VF_E_INVALID, "vf_trace_rays: bad count %lld", (long long)n);
    if (n == 0) return VF_OK;
    if (!rays || !hits) return vf::fail(ctx, VF_E_INVALID, "vf_trace_rays: null buffer");
    if (skip_mat != 0u && skip_mat != 5u) return vf::fail(ctx, VF_E_INVALID, "vf_trace_rays: skip_mat must be 0 or 5");
    const unsigned nb = blocks_for(n);
    if ((rc = begin_timed(ctx))) return rc;
    const bool box = vf::small_frame(ctx, (n + 63) / 64);
    if (skip_mat == 5u) {
        if (ctx->counting)
            hipLaunchKernelGGL((k_rays<5, true>), dim3(nb), dim3(256), 0, ctx->stream, ctx->W, rays, (long long)n, tcap, hits, ctx->d_cnt);
        else if (box)
            hipLaunchKernelGGL((k_rays<5, false, true>), dim3(nb), dim3(256), 0, ctx->stream, ctx->W, rays, (long long)n, tcap, hits, ctx->d_cnt);
        else
            hipLaunchKernelGGL((k_rays<5, false>), dim3(nb), dim3(256), 0, ctx->stream, ctx->W, rays, (long long)n, tcap, hits, ctx->d_cnt);
    } else {
        if (ctx->counting)
            hipLaunchKernelGGL((k_rays<0, true>), dim3(nb), dim3(256), 0, ctx->stream, ctx->W, rays, (long long)n, tcap, hits, ctx->d_cnt);
        else if (box)
            hipLaunchKernelGGL((k_rays<0, false, true>), dim3(nb), dim3(256), 0, ctx->stream, ctx->W, rays, (long long)n, tcap, hits, ctx->d_cnt);
        else
            hipLaunchKernelGGL((k_rays<0, false>), dim3(nb), dim3(256), 0, ctx->stream, ctx->W, rays, (long long)n, tcap, hits, ctx->d_cnt);
    }
    if ((rc = launch_check(ctx, "k_rays"))) return rc;
    return end_timed(ctx);
}

int vf_trace_probes(vf_ctx *ctx, const float origin[3], float spacing, const int32_t dims[3], int32_t n_rays, uint32_t frame,
                    float reach, vf_hit *hits) {
    int rc = check_ready(ctx, "vf_trace_probes");
    if (rc) return rc;
    if (!origin || !dims || !hits) return vf::fail(ctx, VF_E_INVALID, "vf_trace_probes: null argument");
    if (n_rays <= 0 || n_rays > 65536 || dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0)
        return vf::fail(ctx, VF_E_INVALID, "vf_trace_probes: bad grid or ray count");
    ProbeArgs A;
    for (int a = 0; a < 3; ++a) { A.origin[a] = origin[a]; A.dims[a] = dims[a]; }
    A.spacing = spacing;
    A.n_rays = n_rays;
    A.frame = frame;
    A.reach = reach;
    A.n = (long long)dims[0] * dims[1] * dims[2] * n_rays;
    if (A.n > (1ll << 31) * 64) return vf::fail(ctx, VF_E_INVALID, "vf_trace_probes: too many rays");
    if ((rc = vf::check_bodies(ctx, "vf_trace_probes"))) return rc;
    const DevBodies B = vf::dev_bodies(ctx);
    const unsigned nb = blocks_for(A.n);
    if ((rc = begin_timed(ctx))) return rc;
    if (B.n > 0) {
        if (ctx->counting)
            hipLaunchKernelGGL((k_probes<true, true>), dim3(nb), dim3(256), 0, ctx->stream, ctx->W, B, A, hits, ctx->d_cnt);
        else
            hipLaunchKernelGGL((k_probes<false, true>), dim3(nb), dim3(256), 0, ctx->stream, ctx->W, B, A, hits, ctx->d_cnt);
    } else if (ctx->counting) {
        hipLaunchKernelGGL((k_probes<true, false>), dim3(nb), dim3(256), 0, ctx->stream, ctx->W, B, A, hits, ctx->d_cnt);
    } else {
        hipLaunchKernelGGL((k_probes<false, false>), dim3(nb), dim3(256), 0, ctx->stream, ctx->W, B, A, hits, ctx->d_cnt);
    }
    if ((rc = launch_check(ctx, "k_probes"))) return rc;
    return end_timed(ctx);
}

}  // extern "C"
