// Device-side bodies (dynamic voxel objects): traceObject / traceObjectsA / traceSceneS (reference
// shaders/common.glsl:664-720, 787-793).  A body is a brick grid of its own, in the world's atlas, with a
// rigid pose; a ray meets it in the body's voxel space.  Bodies are found through a BVH over their world
// boxes, walked in the reference's order (left child first) without a stack: each node carries the node
// the reference's stack would pop once the node's subtree is done.
#pragma once

#include "vf_device.h"

struct DevBodies {
    int32_t n;              // 0: no bodies (every body function returns at once)
    const float4 *hdr;      // 8 float4 per body, see body_header() on the host
    const uint32_t *grid;   // every body's brick grid
    const float4 *nodes;    // 2 float4 per BVH node: (min.xyz, first bits), (max.xyz, count bits)
    const int32_t *esc;     // per node: the next node after its subtree (-1: the walk ends)
    const uint32_t *prims;  // leaf ranges -> body index
};

// Per body, 8 float4 (128 B):
//   [0..2] rows of worldToObj's top 3x4: (m[0][r], m[1][r], m[2][r], m[3][r]) for r = 0, 1, 2
//   [3..5] columns 0..2 of objToWorld's 3x3 (.xyz), .w = bricks x / y / z (int bits)
//   [6]    world box min (.xyz), .w = offset of the body's grid (int bits)
//   [7]    world box max (.xyz)
enum { kBodyWords = 8 };

// World normal of a body hit: normalize(mat3(objToWorld) * n) with n the body-space face normal
// (common.glsl:678).  The product is ((c0 * n.x + c1 * n.y) + c2 * n.z) per component.  A hit with no
// face (the ray started inside the body's box) has n = 0: the result is NaN, as normalize(0) is.
__device__ __forceinline__ void body_normal(const DevBodies &B, int32_t oi, int32_t face, float out[3]) {
    const float4 *hd = B.hdr + (size_t)oi * kBodyWords;
    float n[3];
    face_normal(face, n);
    const float4 c0 = hd[3], c1 = hd[4], c2 = hd[5];
    out[0] = (c0.x * n[0] + c1.x * n[1]) + c2.x * n[2];
    out[1] = (c0.y * n[0] + c1.y * n[1]) + c2.y * n[2];
    out[2] = (c0.z * n[0] + c1.z * n[1]) + c2.z * n[2];
    normalize3(out);
}

// The hit's world normal: the face normal for the world, the rotated one for a body.
template <bool BODIES>
__device__ __forceinline__ void hit_normal(const DevBodies *B, const Hit &h, float n[3]) {
    if (BODIES && h.obj >= 0) body_normal(*B, h.obj, h.face, n);
    else face_normal(h.face, n);
}

// RGBA8 unorm of n * 0.5 + 0.5 with alpha byte `a` (gbuffer.comp's matFlag); NaN components store 0 (Vulkan's
// float -> unorm rule).
__device__ __forceinline__ uint32_t pack_normal(const float n[3], uint32_t a = 255u) {
    uint32_t v = a << 24;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float x = n[a] * 0.5f + 0.5f;
        v |= (x == x ? unorm8(x) : 0u) << (8 * a);
    }
    return v;
}

// traceVolume over body oi (common.glsl:541-658 with useTex false, skipMat kSkipMat = 5): the box
// [0, dims*8), the body's own grid, plain brick steps (no distance field or coarse skips, no water
// test); detailed bricks are the world atlas's, walked by brick_voxels (LDSM: with the brick's mask in the
// lane's LDS slots, in 256-thread kernels).  ro / rd in body space.
template <bool COUNT, bool LDSM = false>
__device__ bool trace_body(const DevWorld &W, const DevBodies &B, int32_t oi, const float ro[3], const float rd[3],
                           float tcap, Hit &h, Cnt &c) {
    const float4 *hd = B.hdr + (size_t)oi * kBodyWords;
    const int32_t db[3] = {__float_as_int(hd[3].w), __float_as_int(hd[4].w), __float_as_int(hd[5].w)};
    const uint32_t goff = __float_as_uint(hd[6].w);
    h.hit = 0; h.t = -1.0f; h.face = 0; h.mat = 0; h.cell[0] = h.cell[1] = h.cell[2] = 0; h.obj = -1;
    float sd[3], inv[3], up[3], tmin3[3], tmax3[3];
    int32_t st[3], faceA[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        sd[a] = fabsf(rd[a]) < 1e-5f ? 1e-5f : rd[a];
        inv[a] = 1.0f / sd[a];
        const float lo = 0.0f;
        const float dims = (float)(db[a] * 8);
        const float t0 = (lo - ro[a]) * inv[a];
        const float t1 = ((lo + dims) - ro[a]) * inv[a];
        tmin3[a] = gmin(t0, t1);
        tmax3[a] = gmax(t0, t1);
        st[a] = sd[a] > 0.0f ? 1 : -1;
        up[a] = sd[a] < 0.0f ? 0.0f : 1.0f;
        faceA[a] = 2 * a + (st[a] > 0 ? 2 : 1);
    }
    const float tEnter = gmax(gmax(tmin3[0], tmin3[1]), tmin3[2]);
    const float tExit = gmin(gmin(tmax3[0], tmax3[1]), tmax3[2]);
    if (tExit < gmax(tEnter, 0.0f) || tEnter > tcap) return false;
    float t = gmax(tEnter, 0.0f);
    int32_t bc[3];
    float tD8[3], tM8[3];
    {
        const float e = t + 1e-3f;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float entry = ro[a] + sd[a] * e;
            bc[a] = iclamp(f2i(floorf(entry * 0.125f)), 0, db[a] - 1);
            tD8[a] = fabsf(inv[a]) * 8.0f;
            tM8[a] = (((float)bc[a] + up[a]) * 8.0f - ro[a]) * inv[a];
        }
    }
    int32_t face = 0;
    if (tEnter > 0.0f) {
        if (tEnter == tmin3[0]) face = faceA[0];
        else if (tEnter == tmin3[1]) face = faceA[1];
        else face = faceA[2];
    }
    const int32_t max_outer = 3 * (db[0] + db[1] + db[2]) + 4;
    for (int32_t i = 0; i < max_outer; ++i) {
        if (t > tcap) return false;
        const uint32_t g = B.grid[goff + (uint32_t)((bc[2] * db[1] + bc[1]) * db[0] + bc[0])];
        if (COUNT) { c.outer++; c.bytes += 4; }
        if (g >= 256u) {
            if (brick_voxels<COUNT, LDSM>(W, ro, sd, inv, st, up, faceA, bc, g, t, face, h, c, 5)) return true;
        } else if (g != 0u && g != 5u) {
            const float e = t + 1e-3f;
            h.hit = 1; h.t = t; h.face = face; h.mat = g;
#pragma unroll
            for (int a = 0; a < 3; ++a) h.cell[a] = f2i(floorf(ro[a] + sd[a] * e));
            return true;
        }
        const bool a0 = tM8[0] < tM8[1] && tM8[0] < tM8[2];
        const bool a1 = !a0 && tM8[1] < tM8[2];
        const bool a2 = !a0 && !a1;
        t = a0 ? tM8[0] : (a1 ? tM8[1] : tM8[2]);
        face = a0 ? faceA[0] : (a1 ? faceA[1] : faceA[2]);
        const float n0 = tM8[0] + tD8[0], n1 = tM8[1] + tD8[1], n2 = tM8[2] + tD8[2];
        tM8[0] = a0 ? n0 : tM8[0];
        tM8[1] = a1 ? n1 : tM8[1];
        tM8[2] = a2 ? n2 : tM8[2];
        bc[0] += a0 ? st[0] : 0;
        bc[1] += a1 ? st[1] : 0;
        bc[2] += a2 ? st[2] : 0;
        if (bc[0] < 0 || bc[1] < 0 || bc[2] < 0 || bc[0] >= db[0] || bc[1] >= db[1] || bc[2] >= db[2]) return false;
    }
    return false;
}

// Slab test of a world box against the ray (common.glsl:666-671, 701-705).
__device__ __forceinline__ void box_span(const float4 lo, const float4 hi, const float ro[3], const float inv[3],
                                         float &tEnter, float &tExit) {
    const float x0 = (lo.x - ro[0]) * inv[0], x1 = (hi.x - ro[0]) * inv[0];
    const float y0 = (lo.y - ro[1]) * inv[1], y1 = (hi.y - ro[1]) * inv[1];
    const float z0 = (lo.z - ro[2]) * inv[2], z1 = (hi.z - ro[2]) * inv[2];
    tEnter = gmax(gmax(gmin(x0, x1), gmin(y0, y1)), gmin(z0, z1));
    tExit = gmin(gmin(gmax(x0, x1), gmax(y0, y1)), gmax(z0, z1));
}

// traceObject (common.glsl:665-683): the body's world box, then its volume in body space; a hit closer
// than limit becomes the best hit (body-space cell and face, obj = oi) and the new limit.
template <bool COUNT, bool LDSM = false>
__device__ __forceinline__ void trace_object(const DevWorld &W, const DevBodies &B, int32_t oi, const float ro[3],
                                             const float rd[3], const float inv[3], Hit &best, float &limit, Cnt &c) {
    const float4 *hd = B.hdr + (size_t)oi * kBodyWords;
    float tEnter, tExit;
    box_span(hd[6], hd[7], ro, inv, tEnter, tExit);
    if (tExit < gmax(tEnter, 0.0f) || tEnter > limit) return;
    const float4 r0 = hd[0], r1 = hd[1], r2 = hd[2];
    const float roO[3] = {((r0.x * ro[0] + r0.y * ro[1]) + r0.z * ro[2]) + r0.w,
                          ((r1.x * ro[0] + r1.y * ro[1]) + r1.z * ro[2]) + r1.w,
                          ((r2.x * ro[0] + r2.y * ro[1]) + r2.z * ro[2]) + r2.w};
    const float rdO[3] = {(r0.x * rd[0] + r0.y * rd[1]) + r0.z * rd[2],
                          (r1.x * rd[0] + r1.y * rd[1]) + r1.z * rd[2],
                          (r2.x * rd[0] + r2.y * rd[1]) + r2.z * rd[2]};
    Hit oh;
    trace_body<COUNT, LDSM>(W, B, oi, roO, rdO, limit, oh, c);
    if (oh.hit && oh.t < limit) {
        best = oh;
        best.obj = oi;
        limit = oh.t;
    }
}

// traceObjectsA (common.glsl:689-720, USE_OBJ_BVH): the nearest body hit closer than limit merged into
// best; any_hit returns at the first body hit.  The walk visits the nodes in the reference's order
// (a visited interior node continues with its first child; a leaf or a culled node continues with the
// node its stack would pop) and stops after 512 nodes as the reference's loop does.
template <bool COUNT, bool LDSM = false>
__device__ void trace_objects(const DevWorld &W, const DevBodies &B, const float ro[3], const float rd[3], Hit &best,
                              float limit, bool any_hit, Cnt &c) {
    if (B.n == 0) return;
    float inv[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) inv[a] = 1.0f / (fabsf(rd[a]) < 1e-5f ? 1e-5f : rd[a]);
    int32_t node = 0;
    for (int it = 0; it < 512; ++it) {
        const float4 na = B.nodes[2 * node], nb = B.nodes[2 * node + 1];
        float tEnter, tExit;
        box_span(na, nb, ro, inv, tEnter, tExit);
        const bool visit = tExit >= gmax(tEnter, 0.0f) && tEnter <= limit;
        const uint32_t first = __float_as_uint(na.w), count = __float_as_uint(nb.w);
        if (visit && count == 0u) {
            node = (int32_t)first;
            continue;
        }
        if (visit)
            for (uint32_t k = 0; k < count; ++k)
                trace_object<COUNT, LDSM>(W, B, (int32_t)B.prims[first + k], ro, rd, inv, best, limit, c);
        if (any_hit && best.hit) return;
        node = B.esc[node];
        if (node < 0) return;
    }
}

// traceSceneS (common.glsl:787-793) after the world's part: h is the world hit (tCap when it missed).
template <bool COUNT, bool LDSM = false>
__device__ __forceinline__ void add_bodies(const DevWorld &W, const DevBodies &B, const float ro[3], const float rd[3],
                                           float tcap, Hit &h, Cnt &c) {
    h.obj = -1;
    trace_objects<COUNT, LDSM>(W, B, ro, rd, h, h.hit ? h.t : tcap, false, c);
}

// gbuffer.comp's primary ray over the scene (gbuffer.comp:233-266 with traceScenePrimary = traceSceneS):
// a camera inside solid terrain starts the ray where it leaves the solid, the bodies are traced from the
// same origin (closer than the world hit), t counts from the camera, and a ray that only leaves the solid
// out of the world is a miss; a camera in water traces the world with skipMat 5.
template <bool COUNT, bool LDSM = false>
__device__ __forceinline__ void scene_primary_trace(const DevWorld &W, const DevBodies &B, const float cam[3],
                                                    uint32_t cam_mat, const float rd[3], Hit &h, Cnt &c) {
    float ro[3] = {cam[0], cam[1], cam[2]};
    float tskip = 0.0f;
    bool void_ray = false;
    if (cam_mat != 0u && cam_mat != 5u) {
        tskip = exit_solid(W, cam, rd, 2048);
        void_ray = tskip < 0.0f;
        tskip = gmax(tskip, 0.0f);
#pragma unroll
        for (int a = 0; a < 3; ++a) ro[a] = cam[a] + rd[a] * tskip;
    }
    trace_world<-1, COUNT, LDSM>(W, ro, rd, 1e30f, h, c, cam_mat == 5u ? 5 : 0);
    add_bodies<COUNT, LDSM>(W, B, ro, rd, 1e30f, h, c);
    h.t = h.t + tskip;
    if (void_ray) h.hit = 0;
}
