/*
 * CPU oracle for the brick-DDA voxel traversal — TEST INFRASTRUCTURE ONLY (see vf_oracle.h).
 * Built with -O2 -ffp-contract=off -fno-fast-math: every FP32 operation below is rounded
 * once, in the order written, which is the order of the reference GLSL it restates.
 * Reference citations are relative to the reference repository root.
 */
#include "vf_oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- GLSL built-ins with their specified definitions (GLSL 4.60 §8.3) ---- */
static inline float gmin(float x, float y) { return y < x ? y : x; }
static inline float gmax(float x, float y) { return x < y ? y : x; }
static inline int32_t imin(int32_t x, int32_t y) { return y < x ? y : x; }
static inline int32_t imax(int32_t x, int32_t y) { return x < y ? y : x; }
static inline int32_t iclamp(int32_t x, int32_t lo, int32_t hi) { return imin(imax(x, lo), hi); }
static inline float fclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
/* ivec3(float): truncation; saturating and NaN -> 0 like the GPU's conversion instruction. */
static inline int32_t f2i(float x) {
    if (x != x) return 0;
    if (x >= 2147483648.0f) return INT32_MAX;
    if (x <= -2147483648.0f) return INT32_MIN;
    return (int32_t)x;
}
static inline float smoothstepf(float e0, float e1, float x) {
    float t = fclamp((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
static inline float dot3(const float a[3], const float b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
static inline void cross3(const float a[3], const float b[3], float o[3]) {
    float x = a[1] * b[2] - b[1] * a[2];
    float y = a[2] * b[0] - b[2] * a[0];
    float z = a[0] * b[1] - b[0] * a[1];
    o[0] = x; o[1] = y; o[2] = z;
}
/* normalize(v) = v / length(v), IEEE sqrt and division. */
static inline void normalize3(float v[3]) {
    float len = sqrtf(dot3(v, v));
    v[0] = v[0] / len; v[1] = v[1] / len; v[2] = v[2] / len;
}

/* ---- world addressing (common.glsl:134-137 gridWrap / worldMinB) ---- */
static inline int is_ring(const vfo_world *w) { return w->win[3] != 0; }
static inline size_t gidx(const vfo_world *w, int32_t x, int32_t y, int32_t z) {
    if (is_ring(w)) { x &= w->wb[0] - 1; z &= w->wb[2] - 1; }
    return ((size_t)z * (size_t)w->wb[1] + (size_t)y) * (size_t)w->wb[0] + (size_t)x;
}

int32_t vfo_face_code(const int32_t n[3]) {
    if (n[0]) return n[0] > 0 ? 1 : 2;
    if (n[1]) return n[1] > 0 ? 3 : 4;
    if (n[2]) return n[2] > 0 ? 5 : 6;
    return 0;
}

uint32_t vfo_voxel(const vfo_world *w, int32_t x, int32_t y, int32_t z) {
    int32_t mn[3] = {is_ring(w) ? w->win[0] : 0, 0, is_ring(w) ? w->win[2] : 0};
    int32_t v[3] = {x, y, z};
    for (int a = 0; a < 3; ++a)
        if (v[a] < mn[a] * 8 || v[a] >= (mn[a] + w->wb[a]) * 8) return 0;
    uint32_t g = w->grid[gidx(w, x >> 3, y >> 3, z >> 3)];
    if (g == 0) return 0;
    if (g < 256) return g;
    return w->atlas[(size_t)(g - 256) * 512 + (size_t)((z & 7) * 64 + (y & 7) * 8 + (x & 7))];
}

/* ---- brickVoxels (common.glsl:474-539), USE_BRICK_MASKS off (the default, cvars.rs:108) ---- */
static int brick_voxels(const vfo_world *w, const float ro[3], const float sd[3], const float inv[3],
                        const int32_t st[3], const float up[3], const int32_t bc[3], uint32_t g, float t,
                        const int32_t normal[3], uint32_t skip, vfo_hit *h, vfo_count *cnt) {
    const uint8_t *brick = w->atlas + (size_t)(g - 256u) * 512u;
    int32_t bmin[3], vc[3], n2[3];
    float tM[3], tD[3];
    float e = t + 1e-3f;
    for (int a = 0; a < 3; ++a) {
        bmin[a] = bc[a] * 8;
        float q = ro[a] + sd[a] * e;
        vc[a] = iclamp(f2i(floorf(q)), bmin[a], bmin[a] + 7);
        tM[a] = (((float)vc[a] + up[a]) - ro[a]) * inv[a];
        tD[a] = fabsf(inv[a]);
        n2[a] = normal[a];
    }
    float tv = t;
    for (int j = 0; j < 26; ++j) {
        uint32_t m = brick[(vc[2] - bmin[2]) * 64 + (vc[1] - bmin[1]) * 8 + (vc[0] - bmin[0])];
        if (cnt) { cnt->voxel++; cnt->bytes += 1; }
        if (m != 0u && m != skip) {
            h->hit = 1; h->t = tv; h->mat = m;
            for (int a = 0; a < 3; ++a) { h->n[a] = n2[a]; h->cell[a] = vc[a]; }
            return 1;
        }
        int ax;
        if (tM[0] < tM[1] && tM[0] < tM[2]) ax = 0;
        else if (tM[1] < tM[2]) ax = 1;
        else ax = 2;
        vc[ax] += st[ax];
        tv = tM[ax];
        tM[ax] += tD[ax];
        n2[0] = n2[1] = n2[2] = 0;
        n2[ax] = -st[ax];
        if (vc[0] < bmin[0] || vc[1] < bmin[1] || vc[2] < bmin[2] ||
            vc[0] > bmin[0] + 7 || vc[1] > bmin[1] + 7 || vc[2] > bmin[2] + 7) break;
    }
    return 0;
}

/* ---- the library's small-frame mask test, restated to check that it is conservative ----
 * Before a detailed brick's walk the library's small-frame traversal tests the brick's solid voxels (voxel
 * != 0 and != skip) in the rows and columns, over all 8 layers, of the box from the walk's entry voxel to
 * the voxel at the brick DDA's exit time (least tMax), that coordinate taken `margin` voxel further along
 * the ray, and skips a walk with none.  mode 1 drops the entry voxel from the box, a planted mistake the
 * tests must catch.  Returns 1 when the walk may be skipped. */
static int box_may_skip(const vfo_world *w, const float ro[3], const float sd[3], const int32_t st[3], const int32_t bc[3],
                        uint32_t g, float t, const float tM8[3], uint32_t skip, float margin, int mode) {
    const uint8_t *brick = w->atlas + (size_t)(g - 256u) * 512u;
    const float e = t + 1e-3f;
    const float tout = gmin(gmin(tM8[0], tM8[1]), tM8[2]);
    int32_t lo[2], hi[2];
    for (int a = 0; a < 2; ++a) {
        const int32_t bmin = bc[a] * 8;
        const int32_t lv = iclamp(f2i(floorf(ro[a] + sd[a] * e)), bmin, bmin + 7) - bmin;
        const float q = (ro[a] + sd[a] * tout) - (float)bmin;
        const int32_t ev = iclamp(f2i(floorf(st[a] > 0 ? q + margin : q - margin)), 0, 7);
        lo[a] = mode == 1 ? ev : imin(lv, ev);
        hi[a] = mode == 1 ? ev : imax(lv, ev);
    }
    for (int z = 0; z < 8; ++z)
        for (int y = lo[1]; y <= hi[1]; ++y)
            for (int x = lo[0]; x <= hi[0]; ++x) {
                const uint32_t m = brick[z * 64 + y * 8 + x];
                if (m != 0u && m != skip) return 0;
            }
    return 1;
}

/* ---- traceVolumeS over the world (common.glsl:541-658), WATER_CLEAR and USE_DIST_FIELD on ----
 * box_mode < 0: the reference's traversal; else the walks box_may_skip rules out are skipped (counted in
 * box_stats[0], walks run in box_stats[1]). */
static void trace_impl(const vfo_world *w, const float ro[3], const float rd[3], float tcap, uint32_t skip,
                       vfo_hit *h, vfo_count *cnt, float box_margin, int box_mode, int64_t *box_stats);

void vfo_trace(const vfo_world *w, const float ro[3], const float rd[3], float tcap, uint32_t skip,
               vfo_hit *h, vfo_count *cnt) {
    trace_impl(w, ro, rd, tcap, skip, h, cnt, 0.0f, -1, NULL);
}

static void trace_impl(const vfo_world *w, const float ro[3], const float rd[3], float tcap, uint32_t skip,
                       vfo_hit *h, vfo_count *cnt, float box_margin, int box_mode, int64_t *box_stats) {
    memset(h, 0, sizeof(*h));
    h->t = -1.0f;
    if (cnt) cnt->rays++;
    const int ring = is_ring(w);
    const int32_t minB[3] = {ring ? w->win[0] : 0, 0, ring ? w->win[2] : 0};
    float sd[3], inv[3], t0[3], t1[3], tmin3[3], tmax3[3], up[3];
    int32_t st[3];
    for (int a = 0; a < 3; ++a) {
        sd[a] = fabsf(rd[a]) < 1e-5f ? 1e-5f : rd[a];
        inv[a] = 1.0f / sd[a];
        float lo = (float)(minB[a] * 8);
        float dims = (float)(w->wb[a] * 8);
        t0[a] = (lo - ro[a]) * inv[a];
        t1[a] = ((lo + dims) - ro[a]) * inv[a];
        tmin3[a] = gmin(t0[a], t1[a]);
        tmax3[a] = gmax(t0[a], t1[a]);
        st[a] = sd[a] > 0.0f ? 1 : -1;
        up[a] = sd[a] < 0.0f ? 0.0f : 1.0f;
    }
    float tEnter = gmax(gmax(tmin3[0], tmin3[1]), tmin3[2]);
    float tExit = gmin(gmin(tmax3[0], tmax3[1]), tmax3[2]);
    if (tExit < gmax(tEnter, 0.0f) || tEnter > tcap) return;

    float t = gmax(tEnter, 0.0f);
    int32_t bc[3], normal[3] = {0, 0, 0};
    float tD8[3], tM8[3];
    {
        float e = t + 1e-3f;
        for (int a = 0; a < 3; ++a) {
            float entry = ro[a] + sd[a] * e;
            bc[a] = iclamp(f2i(floorf(entry * 0.125f)), minB[a], minB[a] + w->wb[a] - 1);
            tD8[a] = fabsf(inv[a]) * 8.0f;
            tM8[a] = (((float)bc[a] + up[a]) * 8.0f - ro[a]) * inv[a];
        }
    }
    if (tEnter > 0.0f) {
        if (tEnter == tmin3[0]) normal[0] = -st[0];
        else if (tEnter == tmin3[1]) normal[1] = -st[1];
        else normal[2] = -st[2];
    }
    const int maxOuter = 3 * (w->wb[0] + w->wb[1] + w->wb[2]) + 4;
    for (int i = 0; i < maxOuter; ++i) {
        if (t > tcap) return;
        size_t gi = gidx(w, bc[0], bc[1], bc[2]);
        uint32_t g = w->grid[gi];
        if (cnt) { cnt->outer++; cnt->bytes += 4; }
        if (skip == 5u) {   /* WATER_CLEAR: water-only bricks are empty for this ray (:582-584) */
            if (g == 5u) g = 0u;
            else if (g >= 256u) {
                if (cnt) cnt->bytes += 1;
                if (w->dry[gi] > 0) g = 0u;
            }
        }
        if (g >= 256u) {
            if (box_mode >= 0 && box_may_skip(w, ro, sd, st, bc, g, t, tM8, skip, box_margin, box_mode)) {
                box_stats[0]++;
            } else {
                if (box_stats) box_stats[1]++;
                if (brick_voxels(w, ro, sd, inv, st, up, bc, g, t, normal, skip, h, cnt)) return;
            }
        } else if (g != 0u && g != skip) {
            float e = t + 1e-3f;
            h->hit = 1; h->t = t; h->mat = g;
            for (int a = 0; a < 3; ++a) { h->n[a] = normal[a]; h->cell[a] = f2i(floorf(ro[a] + sd[a] * e)); }
            return;
        }
        int stepped = 0;
        if (w->coarse_skip && g == 0u) {
            int32_t cellB[3] = {0, 0, 0};
            int32_t Cf = 0;
            int32_t d = 0;
            if (w->use_dist) {
                d = (int32_t)(skip == 5u ? w->dry[gi] : w->dist[gi]);
                if (cnt) cnt->bytes += 1;
                if (d > 1) {
                    Cf = 2 * d - 1;
                    for (int a = 0; a < 3; ++a) cellB[a] = bc[a] - (d - 1);
                }
            }
            if (!(w->use_dist && d > 1)) {
                int32_t wc[3] = {bc[0], bc[1], bc[2]};
                if (ring) { wc[0] &= w->wb[0] - 1; wc[2] &= w->wb[2] - 1; }
                if (cnt) cnt->bytes += 1;
                size_t i2 = ((size_t)(wc[2] >> 4) * w->c2[1] + (size_t)(wc[1] >> 4)) * w->c2[0] + (size_t)(wc[0] >> 4);
                if (w->coarse2[i2] == 0) Cf = 16;
                else {
                    if (cnt) cnt->bytes += 1;
                    size_t i1 = ((size_t)(wc[2] >> 2) * w->c1[1] + (size_t)(wc[1] >> 2)) * w->c1[0] + (size_t)(wc[0] >> 2);
                    Cf = w->coarse1[i1] == 0 ? 4 : 0;
                }
                if (Cf > 0)
                    for (int a = 0; a < 3; ++a) cellB[a] = bc[a] - (bc[a] & (Cf - 1));
            }
            if (Cf > 0) {
                float tFar[3];
                for (int a = 0; a < 3; ++a) {
                    float cmin = (float)cellB[a] * 8.0f;
                    float cmax = cmin + (float)Cf * 8.0f;
                    tFar[a] = ((up[a] != 0.0f ? cmax : cmin) - ro[a]) * inv[a];
                }
                int ax[3];
                ax[0] = tFar[0] < tFar[1] && tFar[0] < tFar[2];
                ax[1] = !ax[0] && tFar[1] < tFar[2];
                ax[2] = !ax[0] && !ax[1];
                float tx = ax[0] ? tFar[0] : (ax[1] ? tFar[1] : tFar[2]);
                t = tx;
                float e = t + 1e-3f;
                for (int a = 0; a < 3; ++a) {
                    normal[a] = ax[a] ? -st[a] : 0;
                    int32_t v = iclamp(f2i(floorf((ro[a] + sd[a] * e) * 0.125f)), cellB[a], cellB[a] + Cf - 1);
                    if (ax[a]) v = st[a] > 0 ? cellB[a] + Cf : cellB[a] - 1;
                    bc[a] = v;
                    tM8[a] = (((float)bc[a] + up[a]) * 8.0f - ro[a]) * inv[a];
                }
                stepped = 1;
            }
        }
        if (!stepped) {
            int ax;
            if (tM8[0] < tM8[1] && tM8[0] < tM8[2]) ax = 0;
            else if (tM8[1] < tM8[2]) ax = 1;
            else ax = 2;
            bc[ax] += st[ax];
            t = tM8[ax];
            tM8[ax] += tD8[ax];
            normal[0] = normal[1] = normal[2] = 0;
            normal[ax] = -st[ax];
        }
        for (int a = 0; a < 3; ++a)
            if (bc[a] < minB[a] || bc[a] >= minB[a] + w->wb[a]) return;
    }
}

/* ---- bodies: traceObject / traceObjectsA / traceSceneS (common.glsl:664-720, 787-793) ---- */
static inline int32_t fbits(float f) { int32_t v; memcpy(&v, &f, 4); return v; }
static inline float ibits(int32_t v) { float f; memcpy(&f, &v, 4); return f; }
/* ObjHeader as 60 floats: worldToObj [0..15], objToWorld [16..31], prevObjToWorld [32..47] (column-major,
 * m[c][r] = a[4c + r]), dimsBricks [48..51] (int bits), aabbMin [52..55], aabbMax [56..59]. */
static inline const float *obj_header(const vfo_world *w, int32_t oi) { return w->bodies + (size_t)oi * 60; }

/* traceVolumeS with useTex false (common.glsl:541-658): the box [0, dims*8), the body's grid at its offset
 * in objGrid, skipMat kSkipMat = 5, no water test, no distance-field or coarse skips. */
static void trace_body(const vfo_world *w, int32_t oi, const float ro[3], const float rd[3], float tcap, vfo_hit *h,
                       vfo_count *cnt) {
    const float *b = obj_header(w, oi);
    const int32_t db[3] = {fbits(b[48]), fbits(b[49]), fbits(b[50])};
    const uint32_t goff = (uint32_t)fbits(b[51]);
    const uint32_t skip = 5u;
    memset(h, 0, sizeof(*h));
    h->t = -1.0f;
    float sd[3], inv[3], t0[3], t1[3], tmin3[3], tmax3[3], up[3];
    int32_t st[3];
    for (int a = 0; a < 3; ++a) {
        sd[a] = fabsf(rd[a]) < 1e-5f ? 1e-5f : rd[a];
        inv[a] = 1.0f / sd[a];
        float lo = 0.0f;
        float dims = (float)(db[a] * 8);
        t0[a] = (lo - ro[a]) * inv[a];
        t1[a] = ((lo + dims) - ro[a]) * inv[a];
        tmin3[a] = gmin(t0[a], t1[a]);
        tmax3[a] = gmax(t0[a], t1[a]);
        st[a] = sd[a] > 0.0f ? 1 : -1;
        up[a] = sd[a] < 0.0f ? 0.0f : 1.0f;
    }
    float tEnter = gmax(gmax(tmin3[0], tmin3[1]), tmin3[2]);
    float tExit = gmin(gmin(tmax3[0], tmax3[1]), tmax3[2]);
    if (tExit < gmax(tEnter, 0.0f) || tEnter > tcap) return;
    float t = gmax(tEnter, 0.0f);
    int32_t bc[3], normal[3] = {0, 0, 0};
    float tD8[3], tM8[3];
    {
        float e = t + 1e-3f;
        for (int a = 0; a < 3; ++a) {
            float entry = ro[a] + sd[a] * e;
            bc[a] = iclamp(f2i(floorf(entry * 0.125f)), 0, db[a] - 1);
            tD8[a] = fabsf(inv[a]) * 8.0f;
            tM8[a] = (((float)bc[a] + up[a]) * 8.0f - ro[a]) * inv[a];
        }
    }
    if (tEnter > 0.0f) {
        if (tEnter == tmin3[0]) normal[0] = -st[0];
        else if (tEnter == tmin3[1]) normal[1] = -st[1];
        else normal[2] = -st[2];
    }
    const int maxOuter = 3 * (db[0] + db[1] + db[2]) + 4;
    for (int i = 0; i < maxOuter; ++i) {
        if (t > tcap) return;
        uint32_t g = w->obj_grid[goff + (uint32_t)((bc[2] * db[1] + bc[1]) * db[0] + bc[0])];
        if (cnt) { cnt->outer++; cnt->bytes += 4; }
        if (g >= 256u) {
            if (brick_voxels(w, ro, sd, inv, st, up, bc, g, t, normal, skip, h, cnt)) return;
        } else if (g != 0u && g != skip) {
            float e = t + 1e-3f;
            h->hit = 1; h->t = t; h->mat = g;
            for (int a = 0; a < 3; ++a) { h->n[a] = normal[a]; h->cell[a] = f2i(floorf(ro[a] + sd[a] * e)); }
            return;
        }
        int ax;
        if (tM8[0] < tM8[1] && tM8[0] < tM8[2]) ax = 0;
        else if (tM8[1] < tM8[2]) ax = 1;
        else ax = 2;
        bc[ax] += st[ax];
        t = tM8[ax];
        tM8[ax] += tD8[ax];
        normal[0] = normal[1] = normal[2] = 0;
        normal[ax] = -st[ax];
        for (int a = 0; a < 3; ++a)
            if (bc[a] < 0 || bc[a] >= db[a]) return;
    }
}

static void box_span(const float lo[3], const float hi[3], const float ro[3], const float inv[3], float *tEnter,
                     float *tExit) {
    float tmin3[3], tmax3[3];
    for (int a = 0; a < 3; ++a) {
        float t0 = (lo[a] - ro[a]) * inv[a], t1 = (hi[a] - ro[a]) * inv[a];
        tmin3[a] = gmin(t0, t1);
        tmax3[a] = gmax(t0, t1);
    }
    *tEnter = gmax(gmax(tmin3[0], tmin3[1]), tmin3[2]);
    *tExit = gmin(gmin(tmax3[0], tmax3[1]), tmax3[2]);
}

/* traceObject (common.glsl:665-683) */
static void trace_object(const vfo_world *w, int32_t oi, const float ro[3], const float rd[3], const float inv[3],
                         vfo_hit *best, float *limit, vfo_count *cnt) {
    const float *b = obj_header(w, oi);
    float tEnter, tExit;
    box_span(b + 52, b + 56, ro, inv, &tEnter, &tExit);
    if (tExit < gmax(tEnter, 0.0f) || tEnter > *limit) return;
    const float *m = b;   /* worldToObj, m[c][r] = m[4c + r] */
    float roO[3], rdO[3];
    for (int r = 0; r < 3; ++r) {
        roO[r] = ((m[r] * ro[0] + m[4 + r] * ro[1]) + m[8 + r] * ro[2]) + m[12 + r];
        rdO[r] = (m[r] * rd[0] + m[4 + r] * rd[1]) + m[8 + r] * rd[2];
    }
    vfo_hit oh;
    trace_body(w, oi, roO, rdO, *limit, &oh, cnt);
    if (oh.hit && oh.t < *limit) {
        *best = oh;
        best->body = oi + 1;
        *limit = oh.t;
    }
}

void vfo_trace_objects(const vfo_world *w, const float ro[3], const float rd[3], vfo_hit *best, float limit,
                       int32_t any_hit, vfo_count *cnt) {
    if (w->n_bodies == 0) return;
    float inv[3];
    for (int a = 0; a < 3; ++a) inv[a] = 1.0f / (fabsf(rd[a]) < 1e-5f ? 1e-5f : rd[a]);
    if (!w->use_bvh) {   /* common.glsl:719 */
        for (int32_t oi = 0; oi < w->n_bodies && !(any_hit && best->hit); ++oi) trace_object(w, oi, ro, rd, inv, best, &limit, cnt);
        return;
    }
    /* common.glsl:697-716: near child first; subtrees beyond the current limit are culled */
    int32_t stack[24];
    int sp = 0, node = 0;
    for (int it = 0; it < 512; ++it) {
        const float *nd = w->bvh + (size_t)node * 8;
        float tEnter, tExit;
        box_span(nd, nd + 4, ro, inv, &tEnter, &tExit);
        int visit = tExit >= gmax(tEnter, 0.0f) && tEnter <= limit;
        uint32_t first = (uint32_t)fbits(nd[3]), count = (uint32_t)fbits(nd[7]);
        if (visit && count == 0u) {
            if (sp < 24) stack[sp++] = (int32_t)first + 1;
            node = (int32_t)first;
            continue;
        }
        if (visit)
            for (uint32_t k = 0; k < count; ++k) trace_object(w, (int32_t)w->bvh_prims[first + k], ro, rd, inv, best, &limit, cnt);
        if (any_hit && best->hit) return;
        if (sp == 0) break;
        node = stack[--sp];
    }
}

void vfo_trace_scene(const vfo_world *w, const float ro[3], const float rd[3], float tcap, uint32_t skip, vfo_hit *h,
                     vfo_count *cnt) {
    vfo_trace(w, ro, rd, tcap, skip, h, cnt);
    h->body = 0;
    vfo_trace_objects(w, ro, rd, h, h->hit ? h->t : tcap, 0, cnt);
}

void vfo_hit_normal(const vfo_world *w, const vfo_hit *h, float out[3]) {
    const float n[3] = {(float)h->n[0], (float)h->n[1], (float)h->n[2]};
    if (h->body == 0) {
        out[0] = n[0]; out[1] = n[1]; out[2] = n[2];
        return;
    }
    const float *o = obj_header(w, h->body - 1) + 16;   /* objToWorld */
    for (int r = 0; r < 3; ++r) out[r] = (o[r] * n[0] + o[4 + r] * n[1]) + o[8 + r] * n[2];
    normalize3(out);
}

/* build_bvh (gpu_world.rs:550-600): box of the node's bodies; <= 2 bodies: a leaf; else the median of
 * the box centroids (min + max) along the axis of their widest extent (strictly wider than a lower axis
 * to win), children appended as a pair, the left subtree built first.  The median is taken after
 * ordering the range by (centroid, body index), one of the partitions select_nth_unstable_by may give. */
static void bvh_set(float *nd, const float mn[3], const float mx[3], uint32_t a, uint32_t cnt) {
    nd[0] = mn[0]; nd[1] = mn[1]; nd[2] = mn[2]; nd[3] = ibits((int32_t)a);
    nd[4] = mx[0]; nd[5] = mx[1]; nd[6] = mx[2]; nd[7] = ibits((int32_t)cnt);
}
static void bvh_build(const float *bodies, uint32_t *prims, uint32_t count, float *nodes, int32_t *n_nodes, int32_t n,
                      uint32_t first) {
    float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
    for (uint32_t i = 0; i < count; ++i) {
        const float *b = bodies + (size_t)prims[i] * 60;
        for (int k = 0; k < 3; ++k) {
            if (b[52 + k] < mn[k]) mn[k] = b[52 + k];
            if (b[56 + k] > mx[k]) mx[k] = b[56 + k];
        }
    }
    if (count <= 2) {
        bvh_set(nodes + (size_t)n * 8, mn, mx, first, count);
        return;
    }
    float cmn[3] = {1e30f, 1e30f, 1e30f}, cmx[3] = {-1e30f, -1e30f, -1e30f};
    for (uint32_t i = 0; i < count; ++i) {
        const float *b = bodies + (size_t)prims[i] * 60;
        for (int k = 0; k < 3; ++k) {
            float c = b[52 + k] + b[56 + k];
            if (c < cmn[k]) cmn[k] = c;
            if (c > cmx[k]) cmx[k] = c;
        }
    }
    int ax = 0;
    for (int k = 1; k < 3; ++k)
        if (cmx[k] - cmn[k] > cmx[ax] - cmn[ax]) ax = k;
    /* insertion sort of the range by (centroid on ax, index) */
    for (uint32_t i = 1; i < count; ++i) {
        uint32_t p = prims[i];
        float cp = bodies[(size_t)p * 60 + 52 + ax] + bodies[(size_t)p * 60 + 56 + ax];
        uint32_t j = i;
        while (j > 0) {
            uint32_t q = prims[j - 1];
            float cq = bodies[(size_t)q * 60 + 52 + ax] + bodies[(size_t)q * 60 + 56 + ax];
            if (cq < cp || (cq == cp && q < p)) break;
            prims[j] = q;
            --j;
        }
        prims[j] = p;
    }
    uint32_t mid = count / 2;
    int32_t left = *n_nodes;
    *n_nodes += 2;
    bvh_set(nodes + (size_t)n * 8, mn, mx, (uint32_t)left, 0u);
    bvh_build(bodies, prims, mid, nodes, n_nodes, left, first);
    bvh_build(bodies, prims + mid, count - mid, nodes, n_nodes, left + 1, first + mid);
}

void vfo_build_bvh(const float *bodies, int32_t n, float *nodes, int32_t *n_nodes, uint32_t *prims) {
    *n_nodes = 0;
    if (n <= 0) return;
    for (int32_t i = 0; i < n; ++i) prims[i] = (uint32_t)i;
    *n_nodes = 1;
    bvh_build(bodies, prims, (uint32_t)n, nodes, n_nodes, 0, 0u);
}

static void to_rec(const vfo_hit *h, vfo_hitrec *r) {
    memset(r, 0, sizeof(*r));
    if (!h->hit) { r->t = -1.0f; return; }
    r->cell[0] = h->cell[0]; r->cell[1] = h->cell[1]; r->cell[2] = h->cell[2];
    r->t = h->t;
    r->mat = (uint8_t)h->mat;
    r->face = (uint8_t)vfo_face_code(h->n);
    r->hit = 1;
    r->body = h->body != 0;
}

static int nthreads(int threads) {
#ifdef _OPENMP
    return threads > 0 ? threads : omp_get_max_threads();
#else
    (void)threads;
    return 1;
#endif
}

static void add_cnt(vfo_count *dst, const vfo_count *src) {
    dst->rays += src->rays; dst->outer += src->outer; dst->voxel += src->voxel; dst->bytes += src->bytes;
}

void vfo_trace_batch(const vfo_world *w, const vfo_ray *rays, int64_t n, float tcap, uint32_t skip,
                     vfo_hitrec *out, vfo_count *cnt, int threads) {
    vfo_count total = {0, 0, 0, 0};
#pragma omp parallel num_threads(nthreads(threads))
    {
        vfo_count c = {0, 0, 0, 0};
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < n; ++i) {
            vfo_hit h;
            vfo_trace(w, rays[i].o, rays[i].d, tcap, skip, &h, &c);
            to_rec(&h, &out[i]);
        }
#pragma omp critical
        add_cnt(&total, &c);
    }
    if (cnt) add_cnt(cnt, &total);
}

void vfo_trace_box_batch(const vfo_world *w, const vfo_ray *rays, int64_t n, float tcap, uint32_t skip, float margin,
                         int mode, vfo_hitrec *out, int64_t stats[2], int threads) {
    int64_t skipped = 0, walked = 0;
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads(threads)) reduction(+ : skipped, walked)
    for (int64_t i = 0; i < n; ++i) {
        vfo_hit h;
        int64_t st[2] = {0, 0};
        trace_impl(w, rays[i].o, rays[i].d, tcap, skip, &h, NULL, margin, mode, st);
        to_rec(&h, &out[i]);
        skipped += st[0];
        walked += st[1];
    }
    stats[0] = skipped;
    stats[1] = walked;
}

/* traceSceneS per ray (world + bodies), with the body index per ray (-1 = world or miss; nullable). */
void vfo_trace_scene_batch(const vfo_world *w, const vfo_ray *rays, int64_t n, float tcap, uint32_t skip,
                           vfo_hitrec *out, int32_t *body, int threads) {
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads(threads))
    for (int64_t i = 0; i < n; ++i) {
        vfo_hit h;
        vfo_trace_scene(w, rays[i].o, rays[i].d, tcap, skip, &h, NULL);
        to_rec(&h, &out[i]);
        if (body) body[i] = h.hit ? h.body - 1 : -1;
    }
}

/* ---- FP64 plain voxel march: an independent formulation over the same voxels ----
 * No bricks, no skips, no nudges: Amanatides-Woo in double precision along the ray with the
 * same substituted direction (|rd| < 1e-5 -> 1e-5).  Keeps the reference's rule that a brick
 * entered at t <= tCap is searched whole (common.glsl:574, 803-804). */
void vfo_march_f64(const vfo_world *w, const double ro[3], const double rd[3], double tcap, uint32_t skip,
                   int32_t *hit, double *tout, int32_t cell[3], int32_t n[3], uint32_t *mat) {
    *hit = 0; *tout = -1.0; *mat = 0;
    cell[0] = cell[1] = cell[2] = 0; n[0] = n[1] = n[2] = 0;
    const int ring = is_ring(w);
    const int32_t minB[3] = {ring ? w->win[0] : 0, 0, ring ? w->win[2] : 0};
    double sd[3], lo[3], hi[3], tEnter = -1e300, tExit = 1e300;
    int entry_axis = -1;
    int32_t st[3];
    for (int a = 0; a < 3; ++a) {
        float r = (float)rd[a];
        sd[a] = fabsf(r) < 1e-5f ? (double)1e-5f : rd[a];
        lo[a] = (double)(minB[a] * 8);
        hi[a] = (double)((minB[a] + w->wb[a]) * 8);
        double ta = (lo[a] - ro[a]) / sd[a], tb = (hi[a] - ro[a]) / sd[a];
        double tn = ta < tb ? ta : tb, tf = ta < tb ? tb : ta;
        if (tn > tEnter) { tEnter = tn; entry_axis = a; }
        if (tf < tExit) tExit = tf;
        st[a] = sd[a] > 0 ? 1 : -1;
    }
    double t = tEnter > 0 ? tEnter : 0.0;
    if (tExit < t || tEnter > tcap) return;
    int32_t v[3], nn[3] = {0, 0, 0};
    double tM[3], tD[3];
    for (int a = 0; a < 3; ++a) {
        double p = ro[a] + sd[a] * t;
        int32_t c = (int32_t)floor(p);
        if (tEnter > 0 && a == entry_axis) c = st[a] > 0 ? (int32_t)lo[a] : (int32_t)hi[a] - 1;
        if (c < (int32_t)lo[a]) c = (int32_t)lo[a];
        if (c > (int32_t)hi[a] - 1) c = (int32_t)hi[a] - 1;
        v[a] = c;
        tM[a] = ((double)(c + (st[a] > 0 ? 1 : 0)) - ro[a]) / sd[a];
        tD[a] = fabs(1.0 / sd[a]);
    }
    if (tEnter > 0) nn[entry_axis] = -st[entry_axis];
    int32_t cur_brick[3] = {v[0] >> 3, v[1] >> 3, v[2] >> 3};
    for (;;) {
        uint32_t m = vfo_voxel(w, v[0], v[1], v[2]);
        if (m != 0u && m != skip && !(skip == 5u && m == 5u)) {
            *hit = 1; *tout = t; *mat = m;
            for (int a = 0; a < 3; ++a) { cell[a] = v[a]; n[a] = nn[a]; }
            return;
        }
        int ax = tM[0] < tM[1] && tM[0] < tM[2] ? 0 : (tM[1] < tM[2] ? 1 : 2);
        v[ax] += st[ax];
        t = tM[ax];
        tM[ax] += tD[ax];
        nn[0] = nn[1] = nn[2] = 0;
        nn[ax] = -st[ax];
        if (v[ax] < (int32_t)lo[ax] || v[ax] >= (int32_t)hi[ax]) return;
        int32_t nb[3] = {v[0] >> 3, v[1] >> 3, v[2] >> 3};
        if (nb[0] != cur_brick[0] || nb[1] != cur_brick[1] || nb[2] != cur_brick[2]) {
            if (t > tcap) return;
            cur_brick[0] = nb[0]; cur_brick[1] = nb[1]; cur_brick[2] = nb[2];
        }
    }
}

/* ---- the reference's CPU raycast (sparse_volume.rs:573-637), bounded volumes ---- */
void vfo_raycast_cpu(const vfo_world *w, const float ro[3], const float rd[3], float max_t, uint32_t skip,
                     int32_t *hit, float *tout, int32_t cell[3], int32_t n[3], uint32_t *mat) {
    *hit = 0; *tout = 0; *mat = 0;
    float inv[3];
    int32_t dims[3] = {w->wb[0] * 8, w->wb[1] * 8, w->wb[2] * 8};
    float tE = 0.0f, tX = max_t;
    for (int a = 0; a < 3; ++a) {
        float s = fabsf(rd[a]) < 1e-6f ? 1e-6f : rd[a];
        inv[a] = 1.0f / s;
        float a0 = (0.0f - ro[a]) * inv[a], a1 = ((float)dims[a] - ro[a]) * inv[a];
        if (a0 > a1) { float tmp = a0; a0 = a1; a1 = tmp; }
        tE = fmaxf(tE, a0);
        tX = fminf(tX, a1);
    }
    if (tX < tE) return;
    float t = tE + 1e-4f;
    int32_t c[3], st[3], nn[3] = {0, 0, 0};
    float tM[3], tD[3];
    for (int a = 0; a < 3; ++a) {
        c[a] = (int32_t)floorf(ro[a] + rd[a] * t);
        c[a] = iclamp(c[a], 0, dims[a] - 1);
        st[a] = rd[a] < 0.0f ? -1 : 1;
        int32_t b = c[a] + (st[a] > 0 ? 1 : 0);
        tM[a] = ((float)b - ro[a]) * inv[a];
        tD[a] = fabsf(inv[a]);
    }
    int max_steps = 4 * (dims[0] + dims[1] + dims[2]);
    for (int i = 0; i < max_steps; ++i) {
        uint32_t m = vfo_voxel(w, c[0], c[1], c[2]);
        if (m != 0u && m != skip) {
            *hit = 1; *tout = t; *mat = m;
            for (int a = 0; a < 3; ++a) { cell[a] = c[a]; n[a] = nn[a]; }
            return;
        }
        int ax = tM[0] < tM[1] && tM[0] < tM[2] ? 0 : (tM[1] < tM[2] ? 1 : 2);
        c[ax] += st[ax];
        t = tM[ax];
        tM[ax] += tD[ax];
        nn[0] = nn[1] = nn[2] = 0;
        nn[ax] = -st[ax];
        if (t > max_t || c[ax] < 0 || c[ax] >= dims[ax]) return;
    }
}

/* ---- mirror builders ---- */
/* brick_mask (sparse_volume.rs:48-63): octant o = ox | oy<<1 | oz<<2, bit = (x&3) | (y&3)<<2 | (z&3)<<4. */
void vfo_brick_mask(const uint8_t vox[512], uint32_t m[16]) {
    memset(m, 0, 16 * sizeof(uint32_t));
    for (int z = 0; z < 8; ++z)
        for (int y = 0; y < 8; ++y)
            for (int x = 0; x < 8; ++x) {
                if (vox[z * 64 + y * 8 + x] == 0) continue;
                int oct = (x >> 2) | ((y >> 2) << 1) | ((z >> 2) << 2);
                int bit = (x & 3) | ((y & 3) << 2) | ((z & 3) << 4);
                m[oct * 2 + (bit >> 5)] |= 1u << (bit & 31);
            }
}

/* Coarse pyramid (gpu_world.rs:300-313): a cell is 1 iff any brick in it is non-empty. */
void vfo_coarse(const uint32_t *grid, const int32_t wb[3], uint8_t *c1, uint8_t *c2) {
    int32_t d1[3], d2[3];
    for (int a = 0; a < 3; ++a) { d1[a] = (wb[a] + 3) / 4; d2[a] = (wb[a] + 15) / 16; }
    memset(c1, 0, (size_t)d1[0] * d1[1] * d1[2]);
    memset(c2, 0, (size_t)d2[0] * d2[1] * d2[2]);
    for (int32_t z = 0; z < wb[2]; ++z)
        for (int32_t y = 0; y < wb[1]; ++y)
            for (int32_t x = 0; x < wb[0]; ++x) {
                if (grid[((size_t)z * wb[1] + y) * wb[0] + x] == 0) continue;
                c1[((size_t)(z / 4) * d1[1] + y / 4) * d1[0] + x / 4] = 1;
                c2[((size_t)(z / 16) * d2[1] + y / 16) * d2[0] + x / 16] = 1;
            }
}

/* One brickdist.comp dispatch (brickdist.comp:31-68), DIST_CAP 16. */
void vfo_brickdist_pass(const uint32_t *grid, const uint8_t *src, uint8_t *dst, const uint8_t *atlas,
                        const int32_t wb[3], const int32_t bmin[3], const int32_t bmax[3], int32_t axis,
                        int32_t first_pass_raw) {
    const int ring = (first_pass_raw & 4) != 0;
    const int fp = first_pass_raw & 3;
    const int32_t CAP = 16;
    for (int32_t cz = bmin[2]; cz <= bmax[2]; ++cz)
        for (int32_t cy = bmin[1]; cy <= bmax[1]; ++cy)
            for (int32_t cx = bmin[0]; cx <= bmax[0]; ++cx) {
                int32_t cw[3] = {ring ? (cx & (wb[0] - 1)) : cx, cy, ring ? (cz & (wb[2] - 1)) : cz};
                size_t ci = ((size_t)cw[2] * wb[1] + cw[1]) * wb[0] + cw[0];
                if (fp == 2) {
                    uint32_t g = grid[ci];
                    int solid = g != 0u && g != 5u;
                    if (g >= 256u) {
                        const uint8_t *b = atlas + (size_t)(g - 256u) * 512u;
                        solid = 0;
                        for (int i = 0; i < 512 && !solid; ++i) solid = b[i] != 0 && b[i] != 5;
                    }
                    dst[ci] = solid ? 0 : (uint8_t)CAP;
                    continue;
                }
                int32_t c[3] = {cx, cy, cz};
                uint32_t best = (uint32_t)CAP;
                for (int32_t k = -CAP; k <= CAP; ++k) {
                    int32_t q[3] = {c[0], c[1], c[2]};
                    q[axis] += k;
                    if (ring) {
                        if (q[1] < 0 || q[1] >= wb[1]) continue;
                        q[0] &= wb[0] - 1;
                        q[2] &= wb[2] - 1;
                    } else if (q[0] < 0 || q[1] < 0 || q[2] < 0 || q[0] >= wb[0] || q[1] >= wb[1] || q[2] >= wb[2]) {
                        continue;
                    }
                    size_t qi = ((size_t)q[2] * wb[1] + q[1]) * wb[0] + q[0];
                    uint32_t f;
                    if (fp != 0) f = grid[qi] != 0u ? 0u : (uint32_t)CAP;
                    else f = src[qi];
                    uint32_t ak = (uint32_t)(k < 0 ? -k : k);
                    uint32_t m = ak > f ? ak : f;
                    if (m < best) best = m;
                }
                dst[ci] = (uint8_t)best;
            }
}

/* record_distance_field (gpu_world.rs:1376-1428). */
void vfo_record_distance_field(const uint32_t *grid, const uint8_t *atlas, const int32_t wb[3], int32_t ring,
                               const int32_t lo_in[3], const int32_t hi_in[3], const int32_t dlo_in[3],
                               const int32_t dhi_in[3], uint8_t *d0, uint8_t *d1, uint8_t *d2, uint8_t *d3) {
    int32_t lo[3], hi[3], dlo[3], dhi[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = lo_in[k]; hi[k] = hi_in[k]; dlo[k] = dlo_in[k]; dhi[k] = dhi_in[k];
        if (ring && k != 1) {
            dhi[k] = imin(dhi[k], dlo[k] + wb[k] - 1);
            hi[k] = imin(hi[k], lo[k] + wb[k] - 1);
            continue;
        }
        lo[k] = iclamp(lo[k], 0, wb[k] - 1);
        hi[k] = iclamp(hi[k], 0, wb[k] - 1);
        dlo[k] = iclamp(dlo[k], 0, wb[k] - 1);
        dhi[k] = iclamp(dhi[k], 0, wb[k] - 1);
    }
    const int rf = ring ? 4 : 0;
    uint8_t *const srcs[7] = {NULL, d0, d1, NULL, d3, d2, d1};
    uint8_t *const dsts[7] = {d0, d1, d0, d3, d2, d1, d2};
    for (int pass = 0; pass < 7; ++pass) {
        int axis = pass % 3, fp = pass == 0 ? 1 : 0;
        const int32_t *bmin = lo, *bmax = hi;
        if (pass == 3) { fp = 2; bmin = dlo; bmax = dhi; }
        if (pass >= 4) axis = pass - 4;
        vfo_brickdist_pass(grid, srcs[pass], dsts[pass], atlas, wb, bmin, bmax, axis, fp | rf);
    }
}

void vfo_brickdist_brute(const uint32_t *grid, const uint8_t *atlas, const int32_t wb[3], int32_t ring, int32_t dry,
                         uint8_t *out) {
    size_t nb = (size_t)wb[0] * wb[1] * wb[2];
    memset(out, 16, nb);
    for (int32_t z = 0; z < wb[2]; ++z)
        for (int32_t y = 0; y < wb[1]; ++y)
            for (int32_t x = 0; x < wb[0]; ++x) {
                uint32_t g = grid[((size_t)z * wb[1] + y) * wb[0] + x];
                int occ = g != 0u;
                if (dry) {
                    occ = g != 0u && g != 5u;
                    if (g >= 256u) {
                        const uint8_t *b = atlas + (size_t)(g - 256u) * 512u;
                        occ = 0;
                        for (int i = 0; i < 512 && !occ; ++i) occ = b[i] != 0 && b[i] != 5;
                    }
                }
                if (!occ) continue;
                for (int32_t dz = -16; dz <= 16; ++dz)
                    for (int32_t dy = -16; dy <= 16; ++dy)
                        for (int32_t dx = -16; dx <= 16; ++dx) {
                            int32_t q[3] = {x + dx, y + dy, z + dz};
                            if (q[1] < 0 || q[1] >= wb[1]) continue;
                            if (ring) { q[0] &= wb[0] - 1; q[2] &= wb[2] - 1; }
                            else if (q[0] < 0 || q[2] < 0 || q[0] >= wb[0] || q[2] >= wb[2]) continue;
                            int32_t ad = abs(dx);
                            if (abs(dy) > ad) ad = abs(dy);
                            if (abs(dz) > ad) ad = abs(dz);
                            size_t qi = ((size_t)q[2] * wb[1] + q[1]) * wb[0] + q[0];
                            if ((uint8_t)ad < out[qi]) out[qi] = (uint8_t)ad;
                        }
            }
}

void vfo_skytop(const vfo_world *w, const float *kinds, uint16_t *out) {
    const int32_t sx = w->wb[0] * 8, sz = w->wb[2] * 8;
    for (int32_t z = 0; z < sz; ++z)
        for (int32_t x = 0; x < sx; ++x) {
            uint16_t sky = 0;
            for (int32_t y = w->wb[1] * 8 - 1; y >= 0 && sky == 0; --y) {
                uint32_t g = w->grid[((size_t)(z >> 3) * w->wb[1] + (y >> 3)) * w->wb[0] + (x >> 3)];
                uint32_t m = g < 256u ? g : w->atlas[(size_t)(g - 256u) * 512u + ((z & 7) * 64 + (y & 7) * 8 + (x & 7))];
                if (m == 0u) continue;
                float k = kinds[m];
                if (k != 1.0f && k != 3.0f) sky = (uint16_t)(y + 1);
            }
            out[(size_t)z * sx + x] = sky;
        }
}

/* ---- RNG and sampling (common.glsl:1233-1288) ---- */
uint32_t vfo_pcg(uint32_t v) {
    v = v * 747796405u + 2891336453u;
    v = ((v >> ((v >> 28u) + 4u)) ^ v) * 277803737u;
    return (v >> 22u) ^ v;
}
uint32_t vfo_pixel_seed(int32_t px, int32_t py, uint32_t frame, uint32_t salt) {
    return vfo_pcg((uint32_t)px + vfo_pcg((uint32_t)py + vfo_pcg(frame ^ salt)));
}
static inline float fractf(float x) { return x - floorf(x); }
void vfo_ld_sample2(int32_t px, int32_t py, uint32_t k, uint32_t salt, float out[2]) {
    uint32_t h = vfo_pcg((uint32_t)px + vfo_pcg((uint32_t)py + salt));
    float rx = (float)(h >> 8u) * (1.0f / 16777216.0f);
    float ry = (float)(vfo_pcg(h) >> 8u) * (1.0f / 16777216.0f);
    float kf = (float)(k & 0xFFFFu);
    out[0] = fractf(rx + kf * 0.7548776662f);
    out[1] = fractf(ry + kf * 0.5698402910f);
}

/* sin / cos: Cody-Waite reduction by pi/2 (split 1.5703125 + 4.838267948966e-4) and Taylor
 * polynomials on [-pi/4, pi/4]; the same arithmetic runs in the HIP kernels, so both sides
 * produce identical GI / probe directions. */
static void sincos_red(float x, float *s, float *c) {
    float kf = floorf(x * 0.63661977f + 0.5f);
    int32_t k = f2i(kf);
    float r = (x - kf * 1.5703125f) - kf * 4.838267948966e-4f;
    float r2 = r * r;
    float ps = r + r * r2 * (-0.16666667f + r2 * (0.0083333333f + r2 * (-1.9841270e-4f + r2 * 2.7557319e-6f)));
    float pc = 1.0f + r2 * (-0.5f + r2 * (0.041666667f + r2 * (-1.3888889e-3f + r2 * 2.4801587e-5f)));
    switch (k & 3) {
    case 0: *s = ps; *c = pc; break;
    case 1: *s = pc; *c = -ps; break;
    case 2: *s = -ps; *c = -pc; break;
    default: *s = -pc; *c = ps; break;
    }
}
float vfo_sin(float x) { float s, c; sincos_red(x, &s, &c); return s; }
float vfo_cos(float x) { float s, c; sincos_red(x, &s, &c); return c; }

static void make_basis(const float n[3], float t1[3], float t2[3]) {
    float up[3] = {0.0f, 1.0f, 0.0f}, ux[3] = {1.0f, 0.0f, 0.0f};
    cross3(n, fabsf(n[1]) < 0.9f ? up : ux, t1);
    normalize3(t1);
    cross3(n, t1, t2);
}
void vfo_cosine_hemisphere_u(const float n[3], const float u[2], float out[3]) {
    float t1[3], t2[3];
    make_basis(n, t1, t2);
    float r = sqrtf(u[0]);
    float ph = 6.2831853f * u[1];
    float s, c;
    sincos_red(ph, &s, &c);
    float a = r * c, b = r * s, z = sqrtf(1.0f - u[0]);
    for (int i = 0; i < 3; ++i) out[i] = (t1[i] * a + t2[i] * b) + n[i] * z;
    normalize3(out);
}

/* gbuffer.comp:194-202 */
void vfo_primary_dir(const float R[3], const float U[3], const float F[3], float thf, int32_t w, int32_t h,
                     int32_t px, int32_t py, const float jit[2], float out[3]) {
    float sx = ((float)px + 0.5f) + jit[0];
    float sy = ((float)py + 0.5f) + jit[1];
    float nx = (sx / (float)w) * 2.0f - 1.0f;
    float ny = (sy / (float)h) * 2.0f - 1.0f;
    float aspect = (float)w / (float)h;
    float a = (nx * aspect) * thf;
    float b = ny * thf;
    for (int i = 0; i < 3; ++i) out[i] = (F[i] + R[i] * a) + U[i] * b;
    normalize3(out);
}

/* probeRayDir (probes.glsl:155-170) with n_rays directions. */
void vfo_probe_dir(int32_t i, int32_t n_rays, uint32_t frame, float out[3]) {
    const float golden = 2.39996323f;
    float z = 1.0f - (2.0f * (float)i + 1.0f) / (float)n_rays;
    float rr = sqrtf(gmax(0.0f, 1.0f - z * z));
    float ph = (float)i * golden;
    float s, c;
    sincos_red(ph, &s, &c);
    float d[3] = {c * rr, z, s * rr};
    uint32_t h = vfo_pcg(frame * 747796405u + 1u);
    float u1 = (float)(h & 0xFFFFu) / 65535.0f, u2 = (float)(h >> 16) / 65535.0f;
    uint32_t h2 = vfo_pcg(h);
    float u3 = (float)(h2 & 0xFFFFu) / 65535.0f;
    float s2, c2, s3, c3;
    sincos_red(6.2831853f * u2, &s2, &c2);
    sincos_red(6.2831853f * u3, &s3, &c3);
    float a = sqrtf(1.0f - u1), b = sqrtf(u1);
    float q[3] = {a * s2, a * c2, b * s3};
    float qw = b * c3;
    float c1v[3], tmp[3], c2v[3];
    cross3(q, d, c1v);
    for (int k = 0; k < 3; ++k) tmp[k] = c1v[k] + qw * d[k];
    cross3(q, tmp, c2v);
    for (int k = 0; k < 3; ++k) out[k] = d[k] + 2.0f * c2v[k];
}

/* ---- camera inside solid terrain (gbuffer.comp:233-266) ---- */
/* voxelMatAt (common.glsl:976-985) */
uint32_t vfo_voxel_mat_at(const vfo_world *w, const float p[3]) {
    return vfo_voxel(w, f2i(floorf(p[0])), f2i(floorf(p[1])), f2i(floorf(p[2])));
}

/* exitSolid (common.glsl:990-1016) */
float vfo_exit_solid(const vfo_world *w, const float ro[3], const float rd[3], int32_t max_steps) {
    int32_t mn[3] = {is_ring(w) ? w->win[0] : 0, 0, is_ring(w) ? w->win[2] : 0};
    float sd[3], inv[3], up[3], t = 0.0f;
    int32_t st[3], v[3];
    for (int a = 0; a < 3; ++a) {
        sd[a] = fabsf(rd[a]) < 1e-5f ? 1e-5f : rd[a];
        inv[a] = 1.0f / sd[a];
        st[a] = sd[a] > 0.0f ? 1 : (sd[a] < 0.0f ? -1 : 0);
        up[a] = sd[a] < 0.0f ? 0.0f : 1.0f;
        v[a] = f2i(floorf(ro[a]));
    }
    for (int32_t i = 0; i < max_steps; ++i) {
        for (int a = 0; a < 3; ++a)
            if (v[a] < mn[a] * 8 || v[a] >= (mn[a] + w->wb[a]) * 8) return -2.0f;
        uint32_t g = w->grid[gidx(w, v[0] >> 3, v[1] >> 3, v[2] >> 3)];
        if (g == 0u || g == 5u) return t;
        int32_t size = 1, lo[3] = {v[0], v[1], v[2]};
        if (g < 256u) {
            size = 8;
            for (int a = 0; a < 3; ++a) lo[a] = (v[a] >> 3) * 8;
        } else {
            uint32_t m = w->atlas[(size_t)(g - 256u) * 512u + (size_t)((v[2] & 7) * 64 + (v[1] & 7) * 8 + (v[0] & 7))];
            if (m == 0u || m == 5u) return t;
        }
        float tf[3];
        for (int a = 0; a < 3; ++a) tf[a] = (((float)lo[a] + up[a] * (float)size) - ro[a]) * inv[a];
        int ax[3];
        ax[0] = tf[0] < tf[1] && tf[0] < tf[2];
        ax[1] = !ax[0] && tf[1] < tf[2];
        ax[2] = !ax[0] && !ax[1];
        t = ax[0] ? tf[0] : (ax[1] ? tf[1] : tf[2]);
        for (int a = 0; a < 3; ++a) {
            int32_t nv = imin(imax(f2i(floorf(ro[a] + sd[a] * t)), lo[a]), lo[a] + (size - 1));
            v[a] = ax[a] ? (st[a] > 0 ? lo[a] + size : lo[a] - 1) : nv;
        }
    }
    return -1.0f;
}

/* gbuffer.comp:233-266: the primary trace (traceScenePrimary: world, then bodies) with the
 * camera-inside-solid pre-step. */
static void primary_trace(const vfo_world *w, const float cam[3], uint32_t cam_mat, const float rd[3], vfo_hit *h,
                          vfo_count *cnt) {
    float ro[3] = {cam[0], cam[1], cam[2]}, tskip = 0.0f;
    int void_ray = 0;
    if (cam_mat != 0u && cam_mat != 5u) {
        tskip = vfo_exit_solid(w, cam, rd, 2048);
        void_ray = tskip < 0.0f;
        tskip = gmax(tskip, 0.0f);
        for (int a = 0; a < 3; ++a) ro[a] = cam[a] + rd[a] * tskip;
    }
    vfo_trace_scene(w, ro, rd, 1e30f, cam_mat == 5u ? 5u : 0u, h, cnt);
    h->t = h->t + tskip;
    if (void_ray) h->hit = 0;
}

/* ---- passes ---- */
static inline uint32_t unorm8(float v) { return (uint32_t)floorf(fclamp(v, 0.0f, 1.0f) * 255.0f + 0.5f); }
static inline float unorm8_dec(uint32_t b) { return ((float)b / 255.0f) * 2.0f - 1.0f; }
/* gbNormal RGBA8 of the hit's world normal (n * 0.5 + 0.5; a NaN component stores 0, as Vulkan's
 * float -> unorm conversion does) with gbNormal.a = matFlag (gbuffer.comp:524-526, WATER_CLEAR on): 0.8 for
 * a foliage material (palette kind 3), else 1.0. */
static uint32_t pack_normal(const vfo_world *w, const vfo_hit *h) {
    float n[3];
    vfo_hit_normal(w, h, n);
    const float mat_flag = (w->palette && w->palette[(size_t)h->mat * 12 + 6] == 3.0f) ? 0.8f : 1.0f;
    uint32_t v = unorm8(mat_flag) << 24;
    for (int a = 0; a < 3; ++a) {
        float x = n[a] * 0.5f + 0.5f;
        v |= (x == x ? unorm8(x) : 0u) << (8 * a);
    }
    return v;
}

void vfo_pass_primary(const vfo_world *w, const vfo_cam *c, int32_t W, int32_t H, vfo_hitrec *hits,
                      float *gb_pos, uint32_t *gb_normal, int32_t *body, vfo_count *cnt, int threads) {
    vfo_count total = {0, 0, 0, 0};
    const uint32_t cam_mat = vfo_voxel_mat_at(w, c->pos);
#pragma omp parallel num_threads(nthreads(threads))
    {
        vfo_count cc = {0, 0, 0, 0};
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < (int64_t)W * H; ++i) {
            int32_t px = (int32_t)(i % W), py = (int32_t)(i / W);
            float rd[3];
            vfo_primary_dir(c->right, c->up, c->fwd, c->tan_half_fov, W, H, px, py, c->jitter, rd);
            vfo_hit h;
            primary_trace(w, c->pos, cam_mat, rd, &h, &cc);
            if (hits) to_rec(&h, &hits[i]);
            if (gb_pos) {
                if (h.hit) {
                    for (int a = 0; a < 3; ++a) gb_pos[i * 4 + a] = c->pos[a] + rd[a] * h.t;
                    gb_pos[i * 4 + 3] = h.t;
                } else {
                    gb_pos[i * 4 + 0] = gb_pos[i * 4 + 1] = gb_pos[i * 4 + 2] = 0.0f;
                    gb_pos[i * 4 + 3] = -1.0f;
                }
            }
            if (gb_normal) gb_normal[i] = h.hit ? pack_normal(w, &h) : 0u;
            if (body) body[i] = h.hit ? h.body - 1 : -1;
        }
#pragma omp critical
        add_cnt(&total, &cc);
    }
    if (cnt) add_cnt(cnt, &total);
}

/* Rows y0, y0 + ystep, ... (nrows of them) of the primary pass of a width x height image: the CPU
 * baseline's bounded sample.  Outputs are packed row after row. */
void vfo_pass_primary_rows(const vfo_world *w, const vfo_cam *c, int32_t W, int32_t H, int32_t y0, int32_t ystep,
                           int32_t nrows, vfo_hitrec *hits, float *gb_pos, uint32_t *gb_normal, vfo_count *cnt,
                           int threads) {
    vfo_count total = {0, 0, 0, 0};
    const uint32_t cam_mat = vfo_voxel_mat_at(w, c->pos);
#pragma omp parallel num_threads(nthreads(threads))
    {
        vfo_count cc = {0, 0, 0, 0};
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < (int64_t)W * nrows; ++i) {
            int32_t px = (int32_t)(i % W), py = y0 + ystep * (int32_t)(i / W);
            float rd[3];
            vfo_primary_dir(c->right, c->up, c->fwd, c->tan_half_fov, W, H, px, py, c->jitter, rd);
            vfo_hit h;
            primary_trace(w, c->pos, cam_mat, rd, &h, &cc);
            to_rec(&h, &hits[i]);
            if (h.hit) {
                for (int a = 0; a < 3; ++a) gb_pos[i * 4 + a] = c->pos[a] + rd[a] * h.t;
                gb_pos[i * 4 + 3] = h.t;
                gb_normal[i] = pack_normal(w, &h);
            } else {
                gb_pos[i * 4 + 0] = gb_pos[i * 4 + 1] = gb_pos[i * 4 + 2] = 0.0f;
                gb_pos[i * 4 + 3] = -1.0f;
                gb_normal[i] = 0;
            }
        }
#pragma omp critical
        add_cnt(&total, &cc);
    }
    if (cnt) add_cnt(cnt, &total);
}

/* isFoliageFlag / isBladeFlag (common.glsl:1953-1954) of the sampled gbNormal.a, and foliageSunOrigin
 * (common.glsl:1987-1989). */
static inline int foliage_flag(float a) { return a > 0.62f && a < 0.88f; }
static inline int blade_flag(float a) { return a > 0.62f && a < 0.75f; }
static void foliage_sun_origin(const float P[3], const float N[3], const float sun[3], int blade, float o[3]) {
    for (int a = 0; a < 3; ++a) o[a] = blade ? P[a] + sun[a] * 0.05f : (P[a] - N[a] * 0.5f) + sun[a] * 0.9f;
}

void vfo_pass_shadow(const vfo_world *w, const float *palette, const float sun[3], float leaf_light, int32_t W,
                     int32_t H, const float *gb_pos, const uint32_t *gb_normal, float *shadow, vfo_hitrec *hits,
                     vfo_count *cnt, int threads) {
    vfo_count total = {0, 0, 0, 0};
#pragma omp parallel num_threads(nthreads(threads))
    {
        vfo_count cc = {0, 0, 0, 0};
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < (int64_t)W * H; ++i) {
            vfo_hit h;
            memset(&h, 0, sizeof h);
            h.t = -1.0f;
            float vis, tb;
            const float *pt = gb_pos + i * 4;
            if (pt[3] < 0.0f) {
                vis = 1.0f; tb = 0.0f;
            } else {
                uint32_t nb = gb_normal[i];
                float N[3] = {unorm8_dec(nb & 255u), unorm8_dec((nb >> 8) & 255u), unorm8_dec((nb >> 16) & 255u)};
                if (dot3(N, sun) <= 0.0f) {
                    vis = 0.0f; tb = 0.0f;
                    /* FOLIAGE_SSS (shadow_trace.comp:35-46): back-lit foliage with the sun above the horizon
                     * gets sunVisibility from foliageSunOrigin, stored without a blocker distance */
                    const float fa = (float)(nb >> 24) / 255.0f;
                    if (foliage_flag(fa) && sun[1] > 0.0f) {
                        float so[3], tf;
                        uint32_t seed = 0u;
                        foliage_sun_origin(pt, N, sun, blade_flag(fa), so);
                        vis = vfo_sun_visibility_t(w, palette, so, sun, leaf_light, &seed, 0, &tf, &h, &cc);
                    }
                } else {
                    /* FOLIAGE_LIGHT (shadow_trace.comp:96-99): sunVisibilityT, deterministic */
                    float so[3];
                    uint32_t seed = 0u;
                    for (int a = 0; a < 3; ++a) so[a] = pt[a] + N[a] * 3e-2f;
                    vis = vfo_sun_visibility_t(w, palette, so, sun, leaf_light, &seed, 0, &tb, &h, &cc);
                }
            }
            shadow[i * 2 + 0] = vis;
            shadow[i * 2 + 1] = gmin(tb, 60000.0f);
            if (hits) to_rec(&h, &hits[i]);
        }
#pragma omp critical
        add_cnt(&total, &cc);
    }
    if (cnt) add_cnt(cnt, &total);
}

/* colLoaded + skyOpen (common.glsl:232-258); a bounded world has no horizon: outside -> 1. */
float vfo_sky_open(const vfo_world *w, const vfo_sky *s, const float p[3]) {
    if (p[1] < 0.0f) return 0.0f;
    int32_t bx = f2i(floorf(p[0] / 8.0f)), bz = f2i(floorf(p[2] / 8.0f));
    int32_t mnx = is_ring(w) ? w->win[0] : 0, mnz = is_ring(w) ? w->win[2] : 0;
    int loaded = bx >= mnx && bz >= mnz && bx < mnx + w->wb[0] && bz < mnz + w->wb[2] &&
                 w->grid[gidx(w, bx, 0, bz)] != 0u;
    if (!loaded) return 1.0f;
    int32_t sx = w->wb[0] * 8, sz = w->wb[2] * 8;
    int32_t cx = f2i(floorf(p[0])), cz = f2i(floorf(p[2]));
    int32_t ix = ((cx % sx) + sx) % sx, iz = ((cz % sz) + sz) % sz;
    float top = (float)s->skytop[(size_t)iz * sx + ix];
    return smoothstepf(top - 0.5f, top + 1.0f, p[1]);
}

/* giEscape (common.glsl:2131-2148) */
static float gi_escape(const vfo_world *w, const vfo_sky *s, const float o[3], const float d[3], float reach) {
    int32_t mn[3] = {is_ring(w) ? w->win[0] : 0, 0, is_ring(w) ? w->win[2] : 0};
    float lo[3], hi[3], tf[3];
    for (int a = 0; a < 3; ++a) {
        lo[a] = (float)(mn[a] * 8);
        hi[a] = (float)((mn[a] + w->wb[a]) * 8);
        float sd = fabsf(d[a]) < 1e-6f ? 1e-6f : d[a];
        float t1 = (lo[a] - o[a]) / sd, t2 = (hi[a] - o[a]) / sd;
        tf[a] = gmax(t1, t2);
    }
    float tExit = gmax(gmin(gmin(tf[0], tf[1]), tf[2]), 0.0f);
    if (tExit > reach) {
        float e[3], e1[3], e2[3];
        for (int a = 0; a < 3; ++a) {
            e[a] = o[a] + d[a] * reach;
        }
        float so = vfo_sky_open(w, s, e);
        if (so > 0.0f) {
            for (int a = 0; a < 3; ++a) { e1[a] = e[a] + d[a] * 16.0f; e2[a] = e[a] + d[a] * 32.0f; }
            so *= vfo_sky_open(w, s, e1) * vfo_sky_open(w, s, e2);
        }
        return so;
    }
    float f[3], f2[3];
    for (int a = 0; a < 3; ++a) f[a] = o[a] + d[a] * tExit;
    if (f[1] <= lo[1] + 1.0f) return 0.0f;
    if (f[1] >= hi[1] - 1.0f) return 1.0f;
    for (int a = 0; a < 3; ++a) f2[a] = f[a] + d[a] * 2.0f;
    return vfo_sky_open(w, s, f2);
}

/* skyColor(rd, sunDisk = 0) (common.glsl:1082-1099) */
static void sky_color0(const vfo_sky *s, const float rd[3], float out[3]) {
    float horizon = fclamp(rd[1] * 0.5f + 0.5f, 0.0f, 1.0f);
    float mu = gmax(dot3(rd, s->sun_dir), 0.0f);
    float sunAmt = 0.0f;   /* pow(mu, 800) * sunDisk(0) */
    float glow = powf(mu, 8.0f) * (1.0f - 0.9f * s->night);
    float ph = powf(horizon, 0.7f);
    const float tint[3] = {1.0f, 0.9f, 0.7f};
    if (s->sky_model == 0) {
        const float a[3] = {0.75f, 0.82f, 0.90f}, b[3] = {0.25f, 0.45f, 0.85f}, nt[3] = {0.012f, 0.02f, 0.045f};
        for (int i = 0; i < 3; ++i) {
            float sky = a[i] * (1.0f - ph) + b[i] * ph;
            sky = sky * (1.0f - s->night) + nt[i] * s->night;
            out[i] = (sky + tint[i] * (sunAmt * 20.0f + glow * 0.25f) * (1.0f - 0.7f * s->night)) * s->sky_scale;
        }
        return;
    }
    /* sunTint = sunRGB / max(max(sunRGB.r, sunRGB.g), 1e-3) */
    float m = gmax(gmax(s->sun_rgb[0], s->sun_rgb[1]), 1e-3f);
    float k1 = powf(mu, 3.0f) * expf(-gmax(rd[1], 0.0f) * 5.0f) + 0.25f * expf(-fabsf(rd[1]) * 9.0f);
    for (int i = 0; i < 3; ++i) {
        float sky = s->sky_horizon[i] * (1.0f - ph) + s->sky_zenith[i] * ph;
        sky += s->sunset_rgb[i] * k1;
        float st = s->sun_rgb[i] / m;
        out[i] = (sky + st * tint[i] * (sunAmt * 20.0f + glow * 0.25f) * (1.0f - 0.7f * s->night)) * s->sky_scale;
    }
}

/* rnd (common.glsl:1246-1251) and cosineHemisphere (common.glsl:1271-1274). */
static float rnd(uint32_t *state) {
    *state = vfo_pcg(*state);
    return (float)(*state >> 8u) * (1.0f / 16777216.0f);
}
void vfo_cosine_hemisphere(const float n[3], uint32_t *state, float out[3]) {
    float u[2];
    u[0] = rnd(state);
    u[1] = rnd(state);
    vfo_cosine_hemisphere_u(n, u, out);
}

/* sunVisibilityT (common.glsl:1964-1981) over the scene (a body always blocks): leaves (palette kind 3) let leaf_light of the
 * light through (stochastic: with that probability), up to 6; the far-horizon and cloud factors are 1
 * (neither is traced here).  first: the first segment's hit; the ray counter counts the ray once. */
float vfo_sun_visibility_t(const vfo_world *w, const float *palette, const float o0[3], const float d[3],
                           float leaf_light, uint32_t *seed, int32_t stochastic, float *tfirst, vfo_hit *first,
                           vfo_count *cnt) {
    float T = 1.0f, o[3] = {o0[0], o0[1], o0[2]};
    *tfirst = 0.0f;
    for (int i = 0; i < 6; ++i) {
        vfo_hit h;
        vfo_trace_scene(w, o, d, 1e30f, 5u, &h, cnt);   /* traceSceneShadow */
        if (cnt && i > 0) cnt->rays--;
        if (i == 0 && first) *first = h;
        if (!h.hit) return T;
        float p[3];
        for (int a = 0; a < 3; ++a) p[a] = o[a] + d[a] * h.t;
        if (i == 0) *tfirst = h.t;
        if (h.body != 0 || palette[(size_t)h.mat * 12 + 6] != 3.0f || leaf_light <= 0.0f) return 0.0f;
        if (stochastic) {
            if (rnd(seed) >= leaf_light) return 0.0f;
        } else {
            T = T * leaf_light;
            if (T < 0.03f) return 0.0f;
        }
        float tx[3];
        for (int a = 0; a < 3; ++a) {
            float sd = fabsf(d[a]) < 1e-6f ? 1e-6f : d[a];
            tx[a] = (((float)h.cell[a] + (sd < 0.0f ? 0.0f : 1.0f)) - p[a]) / sd;
        }
        float tm = gmin(gmin(tx[0], tx[1]), tx[2]) + 1e-3f;
        for (int a = 0; a < 3; ++a) o[a] = p[a] + d[a] * tm;
    }
    return 0.0f;
}

/* sampleCone (common.glsl:1259-1265) */
void vfo_sample_cone(const float dir[3], float ang, uint32_t *state, float out[3]) {
    float t1[3], t2[3];
    make_basis(dir, t1, t2);
    float r = ang * sqrtf(rnd(state));
    float ph = 6.2831853f * rnd(state);
    float s, c;
    sincos_red(ph, &s, &c);
    for (int a = 0; a < 3; ++a) out[a] = (dir[a] + t1[a] * (r * c)) + t2[a] * (r * s);
    normalize3(out);
}

/* lighting.comp's sun sample (lighting.comp:44-84): 1 sky; a lit texel, or a back-lit foliage texel with the
 * sun above the horizon (FOLIAGE_SSS, from foliageSunOrigin), gets a stochastic sunVisibility along
 * sampleCone(sunDir, sun_angle), seed pixelSeed(pix, frame, 0); any other texel facing away 0. */
void vfo_pass_sun_cone(const vfo_world *w, const float *palette, const float sun[3], float sun_angle, float leaf_light,
                       uint32_t frame, int32_t W, int32_t H, const float *gb_pos, const uint32_t *gb_normal, float *out,
                       vfo_hitrec *hits, int threads) {
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads(threads))
    for (int64_t i = 0; i < (int64_t)W * H; ++i) {
        int32_t px = (int32_t)(i % W), py = (int32_t)(i / W);
        vfo_hit h;
        memset(&h, 0, sizeof h);
        h.t = -1.0f;
        float v = 1.0f;
        const float *pt = gb_pos + i * 4;
        if (pt[3] >= 0.0f) {
            uint32_t nb = gb_normal[i];
            float N[3] = {unorm8_dec(nb & 255u), unorm8_dec((nb >> 8) & 255u), unorm8_dec((nb >> 16) & 255u)};
            uint32_t seed = vfo_pixel_seed(px, py, frame, 0u);
            v = 0.0f;
            const int lit = gmax(dot3(N, sun), 0.0f) > 0.0f;
            int ray = lit;
            float so[3];
            for (int a = 0; a < 3; ++a) so[a] = pt[a] + N[a] * 3e-2f;
            if (!lit) {   /* FOLIAGE_SSS (lighting.comp:66-71) */
                const float fa = (float)(nb >> 24) / 255.0f;
                if (foliage_flag(fa) && sun[1] > 0.0f) {
                    ray = 1;
                    foliage_sun_origin(pt, N, sun, blade_flag(fa), so);
                }
            }
            if (ray) {
                float sd[3], tf;
                vfo_sample_cone(sun, sun_angle, &seed, sd);
                v = vfo_sun_visibility_t(w, palette, so, sd, leaf_light, &seed, 1, &tf, &h, NULL);
            }
        }
        out[i] = v;
        if (hits) to_rec(&h, &hits[i]);
    }
}

/* giPathImpl (common.glsl:2250-2366) with GI_OBJECTS, without NEE lamps or the hash cache, for lastStage 1 (one
 * bounce) or 3 (two bounces); traceOn = the rising-ray trace-on past the reach (common.glsl:2309-2320).
 * seed: the path's RNG state (bounce 2's direction, cosineHemisphere(hitN, seed)).  With tail non-null
 * the path stops where bounce 2 would start (after `thr *= hitAlb * giScale`, common.glsl:2277 and
 * :2339) and reports that point; L then holds the light gathered before it. */
static void gi_path_impl(const vfo_world *w, const vfo_sky *s, const float P[3], const float N[3], const float dir[3],
                         uint32_t *seed, int32_t last_stage, int32_t trace_on, float L[3], vfo_hit *bounce, vfo_tail *tail,
                         int tail_any, vfo_count *cnt) {
    float thr[3] = {1.0f, 1.0f, 1.0f}, ro[3], rdir[3], hitAlb[3] = {0, 0, 0}, hitP[3] = {0, 0, 0};
    float hitN[3] = {0.0f, 1.0f, 0.0f}, hitD[3] = {0, 0, 0};
    float hitNdl = 0.0f, reach = s->gi_reach;
    int stage = 0, extended = 0, hit_body = 0;
    L[0] = L[1] = L[2] = 0.0f;
    if (bounce) { memset(bounce, 0, sizeof *bounce); bounce->t = -1.0f; }
    if (tail) memset(tail, 0, sizeof *tail);
    for (int a = 0; a < 3; ++a) { ro[a] = P[a] + N[a] * 3e-2f; rdir[a] = dir[a]; }
    for (int it = 0; it < 6; ++it) {
        int sunRay = (stage & 1) == 1;
        vfo_hit h;
        vfo_trace(w, ro, rdir, sunRay ? 1e30f : reach, 0u, &h, cnt);
        if (!(sunRay && h.hit)) {   /* GI_OBJECTS (common.glsl:2267-2271) */
            if (!sunRay) h.body = 0;
            vfo_trace_objects(w, ro, rdir, &h, sunRay ? 1e30f : gmin(h.hit ? h.t : reach, 16.0f), sunRay, cnt);
        }
        int next_bounce = 0;
        if (sunRay) {
            if (!h.hit)
                for (int a = 0; a < 3; ++a) L[a] += thr[a] * hitAlb[a] * s->sun_rgb[a] * hitNdl;
            if (stage >= last_stage) break;
            next_bounce = 1;
        } else {
            if (stage == 0 && bounce) *bounce = h;
            if (h.hit && h.mat == 5u) h.hit = 0;   /* WATER_CLEAR: water ends a bounce ray as a miss */
            if (!h.hit) {
                float esc = gi_escape(w, s, ro, rdir, reach);
                if (trace_on && esc < 0.999f && !extended && stage == 0 && rdir[1] > 0.3f) {
                    float e[3], p1[3], p2[3], p3[3];
                    for (int a = 0; a < 3; ++a) {
                        e[a] = ro[a] + rdir[a] * reach;
                        p1[a] = e[a] + rdir[a] * 32.0f;
                        p2[a] = e[a] + rdir[a] * 64.0f;
                        p3[a] = e[a] + rdir[a] * 96.0f;
                    }
                    float m = gmax(vfo_sky_open(w, s, p1), gmax(vfo_sky_open(w, s, p2), vfo_sky_open(w, s, p3)));
                    if (m > 0.0f) {
                        for (int a = 0; a < 3; ++a) ro[a] = e[a];
                        reach = 96.0f;
                        extended = 1;
                        continue;
                    }
                }
                float sky[3];
                sky_color0(s, rdir, sky);
                for (int a = 0; a < 3; ++a) L[a] += thr[a] * gmin(sky[a], 3.0f) * esc;
                break;
            }
            const float *mat = s->palette + (size_t)h.mat * 12;
            float emit = mat[3] * 8.0f;
            if (s->lights_nee && mat[3] >= 0.5f) emit = 0.0f;
            vfo_hit_normal(w, &h, hitN);
            if (stage == 0) hit_body = h.body != 0;   /* common.glsl:2327 */
            for (int a = 0; a < 3; ++a) {
                L[a] += thr[a] * mat[a] * emit;
                hitP[a] = ro[a] + rdir[a] * h.t;
                hitD[a] = rdir[a];
                hitAlb[a] = mat[a];
            }
            reach = s->gi_reach;
            extended = 0;
            hitNdl = gmax(dot3(hitN, s->sun_dir), 0.0f);
            if (hitNdl > 0.0f) {
                for (int a = 0; a < 3; ++a) { ro[a] = hitP[a] + hitN[a] * 3e-2f; rdir[a] = s->sun_dir[a]; }
                stage += 1;
                continue;
            }
            if (stage + 1 >= last_stage) break;
            next_bounce = 1;
        }
        if (next_bounce) {   /* bounce 2 from the first hit */
            for (int a = 0; a < 3; ++a) thr[a] = thr[a] * (hitAlb[a] * s->gi_scale);
            if (tail && (tail_any || !hit_body)) {   /* a body's bounce-1 hit is never a cache query (common.glsl:2283) */
                tail->valid = 1;
                tail->body = hit_body;
                for (int a = 0; a < 3; ++a) {
                    tail->p[a] = hitP[a];
                    tail->n[a] = hitN[a];
                    tail->dir[a] = hitD[a];
                    tail->thr[a] = thr[a];
                }
                break;
            }
            vfo_cosine_hemisphere(hitN, seed, rdir);
            for (int a = 0; a < 3; ++a) ro[a] = hitP[a] + hitN[a] * 3e-2f;
            stage = 2;
            reach = s->gi_reach;
            extended = 0;
        }
    }
}

void vfo_gi_path_ex(const vfo_world *w, const vfo_sky *s, const float P[3], const float N[3], const float dir[3],
                    uint32_t *seed, int32_t last_stage, int32_t trace_on, float L[3], vfo_hit *bounce, vfo_tail *tail,
                    vfo_count *cnt) {
    gi_path_impl(w, s, P, N, dir, seed, last_stage, trace_on, L, bounce, tail, 0, cnt);
}

/* The rest of a path from its tail: a two-bounce path from the bounce-1 hit (its first direction the next
 * cosine draw of the seed, no trace-on), added with the tail's throughput: L += thr * L_rest. */
static void continue_tail(const vfo_world *w, const vfo_sky *s, const vfo_tail *t, uint32_t *seed, float L[3],
                          vfo_count *cnt) {
    float d2[3], L2[3];
    vfo_cosine_hemisphere(t->n, seed, d2);
    gi_path_impl(w, s, t->p, t->n, d2, seed, 3, 0, L2, NULL, NULL, 0, cnt);
    for (int a = 0; a < 3; ++a) L[a] += t->thr[a] * L2[a];
}

void vfo_gi_path3(const vfo_world *w, const vfo_sky *s, const float P[3], const float N[3], const float dir[3],
                  uint32_t *seed, float L[3], vfo_count *cnt) {
    vfo_tail t;
    gi_path_impl(w, s, P, N, dir, seed, 3, 1, L, NULL, &t, 1, cnt);
    if (t.valid) continue_tail(w, s, &t, seed, L, cnt);
}

void vfo_gi_path(const vfo_world *w, const vfo_sky *s, const float P[3], const float N[3], const float dir[3],
                 float L[3], vfo_hit *bounce, vfo_count *cnt) {
    uint32_t seed = 0u;
    vfo_gi_path_ex(w, s, P, N, dir, &seed, 1, 1, L, bounce, NULL, cnt);
}

static void gbuffer_point(const float *gb_pos, const uint32_t *gb_normal, size_t i, float P[3], float N[3]) {
    uint32_t nb = gb_normal[i];
    for (int a = 0; a < 3; ++a) P[a] = gb_pos[i * 4 + a];
    N[0] = unorm8_dec(nb & 255u);
    N[1] = unorm8_dec((nb >> 8) & 255u);
    N[2] = unorm8_dec((nb >> 16) & 255u);
}

/* Bounce-1 direction and path seed of sample index k of pixel (px, py): gi.comp:62-72 with the
 * sample index in place of the frame. */
static void gi_sample_start(const float N[3], int32_t px, int32_t py, uint32_t k, uint32_t salt, float dir[3],
                            uint32_t *seed) {
    float u[2];
    vfo_ld_sample2(px, py, k, salt, u);
    vfo_cosine_hemisphere_u(N, u, dir);
    *seed = vfo_pixel_seed(px, py, k, salt);
}

void vfo_pass_gi_path(const vfo_world *w, const vfo_sky *s, int32_t W, int32_t H, const float *gb_pos,
                      const uint32_t *gb_normal, uint32_t frame, int32_t spp, uint32_t salt, int32_t bounces, float *out,
                      int threads) {
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads(threads))
    for (int64_t i = 0; i < (int64_t)W * H; ++i) {
        int32_t px = (int32_t)(i % W), py = (int32_t)(i / W);
        float sum[3] = {0, 0, 0};
        int sky = gb_pos[i * 4 + 3] < 0.0f;
        if (!sky) {
            float P[3], N[3];
            gbuffer_point(gb_pos, gb_normal, (size_t)i, P, N);
            for (int32_t sp = 0; sp < spp; ++sp) {
                float dir[3], L[3];
                uint32_t seed;
                gi_sample_start(N, px, py, frame * (uint32_t)spp + (uint32_t)sp, salt, dir, &seed);
                if (bounces == 3) vfo_gi_path3(w, s, P, N, dir, &seed, L, NULL);
                else vfo_gi_path_ex(w, s, P, N, dir, &seed, 3, 1, L, NULL, NULL, NULL);
                for (int a = 0; a < 3; ++a) sum[a] += L[a];
            }
        }
        for (int a = 0; a < 3; ++a) out[i * 4 + a] = sky ? 0.0f : sum[a] / (float)spp;
        out[i * 4 + 3] = sky ? 0.0f : 1.0f;
    }
}

/* The cache render's traced part with the tail at the bounce-1 hit: per pixel pre = the light gathered
 * before the tail summed over the spp samples (a = 1, or 0 for sky), and per slot pixel*spp + s the
 * query record (tag = slot, or 0xFFFFFFFF when the sample has no tail). */
void vfo_gicache_prepass(const vfo_world *w, const vfo_sky *s, int32_t W, int32_t H, const float *gb_pos,
                         const uint32_t *gb_normal, uint32_t frame, int32_t spp, uint32_t salt, int32_t bounces,
                         float *pre, vfo_record *queries, int threads) {
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads(threads))
    for (int64_t i = 0; i < (int64_t)W * H; ++i) {
        int32_t px = (int32_t)(i % W), py = (int32_t)(i / W);
        float sum[3] = {0, 0, 0};
        int sky = gb_pos[i * 4 + 3] < 0.0f;
        float P[3] = {0, 0, 0}, N[3] = {0, 0, 0};
        if (!sky) gbuffer_point(gb_pos, gb_normal, (size_t)i, P, N);
        for (int32_t sp = 0; sp < spp; ++sp) {
            vfo_record *q = &queries[i * spp + sp];
            memset(q, 0, sizeof *q);
            q->tag = 0xFFFFFFFFu;
            if (sky) continue;
            float dir[3], L[3];
            uint32_t seed;
            vfo_tail t;
            gi_sample_start(N, px, py, frame * (uint32_t)spp + (uint32_t)sp, salt, dir, &seed);
            gi_path_impl(w, s, P, N, dir, &seed, 3, 1, L, NULL, &t, bounces == 3, NULL);
            if (t.valid && t.body) {   /* three bounces: a body's bounce-1 hit is traced on, never queried */
                continue_tail(w, s, &t, &seed, L, NULL);
                t.valid = 0;
            }
            for (int a = 0; a < 3; ++a) sum[a] += L[a];
            if (t.valid) {
                q->tag = (uint32_t)(i * spp + sp);
                for (int a = 0; a < 3; ++a) {
                    q->pos[a] = t.p[a];
                    q->n[a] = t.n[a];
                    q->view[a] = -t.dir[a];
                    q->rgb[a] = t.thr[a];
                }
            }
        }
        for (int a = 0; a < 3; ++a) pre[i * 4 + a] = sum[a];
        pre[i * 4 + 3] = sky ? 0.0f : 1.0f;
    }
}

/* Training samples of one frame: sample i takes pixel floor((i + r) * W*H / n) (r = rnd of its own stream
 * pcg(i + pcg(frame ^ salt))), a cosine-hemisphere bounce-1 direction from the same stream and the
 * two-bounce path.  terminate 0: the record is the G-buffer point (view = towards the camera) with the
 * path's radiance; terminate 1: the bounce-1 hit (view = back along bounce 1) with the radiance of the
 * path's remainder from there (giPathImpl(hit, n, cosineHemisphere(n, seed), seed, 1, false): bounce 2 and
 * its sun ray).  bounces 3: the path is vfo_gi_path3 and the remainder has two bounces (lastStage 3).  tag = 0
 * for a sky pixel or a sample without that hit. */
void vfo_gicache_samples(const vfo_world *w, const vfo_sky *s, const float cam_pos[3], int32_t W, int32_t H,
                         const float *gb_pos, const uint32_t *gb_normal, uint32_t frame, int32_t n,
                         int32_t terminate, uint32_t salt, int32_t bounces, vfo_record *out, int threads) {
    const float stratum = (float)((int64_t)W * H) / (float)n;
    const uint32_t base = vfo_pcg(frame ^ salt);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads(threads))
    for (int32_t i = 0; i < n; ++i) {
        vfo_record *r = &out[i];
        memset(r, 0, sizeof *r);
        uint32_t seed = vfo_pcg((uint32_t)i + base);
        float u = rnd(&seed);
        int64_t pix = (int64_t)floorf(((float)i + u) * stratum);
        if (pix > (int64_t)W * H - 1) pix = (int64_t)W * H - 1;
        if (gb_pos[pix * 4 + 3] < 0.0f) continue;
        float P[3], N[3], dir[3], L[3];
        gbuffer_point(gb_pos, gb_normal, (size_t)pix, P, N);
        vfo_cosine_hemisphere(N, &seed, dir);
        if (terminate == 0) {
            if (bounces == 3) vfo_gi_path3(w, s, P, N, dir, &seed, L, NULL);
            else vfo_gi_path_ex(w, s, P, N, dir, &seed, 3, 1, L, NULL, NULL, NULL);
            float v[3] = {cam_pos[0] - P[0], cam_pos[1] - P[1], cam_pos[2] - P[2]};
            normalize3(v);
            r->tag = 1u;
            for (int a = 0; a < 3; ++a) { r->pos[a] = P[a]; r->n[a] = N[a]; r->view[a] = v[a]; r->rgb[a] = L[a]; }
            continue;
        }
        vfo_tail t;
        vfo_gi_path_ex(w, s, P, N, dir, &seed, 3, 1, L, NULL, &t, NULL);
        if (!t.valid) continue;
        float d2[3], L2[3];
        vfo_cosine_hemisphere(t.n, &seed, d2);
        vfo_gi_path_ex(w, s, t.p, t.n, d2, &seed, bounces == 3 ? 3 : 1, 0, L2, NULL, NULL, NULL);
        r->tag = 1u;
        for (int a = 0; a < 3; ++a) { r->pos[a] = t.p[a]; r->n[a] = t.n[a]; r->view[a] = -t.dir[a]; r->rgb[a] = L2[a]; }
    }
}

void vfo_pass_gi(const vfo_world *w, const vfo_sky *s, uint32_t frame, int32_t W, int32_t H, int32_t half_res,
                 const float *gb_pos, const uint32_t *gb_normal, float *radiance, vfo_hitrec *bounce,
                 vfo_count *cnt, int threads) {
    const int32_t OW = half_res ? (W + 1) / 2 : W, OH = half_res ? (H + 1) / 2 : H;
    const uint32_t salt = 0x68e31da4u;
    vfo_count total = {0, 0, 0, 0};
#pragma omp parallel num_threads(nthreads(threads))
    {
        vfo_count cc = {0, 0, 0, 0};
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < (int64_t)OW * OH; ++i) {
            int32_t px = (int32_t)(i % OW), py = (int32_t)(i / OW);
            int32_t fx = px, fy = py;
            vfo_hit bh;
            memset(&bh, 0, sizeof bh);
            bh.t = -1.0f;
            float Lo[3] = {0, 0, 0};
            int sky = 0;
            if (half_res) {
                /* rotating sub-texel (gi.comp:43-44) and repTexel (common.glsl:1213-1219) */
                int32_t ox = (int32_t)(frame & 1u), oy = (int32_t)((frame >> 1u) & 1u);
                fx = imin(px * 2 + ox, W - 1);
                fy = imin(py * 2 + oy, H - 1);
                int32_t gx = imin(px * 2, W - 1), gy = imin(py * 2, H - 1);
                for (int k = 0; k < 4; ++k) {
                    int32_t qx = imin(px * 2 + (k & 1), W - 1), qy = imin(py * 2 + (k >> 1), H - 1);
                    if (gb_pos[((size_t)qy * W + qx) * 4 + 3] >= 0.0f) { gx = qx; gy = qy; break; }
                }
                if (gb_pos[((size_t)gy * W + gx) * 4 + 3] < 0.0f) sky = 1;
                else if (gb_pos[((size_t)fy * W + fx) * 4 + 3] < 0.0f) { fx = gx; fy = gy; }
            } else if (gb_pos[((size_t)fy * W + fx) * 4 + 3] < 0.0f) {
                sky = 1;
            }
            if (!sky) {
                size_t fi = (size_t)fy * W + fx;
                const float *P = gb_pos + fi * 4;
                uint32_t nb = gb_normal[fi];
                float N[3] = {unorm8_dec(nb & 255u), unorm8_dec((nb >> 8) & 255u), unorm8_dec((nb >> 16) & 255u)};
                float u[2], dir[3];
                vfo_ld_sample2(fx, fy, frame >> 2u, salt, u);
                vfo_cosine_hemisphere_u(N, u, dir);
                vfo_gi_path(w, s, P, N, dir, Lo, &bh, &cc);
            }
            radiance[i * 4 + 0] = Lo[0];
            radiance[i * 4 + 1] = Lo[1];
            radiance[i * 4 + 2] = Lo[2];
            radiance[i * 4 + 3] = sky ? 0.0f : 1.0f;
            if (bounce) to_rec(&bh, &bounce[i]);
        }
#pragma omp critical
        add_cnt(&total, &cc);
    }
    if (cnt) add_cnt(cnt, &total);
}

/* The GI rays of one full-resolution G-buffer (gi.comp:55-79 with fpix = pix): for each non-sky pixel,
 * origin P + N*0.03 and direction cosineHemisphereU(N, ldSample2(pix, frame >> 2, 0x68e31da4)) from
 * the quantised normal; rays[i] for pixel i, valid[i] = 0 for sky. */
void vfo_gi_rays(int32_t W, int32_t H, uint32_t frame, const float *gb_pos, const uint32_t *gb_normal, vfo_ray *rays,
                 uint8_t *valid) {
    for (int64_t i = 0; i < (int64_t)W * H; ++i) {
        int32_t px = (int32_t)(i % W), py = (int32_t)(i / W);
        const float *P = gb_pos + i * 4;
        valid[i] = P[3] >= 0.0f;
        memset(&rays[i], 0, sizeof(rays[i]));
        if (!valid[i]) continue;
        uint32_t nb = gb_normal[i];
        float N[3] = {unorm8_dec(nb & 255u), unorm8_dec((nb >> 8) & 255u), unorm8_dec((nb >> 16) & 255u)};
        float u[2];
        vfo_ld_sample2(px, py, frame >> 2u, 0x68e31da4u, u);
        vfo_cosine_hemisphere_u(N, u, rays[i].d);
        for (int a = 0; a < 3; ++a) rays[i].o[a] = P[a] + N[a] * 3e-2f;
    }
}
