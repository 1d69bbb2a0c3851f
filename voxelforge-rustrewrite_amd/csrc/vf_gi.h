// Device code of the GI paths (shared by the GI pass, the two-bounce path and the learned cache):
// skyOpen, giEscape, skyColor, the PCG streams and giPathImpl.
#pragma once

#include "vf_bodies.h"
#include "vf_device.h"

// skyOpen with colLoaded (common.glsl:232-258); a bounded world has no horizon (outside -> 1).
__device__ __forceinline__ float sky_open(const DevWorld &W, const float p[3]) {
    if (p[1] < 0.0f) return 0.0f;
    const int32_t bx = f2i(floorf(p[0] / 8.0f)), bz = f2i(floorf(p[2] / 8.0f));
    const bool inside = bx >= W.minb[0] && bz >= W.minb[2] && bx < W.minb[0] + W.wb[0] && bz < W.minb[2] + W.wb[2];
    if (!inside || W.cells[cell_index(W, bx, 0, bz)].x == 0u) return 1.0f;
    const int32_t sxn = W.wb[0] * 8, szn = W.wb[2] * 8;
    const int32_t cx = f2i(floorf(p[0])), cz = f2i(floorf(p[2]));
    const int32_t ix = ((cx % sxn) + sxn) % sxn, iz = ((cz % szn) + szn) % szn;
    const float top = (float)W.skytop[(size_t)iz * (size_t)sxn + (size_t)ix];
    const float e0 = top - 0.5f, e1 = top + 1.0f;
    const float tt = fclampg((p[1] - e0) / (e1 - e0), 0.0f, 1.0f);
    return tt * tt * (3.0f - 2.0f * tt);
}

// giEscape (common.glsl:2131-2148)
__device__ float gi_escape(const DevWorld &W, const float o[3], const float d[3], float reach) {
    float lo[3], hi[3], tf[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = (float)(W.minb[a] * 8);
        hi[a] = (float)((W.minb[a] + W.wb[a]) * 8);
        const float s = fabsf(d[a]) < 1e-6f ? 1e-6f : d[a];
        const float t1 = (lo[a] - o[a]) / s, t2 = (hi[a] - o[a]) / s;
        tf[a] = gmax(t1, t2);
    }
    const float tExit = gmax(gmin(gmin(tf[0], tf[1]), tf[2]), 0.0f);
    if (tExit > reach) {
        float e[3], e1[3], e2[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) e[a] = o[a] + d[a] * reach;
        float so = sky_open(W, e);
        if (so > 0.0f) {
#pragma unroll
            for (int a = 0; a < 3; ++a) { e1[a] = e[a] + d[a] * 16.0f; e2[a] = e[a] + d[a] * 32.0f; }
            so *= sky_open(W, e1) * sky_open(W, e2);
        }
        return so;
    }
    float f[3], f2[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) f[a] = o[a] + d[a] * tExit;
    if (f[1] <= lo[1] + 1.0f) return 0.0f;
    if (f[1] >= hi[1] - 1.0f) return 1.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) f2[a] = f[a] + d[a] * 2.0f;
    return sky_open(W, f2);
}

// skyColor(rd, sunDisk 0) (common.glsl:1082-1099)
__device__ void sky_color0(const SkyParams &S, const float rd[3], float out[3]) {
    const float horizon = fclampg(rd[1] * 0.5f + 0.5f, 0.0f, 1.0f);
    const float mu = gmax(dot3(rd, S.sun_dir), 0.0f);
    const float glow = powf(mu, 8.0f) * (1.0f - 0.9f * S.night);
    const float ph = powf(horizon, 0.7f);
    const float tint[3] = {1.0f, 0.9f, 0.7f};
    if (S.sky_model == 0) {
        const float a[3] = {0.75f, 0.82f, 0.90f}, b[3] = {0.25f, 0.45f, 0.85f}, nt[3] = {0.012f, 0.02f, 0.045f};
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            float sky = a[i] * (1.0f - ph) + b[i] * ph;
            sky = sky * (1.0f - S.night) + nt[i] * S.night;
            out[i] = (sky + tint[i] * (glow * 0.25f) * (1.0f - 0.7f * S.night)) * S.sky_scale;
        }
        return;
    }
    const float m = gmax(gmax(S.sun_rgb[0], S.sun_rgb[1]), 1e-3f);
    const float k1 = powf(mu, 3.0f) * expf(-gmax(rd[1], 0.0f) * 5.0f) + 0.25f * expf(-fabsf(rd[1]) * 9.0f);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        float sky = S.sky_horizon[i] * (1.0f - ph) + S.sky_zenith[i] * ph;
        sky += S.sunset_rgb[i] * k1;
        out[i] = (sky + (S.sun_rgb[i] / m) * tint[i] * (glow * 0.25f) * (1.0f - 0.7f * S.night)) * S.sky_scale;
    }
}

// rnd (common.glsl:1246-1251)
__device__ __forceinline__ float rnd(uint32_t &state) {
    state = pcg(state);
    return (float)(state >> 8u) * (1.0f / 16777216.0f);
}
// cosineHemisphere (common.glsl:1271-1274): u first, then v.
__device__ __forceinline__ void cosine_hemisphere(const float n[3], uint32_t &state, float out[3]) {
    const float u0 = rnd(state);
    const float u1 = rnd(state);
    cosine_hemisphere_u(n, u0, u1, out);
}
// pixelSeed (common.glsl:1242)
__device__ __forceinline__ uint32_t pixel_seed(int32_t px, int32_t py, uint32_t frame, uint32_t salt) {
    return pcg((uint32_t)px + pcg((uint32_t)py + pcg(frame ^ salt)));
}
// ldSample2 (common.glsl:1284-1288)
__device__ __forceinline__ void ld_sample2(int32_t px, int32_t py, uint32_t k, uint32_t salt, float &u0, float &u1) {
    const uint32_t hh = pcg((uint32_t)px + pcg((uint32_t)py + salt));
    const float rx = (float)(hh >> 8u) * (1.0f / 16777216.0f);
    const float ry = (float)(pcg(hh) >> 8u) * (1.0f / 16777216.0f);
    const float kf = (float)(k & 0xFFFFu);
    u0 = fractg(rx + kf * 0.7548776662f);
    u1 = fractg(ry + kf * 0.5698402910f);
}

// Where a two-bounce path's bounce 2 would start: the bounce-1 hit, its normal, the bounce-1 ray's
// direction and the throughput after `thr *= hitAlb * giScale`.
struct GiTail {
    bool valid;
    bool body;   // the bounce-1 hit is a body (reported with TAIL_ANY only)
    float p[3], n[3], dir[3], thr[3];
};

struct GiPathCfg {
    float reach;        // GI_REACH
    int32_t lights_nee; // LIGHTS_NEE
};

// giPathImpl (common.glsl:2250-2366) without lamps' NEE or the hash cache.  LAST_STAGE 1: one
// bounce (+ its sun ray); 3: two bounces.  trace_on: the rising-ray trace-on past the reach
// (common.glsl:2309-2320).  TAIL: stop where bounce 2 would start and report it in tail, unless the
// bounce-1 hit is a body (a moving surface is never answered by a cache, common.glsl:2327, 2283): that
// path is traced to its end; with TAIL_ANY it stops there too).  BODIES (GI_OBJECTS, common.glsl:2267-2271): a
// bounce ray also meets the bodies nearer than min(its world hit or reach, 16), a sun ray that missed the world
// any body.
template <int LAST_STAGE, bool TAIL, bool COUNT, bool BODIES = false, bool LDSM = false, bool TAIL_ANY = false>
__device__ void gi_path(const DevWorld &W, const SkyParams &S, const GiPathCfg &G, const float P[3], const float N[3],
                        const float dir0[3], uint32_t &seed, bool trace_on, float L[3], Hit *bounce, GiTail *tail,
                        Cnt &c, const DevBodies *B = nullptr) {
    float thr[3] = {1.0f, 1.0f, 1.0f}, ro[3], rdir[3];
    float hitAlb[3] = {0.0f, 0.0f, 0.0f}, hitP[3] = {0.0f, 0.0f, 0.0f}, hitN[3] = {0.0f, 1.0f, 0.0f};
    float hitD[3] = {0.0f, 0.0f, 0.0f};
    float hitNdl = 0.0f, reach = G.reach;
    int stage = 0;
    bool extended = false, hit_body = false;
    L[0] = L[1] = L[2] = 0.0f;
    if (TAIL) {
        tail->valid = false;
        tail->body = false;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) { ro[a] = P[a] + N[a] * 3e-2f; rdir[a] = dir0[a]; }
    for (int it = 0; it < 6; ++it) {
        const bool sunRay = (stage & 1) == 1;
        Hit h;
        trace_world<0, COUNT, LDSM>(W, ro, rdir, sunRay ? 1e30f : reach, h, c);
        if (BODIES) {
            h.obj = -1;
            if (!(sunRay && h.hit))
                trace_objects<COUNT>(W, *B, ro, rdir, h, sunRay ? 1e30f : gmin(h.hit ? h.t : reach, 16.0f), sunRay, c);
        }
        if (sunRay) {
            if (!h.hit) {
#pragma unroll
                for (int a = 0; a < 3; ++a) L[a] += thr[a] * hitAlb[a] * S.sun_rgb[a] * hitNdl;
            }
            if (stage >= LAST_STAGE) break;
        } else {
            if (stage == 0 && bounce) *bounce = h;
            if (h.hit && h.mat == 5u) h.hit = 0;   // WATER_CLEAR: water ends a bounce ray as a miss
            if (!h.hit) {
                const float esc = gi_escape(W, ro, rdir, reach);
                if (trace_on && esc < 0.999f && !extended && stage == 0 && rdir[1] > 0.3f) {
                    float e[3], p1[3], p2[3], p3[3];
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        e[a] = ro[a] + rdir[a] * reach;
                        p1[a] = e[a] + rdir[a] * 32.0f;
                        p2[a] = e[a] + rdir[a] * 64.0f;
                        p3[a] = e[a] + rdir[a] * 96.0f;
                    }
                    const float m = gmax(sky_open(W, p1), gmax(sky_open(W, p2), sky_open(W, p3)));
                    if (m > 0.0f) {
#pragma unroll
                        for (int a = 0; a < 3; ++a) ro[a] = e[a];
                        reach = 96.0f;
                        extended = true;
                        continue;
                    }
                }
                float skyc[3];
                sky_color0(S, rdir, skyc);
#pragma unroll
                for (int a = 0; a < 3; ++a) L[a] += thr[a] * gmin(skyc[a], 3.0f) * esc;
                break;
            }
            const float4 alb4 = W.palette[h.mat * 3u];
            float emit = alb4.w * 8.0f;
            if (G.lights_nee && alb4.w >= 0.5f) emit = 0.0f;
            const float alb[3] = {alb4.x, alb4.y, alb4.z};
            hit_normal<BODIES>(B, h, hitN);
            if (BODIES && stage == 0) hit_body = h.obj >= 0;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                L[a] += thr[a] * alb[a] * emit;
                hitP[a] = ro[a] + rdir[a] * h.t;
                hitD[a] = rdir[a];
                hitAlb[a] = alb[a];
            }
            reach = G.reach;
            extended = false;
            hitNdl = gmax(dot3(hitN, S.sun_dir), 0.0f);
            if (hitNdl > 0.0f) {   // its direct sun next
#pragma unroll
                for (int a = 0; a < 3; ++a) { ro[a] = hitP[a] + hitN[a] * 3e-2f; rdir[a] = S.sun_dir[a]; }
                stage += 1;
                continue;
            }
            if (stage + 1 >= LAST_STAGE) break;
        }
        // bounce 2 from the first hit
#pragma unroll
        for (int a = 0; a < 3; ++a) thr[a] = thr[a] * (hitAlb[a] * S.gi_scale);
        if (TAIL && (TAIL_ANY || !hit_body)) {
            tail->valid = true;
            tail->body = hit_body;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                tail->p[a] = hitP[a];
                tail->n[a] = hitN[a];
                tail->dir[a] = hitD[a];
                tail->thr[a] = thr[a];
            }
            break;
        }
        cosine_hemisphere(hitN, seed, rdir);
#pragma unroll
        for (int a = 0; a < 3; ++a) ro[a] = hitP[a] + hitN[a] * 3e-2f;
        stage = 2;
        reach = G.reach;
        extended = false;
    }
}

// A three-bounce path (the learned cache's three-bounce samples; the reference's giPath has two bounces at
// most): the two-bounce path up to where its bounce 2 would start, then a two-bounce path from there (bounce 2
// as a path's first bounce, its direction the next draw of the same seed, never traced on past the reach), its
// light weighted by the throughput reached: L = L_head + thr * L_rest.
template <bool COUNT, bool BODIES = false, bool LDSM = false>
__device__ void gi_path3(const DevWorld &W, const SkyParams &S, const GiPathCfg &G, const float P[3], const float N[3],
                         const float dir0[3], uint32_t &seed, float L[3], Cnt &c, const DevBodies *B = nullptr) {
    GiTail t;
    gi_path<3, true, COUNT, BODIES, LDSM, true>(W, S, G, P, N, dir0, seed, true, L, nullptr, &t, c, B);
    if (!t.valid) return;
    float d2[3], L2[3];
    cosine_hemisphere(t.n, seed, d2);
    gi_path<3, false, COUNT, BODIES, LDSM>(W, S, G, t.p, t.n, d2, seed, false, L2, nullptr, nullptr, c, B);
#pragma unroll
    for (int a = 0; a < 3; ++a) L[a] += t.thr[a] * L2[a];
}

// sampleCone (common.glsl:1259-1265): dir jittered within a cone of angular radius ang.
__device__ __forceinline__ void sample_cone(const float dir[3], float ang, uint32_t &state, float out[3]) {
    float t1[3], t2[3];
    const float upv[3] = {0.0f, 1.0f, 0.0f}, uxv[3] = {1.0f, 0.0f, 0.0f};
    cross3(dir, fabsf(dir[1]) < 0.9f ? upv : uxv, t1);
    normalize3(t1);
    cross3(dir, t1, t2);
    const float r = ang * sqrtf(rnd(state));
    const float ph = 6.2831853f * rnd(state);
    float s, c;
    sincos_red(ph, s, c);
#pragma unroll
    for (int a = 0; a < 3; ++a) out[a] = (dir[a] + t1[a] * (r * c)) + t2[a] * (r * s);
    normalize3(out);
}

// gbNormal.a (gbuffer.comp:524-526, WATER_CLEAR on): 0.8 for a foliage material (palette kind 3), else 1;
// stored as RGBA8 unorm, so 204 or 255.
constexpr uint32_t kFoliageAlpha = 204u;   // unorm8(0.8f)
__device__ __forceinline__ uint32_t mat_flag_byte(const DevWorld &W, uint32_t mat) {
    return W.palette[mat * 3u + 1u].z == 3.0f ? kFoliageAlpha : 255u;
}
// isFoliageFlag / isBladeFlag (common.glsl:1953-1954) of a gbNormal texel's alpha as the sampler returns it.
__device__ __forceinline__ float gbn_alpha(uint32_t nq) { return (float)(nq >> 24) / 255.0f; }
__device__ __forceinline__ bool foliage_flag(float a) { return a > 0.62f && a < 0.88f; }
__device__ __forceinline__ bool blade_flag(float a) { return a > 0.62f && a < 0.75f; }
// foliageSunOrigin (common.glsl:1987-1989): a leaf voxel is lit through itself from its centre stepped out
// toward the sun; a blade from just off its surface.
__device__ __forceinline__ void foliage_sun_origin(const float P[3], const float N[3], const float sun[3], bool blade,
                                                   float o[3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) o[a] = blade ? P[a] + sun[a] * 0.05f : (P[a] - N[a] * 0.5f) + sun[a] * 0.9f;
}

// sunVisibilityT (common.glsl:1964-1981) over the world (and the bodies: traceSceneShadow): a leaf voxel (palette kind 3) lets the light
// through with probability (STOCHASTIC) or fraction leaf_light, and the ray goes on from where it
// leaves that voxel; up to 6 leaves.  tfirst = distance to the first blocker (0 = none).  The far
// horizon's and the clouds' factors (horizonSunVisibility, cloudShadow) are 1 here: neither the
// horizon heightfield nor clouds are traced by this library.  The first
// segment's hit record at first_base[idx] (first_base nullable); the ray counter counts the whole
// visibility ray once.
template <bool STOCHASTIC, bool COUNT, bool BODIES = false, bool LDSM = false, int LS = 256, bool PRE = false, bool BOX = false>
__device__ __forceinline__ float sun_visibility_t(const DevWorld &W, const float o0[3], const float d[3], float leaf_light,
                                                  uint32_t &seed, float &tfirst, vf_hit *first_base, size_t idx, Cnt &c,
                                                  const DevBodies *B = nullptr, const DirPre *pre = nullptr) {
    // the first segment: the plain sun ray (most rays end here: no hit, or a hit that is not a leaf)
    Hit h;
    trace_world<5, COUNT, LDSM, LS, PRE, BOX>(W, o0, d, 1e30f, h, c, 0, pre);
    if (BODIES) add_bodies<COUNT, LDSM && LS == 256>(W, *B, o0, d, 1e30f, h, c);   // traceSceneShadow: a body always blocks
    if (first_base) store_hit(first_base + idx, h, BODIES && h.obj >= 0);
    tfirst = 0.0f;
    if (!h.hit) return 1.0f;
    tfirst = h.t;
    if ((BODIES && h.obj >= 0) || W.palette[h.mat * 3u + 1u].z != 3.0f || leaf_light <= 0.0f) return 0.0f;
    // through leaves: leave each leaf voxel across its exit face along d and trace on
    float T = 1.0f;
    float o[3] = {o0[0], o0[1], o0[2]};
#pragma unroll 1
    for (int i = 1; i <= 6; ++i) {
        if (STOCHASTIC) {
            if (rnd(seed) >= leaf_light) return 0.0f;
        } else {
            T = T * leaf_light;
            if (T < 0.03f) return 0.0f;
        }
        if (i == 6) return 0.0f;   // six leaves: the loop of common.glsl:1967 ends without a pass
        float p[3], tx[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            p[a] = o[a] + d[a] * h.t;
            const float sd = fabsf(d[a]) < 1e-6f ? 1e-6f : d[a];
            tx[a] = (((float)h.cell[a] + (sd < 0.0f ? 0.0f : 1.0f)) - p[a]) / sd;
        }
        const float tm = gmin(gmin(tx[0], tx[1]), tx[2]) + 1e-3f;
#pragma unroll
        for (int a = 0; a < 3; ++a) o[a] = p[a] + d[a] * tm;
        trace_world<5, COUNT, LDSM, LS, PRE, BOX>(W, o, d, 1e30f, h, c, 0, pre);
        if (BODIES) add_bodies<COUNT, LDSM && LS == 256>(W, *B, o, d, 1e30f, h, c);
        if (COUNT) c.rays--;
        if (!h.hit) return T;
        if ((BODIES && h.obj >= 0) || W.palette[h.mat * 3u + 1u].z != 3.0f) return 0.0f;
    }
    return 0.0f;
}
