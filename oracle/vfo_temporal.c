/*
 * CPU oracle for the temporal consumers of the traced output — TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of:
 *   - gbuffer.comp's motion outputs (shaders/gbuffer.comp:84-104 prevWorldPos / reprojMotion /
 *     STORE_MOTION, 198-202 the sample position, 409-410 the sky's motion);
 *   - gi.comp's accumulation half (shaders/gi.comp:37-58 filter guide and all-sky texels,
 *     93-184 reprojection, bilinear or Catmull-Rom history, firefly cap, anti-lag, moments);
 *   - lighting.comp's sun-shadow history (shaders/lighting.comp:36-42 sky, 92-139);
 *   - one SVGF a-trous level over the half-res GI (shaders/atrous.comp:34-154, GI_HISTFIX on).
 * Used only by tests/, the smoke check and the benchmark's CPU-baseline leg, as the checker.
 *
 * Numerics: FP32 in the shaders' operation order (-ffp-contract=off).  Two functions have no
 * IEEE-exact GLSL definition and are fixed here, identically in the HIP kernels:
 *   - exp(x): Cody-Waite reduction x = k ln2 + r (ln2 split in a 16-bit head and a tail) and the
 *     degree-7 Taylor polynomial of e^r, |r| <= ln2/2, scaled by 2^k; x < -86 gives 0, so no
 *     result is subnormal.  Error < 2 ulp (GLSL allows 3 + 2|x| ulp).
 *   - pow(x, 16) and pow(x, 64) (the a-trous normal weights): 4 and 6 squarings.
 * RGBA16F stores round to nearest even; RGBA8 normals decode as c/255*2-1.  In the a-trous weights
 * a division by the centre's depth scale or luminance scale is a multiplication by its reciprocal
 * (computed once per centre), which GLSL's a/b precision (2.5 ulp) allows.
 *
 * Pinning (tests/test_temporal_oracle.py): exp against float64 within 2 ulp; the half
 * conversions against numpy's; closed forms for a still camera (running mean, history count
 * 1..49), a whole-texel and a half-texel reprojection (bilinear and Catmull-Rom weights),
 * tap rejection by plane and normal, the firefly cap, the anti-lag clamp, the guide bits, and the
 * a-trous filter on planes (identity on a constant field, weight-squared variance, edge stops,
 * the history fix).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "vf_oracle.h"

static inline float tmin(float x, float y) { return y < x ? y : x; }   /* GLSL min */
static inline float tmax(float x, float y) { return x < y ? y : x; }   /* GLSL max */
static inline float tclamp(float x, float lo, float hi) { return tmin(tmax(x, lo), hi); }
static inline int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }
static inline float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }   /* GLSL mix */
static inline float dot3(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline float luminance(const float c[3]) { return c[0] * 0.2126f + c[1] * 0.7152f + c[2] * 0.0722f; }
static inline void dec_normal(uint32_t nb, float n[3]) {
    for (int a = 0; a < 3; ++a) n[a] = ((float)((nb >> (8 * a)) & 255u) / 255.0f) * 2.0f - 1.0f;
}

float vfo_exp(float x) {
    if (x != x) return x;
    if (x < -86.0f) return 0.0f;
    if (x > 88.0f) return INFINITY;
    const float k = rintf(x * 1.44269504f);
    const float r = (x - k * 0.693145751953125f) - k * 1.42860677e-6f;
    float p = 1.0f / 5040.0f;
    p = p * r + 1.0f / 720.0f;
    p = p * r + 1.0f / 120.0f;
    p = p * r + 1.0f / 24.0f;
    p = p * r + 1.0f / 6.0f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    union { uint32_t u; float f; } s;
    s.u = (uint32_t)((int32_t)k + 127) << 23;
    return p * s.f;
}

float vfo_pow16(float x) { float y = x * x; y = y * y; y = y * y; return y * y; }
float vfo_pow64(float x) { float y = vfo_pow16(x); y = y * y; return y * y; }

float vfo_f16_to_f32(uint16_t h) {
    const uint32_t s = (uint32_t)(h >> 15) << 31, e = (h >> 10) & 31u, m = h & 1023u;
    union { uint32_t u; float f; } o;
    if (e == 0) {
        const float v = (float)m * 5.9604644775390625e-8f;   /* m * 2^-24, exact */
        return s ? -v : v;
    }
    if (e == 31) o.u = s | 0x7f800000u | (m << 13);
    else o.u = s | ((e + 112u) << 23) | (m << 13);
    return o.f;
}

uint16_t vfo_f32_to_f16(float f) {
    union { float f; uint32_t u; } i;
    i.f = f;
    const uint32_t s = (i.u >> 16) & 0x8000u, a = i.u & 0x7fffffffu;
    if (a > 0x7f800000u) return (uint16_t)(s | 0x7e00u | ((a >> 13) & 0x1ffu));   /* NaN stays NaN */
    if (a >= 0x477ff000u) return (uint16_t)(s | 0x7c00u);                        /* >= 65520: inf */
    if (a < 0x38800000u) {                                                         /* below 2^-14 */
        if (a < 0x33000000u) return (uint16_t)s;                                   /* below 2^-25: 0 */
        const uint32_t e = a >> 23, m = (a & 0x7fffffu) | 0x800000u;
        const uint32_t sh = 126u - e;                                              /* 14..24 */
        uint32_t q = m >> sh;
        const uint32_t rem = m & ((1u << sh) - 1u), half = 1u << (sh - 1u);
        if (rem > half || (rem == half && (q & 1u))) ++q;
        return (uint16_t)(s | q);
    }
    uint32_t q = ((a >> 13) - (112u << 10));
    const uint32_t rem = a & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) ++q;
    return (uint16_t)(s | q);
}

/* repTexel (common.glsl:1212-1218): the first non-sky texel of the 2x2 block, else h*2. */
static void rep_texel(const float *pos, int32_t w, int32_t h, int32_t hx, int32_t hy, int32_t *qx, int32_t *qy) {
    for (int k = 0; k < 4; ++k) {
        const int32_t x = imin(hx * 2 + (k & 1), w - 1), y = imin(hy * 2 + (k >> 1), h - 1);
        if (pos[((size_t)y * w + x) * 4 + 3] >= 0.0f) { *qx = x; *qy = y; return; }
    }
    *qx = imin(hx * 2, w - 1);
    *qy = imin(hy * 2, h - 1);
}

/* The 4-tap bilinear reprojection shared by gi.comp:108-127 and lighting.comp:104-122.  (fpx, fpy) =
 * puv * res - 0.5; taps outside [0, res) are skipped, so a position below -1 or at res or beyond has
 * none (the GLSL float-to-int conversion of such values is not needed). */
typedef struct { int32_t bx, by, any; float frx, fry; } tap_base;
static tap_base tap_base_of(float fpx, float fpy, int32_t rx, int32_t ry) {
    tap_base t = {0, 0, 0, 0.0f, 0.0f};
    if (!(fpx >= -1.0f && fpx < (float)rx && fpy >= -1.0f && fpy < (float)ry)) return t;
    const float flx = floorf(fpx), fly = floorf(fpy);
    t.bx = (int32_t)flx; t.by = (int32_t)fly;
    t.frx = fpx - flx; t.fry = fpy - fly;
    t.any = 1;
    return t;
}

void vfo_gi_accumulate(int32_t w, int32_t h, uint32_t frame_idx, const float shift[3], int32_t bicubic,
                       const float *gb_pos, const uint32_t *gb_normal, const float *gb_pos_prev,
                       const uint32_t *gb_normal_prev, const uint32_t *gb_motion, const uint16_t *gb_delta,
                       const float *indirect, const uint16_t *hist_prev, const float *mom_prev, uint16_t *hist_cur,
                       float *mom_cur, uint32_t *guide) {
    const int32_t rx = (w + 1) / 2, ry = (h + 1) / 2;
#pragma omp parallel for schedule(dynamic, 4)
    for (int32_t py = 0; py < ry; ++py)
        for (int32_t px = 0; px < rx; ++px) {
            const size_t o = (size_t)py * rx + px;
            int32_t gx, gy;
            rep_texel(gb_pos, w, h, px, py, &gx, &gy);                             /* gi.comp:50-58 */
            const float *gp = gb_pos + ((size_t)gy * w + gx) * 4;
            uint32_t *gd = guide + o * 4;
            memcpy(gd, gp, 12);
            gd[3] = gp[3] < 0.0f ? 0u : gb_normal[(size_t)gy * w + gx];          /* packUnorm4x8 of c/255 = c */
            if (gp[3] < 0.0f) {                                                    /* gi.comp:59-63 */
                for (int c = 0; c < 4; ++c) { hist_cur[o * 4 + c] = 0; mom_cur[o * 4 + c] = 0.0f; }
                continue;
            }
            int32_t fx = imin(px * 2 + (int32_t)(frame_idx & 1u), w - 1);        /* gi.comp:45-46 */
            int32_t fy = imin(py * 2 + (int32_t)((frame_idx >> 1) & 1u), h - 1);
            const float *posT = gb_pos + ((size_t)fy * w + fx) * 4;
            if (posT[3] < 0.0f) { fx = gx; fy = gy; posT = gp; }                 /* gi.comp:67-68 */
            const size_t fi = (size_t)fy * w + fx;
            const float P[3] = {posT[0], posT[1], posT[2]};
            float N[3];
            dec_normal(gb_normal[fi], N);
            float ind[3] = {indirect[o * 4], indirect[o * 4 + 1], indirect[o * 4 + 2]};
            float lum = luminance(ind);                                            /* gi.comp:103-106 */
            float count = 0.0f;
            float acc[3] = {ind[0], ind[1], ind[2]};
            float mom[2] = {lum, lum * lum};
            const float mx = vfo_f16_to_f32((uint16_t)(gb_motion[fi] & 0xffffu));
            const float my = vfo_f16_to_f32((uint16_t)(gb_motion[fi] >> 16));
            const float puvx = ((float)px + 0.5f) / (float)rx + mx, puvy = ((float)py + 0.5f) / (float)ry + my;
            float Pp[3];
            for (int a = 0; a < 3; ++a) Pp[a] = P[a] - vfo_f16_to_f32(gb_delta[fi * 4 + a]);
            const tap_base tb = tap_base_of(puvx * (float)rx - 0.5f, puvy * (float)ry - 0.5f, rx, ry);
            float hs[4] = {0, 0, 0, 0}, ms[3] = {0, 0, 0}, ws = 0.0f;
            const float tol = 0.5f + posT[3] * 0.004f;
            for (int k = 0; k < 4 && tb.any; ++k) {                                /* gi.comp:118-132 */
                const int32_t qx = tb.bx + (k & 1), qy = tb.by + (k >> 1);
                if (qx < 0 || qy < 0 || qx >= rx || qy >= ry) continue;
                int32_t ux, uy;
                rep_texel(gb_pos_prev, w, h, qx, qy, &ux, &uy);
                const float *pp = gb_pos_prev + ((size_t)uy * w + ux) * 4;
                const float prevP[3] = {pp[0] + shift[0], pp[1] + shift[1], pp[2] + shift[2]};
                if (pp[3] <= 0.0f) continue;
                float prevN[3];
                dec_normal(gb_normal_prev[(size_t)uy * w + ux], prevN);
                const float d[3] = {prevP[0] - Pp[0], prevP[1] - Pp[1], prevP[2] - Pp[2]};
                if (fabsf(dot3(d, N)) >= tol || dot3(prevN, N) <= 0.85f) continue;
                const float wt = ((k & 1) ? tb.frx : 1.0f - tb.frx) * ((k >> 1) ? tb.fry : 1.0f - tb.fry);
                const size_t q = (size_t)qy * rx + qx;
                for (int c = 0; c < 4; ++c) hs[c] = hs[c] + vfo_f16_to_f32(hist_prev[q * 4 + c]) * wt;
                for (int c = 0; c < 3; ++c) ms[c] = ms[c] + mom_prev[q * 4 + c] * wt;
                ws = ws + wt;
            }
            float fast = lum;
            if (bicubic && ws > 0.999f && (fabsf(tb.frx) > 1e-4f || fabsf(tb.fry) > 1e-4f)) {   /* gi.comp:134-157 */
                const float fr[2] = {tb.frx, tb.fry};
                float wk[4][2];
                for (int a = 0; a < 2; ++a) {
                    const float f = fr[a];
                    wk[0][a] = f * (-0.5f + f * (1.0f - 0.5f * f));
                    wk[1][a] = 1.0f + f * f * (-2.5f + 1.5f * f);
                    wk[2][a] = f * (0.5f + f * (2.0f - 1.5f * f));
                    wk[3][a] = f * f * (-0.5f + 0.5f * f);
                }
                float cr[4] = {0, 0, 0, 0}, lo[4], hi[4];
                for (int c = 0; c < 4; ++c) { lo[c] = 1e30f; hi[c] = -1e30f; }
                for (int j = 0; j < 4; ++j)
                    for (int i = 0; i < 4; ++i) {
                        int32_t qx = tb.bx + i - 1, qy = tb.by + j - 1;
                        qx = qx < 0 ? 0 : (qx > rx - 1 ? rx - 1 : qx);
                        qy = qy < 0 ? 0 : (qy > ry - 1 ? ry - 1 : qy);
                        const size_t q = (size_t)qy * rx + qx;
                        const float wxy = wk[i][0] * wk[j][1];
                        for (int c = 0; c < 4; ++c) {
                            const float hv = vfo_f16_to_f32(hist_prev[q * 4 + c]);
                            cr[c] = cr[c] + hv * wxy;
                            if ((i == 1 || i == 2) && (j == 1 || j == 2)) { lo[c] = tmin(lo[c], hv); hi[c] = tmax(hi[c], hv); }
                        }
                    }
                const float ha = hs[3] / ws;
                for (int c = 0; c < 3; ++c) hs[c] = tclamp(cr[c], lo[c], hi[c]) * ws;
                hs[3] = ha * ws;
            }
            if (ws > 1e-3f) {                                                      /* gi.comp:158-181 */
                float hp[4], mp[3];
                for (int c = 0; c < 4; ++c) hp[c] = hs[c] / ws;
                for (int c = 0; c < 3; ++c) mp[c] = ms[c] / ws;
                count = tmin(hp[3], 48.0f);
                const float sigma = sqrtf(tmax(mp[1] - mp[0] * mp[0], 0.0f));
                if (count >= 4.0f) {
                    const float cap = mp[0] + 4.0f * sigma + 0.05f;
                    if (lum > cap) {
                        const float s = cap / lum;
                        for (int c = 0; c < 3; ++c) ind[c] = ind[c] * s;
                        lum = cap;
                    }
                }
                const float wt = 1.0f / (count + 1.0f);
                fast = mixf(mp[2], lum, tmax(wt, 0.25f));
                float slow[3] = {hp[0], hp[1], hp[2]};
                const float sl = luminance(slow), band = 3.0f * sigma * 0.378f + 0.02f;
                const float cl = tclamp(sl, fast - band, fast + band);
                if (sl > 1e-4f) {
                    const float s = cl / sl;
                    for (int c = 0; c < 3; ++c) slow[c] = slow[c] * s;
                }
                for (int c = 0; c < 3; ++c) acc[c] = mixf(slow[c], ind[c], wt);
                mom[0] = mixf(mp[0], lum, wt);
                mom[1] = mixf(mp[1], lum * lum, wt);
            }
            for (int c = 0; c < 3; ++c) hist_cur[o * 4 + c] = vfo_f32_to_f16(acc[c]);
            hist_cur[o * 4 + 3] = vfo_f32_to_f16(count + 1.0f);
            mom_cur[o * 4] = mom[0];
            mom_cur[o * 4 + 1] = mom[1];
            mom_cur[o * 4 + 2] = fast;
            mom_cur[o * 4 + 3] = 0.0f;
        }
}

void vfo_light_accumulate(int32_t w, int32_t h, const float shift[3], const float *gb_pos, const uint32_t *gb_normal,
                          const float *gb_pos_prev, const uint32_t *gb_normal_prev, const uint32_t *gb_motion,
                          const uint16_t *gb_delta, const float *shadow, const float *hist_prev, float *hist_cur) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int32_t py = 0; py < h; ++py)
        for (int32_t px = 0; px < w; ++px) {
            const size_t i = (size_t)py * w + px;
            const float *posT = gb_pos + i * 4;
            float *out = hist_cur + i * 4;
            if (posT[3] < 0.0f) {                                                  /* lighting.comp:39-42 */
                out[0] = 1.0f; out[1] = 1.0f; out[2] = 0.0f; out[3] = 0.0f;
                continue;
            }
            const float P[3] = {posT[0], posT[1], posT[2]};
            float N[3];
            dec_normal(gb_normal[i], N);
            const float s = shadow[i];
            float count = 0.0f, sAcc = s, sM2 = s * s, sFast = s;                /* lighting.comp:98-101 */
            const float mx = vfo_f16_to_f32((uint16_t)(gb_motion[i] & 0xffffu));
            const float my = vfo_f16_to_f32((uint16_t)(gb_motion[i] >> 16));
            const float puvx = ((float)px + 0.5f) / (float)w + mx, puvy = ((float)py + 0.5f) / (float)h + my;
            float Pp[3];
            for (int a = 0; a < 3; ++a) Pp[a] = P[a] - vfo_f16_to_f32(gb_delta[i * 4 + a]);
            const tap_base tb = tap_base_of(puvx * (float)w - 0.5f, puvy * (float)h - 0.5f, w, h);
            float hs[4] = {0, 0, 0, 0}, ws = 0.0f;
            const float tol = 0.5f + posT[3] * 0.004f;
            for (int k = 0; k < 4 && tb.any; ++k) {                                /* lighting.comp:113-123 */
                const int32_t qx = tb.bx + (k & 1), qy = tb.by + (k >> 1);
                if (qx < 0 || qy < 0 || qx >= w || qy >= h) continue;
                const size_t q = (size_t)qy * w + qx;
                const float *pp = gb_pos_prev + q * 4;
                const float prevP[3] = {pp[0] + shift[0], pp[1] + shift[1], pp[2] + shift[2]};
                if (pp[3] <= 0.0f) continue;
                float prevN[3];
                dec_normal(gb_normal_prev[q], prevN);
                const float d[3] = {prevP[0] - Pp[0], prevP[1] - Pp[1], prevP[2] - Pp[2]};
                if (fabsf(dot3(d, N)) >= tol || dot3(prevN, N) <= 0.85f) continue;
                const float wt = ((k & 1) ? tb.frx : 1.0f - tb.frx) * ((k >> 1) ? tb.fry : 1.0f - tb.fry);
                for (int c = 0; c < 4; ++c) hs[c] = hs[c] + hist_prev[q * 4 + c] * wt;
                ws = ws + wt;
            }
            if (ws > 1e-3f) {                                                      /* lighting.comp:124-137 */
                float hist[4];
                for (int c = 0; c < 4; ++c) hist[c] = hs[c] / ws;
                count = tmin(hist[2], 48.0f);
                const float wt = 1.0f / (count + 1.0f);
                sFast = mixf(hist[3], s, tmax(wt, 0.25f));
                const float sigma = sqrtf(tmax(hist[1] - hist[0] * hist[0], 0.0f));
                const float band = 3.0f * sigma * 0.378f + 0.02f;
                const float slow = tclamp(hist[0], sFast - band, sFast + band);
                sAcc = mixf(slow, s, wt);
                sM2 = mixf(hist[1], s * s, wt);
            }
            out[0] = sAcc; out[1] = sM2; out[2] = count + 1.0f; out[3] = sFast;
        }
}

/* readInput (atrous.comp:34-45). */
static void read_input(int32_t it, size_t q, const uint16_t *hist, const float *mom, const uint16_t *in_cv, float col[3],
                       float *var) {
    if (it == 0) {
        for (int c = 0; c < 3; ++c) col[c] = vfo_f16_to_f32(hist[q * 4 + c]);
        const float m0 = mom[q * 4], m1 = mom[q * 4 + 1];
        *var = tmax(m1 - m0 * m0, 0.0f);
    } else {
        for (int c = 0; c < 3; ++c) col[c] = vfo_f16_to_f32(in_cv[q * 4 + c]);
        *var = vfo_f16_to_f32(in_cv[q * 4 + 3]);
    }
}

void vfo_gi_atrous_level(int32_t w, int32_t h, int32_t it, int32_t step, int32_t histfix, const float *gb_pos,
                         const uint32_t *gb_normal, const float *gi_mom, const uint16_t *gi_hist, const uint32_t *guide,
                         const uint16_t *in_cv, uint16_t *out_cv) {
    const int32_t rx = (w + 1) / 2, ry = (h + 1) / 2;
    static const float h5[3] = {0.375f, 0.25f, 0.0625f};
#pragma omp parallel for schedule(dynamic, 4)
    for (int32_t py = 0; py < ry; ++py)
        for (int32_t px = 0; px < rx; ++px) {
            const size_t o = (size_t)py * rx + px;
            int32_t fx, fy;
            rep_texel(gb_pos, w, h, px, py, &fx, &fy);                             /* atrous.comp:76-81 */
            const float *posT = gb_pos + ((size_t)fy * w + fx) * 4;
            if (posT[3] < 0.0f) {
                for (int c = 0; c < 4; ++c) out_cv[o * 4 + c] = 0;
                continue;
            }
            const float P[3] = {posT[0], posT[1], posT[2]};
            float N[3];
            dec_normal(gb_normal[(size_t)fy * w + fx], N);
            float cCol[3], cVar;
            read_input(it, o, gi_hist, gi_mom, in_cv, cCol, &cVar);
            if (histfix && it == 0) {                                              /* atrous.comp:87-112 */
                const float cLen = vfo_f16_to_f32(gi_hist[o * 4 + 3]);
                if (cLen < 4.0f) {
                    float acc[3] = {cCol[0] * cLen, cCol[1] * cLen, cCol[2] * cLen};
                    float ws = cLen;
                    const float dinv = 1.0f / (1.0f + posT[3] * 0.01f);
                    for (int dy = -3; dy <= 3; ++dy)
                        for (int dx = -3; dx <= 3; ++dx) {
                            if (dx == 0 && dy == 0) continue;
                            const int32_t qx = px + dx * 3, qy = py + dy * 3;
                            if (qx < 0 || qy < 0 || qx >= rx || qy >= ry) continue;
                            const size_t q = (size_t)qy * rx + qx;
                            const uint32_t *g = guide + q * 4;
                            if (g[3] == 0u) continue;
                            float gP[3], gN[3];
                            memcpy(gP, g, 12);
                            dec_normal(g[3], gN);
                            const float d[3] = {gP[0] - P[0], gP[1] - P[1], gP[2] - P[2]};
                            const float ha = vfo_f16_to_f32(gi_hist[q * 4 + 3]);
                            const float wt = vfo_pow16(tmax(dot3(N, gN), 0.0f)) * vfo_exp(-fabsf(dot3(d, N)) * dinv) * ha;
                            for (int c = 0; c < 3; ++c) acc[c] = acc[c] + vfo_f16_to_f32(gi_hist[q * 4 + c]) * wt;
                            ws = ws + wt;
                        }
                    if (ws > 0.0f)
                        for (int c = 0; c < 3; ++c) cCol[c] = acc[c] / ws;
                }
            }
            const float cLum = luminance(cCol);
            float vg;                                                              /* centreVariance, atrous.comp:51-71 */
            if (it != 0) {
                vg = vfo_f16_to_f32(in_cv[o * 4 + 3]);
            } else if (vfo_f16_to_f32(gi_hist[o * 4 + 3]) >= 4.0f) {
                const float m0 = gi_mom[o * 4], m1 = gi_mom[o * 4 + 1];
                vg = tmax(m1 - m0 * m0, 0.0f);
            } else {
                float s = 0.0f, s2 = 0.0f, n = 0.0f;
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        int32_t qx = px + dx, qy = py + dy;
                        qx = qx < 0 ? 0 : (qx > rx - 1 ? rx - 1 : qx);
                        qy = qy < 0 ? 0 : (qy > ry - 1 ? ry - 1 : qy);
                        const size_t q = (size_t)qy * rx + qx;
                        const float c3[3] = {vfo_f16_to_f32(gi_hist[q * 4]), vfo_f16_to_f32(gi_hist[q * 4 + 1]),
                                             vfo_f16_to_f32(gi_hist[q * 4 + 2])};
                        const float l = luminance(c3);
                        s = s + l;
                        s2 = s2 + l * l;
                        n = n + 1.0f;
                    }
                const float mean = s / n;
                vg = (tmax(s2 / n - mean * mean, 0.0f) + 1e-3f) * 2.0f;
            }
            const int fresh0 = it == 0 && vfo_f16_to_f32(gi_hist[o * 4 + 3]) < 4.0f;   /* atrous.comp:120-124 */
            if (fresh0) cVar = vg;
            const float lumPhi = 4.0f * sqrtf(tmax(vg, 0.0f)) + 1e-4f;
            const float k00 = h5[0] * h5[0];
            float cAcc[3] = {cCol[0] * k00, cCol[1] * k00, cCol[2] * k00};
            float vAcc = cVar * k00 * k00;
            float wSum = k00;
            const float depthScale = 1.0f * (1.0f + posT[3] * 0.01f);
            const float zinv = 1.0f / depthScale, linv = 1.0f / lumPhi;
            for (int dy = -2; dy <= 2; ++dy)                                       /* atrous.comp:132-151 */
                for (int dx = -2; dx <= 2; ++dx) {
                    if (dx == 0 && dy == 0) continue;
                    const int32_t qx = px + dx * step, qy = py + dy * step;
                    if (qx < 0 || qy < 0 || qx >= rx || qy >= ry) continue;
                    const size_t q = (size_t)qy * rx + qx;
                    const uint32_t *g = guide + q * 4;
                    if (g[3] == 0u) continue;
                    float qp[3], qn[3];
                    memcpy(qp, g, 12);
                    dec_normal(g[3], qn);
                    float qc[3], qv;
                    read_input(it, q, gi_hist, gi_mom, in_cv, qc, &qv);
                    if (it == 0 && vfo_f16_to_f32(gi_hist[q * 4 + 3]) < 4.0f) qv = tmax(qv, vg);
                    const float qLum = luminance(qc);
                    const float kern = h5[dx < 0 ? -dx : dx] * h5[dy < 0 ? -dy : dy];
                    const float wN = vfo_pow64(tmax(dot3(N, qn), 0.0f));
                    const float d[3] = {qp[0] - P[0], qp[1] - P[1], qp[2] - P[2]};
                    const float wZ = vfo_exp(-fabsf(dot3(d, N)) * zinv);
                    const float wL = vfo_exp(-fabsf(cLum - qLum) * linv);
                    const float wt = kern * wN * wZ * wL;
                    for (int c = 0; c < 3; ++c) cAcc[c] = cAcc[c] + wt * qc[c];
                    vAcc = vAcc + wt * wt * qv;
                    wSum = wSum + wt;
                }
            const float inv = 1.0f / tmax(wSum, 1e-6f);
            for (int c = 0; c < 3; ++c) out_cv[o * 4 + c] = vfo_f32_to_f16(cAcc[c] * inv);
            out_cv[o * 4 + 3] = vfo_f32_to_f16(vAcc * inv * inv);
        }
}

void vfo_gi_atrous(int32_t w, int32_t h, int32_t levels, int32_t histfix, const float *gb_pos, const uint32_t *gb_normal,
                   const float *gi_mom, const uint16_t *gi_hist, const uint32_t *guide, uint16_t *filt0, uint16_t *filt1) {
    /* record.rs:372-376 + mod.rs:1062-1074: level i writes filt[i & 1] and reads filt[(i + 1) & 1]. */
    for (int32_t i = 0; i < levels; ++i) {
        uint16_t *dst = (i & 1) ? filt1 : filt0;
        const uint16_t *src = (i & 1) ? filt0 : filt1;
        vfo_gi_atrous_level(w, h, i, 1 << i, histfix, gb_pos, gb_normal, gi_mom, gi_hist, guide, i == 0 ? NULL : src, dst);
    }
}

/* reprojMotion (gbuffer.comp:91-99): prevUV - curUV of a previous-frame world point. */
static void reproj_motion(const float Pp[3], float scx, float scy, int32_t w, int32_t h, const float pc[3],
                          const float pr[3], const float pu[3], const float pf[3], float pthf, float out[2]) {
    const float aspect = (float)w / (float)h;
    const float dp[3] = {Pp[0] - pc[0], Pp[1] - pc[1], Pp[2] - pc[2]};
    const float pz = dot3(dp, pf);
    if (pz <= 0.0f) { out[0] = 0.0f; out[1] = 0.0f; return; }
    const float u = dot3(dp, pr) / (pz * pthf * aspect), v = dot3(dp, pu) / (pz * pthf);
    out[0] = (u * 0.5f + 0.5f) - scx / (float)w;
    out[1] = (v * 0.5f + 0.5f) - scy / (float)h;
}

void vfo_gbuffer_motion(int32_t w, int32_t h, const float cam_right[3], const float cam_up[3], const float cam_fwd[3],
                        float thf, const float jit[2], const float prev_pos[3], const float prev_right[3],
                        const float prev_up[3], const float prev_fwd[3], float prev_thf, const float *gb_pos,
                        const int32_t *body, const float *bodies, int32_t n_bodies, uint32_t *motion, uint16_t *delta) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int32_t py = 0; py < h; ++py)
        for (int32_t px = 0; px < w; ++px) {
            const size_t i = (size_t)py * w + px;
            const float *p = gb_pos + i * 4;
            const float scx = ((float)px + 0.5f) + jit[0], scy = ((float)py + 0.5f) + jit[1];
            float Pp[3], m[2], d[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            if (p[3] < 0.0f) {                                                     /* gbuffer.comp:405-410 */
                float rd[3];
                vfo_primary_dir(cam_right, cam_up, cam_fwd, thf, w, h, px, py, jit, rd);
                for (int a = 0; a < 3; ++a) Pp[a] = prev_pos[a] + rd[a] * 1e6f;
            } else {
                const int32_t ob = body ? body[i] : -1;
                if (ob >= 0 && ob < n_bodies) {                                    /* prevWorldPos, gbuffer.comp:84-88 */
                    const float *wo = bodies + (size_t)ob * 60, *po = wo + 32;
                    float l[3];
                    for (int r = 0; r < 3; ++r) l[r] = ((wo[r] * p[0] + wo[4 + r] * p[1]) + wo[8 + r] * p[2]) + wo[12 + r];
                    for (int r = 0; r < 3; ++r) Pp[r] = ((po[r] * l[0] + po[4 + r] * l[1]) + po[8 + r] * l[2]) + po[12 + r];
                    d[3] = (float)(ob + 1);
                } else {
                    Pp[0] = p[0]; Pp[1] = p[1]; Pp[2] = p[2];
                }
                for (int a = 0; a < 3; ++a) d[a] = p[a] - Pp[a];
            }
            reproj_motion(Pp, scx, scy, w, h, prev_pos, prev_right, prev_up, prev_fwd, prev_thf, m);
            motion[i] = (uint32_t)vfo_f32_to_f16(m[0]) | ((uint32_t)vfo_f32_to_f16(m[1]) << 16);
            for (int c = 0; c < 4; ++c) delta[i * 4 + c] = vfo_f32_to_f16(d[c]);
        }
}
