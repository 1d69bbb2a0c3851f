// The temporal consumers of the traced output (include/vf_temporal.h): gi.comp's reprojected
// accumulation, lighting.comp's sun-shadow history and the SVGF a-trous levels over the half-res GI.
//
// These passes do a few dozen flops per texel against ~100-250 bytes of G-buffer and history
// traffic, so they are HBM-bound gathers: one thread per output texel, 256-thread workgroups over
// 16x16 texel tiles (8x8 per wavefront, so a wavefront's 2x2 full-res reads cover one contiguous
// 16x16 full-res patch), tiles dealt to the 8 XCDs in interleaved runs, 16-byte records read and
// written whole.  The arithmetic follows the shaders' FP32 operation order (no contraction); exp and
// the integer powers are fixed functions (below) so the results are reproducible bit for bit.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/vf_temporal.h"
#include "vf_ctx.h"

namespace {

// exp(x): Cody-Waite reduction x = k ln2 + r with ln2 split into a 16-bit head and a tail, the
// degree-7 Taylor polynomial of e^r (|r| <= ln2/2, truncation < 6e-9 relative), times 2^k built in
// the exponent field.  x < -86 gives 0 (no subnormal results); error < 2 ulp.  The filters only take
// exp of a non-positive argument, so the polynomial runs without a branch on x clamped to [-86, 0]
// (GLSL-style min / max: a NaN passes through and comes out NaN) and is multiplied by 0 below -86
// (x > 0 is outside this function's domain).
__device__ __forceinline__ float t_exp(float x) {
    const float xc = gmin(gmax(x, -86.0f), 0.0f);
    const float k = __builtin_rintf(xc * 1.44269504f);
    const float r = (xc - k * 0.693145751953125f) - k * 1.42860677e-6f;
    float p = 1.0f / 5040.0f;
    p = p * r + 1.0f / 720.0f;
    p = p * r + 1.0f / 120.0f;
    p = p * r + 1.0f / 24.0f;
    p = p * r + 1.0f / 6.0f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    const float e = p * __int_as_float(((int32_t)k + 127) << 23);
    return e * (x < -86.0f ? 0.0f : 1.0f);
}
// pow(x, 16) and pow(x, 64) by squaring.
__device__ __forceinline__ float t_pow16(float x) { float y = x * x; y = y * y; y = y * y; return y * y; }
__device__ __forceinline__ float t_pow64(float x) { float y = t_pow16(x); y = y * y; return y * y; }

__device__ __forceinline__ float h2f(uint32_t bits16) { return (float)__builtin_bit_cast(_Float16, (uint16_t)bits16); }
// f32 -> f16 of a value already rounded to f32.  The empty asm keeps the compiler from fusing the
// producing multiply or add with the conversion (v_fma_mixlo_f16 rounds the exact product once, which
// differs from the shader's f32 product stored to an RGBA16F image whenever that product is a tie).
__device__ __forceinline__ uint32_t f2h(float x) {
    asm("" : "+v"(x));
    return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)x);
}
__device__ __forceinline__ void h4(uint2 v, float o[4]) {
    o[0] = h2f(v.x & 0xffffu); o[1] = h2f(v.x >> 16); o[2] = h2f(v.y & 0xffffu); o[3] = h2f(v.y >> 16);
}
__device__ __forceinline__ uint2 pack_h4(float a, float b, float c, float d) {
    return make_uint2(f2h(a) | (f2h(b) << 16), f2h(c) | (f2h(d) << 16));
}
__device__ __forceinline__ float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
__device__ __forceinline__ float lumi(float r, float g, float b) { return (r * 0.2126f + g * 0.7152f) + b * 0.0722f; }
__device__ __forceinline__ void dec_n(uint32_t nb, float n[3]) {
    n[0] = unorm8_dec(nb & 255u); n[1] = unorm8_dec((nb >> 8) & 255u); n[2] = unorm8_dec((nb >> 16) & 255u);
}

struct TArgs {
    int32_t w, h;        // full res
    int32_t rx, ry;      // output res (half res, or full res for the lighting history)
    int32_t tiles_x, ntiles;
    uint32_t frame;
    float shift[3];
    int32_t it, step;    // a-trous level and tap spacing
};

// Output texel of this thread: 16x16 tiles, 8x8 per wavefront, XCD-interleaved tile order.
__device__ __forceinline__ bool texel_of(const TArgs &A, int32_t &px, int32_t &py) {
    const int32_t L = xcd_remap((int32_t)blockIdx.x, (int32_t)gridDim.x);
    const int32_t tx = L % A.tiles_x, ty = L / A.tiles_x;
    const int32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    px = tx * 16 + (wv & 1) * 8 + (lane & 7);
    py = ty * 16 + (wv >> 1) * 8 + (lane >> 3);
    return px < A.rx && py < A.ry;
}

// repTexel (common.glsl:1212-1218): the first non-sky texel of the 2x2 block, else h*2.  Written as a
// search for the index (hit distances only) followed by one load of the chosen texel: the form that
// returned each loaded texel from inside the search was compiled (hipcc, ROCm 7.2, gfx950) so that an
// all-sky block returned the last texel's position with the first texel's distance.
__device__ __forceinline__ float4 rep_texel(const float4 *pos, int32_t w, int32_t h, int32_t hx, int32_t hy, int32_t &qx,
                                            int32_t &qy) {
    int32_t gx = imin(hx * 2, w - 1), gy = imin(hy * 2, h - 1);
    for (int k = 0; k < 4; ++k) {
        const int32_t x = imin(hx * 2 + (k & 1), w - 1), y = imin(hy * 2 + (k >> 1), h - 1);
        if (pos[(size_t)y * w + x].w >= 0.0f) { gx = x; gy = y; break; }
    }
    qx = gx; qy = gy;
    return pos[(size_t)gy * w + gx];
}

// floor / fraction of the reprojected texel position; false when no bilinear tap can be inside
// [0, res) (also for NaN), which is also where the shader's float-to-int conversion is not needed.
__device__ __forceinline__ bool tap_base(float fpx, float fpy, int32_t rx, int32_t ry, int32_t &bx, int32_t &by, float &frx,
                                         float &fry) {
    if (!(fpx >= -1.0f && fpx < (float)rx && fpy >= -1.0f && fpy < (float)ry)) return false;
    const float flx = floorf(fpx), fly = floorf(fpy);
    bx = (int32_t)flx; by = (int32_t)fly;
    frx = fpx - flx; fry = fpy - fly;
    return true;
}

// The tap test shared by gi.comp:121-127 and lighting.comp:115-120.
__device__ __forceinline__ bool tap_ok(float4 pp, uint32_t nprev, const TArgs &A, const float Pp[3], const float N[3],
                                       float tol) {
    const float prevP[3] = {pp.x + A.shift[0], pp.y + A.shift[1], pp.z + A.shift[2]};
    if (pp.w <= 0.0f) return false;
    float prevN[3];
    dec_n(nprev, prevN);
    const float d[3] = {prevP[0] - Pp[0], prevP[1] - Pp[1], prevP[2] - Pp[2]};
    return !(fabsf(dot3(d, N)) >= tol || dot3(prevN, N) <= 0.85f);
}

// gi.comp:37-63, 93-184 for one half-res texel.
template <bool BICUBIC>
__global__ void __launch_bounds__(256) k_gi_accumulate(TArgs A, const float4 *__restrict__ gb_pos, const uint32_t *__restrict__ gb_n,
                                                       const float4 *__restrict__ pos_prev, const uint32_t *__restrict__ n_prev,
                                                       const uint32_t *__restrict__ motion, const uint2 *__restrict__ delta,
                                                       const float4 *__restrict__ indirect, const uint2 *__restrict__ hist_prev,
                                                       const float4 *__restrict__ mom_prev, uint2 *__restrict__ hist_cur,
                                                       float4 *__restrict__ mom_cur, uint4 *__restrict__ guide) {
    int32_t px, py;
    if (!texel_of(A, px, py)) return;
    const size_t o = (size_t)py * A.rx + px;
    int32_t gx, gy;
    const float4 gp = rep_texel(gb_pos, A.w, A.h, px, py, gx, gy);
    const bool sky = gp.w < 0.0f;
    guide[o] = make_uint4(__float_as_uint(gp.x), __float_as_uint(gp.y), __float_as_uint(gp.z),
                          sky ? 0u : gb_n[(size_t)gy * A.w + gx]);
    if (sky) {
        hist_cur[o] = make_uint2(0u, 0u);
        mom_cur[o] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    int32_t fx = imin(px * 2 + (int32_t)(A.frame & 1u), A.w - 1);
    int32_t fy = imin(py * 2 + (int32_t)((A.frame >> 1) & 1u), A.h - 1);
    float4 posT = gb_pos[(size_t)fy * A.w + fx];
    if (posT.w < 0.0f) { fx = gx; fy = gy; posT = gp; }
    const size_t fi = (size_t)fy * A.w + fx;
    const float P[3] = {posT.x, posT.y, posT.z};
    float N[3];
    dec_n(gb_n[fi], N);
    const float4 ind4 = indirect[o];
    float ind[3] = {ind4.x, ind4.y, ind4.z};
    float lum = lumi(ind[0], ind[1], ind[2]);
    float count = 0.0f;
    float acc[3] = {ind[0], ind[1], ind[2]};
    float mom0 = lum, mom1 = lum * lum;
    const uint32_t mv = motion[fi];
    const uint2 dv = delta[fi];
    const float puvx = ((float)px + 0.5f) / (float)A.rx + h2f(mv & 0xffffu);
    const float puvy = ((float)py + 0.5f) / (float)A.ry + h2f(mv >> 16);
    const float Pp[3] = {P[0] - h2f(dv.x & 0xffffu), P[1] - h2f(dv.x >> 16), P[2] - h2f(dv.y & 0xffffu)};
    int32_t bx = 0, by = 0;
    float frx = 0.0f, fry = 0.0f;
    const bool any = tap_base(puvx * (float)A.rx - 0.5f, puvy * (float)A.ry - 0.5f, A.rx, A.ry, bx, by, frx, fry);
    float hs[4] = {0.0f, 0.0f, 0.0f, 0.0f}, ms[3] = {0.0f, 0.0f, 0.0f}, ws = 0.0f;
    const float tol = 0.5f + posT.w * 0.004f;
    if (any) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int32_t qx = bx + (k & 1), qy = by + (k >> 1);
            if (qx < 0 || qy < 0 || qx >= A.rx || qy >= A.ry) continue;
            int32_t ux, uy;
            const float4 pp = rep_texel(pos_prev, A.w, A.h, qx, qy, ux, uy);
            if (!tap_ok(pp, n_prev[(size_t)uy * A.w + ux], A, Pp, N, tol)) continue;
            const float wt = ((k & 1) ? frx : 1.0f - frx) * ((k >> 1) ? fry : 1.0f - fry);
            const size_t q = (size_t)qy * A.rx + qx;
            float hv[4];
            h4(hist_prev[q], hv);
            const float4 mv4 = mom_prev[q];
            hs[0] = hs[0] + hv[0] * wt; hs[1] = hs[1] + hv[1] * wt; hs[2] = hs[2] + hv[2] * wt; hs[3] = hs[3] + hv[3] * wt;
            ms[0] = ms[0] + mv4.x * wt; ms[1] = ms[1] + mv4.y * wt; ms[2] = ms[2] + mv4.z * wt;
            ws = ws + wt;
        }
    }
    float fast = lum;
    if (BICUBIC && ws > 0.999f && (fabsf(frx) > 1e-4f || fabsf(fry) > 1e-4f)) {   // gi.comp:134-157
        float wk[4][2];
        const float fr[2] = {frx, fry};
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const float f = fr[a];
            wk[0][a] = f * (-0.5f + f * (1.0f - 0.5f * f));
            wk[1][a] = 1.0f + f * f * (-2.5f + 1.5f * f);
            wk[2][a] = f * (0.5f + f * (2.0f - 1.5f * f));
            wk[3][a] = f * f * (-0.5f + 0.5f * f);
        }
        float cr[3] = {0.0f, 0.0f, 0.0f}, lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int32_t qy = imin(imax(by + j - 1, 0), A.ry - 1);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int32_t qx = imin(imax(bx + i - 1, 0), A.rx - 1);
                float hv[4];
                h4(hist_prev[(size_t)qy * A.rx + qx], hv);
                const float wxy = wk[i][0] * wk[j][1];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    cr[c] = cr[c] + hv[c] * wxy;
                    if ((i == 1 || i == 2) && (j == 1 || j == 2)) { lo[c] = gmin(lo[c], hv[c]); hi[c] = gmax(hi[c], hv[c]); }
                }
            }
        }
        const float ha = hs[3] / ws;
#pragma unroll
        for (int c = 0; c < 3; ++c) hs[c] = fclampg(cr[c], lo[c], hi[c]) * ws;
        hs[3] = ha * ws;
    }
    if (ws > 1e-3f) {   // gi.comp:158-181
        const float hp[4] = {hs[0] / ws, hs[1] / ws, hs[2] / ws, hs[3] / ws};
        const float mp[3] = {ms[0] / ws, ms[1] / ws, ms[2] / ws};
        count = gmin(hp[3], 48.0f);
        const float sigma = sqrtf(gmax(mp[1] - mp[0] * mp[0], 0.0f));
        if (count >= 4.0f) {
            const float cap = mp[0] + 4.0f * sigma + 0.05f;
            if (lum > cap) {
                const float s = cap / lum;
                ind[0] = ind[0] * s; ind[1] = ind[1] * s; ind[2] = ind[2] * s;
                lum = cap;
            }
        }
        const float wt = 1.0f / (count + 1.0f);
        fast = mixf(mp[2], lum, gmax(wt, 0.25f));
        float slow[3] = {hp[0], hp[1], hp[2]};
        const float sl = lumi(slow[0], slow[1], slow[2]), band = 3.0f * sigma * 0.378f + 0.02f;
        const float cl = fclampg(sl, fast - band, fast + band);
        if (sl > 1e-4f) {
            const float s = cl / sl;
            slow[0] = slow[0] * s; slow[1] = slow[1] * s; slow[2] = slow[2] * s;
        }
        acc[0] = mixf(slow[0], ind[0], wt); acc[1] = mixf(slow[1], ind[1], wt); acc[2] = mixf(slow[2], ind[2], wt);
        mom0 = mixf(mp[0], lum, wt);
        mom1 = mixf(mp[1], lum * lum, wt);
    }
    hist_cur[o] = pack_h4(acc[0], acc[1], acc[2], count + 1.0f);
    mom_cur[o] = make_float4(mom0, mom1, fast, 0.0f);
}

// lighting.comp:36-42, 92-139 for one full-res pixel.
__global__ void __launch_bounds__(256) k_light_accumulate(TArgs A, const float4 *__restrict__ gb_pos, const uint32_t *__restrict__ gb_n,
                                                          const float4 *__restrict__ pos_prev, const uint32_t *__restrict__ n_prev,
                                                          const uint32_t *__restrict__ motion, const uint2 *__restrict__ delta,
                                                          const float *__restrict__ shadow, const float4 *__restrict__ hist_prev,
                                                          float4 *__restrict__ hist_cur) {
    int32_t px, py;
    if (!texel_of(A, px, py)) return;
    const size_t i = (size_t)py * A.w + px;
    const float4 posT = gb_pos[i];
    if (posT.w < 0.0f) {
        hist_cur[i] = make_float4(1.0f, 1.0f, 0.0f, 0.0f);
        return;
    }
    const float P[3] = {posT.x, posT.y, posT.z};
    float N[3];
    dec_n(gb_n[i], N);
    const float s = shadow[i];
    float count = 0.0f, sAcc = s, sM2 = s * s, sFast = s;
    const uint32_t mv = motion[i];
    const uint2 dv = delta[i];
    const float puvx = ((float)px + 0.5f) / (float)A.w + h2f(mv & 0xffffu);
    const float puvy = ((float)py + 0.5f) / (float)A.h + h2f(mv >> 16);
    const float Pp[3] = {P[0] - h2f(dv.x & 0xffffu), P[1] - h2f(dv.x >> 16), P[2] - h2f(dv.y & 0xffffu)};
    int32_t bx = 0, by = 0;
    float frx = 0.0f, fry = 0.0f;
    const bool any = tap_base(puvx * (float)A.w - 0.5f, puvy * (float)A.h - 0.5f, A.w, A.h, bx, by, frx, fry);
    float hs[4] = {0.0f, 0.0f, 0.0f, 0.0f}, ws = 0.0f;
    const float tol = 0.5f + posT.w * 0.004f;
    if (any) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int32_t qx = bx + (k & 1), qy = by + (k >> 1);
            if (qx < 0 || qy < 0 || qx >= A.w || qy >= A.h) continue;
            const size_t q = (size_t)qy * A.w + qx;
            if (!tap_ok(pos_prev[q], n_prev[q], A, Pp, N, tol)) continue;
            const float wt = ((k & 1) ? frx : 1.0f - frx) * ((k >> 1) ? fry : 1.0f - fry);
            const float4 hv = hist_prev[q];
            hs[0] = hs[0] + hv.x * wt; hs[1] = hs[1] + hv.y * wt; hs[2] = hs[2] + hv.z * wt; hs[3] = hs[3] + hv.w * wt;
            ws = ws + wt;
        }
    }
    if (ws > 1e-3f) {
        const float hist[4] = {hs[0] / ws, hs[1] / ws, hs[2] / ws, hs[3] / ws};
        count = gmin(hist[2], 48.0f);
        const float wt = 1.0f / (count + 1.0f);
        sFast = mixf(hist[3], s, gmax(wt, 0.25f));
        const float sigma = sqrtf(gmax(hist[1] - hist[0] * hist[0], 0.0f));
        const float band = 3.0f * sigma * 0.378f + 0.02f;
        const float slow = fclampg(hist[0], sFast - band, sFast + band);
        sAcc = mixf(slow, s, wt);
        sM2 = mixf(hist[1], s * s, wt);
    }
    hist_cur[i] = make_float4(sAcc, sM2, count + 1.0f, sFast);
}

// readInput (atrous.comp:34-45).
template <bool ITER0>
__device__ __forceinline__ void read_input(size_t q, const uint2 *hist, const float4 *mom, const uint2 *in_cv, float col[3],
                                           float &var, float &len) {
    float v[4];
    if (ITER0) {
        h4(hist[q], v);
        const float4 m = mom[q];
        var = gmax(m.y - m.x * m.x, 0.0f);
        len = v[3];
    } else {
        h4(in_cv[q], v);
        var = v[3];
        len = 0.0f;
    }
    col[0] = v[0]; col[1] = v[1]; col[2] = v[2];
}

// GI_HISTFIX (atrous.comp:87-112): a texel with fewer than 4 frames of history takes the history-length
// weighted mean of converged texels on its surface, 7x7 taps 3 texels apart.  Rare (disocclusions), but
// a wavefront holding one such texel runs its 48 taps (~5000 instructions and their loads) on top of
// the filter; such wavefronts set the first level's tail.
__device__ __forceinline__ void history_fix(const TArgs &A, int32_t px, int32_t py, float posTw, const float P[3],
                                            const float N[3], const uint2 *hist, const uint4 *guide, float cLen,
                                            float cCol[3]) {
    float acc[3] = {cCol[0] * cLen, cCol[1] * cLen, cCol[2] * cLen};
    float ws = cLen;
    const float dinv = 1.0f / (1.0f + posTw * 0.01f);
    for (int dy = -3; dy <= 3; ++dy) {
        const int32_t qy = py + dy * 3;
        if (qy < 0 || qy >= A.ry) continue;
        for (int dx = -3; dx <= 3; ++dx) {
            const int32_t qx = px + dx * 3;
            if ((dx == 0 && dy == 0) || qx < 0 || qx >= A.rx) continue;
            const size_t q = (size_t)qy * A.rx + qx;
            const uint4 g = guide[q];
            if (g.w == 0u) continue;
            float gN[3];
            dec_n(g.w, gN);
            const float d[3] = {__uint_as_float(g.x) - P[0], __uint_as_float(g.y) - P[1], __uint_as_float(g.z) - P[2]};
            float hv[4];
            h4(hist[q], hv);
            const float wt = t_pow16(gmax(dot3(N, gN), 0.0f)) * t_exp(-fabsf(dot3(d, N)) * dinv) * hv[3];
            acc[0] = acc[0] + hv[0] * wt; acc[1] = acc[1] + hv[1] * wt; acc[2] = acc[2] + hv[2] * wt;
            ws = ws + wt;
        }
    }
    if (ws > 0.0f) { cCol[0] = acc[0] / ws; cCol[1] = acc[1] / ws; cCol[2] = acc[2] / ws; }
}

// One a-trous level (atrous.comp:73-154) per 16x16 lattice of texels: a workgroup takes the texels
// (X0 + i*step, Y0 + j*step), i, j < 16, of one phase (X0, Y0) of a 16*step-texel super-tile, so every
// tap of every centre lies on the same lattice and the 20x20 lattice points around the tile (2 on each
// side) hold all of them, whatever the level's tap spacing.  Those are staged in LDS first, per texel:
// the guide's position and decoded normal, the level's input colour, variance, luminance, history
// length (level 0) and a sky flag, computed once instead of by each of the 24 centres that tap it.
// A tap's LDS address is the centre's plus a constant.
template <bool ITER0> constexpr int stage_fields() { return ITER0 ? 13 : 12; }
constexpr int32_t kStageW = 20;   // lattice points per staged row (16 + 2 + 2)
// Row pitch of the staged fields: an odd multiple of 8 floats, so the 8 rows of a wavefront's 8x8
// texels fall in 8 different groups of 8 banks.
constexpr int32_t kStageP = 24;
constexpr int32_t kStageT = kStageP * kStageW;   // floats per staged field

struct Lattice { int32_t stx, phases, nwg; };
Lattice lattice_of(int32_t rx, int32_t ry, int32_t step) {
    Lattice g;
    const int32_t span = 16 * step;
    g.stx = (rx + span - 1) / span;
    g.phases = step * step;
    g.nwg = g.stx * ((ry + span - 1) / span) * g.phases;
    return g;
}

template <bool ITER0>
__global__ void __launch_bounds__(256) k_atrous(TArgs A, int32_t stx, int32_t histfix, const float4 *__restrict__ gb_pos,
                                                const uint32_t *__restrict__ gb_n, const float4 *__restrict__ mom,
                                                const uint2 *__restrict__ hist, const uint4 *__restrict__ guide,
                                                const uint2 *__restrict__ in_cv, uint2 *__restrict__ out_cv) {
    constexpr int F = stage_fields<ITER0>();
    constexpr int32_t T = kStageT;
    __shared__ float st[F * kStageT];
    const int32_t step = A.step, phases = step * step;
    const int32_t L = xcd_remap((int32_t)blockIdx.x, (int32_t)gridDim.x);
    const int32_t ph = L % phases, stile = L / phases;
    const int32_t X0 = (stile % stx) * 16 * step + ph % step, Y0 = (stile / stx) * 16 * step + ph / step;
    for (int32_t j = (int32_t)threadIdx.x; j < kStageW * kStageW; j += 256) {
        const int32_t cx = j % kStageW, cy = j / kStageW, i = cy * kStageP + cx;
        const int32_t qx = X0 + (cx - 2) * step, qy = Y0 + (cy - 2) * step;
        float v[F];
#pragma unroll
        for (int k = 0; k < F; ++k) v[k] = 0.0f;
        v[11] = 1.0f;   // outside the image: like sky, never a tap
        if (qx >= 0 && qy >= 0 && qx < A.rx && qy < A.ry) {
            const size_t q = (size_t)qy * A.rx + qx;
            const uint4 g = guide[q];
            float col[3], var, len;
            read_input<ITER0>(q, hist, mom, in_cv, col, var, len);
            v[0] = __uint_as_float(g.x); v[1] = __uint_as_float(g.y); v[2] = __uint_as_float(g.z);
            dec_n(g.w, &v[3]);
            v[6] = col[0]; v[7] = col[1]; v[8] = col[2];
            v[9] = var;
            v[10] = lumi(col[0], col[1], col[2]);
            v[11] = g.w == 0u ? 1.0f : 0.0f;
            if (ITER0) v[F - 1] = len;
        }
#pragma unroll
        for (int k = 0; k < F; ++k) st[k * T + i] = v[k];
    }
    __syncthreads();
    const int32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int32_t lx = (wv & 1) * 8 + (lane & 7), ly = (wv >> 1) * 8 + (lane >> 3);
    const int32_t px = X0 + lx * step, py = Y0 + ly * step;
    const bool inside = px < A.rx && py < A.ry;
    const size_t o = (size_t)py * A.rx + px;
    int32_t fx = 0, fy = 0;
    float4 posT = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    if (inside) posT = rep_texel(gb_pos, A.w, A.h, px, py, fx, fy);
    const bool live = inside && posT.w >= 0.0f;
    const float P[3] = {posT.x, posT.y, posT.z};
    float N[3] = {0.0f, 0.0f, 0.0f};
    if (live) dec_n(gb_n[(size_t)fy * A.w + fx], N);
    const int32_t ci = (ly + 2) * kStageP + (lx + 2);
    float cCol[3] = {st[6 * T + ci], st[7 * T + ci], st[8 * T + ci]};
    float cVar = st[9 * T + ci];
    const float cLen = ITER0 ? st[12 * T + ci] : 0.0f;
    if (ITER0 && histfix && live && cLen < 4.0f) history_fix(A, px, py, posT.w, P, N, hist, guide, cLen, cCol);
    if (!live) {
        if (inside) out_cv[o] = make_uint2(0u, 0u);
        return;
    }
    const float cLum = lumi(cCol[0], cCol[1], cCol[2]);
    float vg;   // centreVariance (atrous.comp:51-71); level 0 has step 1, so the 3x3 texels are staged
    if (!ITER0 || cLen >= 4.0f) {
        vg = cVar;
    } else {
        float s = 0.0f, s2 = 0.0f, n = 0.0f;
        for (int dy = -1; dy <= 1; ++dy) {
            const int32_t qy = imin(imax(py + dy, 0), A.ry - 1) - Y0 + 2;
            for (int dx = -1; dx <= 1; ++dx) {
                const int32_t qx = imin(imax(px + dx, 0), A.rx - 1) - X0 + 2;
                const float l = st[10 * T + qy * kStageP + qx];
                s = s + l;
                s2 = s2 + l * l;
                n = n + 1.0f;
            }
        }
        const float mean = s / n;
        vg = (gmax(s2 / n - mean * mean, 0.0f) + 1e-3f) * 2.0f;
    }
    if (ITER0 && cLen < 4.0f) cVar = vg;
    const float lumPhi = 4.0f * sqrtf(gmax(vg, 0.0f)) + 1e-4f;
    const float k00 = 0.375f * 0.375f;
    float cAcc[3] = {cCol[0] * k00, cCol[1] * k00, cCol[2] * k00};
    float vAcc = cVar * k00 * k00;
    float wSum = k00;
    const float depthScale = 1.0f * (1.0f + posT.w * 0.01f);
    const float zinv = 1.0f / depthScale, linv = 1.0f / lumPhi;
    const float h5[3] = {0.375f, 0.25f, 0.0625f};
#pragma unroll
    for (int dy = -2; dy <= 2; ++dy) {
#pragma unroll
        for (int dx = -2; dx <= 2; ++dx) {
            if (dx == 0 && dy == 0) continue;
            const int32_t li = ci + dy * kStageP + dx;
            if (st[11 * T + li] != 0.0f) continue;   // outside the image or a sky guide
            const float qn[3] = {st[3 * T + li], st[4 * T + li], st[5 * T + li]};
            const float qc[3] = {st[6 * T + li], st[7 * T + li], st[8 * T + li]};
            float qv = st[9 * T + li];
            if (ITER0 && st[12 * T + li] < 4.0f) qv = gmax(qv, vg);
            const float qLum = st[10 * T + li];
            const float kern = h5[dx < 0 ? -dx : dx] * h5[dy < 0 ? -dy : dy];
            const float wN = t_pow64(gmax(dot3(N, qn), 0.0f));
            const float d[3] = {st[li] - P[0], st[T + li] - P[1], st[2 * T + li] - P[2]};
            const float wZ = t_exp(-fabsf(dot3(d, N)) * zinv);
            const float wL = t_exp(-fabsf(cLum - qLum) * linv);
            const float wt = kern * wN * wZ * wL;
            cAcc[0] = cAcc[0] + wt * qc[0]; cAcc[1] = cAcc[1] + wt * qc[1]; cAcc[2] = cAcc[2] + wt * qc[2];
            vAcc = vAcc + wt * wt * qv;
            wSum = wSum + wt;
        }
    }
    const float inv = 1.0f / gmax(wSum, 1e-6f);
    out_cv[o] = pack_h4(cAcc[0] * inv, cAcc[1] * inv, cAcc[2] * inv, vAcc * inv * inv);
}

struct MotionArgs {
    float right[3], up[3], fwd[3], thf, jit[2];               // this frame's camera (the sample's ray)
    float ppos[3], pright[3], pup[3], pfwd[3], pthf;          // last frame's camera
    int32_t w, h, tiles_x, ntiles, n_bodies;
};

// reprojMotion (gbuffer.comp:91-99): prevUV - curUV of a previous-frame world point.
__device__ __forceinline__ void reproj_motion(const MotionArgs &A, const float Pp[3], float scx, float scy, float &mx,
                                              float &my) {
    const float aspect = (float)A.w / (float)A.h;
    const float dp[3] = {Pp[0] - A.ppos[0], Pp[1] - A.ppos[1], Pp[2] - A.ppos[2]};
    const float pz = dot3(dp, A.pfwd);
    if (pz <= 0.0f) { mx = 0.0f; my = 0.0f; return; }
    const float u = dot3(dp, A.pright) / (pz * A.pthf * aspect), v = dot3(dp, A.pup) / (pz * A.pthf);
    mx = (u * 0.5f + 0.5f) - scx / (float)A.w;
    my = (v * 0.5f + 0.5f) - scy / (float)A.h;
}

// gbuffer.comp's STORE_MOTION (gbuffer.comp:84-104) and the sky's motion (405-410) from the G-buffer
// just written: one thread per pixel.  w2o: rows of each body's worldToObj (3 float4 of 8 per body);
// prev: rows of its prevObjToWorld (3 float4 per body).
template <bool BODIES>
__global__ void __launch_bounds__(256) k_motion(MotionArgs A, const float4 *__restrict__ gb_pos, const int32_t *__restrict__ body,
                                                const float4 *__restrict__ hdr, const float4 *__restrict__ prev,
                                                uint32_t *__restrict__ motion, uint2 *__restrict__ delta) {
    const int32_t L = xcd_remap((int32_t)blockIdx.x, (int32_t)gridDim.x);
    const int32_t tx = L % A.tiles_x, ty = L / A.tiles_x;
    const int32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int32_t px = tx * 16 + (wv & 1) * 8 + (lane & 7);
    const int32_t py = ty * 16 + (wv >> 1) * 8 + (lane >> 3);
    if (px >= A.w || py >= A.h) return;
    const size_t i = (size_t)py * A.w + px;
    const float4 p = gb_pos[i];
    const float scx = ((float)px + 0.5f) + A.jit[0], scy = ((float)py + 0.5f) + A.jit[1];
    float Pp[3], d[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (p.w < 0.0f) {
        // the sample's view ray (gbuffer.comp:198-202), reprojected from infinity
        const float nx = (scx / (float)A.w) * 2.0f - 1.0f, ny = (scy / (float)A.h) * 2.0f - 1.0f;
        const float aspect = (float)A.w / (float)A.h;
        const float a = (nx * aspect) * A.thf, b = ny * A.thf;
        float rd[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) rd[k] = (A.fwd[k] + A.right[k] * a) + A.up[k] * b;
        normalize3(rd);
#pragma unroll
        for (int k = 0; k < 3; ++k) Pp[k] = A.ppos[k] + rd[k] * 1e6f;
    } else {
        Pp[0] = p.x; Pp[1] = p.y; Pp[2] = p.z;
        if (BODIES) {
            const int32_t ob = body[i];
            if (ob >= 0 && ob < A.n_bodies) {   // prevWorldPos (gbuffer.comp:84-88)
                const float4 *r = hdr + (size_t)ob * kBodyWords, *q = prev + (size_t)ob * 3;
                float l[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) l[k] = ((r[k].x * p.x + r[k].y * p.y) + r[k].z * p.z) + r[k].w;
#pragma unroll
                for (int k = 0; k < 3; ++k) Pp[k] = ((q[k].x * l[0] + q[k].y * l[1]) + q[k].z * l[2]) + q[k].w;
                d[3] = (float)(ob + 1);
            }
        }
        d[0] = p.x - Pp[0]; d[1] = p.y - Pp[1]; d[2] = p.z - Pp[2];
    }
    float mx, my;
    reproj_motion(A, Pp, scx, scy, mx, my);
    motion[i] = f2h(mx) | (f2h(my) << 16);
    delta[i] = pack_h4(d[0], d[1], d[2], d[3]);
}

TArgs targs(int32_t w, int32_t h, int32_t rx, int32_t ry) {
    TArgs A{};
    A.w = w; A.h = h; A.rx = rx; A.ry = ry;
    A.tiles_x = (rx + 15) / 16;
    A.ntiles = A.tiles_x * ((ry + 15) / 16);
    return A;
}

int check_ctx(vf_ctx *ctx, const char *what, int32_t w, int32_t h) {
    if (!ctx) return vf::fail(nullptr, VF_E_INVALID, "%s: null context", what);
    if (w <= 0 || h <= 0 || w > 32768 || h > 32768) return vf::fail(ctx, VF_E_INVALID, "%s: bad size %dx%d", what, w, h);
    return vf::set_device(ctx);
}

}  // namespace

extern "C" int vf_gbuffer_motion(vf_ctx *ctx, const vf_pass_ubo *pass, int32_t w, int32_t h, const float *gb_pos,
                                 const int32_t *body, uint32_t *gb_motion, uint16_t *gb_delta) {
    int rc = check_ctx(ctx, "vf_gbuffer_motion", w, h);
    if (rc) return rc;
    if (!pass || !gb_pos || !gb_motion || !gb_delta) return vf::fail(ctx, VF_E_INVALID, "vf_gbuffer_motion: null argument");
    MotionArgs A{};
    for (int k = 0; k < 3; ++k) {
        A.right[k] = pass->cam_right[k]; A.up[k] = pass->cam_up[k]; A.fwd[k] = pass->cam_fwd[k];
        A.ppos[k] = pass->prev_cam_pos[k]; A.pright[k] = pass->prev_cam_right[k];
        A.pup[k] = pass->prev_cam_up[k]; A.pfwd[k] = pass->prev_cam_fwd[k];
    }
    A.thf = pass->tan_half_fov;
    A.pthf = pass->prev_tan_half_fov;
    A.jit[0] = pass->jitter[0]; A.jit[1] = pass->jitter[1];
    A.w = w; A.h = h;
    A.tiles_x = (w + 15) / 16;
    A.ntiles = A.tiles_x * ((h + 15) / 16);
    A.n_bodies = ctx->n_bodies;
    if ((rc = vf::begin_timed(ctx))) return rc;
    if (body && ctx->n_bodies > 0)
        hipLaunchKernelGGL(k_motion<true>, dim3(A.ntiles), dim3(256), 0, ctx->stream, A, (const float4 *)gb_pos, body,
                           (const float4 *)ctx->b_hdr, (const float4 *)ctx->b_prev, gb_motion, (uint2 *)gb_delta);
    else
        hipLaunchKernelGGL(k_motion<false>, dim3(A.ntiles), dim3(256), 0, ctx->stream, A, (const float4 *)gb_pos,
                           (const int32_t *)nullptr, (const float4 *)nullptr, (const float4 *)nullptr, gb_motion,
                           (uint2 *)gb_delta);
    if ((rc = vf::launch_check(ctx, "k_motion"))) return rc;
    return vf::end_timed(ctx);
}

extern "C" int vf_gi_accumulate(vf_ctx *ctx, const vf_pass_ubo *pass, int32_t w, int32_t h, int32_t bicubic, const float *gb_pos,
                                const uint32_t *gb_normal, const float *gb_pos_prev, const uint32_t *gb_normal_prev,
                                const uint32_t *gb_motion, const uint16_t *gb_delta, const float *indirect,
                                const uint16_t *hist_prev, const float *mom_prev, uint16_t *hist_cur, float *mom_cur,
                                uint32_t *guide) {
    int rc = check_ctx(ctx, "vf_gi_accumulate", w, h);
    if (rc) return rc;
    if (!pass || !gb_pos || !gb_normal || !gb_pos_prev || !gb_normal_prev || !gb_motion || !gb_delta || !indirect ||
        !hist_prev || !mom_prev || !hist_cur || !mom_cur || !guide)
        return vf::fail(ctx, VF_E_INVALID, "vf_gi_accumulate: null argument");
    if ((const void *)hist_prev == (const void *)hist_cur || (const void *)mom_prev == (const void *)mom_cur)
        return vf::fail(ctx, VF_E_INVALID, "vf_gi_accumulate: the previous and current history must be different images");
    TArgs A = targs(w, h, (w + 1) / 2, (h + 1) / 2);
    A.frame = pass->frame_idx;
    for (int a = 0; a < 3; ++a) A.shift[a] = pass->shift_p[a];
    if ((rc = vf::begin_timed(ctx))) return rc;
    if (bicubic)
        hipLaunchKernelGGL(k_gi_accumulate<true>, dim3(A.ntiles), dim3(256), 0, ctx->stream, A, (const float4 *)gb_pos, gb_normal,
                           (const float4 *)gb_pos_prev, gb_normal_prev, gb_motion, (const uint2 *)gb_delta,
                           (const float4 *)indirect, (const uint2 *)hist_prev, (const float4 *)mom_prev, (uint2 *)hist_cur,
                           (float4 *)mom_cur, (uint4 *)guide);
    else
        hipLaunchKernelGGL(k_gi_accumulate<false>, dim3(A.ntiles), dim3(256), 0, ctx->stream, A, (const float4 *)gb_pos, gb_normal,
                           (const float4 *)gb_pos_prev, gb_normal_prev, gb_motion, (const uint2 *)gb_delta,
                           (const float4 *)indirect, (const uint2 *)hist_prev, (const float4 *)mom_prev, (uint2 *)hist_cur,
                           (float4 *)mom_cur, (uint4 *)guide);
    if ((rc = vf::launch_check(ctx, "k_gi_accumulate"))) return rc;
    return vf::end_timed(ctx);
}

extern "C" int vf_light_accumulate(vf_ctx *ctx, const vf_pass_ubo *pass, int32_t w, int32_t h, const float *gb_pos,
                                   const uint32_t *gb_normal, const float *gb_pos_prev, const uint32_t *gb_normal_prev,
                                   const uint32_t *gb_motion, const uint16_t *gb_delta, const float *shadow,
                                   const float *hist_prev, float *hist_cur) {
    int rc = check_ctx(ctx, "vf_light_accumulate", w, h);
    if (rc) return rc;
    if (!pass || !gb_pos || !gb_normal || !gb_pos_prev || !gb_normal_prev || !gb_motion || !gb_delta || !shadow || !hist_prev ||
        !hist_cur)
        return vf::fail(ctx, VF_E_INVALID, "vf_light_accumulate: null argument");
    if (hist_prev == hist_cur)
        return vf::fail(ctx, VF_E_INVALID, "vf_light_accumulate: the previous and current history must be different images");
    TArgs A = targs(w, h, w, h);
    A.frame = pass->frame_idx;
    for (int a = 0; a < 3; ++a) A.shift[a] = pass->shift_p[a];
    if ((rc = vf::begin_timed(ctx))) return rc;
    hipLaunchKernelGGL(k_light_accumulate, dim3(A.ntiles), dim3(256), 0, ctx->stream, A, (const float4 *)gb_pos, gb_normal,
                       (const float4 *)gb_pos_prev, gb_normal_prev, gb_motion, (const uint2 *)gb_delta, shadow,
                       (const float4 *)hist_prev, (float4 *)hist_cur);
    if ((rc = vf::launch_check(ctx, "k_light_accumulate"))) return rc;
    return vf::end_timed(ctx);
}

extern "C" int vf_gi_atrous(vf_ctx *ctx, int32_t w, int32_t h, int32_t levels, int32_t histfix, const float *gb_pos,
                            const uint32_t *gb_normal, const float *gi_mom, const uint16_t *gi_hist, const uint32_t *guide,
                            uint16_t *filt0, uint16_t *filt1) {
    int rc = check_ctx(ctx, "vf_gi_atrous", w, h);
    if (rc) return rc;
    if (levels < 1 || levels > 5) return vf::fail(ctx, VF_E_INVALID, "vf_gi_atrous: levels must be 1..5");
    if (!gb_pos || !gb_normal || !gi_mom || !gi_hist || !guide || !filt0 || (levels > 1 && !filt1))
        return vf::fail(ctx, VF_E_INVALID, "vf_gi_atrous: null argument");
    if (filt0 == filt1 || (const void *)filt0 == (const void *)gi_hist || (const void *)filt1 == (const void *)gi_hist)
        return vf::fail(ctx, VF_E_INVALID, "vf_gi_atrous: the two filter images and the history must be different images");
    TArgs A = targs(w, h, (w + 1) / 2, (h + 1) / 2);
    if ((rc = vf::begin_timed(ctx))) return rc;
    for (int32_t i = 0; i < levels; ++i) {
        A.it = i;
        A.step = 1 << i;
        uint2 *dst = (uint2 *)((i & 1) ? filt1 : filt0);
        const uint2 *src = (const uint2 *)((i & 1) ? filt0 : filt1);
        const Lattice g = lattice_of(A.rx, A.ry, A.step);
        const int32_t hf = histfix ? 1 : 0;
        const float4 *gp = (const float4 *)gb_pos, *gm = (const float4 *)gi_mom;
        const uint2 *gh = (const uint2 *)gi_hist;
        const uint4 *gg = (const uint4 *)guide;
        if (i == 0)
            hipLaunchKernelGGL(k_atrous<true>, dim3(g.nwg), dim3(256), 0, ctx->stream, A, g.stx, hf, gp, gb_normal, gm, gh, gg,
                               (const uint2 *)nullptr, dst);
        else
            hipLaunchKernelGGL(k_atrous<false>, dim3(g.nwg), dim3(256), 0, ctx->stream, A, g.stx, hf, gp, gb_normal, gm, gh, gg,
                               src, dst);
        if ((rc = vf::launch_check(ctx, "k_atrous"))) return rc;
    }
    return vf::end_timed(ctx);
}
