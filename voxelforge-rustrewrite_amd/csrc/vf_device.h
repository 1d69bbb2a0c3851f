// Device-side traversal shared by the pass kernels (vf_trace.hip) and the queue kernels (vf_queue.hip):
// the world view, traceVolumeS / brickVoxels (reference shaders/common.glsl:474-658) in the reference's
// FP32 operation order, and the RNG / sampling helpers of the passes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/vf_trace.h"

// ------------------------------------------------------------------------------------------------
// device-side world
// ------------------------------------------------------------------------------------------------
struct DevWorld {
    int32_t wb[3];
    int32_t minb[3];     // worldMinB(): (win.x, 0, win.z) in ring mode, else 0
    int32_t ring;
    int32_t use_dist;    // USE_DIST_FIELD
    int32_t max_outer;   // 3 * (wbx + wby + wbz) + 4
    const uint2 *cells;  // .x grid value, .y meta: dist | dry << 8 | c2empty << 16 | c1empty << 17
    const uint8_t *atlas;
    const uint4 *occ;     // per slot 4 x uint4: bit (lz*64 + ly*8 + lx) = voxel != 0
    const uint4 *occdry;  // per slot: voxel != 0 && voxel != 5
    const uint16_t *skytop;
    const float4 *palette;  // 256 x 3 float4 (Material)
};

struct SkyParams {
    float sun_dir[3], sun_rgb[3], sky_zenith[3], sky_horizon[3], sunset_rgb[3];
    float sky_scale, gi_scale, night;
    int32_t sky_model;
};

struct Cnt {
    unsigned long long rays, outer, voxel, bytes;
};

struct Hit {
    int32_t hit;
    float t;
    int32_t face;
    uint32_t mat;
    int32_t cell[3];
    int32_t obj;   // the body hit (-1 = the world); set and read only by the body-aware paths
};

// GLSL built-ins with their specified definitions (signed zeros matter for bit equality).
__device__ __forceinline__ float gmin(float x, float y) { return y < x ? y : x; }
__device__ __forceinline__ float gmax(float x, float y) { return x < y ? y : x; }
__device__ __forceinline__ int32_t imin(int32_t x, int32_t y) { return y < x ? y : x; }
__device__ __forceinline__ int32_t imax(int32_t x, int32_t y) { return x < y ? y : x; }
__device__ __forceinline__ int32_t iclamp(int32_t x, int32_t lo, int32_t hi) { return imin(imax(x, lo), hi); }
__device__ __forceinline__ float fclampg(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
// float -> int32 truncating, NaN -> 0, saturating at both ends: exactly what v_cvt_i32_f32 does, as one
// instruction (a C++ cast is undefined out of range, and the checks written out compiled to three nested
// branches per conversion).
__device__ __forceinline__ int32_t f2i(float x) {
    int32_t r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}
// 1 / x rounded to nearest, for |x| >= 1e-5 (the traversal clamps direction components there): the hardware
// reciprocal refined by one Newton-Raphson step with fused multiply-adds (e = 1 - x*y, y + y*e).  For every
// float with 1e-5 <= |x| < 2^126, both signs, this equals the IEEE quotient 1.0f / x bit for bit (checked
// exhaustively on gfx950: 1.2 G values); from 2^126 on, where 1/x is subnormal, it does not, and the
// division is taken instead.  Three instructions in place of the division's eleven.
__device__ __forceinline__ float rcp_rn(float x) {
    const float y = __builtin_amdgcn_rcpf(x);
    float r = __builtin_fmaf(__builtin_fmaf(-x, y, 1.0f), y, y);
    if (__builtin_expect(fabsf(x) >= 0x1p126f, 0)) r = 1.0f / x;
    return r;
}
__device__ __forceinline__ float dot3(const float a[3], const float b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
__device__ __forceinline__ void cross3(const float a[3], const float b[3], float o[3]) {
    float x = a[1] * b[2] - b[1] * a[2];
    float y = a[2] * b[0] - b[2] * a[0];
    float z = a[0] * b[1] - b[0] * a[1];
    o[0] = x; o[1] = y; o[2] = z;
}
__device__ __forceinline__ void normalize3(float v[3]) {
    float len = sqrtf(dot3(v, v));
    v[0] = v[0] / len; v[1] = v[1] / len; v[2] = v[2] / len;
}

// Cell word index of absolute brick (x, y, z) inside the window (gridWrap, common.glsl:134-136): the grid's
// own x-fastest order.  (4x4x4-brick tiles cost 10 more VALU instructions per brick step for no
// measurable cache gain: the traversal kernels are issue-bound and their world reads hit the L1.)
// Both products are of factors below 2^24 (brick dims <= 4096, so z * wby + y < 2^24): 24-bit multiply-adds
// (one full-rate instruction each, written as the instruction: the compiler widens the first to a 64-bit
// multiply-add), the same low 32 bits as 32-bit multiplies.
__device__ __forceinline__ uint32_t mad_u24(uint32_t a, uint32_t b, uint32_t c) {
    uint32_t r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
    return r;
}
__device__ __forceinline__ uint32_t cell_index(const DevWorld &W, int32_t x, int32_t y, int32_t z) {
    if (W.ring) { x &= W.wb[0] - 1; z &= W.wb[2] - 1; }
    return mad_u24(mad_u24((uint32_t)z, (uint32_t)W.wb[1], (uint32_t)y), (uint32_t)W.wb[0], (uint32_t)x);
}

// voxelMatAt (common.glsl:976-985): the world voxel's material at p (0 = air or outside the loaded box).
__device__ __forceinline__ uint32_t voxel_mat_at(const DevWorld &W, const float p[3]);

// 64 occupancy bits of brick layer lz (two 32-bit words) out of the 512-bit mask.
__device__ __forceinline__ void mask_layer(const uint4 &m0, const uint4 &m1, const uint4 &m2, const uint4 &m3, int lz,
                                           uint32_t &lo, uint32_t &hi) {
    uint4 a = (lz & 2) ? m1 : m0;
    uint4 b = (lz & 2) ? m3 : m2;
    uint4 q = (lz & 4) ? b : a;
    lo = (lz & 1) ? q.z : q.x;
    hi = (lz & 1) ? q.w : q.y;
}

// Per-lane copy of the current brick's 512-bit mask for the LDS variant of the voxel DDA: word k of the
// lane with thread id t at vf_vmask[k * 256 + t] (consecutive lanes in consecutive banks).  Only
// kernels of 256 threads that instantiate brick_voxels<., true> reference it (16 KB per workgroup).
static __shared__ uint32_t vf_vmask[16 * 256];
// The same for kernels of 64 threads (LS = 64: 4 KB per workgroup).
static __shared__ uint32_t vf_vmask64[16 * 64];

// Occupancy bit of voxel (lx, ly) in a 64-bit layer (lo, hi), for q = lx | ly << 4 (higher bits ignored): bit
// lx + 8 ly.  The shift is written as the instruction, which takes its amount modulo 64, so the index is
// q's bits 0-2 merged with bits 3-5 of q >> 1 (one bitfield insert) with nothing masked off above them.
__device__ __forceinline__ uint32_t layer_bit(uint32_t lo, uint32_t hi, uint32_t q) {
    uint32_t idx;
    asm("v_bfi_b32 %0, 7, %1, %2" : "=v"(idx) : "v"(q), "v"(q >> 1));   // (q & 7) | ((q >> 1) & ~7)
    uint64_t r;
    asm("v_lshrrev_b64 %0, %1, %2" : "=v"(r) : "v"(idx), "v"(((uint64_t)hi << 32) | (uint64_t)lo));
    return (uint32_t)r & 1u;
}

// brickVoxels (common.glsl:474-539), USE_BRICK_MASKS off: the same voxel sequence and tMax
// arithmetic; solidity comes from the brick's occupancy mask instead of one atlas fetch per step.
// LDSM: the mask goes to the lane's LDS slots and a step along z reads the new 64-bit layer from
// there (one LDS read instead of a 16-to-2 word select in registers, which a wave executes on almost
// every step since some lane steps along z).
// BOX (the small-frame traversal, LDSM only): before the walk, the brick's mask is tested over the box of
// voxels the walk can visit; tout = the brick DDA's time of leaving the brick (the least of its tMax).  A
// walk that cannot hit is skipped: its result is the same miss, after which the brick DDA goes on from its
// own state as it does after any walk.
template <bool COUNT, bool LDSM = false, int LS = 256, bool BOX = false, bool UNI = false>
__device__ __forceinline__ bool brick_voxels(const DevWorld &W, const float ro[3], const float sd[3], const float inv[3],
                                             const int32_t st[3], const float up[3], const int32_t faceA[3],
                                             const int32_t bc[3], uint32_t g, float t, int32_t face, Hit &h, Cnt &c,
                                             const int skip, const float tout = 0.0f) {
    const uint32_t slot = g - 256u;
    const uint4 *mp = (skip == 5 ? W.occdry : W.occ) + (size_t)slot * 4u;
    const uint4 m0 = mp[0], m1 = mp[1], m2 = mp[2], m3 = mp[3];
    uint32_t *sm = (LS == 64 ? vf_vmask64 : vf_vmask) + threadIdx.x;
    if (LDSM && !BOX) {
        sm[0 * LS] = m0.x; sm[1 * LS] = m0.y; sm[2 * LS] = m0.z; sm[3 * LS] = m0.w;
        sm[4 * LS] = m1.x; sm[5 * LS] = m1.y; sm[6 * LS] = m1.z; sm[7 * LS] = m1.w;
        sm[8 * LS] = m2.x; sm[9 * LS] = m2.y; sm[10 * LS] = m2.z; sm[11 * LS] = m2.w;
        sm[12 * LS] = m3.x; sm[13 * LS] = m3.y; sm[14 * LS] = m3.z; sm[15 * LS] = m3.w;
    }
    int32_t bmin[3], vc[3];
    float tM[3], tD[3];
    const float e = t + 1e-3f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        bmin[a] = bc[a] * 8;
        float q = ro[a] + sd[a] * e;
        vc[a] = iclamp(f2i(floorf(q)), bmin[a], bmin[a] + 7);
        tM[a] = (((float)vc[a] + up[a]) - ro[a]) * inv[a];
        tD[a] = fabsf(inv[a]);
    }
    int32_t lx = vc[0] - bmin[0], ly = vc[1] - bmin[1], lz = vc[2] - bmin[2];
    if (BOX) {
        // Every step of the walk moves one coordinate by the ray's fixed sign, from the entry voxel to the
        // voxel where the ray leaves the brick: per axis between lv and the coordinate at tout.  That
        // coordinate is taken 0.05 voxel further along the ray (the walk's and the brick DDA's tMax
        // differ by rounding only, ~1e-6 of t) and clamped into the brick.  No solid voxel in the box's rows
        // and columns of any of the 8 layers: the walk cannot hit.  (Its z range is not applied: testing
        // every layer costs fewer instructions than the range, and measured faster.)
        int32_t blo[2], bhi[2];
        const int32_t lv[2] = {lx, ly};
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const float q = (ro[a] + sd[a] * tout) - (float)bmin[a];
            const int32_t ev = iclamp(f2i(floorf(st[a] > 0 ? q + 0.05f : q - 0.05f)), 0, 7);
            blo[a] = imin(lv[a], ev);
            bhi[a] = imax(lv[a], ev);
        }
        const uint32_t rep = ((2u << bhi[0]) - (1u << blo[0])) * 0x01010101u;   // columns, in every row
        const uint64_t top = bhi[1] == 7 ? ~0ull : ((1ull << (8 * (bhi[1] + 1))) - 1ull);
        const uint64_t rows = top & ~((1ull << (8 * blo[1])) - 1ull);
        const uint32_t rlo = rep & (uint32_t)rows, rhi = rep & (uint32_t)(rows >> 32);
        const uint32_t w[16] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w,
                                m2.x, m2.y, m2.z, m2.w, m3.x, m3.y, m3.z, m3.w};
        uint32_t any = 0u;
#pragma unroll
        for (int z = 0; z < 8; ++z) any |= (w[2 * z] & rlo) | (w[2 * z + 1] & rhi);
        if (any == 0u) return false;
        sm[0 * LS] = m0.x; sm[1 * LS] = m0.y; sm[2 * LS] = m0.z; sm[3 * LS] = m0.w;
        sm[4 * LS] = m1.x; sm[5 * LS] = m1.y; sm[6 * LS] = m1.z; sm[7 * LS] = m1.w;
        sm[8 * LS] = m2.x; sm[9 * LS] = m2.y; sm[10 * LS] = m2.z; sm[11 * LS] = m2.w;
        sm[12 * LS] = m3.x; sm[13 * LS] = m3.y; sm[14 * LS] = m3.z; sm[15 * LS] = m3.w;
    }
    uint32_t wlo, whi;
    if (LDSM) {
        wlo = sm[(lz * 2) * LS];
        whi = sm[(lz * 2 + 1) * LS];
    } else {
        mask_layer(m0, m1, m2, m3, lz, wlo, whi);
    }
    float tv = t;
    int32_t f2 = face;
    bool hit = false;
    // Both walks take one DDA step per trip, written with selects instead of a 3-way branch: the axis
    // choice and the updates are the reference's (x when its tMax is below both others, else y when below
    // z's, else z: strict <, so a tie goes to the later axis; tMax += tDelta).  Scalars, not arrays, so
    // nothing is indexed.  A step out of the brick ends the walk without a hit (the caller's brick DDA goes
    // on from its own state).  The reference caps the walk at 26 trips; it cannot reach the cap: every step
    // moves one coordinate by the ray's fixed sign, so each axis leaves [0, 7] after at most 8 steps and
    // the walk ends within 22 steps.  Each loop therefore has the single exit hit | out.
    float tMx = tM[0], tMy = tM[1], tMz = tM[2];
    const float tDx = tD[0], tDy = tD[1], tDz = tD[2];
    const int32_t fA0 = faceA[0], fA1 = faceA[1], fA2 = faceA[2];
    const int32_t s0 = st[0], s1 = st[1], s2 = st[2];
    if (LDSM) {
        // The walk in mirrored coordinates: u_a = l_a when the ray steps +1 along axis a, 7 - l_a when it
        // steps -1, so every step adds 1 to one u_a.  The three are packed in one word (u_x | u_y << 4 |
        // u_z << 8): a step is one addition, the step that leaves the brick is the one that takes a field
        // to 8 (bit 3 of its nibble), and the voxel's own coordinates are q = u ^ mir.  A trip takes a step,
        // then tests the voxel it stepped into; the entry voxel is tested before the first trip.  The
        // entry time and face of a hit come from the last step's axis (its increment d) and the tMax
        // values, which are advanced only when the walk goes on.  Steps, comparisons and additions are
        // those of the one-step loop below.
        const uint32_t mir = (s0 > 0 ? 0u : 7u) | (s1 > 0 ? 0u : 0x70u) | (s2 > 0 ? 0u : 0x700u);
        uint32_t q = (uint32_t)lx | ((uint32_t)ly << 4) | ((uint32_t)lz << 8);
        uint32_t u = q ^ mir, d = 0u;
        if (COUNT) { c.voxel++; c.bytes += 1; }
        hit = layer_bit(wlo, whi, q) != 0u;
        if (!hit) {
            // One basic block per trip: the loop's only branch is its exit, the tMax values are advanced by
            // selects that keep them on the exit trip, and hit / out are values in registers.  UNI (a
            // direction shared by the launch, the sun's): the tDelta values are scalars, and a select cannot
            // take a scalar operand beside its scalar condition, so they are copied to vector registers once
            // here rather than on every trip.
            float tDxv = tDx, tDyv = tDy, tDzv = tDz;
            if (UNI) asm volatile("" : "+v"(tDxv), "+v"(tDyv), "+v"(tDzv));
            uint32_t hb, ob;
#pragma unroll 1
            for (;;) {
                const bool a0 = tMx < tMy && tMx < tMz;
                const bool a1 = !a0 && tMy < tMz;
                const bool a2 = !a0 && !a1;
                d = a0 ? 1u : (a1 ? 16u : 256u);
                u += d;
                q = u ^ mir;
                // the voxel's layer (its slot index wraps into the lane's 8 layers when the step left the
                // brick; that value is then not used)
                const uint32_t nz = (q >> 8) & 7u;
                const uint32_t lo = sm[(nz * 2) * LS], hi2 = sm[(nz * 2 + 1) * LS];
                hb = layer_bit(lo, hi2, q);
                ob = u & 0x888u;
                const bool stop = (hb | ob) != 0u;
                if (COUNT && ob == 0u) { c.voxel++; c.bytes += 1; }   // a voxel inside the brick was tested
                // adding -0 leaves every value as it is (-0 included)
                tMx = tMx + ((a0 && !stop) ? tDxv : -0.0f);
                tMy = tMy + ((a1 && !stop) ? tDyv : -0.0f);
                tMz = tMz + ((a2 && !stop) ? tDzv : -0.0f);
                if (stop) break;
            }
            asm volatile("" : "+v"(d));   // read d as the register it is (not as the loop's axis conditions)
            hit = ob == 0u;   // the loop stopped on a solid voxel inside the brick
            if (hit) {
                // the step into the hit voxel: its axis from d; tMax of that axis is still its value before
                // the step (the exit trip does not advance it)
                tv = d == 1u ? tMx : (d == 16u ? tMy : tMz);
                f2 = d == 1u ? fA0 : (d == 16u ? fA1 : fA2);
            }
        }
        if (hit) { lx = (int32_t)(q & 15u); ly = (int32_t)((q >> 4) & 15u); lz = (int32_t)(q >> 8); }
    } else
#pragma unroll 1
    for (;;) {   // registers only: the layer is re-selected from the mask words after a step along z
        if (COUNT) { c.voxel++; c.bytes += 1; }
        const uint32_t idx = ((uint32_t)ly << 3) | (uint32_t)lx;   // bit of (lx, ly) in the 64-bit layer
        const uint64_t layer = ((uint64_t)whi << 32) | (uint64_t)wlo;
        hit = ((layer >> idx) & 1ull) != 0ull;
        const bool lt_yz = tMy < tMz;
        const bool a0 = tMx < tMy && tMx < tMz;
        const bool a1 = !a0 && lt_yz;
        const bool a2 = !a0 && !lt_yz;
        const int32_t nlx = lx + (a0 ? s0 : 0), nly = ly + (a1 ? s1 : 0), nlz = lz + (a2 ? s2 : 0);
        // the next voxel's layer, read as soon as its z is known (the slot index wraps to stay in the
        // lane's 8 layers when the step leaves the brick; that value is never used, and the layer is
        // not needed after the walk, so it is replaced before the exit test)
        if (LDSM) {
            wlo = sm[((nlz & 7) * 2) * LS];
            whi = sm[((nlz & 7) * 2 + 1) * LS];
        }
        const bool out = ((uint32_t)nlx | (uint32_t)nly | (uint32_t)nlz) > 7u;
        if (hit | out) break;
        tv = a0 ? tMx : (a1 ? tMy : tMz);
        f2 = a0 ? fA0 : (a1 ? fA1 : fA2);
        const float nx = tMx + tDx, ny = tMy + tDy, nz = tMz + tDz;
        tMx = a0 ? nx : tMx;
        tMy = a1 ? ny : tMy;
        tMz = a2 ? nz : tMz;
        lx = nlx; ly = nly; lz = nlz;
        if (!LDSM && a2) mask_layer(m0, m1, m2, m3, lz, wlo, whi);
    }
    if (hit) {
        h.hit = 1;
        h.t = tv;
        h.face = f2;
        h.mat = W.atlas[(size_t)slot * 512u + (uint32_t)(lz * 64 + ly * 8 + lx)];
        h.cell[0] = bmin[0] + lx; h.cell[1] = bmin[1] + ly; h.cell[2] = bmin[2] + lz;
        return true;
    }
    return false;
}

// The per-direction set-up of trace_world for a direction shared by every ray of a launch (the sun's),
// computed once on the host by dir_pre with trace_world's own operations: the clamped direction, its
// reciprocal (IEEE float division), the step sign, the upper-plane flag, the face code of each axis
// (normal -st) and the brick tDelta.
struct DirPre {
    float sd[3], inv[3], up[3], tD8[3];
    int32_t st[3], faceA[3];
};
__host__ __device__ inline DirPre dir_pre(const float d[3]) {
    DirPre p;
    for (int a = 0; a < 3; ++a) {
        p.sd[a] = fabsf(d[a]) < 1e-5f ? 1e-5f : d[a];
        p.inv[a] = 1.0f / p.sd[a];
        p.st[a] = p.sd[a] > 0.0f ? 1 : -1;
        p.up[a] = p.sd[a] < 0.0f ? 0.0f : 1.0f;
        p.faceA[a] = 2 * a + (p.st[a] > 0 ? 2 : 1);
        p.tD8[a] = fabsf(p.inv[a]) * 8.0f;
    }
    return p;
}

// traceVolumeS over the world (common.glsl:541-658) with WATER_CLEAR; SKIP is skipMat (0 or 5), or -1 to
// take it from skip_rt (one inlined copy serving both).
// PRE: pre holds the set-up of rd (dir_pre), so rays of a shared direction skip three divisions and the
// per-axis selects.
// BOX: the small-frame traversal (LDSM only, never counting): each detailed brick's walk is first tested
// against its mask over the voxels it can visit, and a wave raises its issue priority after 4, 10 and 16
// brick steps, so the few long rays that set a small frame's time are not held back by the frame's other
// waves (the priority stays raised for the wave's later rays).
template <int SKIP, bool COUNT, bool LDSM = false, int LS = 256, bool PRE = false, bool BOX = false>
__device__ __forceinline__ void trace_world(const DevWorld &W, const float ro[3], const float rd[3], float tcap, Hit &h,
                                            Cnt &c, const int skip_rt = 0, const DirPre *pre = nullptr) {
    const int sk = SKIP >= 0 ? SKIP : skip_rt;
    h.hit = 0; h.t = -1.0f; h.face = 0; h.mat = 0; h.cell[0] = h.cell[1] = h.cell[2] = 0;
    if (COUNT) c.rays++;
    float sd[3], inv[3], up[3], tmin3[3], tmax3[3];
    int32_t st[3], faceA[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        sd[a] = PRE ? pre->sd[a] : (fabsf(rd[a]) < 1e-5f ? 1e-5f : rd[a]);
        inv[a] = PRE ? pre->inv[a] : rcp_rn(sd[a]);
        const float lo = (float)(W.minb[a] * 8);
        const float dims = (float)(W.wb[a] * 8);
        const float t0 = (lo - ro[a]) * inv[a];
        const float t1 = ((lo + dims) - ro[a]) * inv[a];
        tmin3[a] = gmin(t0, t1);
        tmax3[a] = gmax(t0, t1);
        st[a] = PRE ? pre->st[a] : (sd[a] > 0.0f ? 1 : -1);
        up[a] = PRE ? pre->up[a] : (sd[a] < 0.0f ? 0.0f : 1.0f);
        faceA[a] = PRE ? pre->faceA[a] : 2 * a + (st[a] > 0 ? 2 : 1);   // face code of normal -st on axis a
    }
    const float tEnter = gmax(gmax(tmin3[0], tmin3[1]), tmin3[2]);
    const float tExit = gmin(gmin(tmax3[0], tmax3[1]), tmax3[2]);
    if (tExit < gmax(tEnter, 0.0f) || tEnter > tcap) return;

    float t = gmax(tEnter, 0.0f);
    int32_t bc[3];
    float tD8[3], tM8[3];
    {
        const float e = t + 1e-3f;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float entry = ro[a] + sd[a] * e;
            bc[a] = iclamp(f2i(floorf(entry * 0.125f)), W.minb[a], W.minb[a] + W.wb[a] - 1);
            tD8[a] = PRE ? pre->tD8[a] : fabsf(inv[a]) * 8.0f;
            tM8[a] = (((float)bc[a] + up[a]) * 8.0f - ro[a]) * inv[a];
        }
    }
    int32_t face = 0;
    if (tEnter > 0.0f) {
        if (tEnter == tmin3[0]) face = faceA[0];
        else if (tEnter == tmin3[1]) face = faceA[1];
        else face = faceA[2];
    }
    static_assert(!BOX || (LDSM && !COUNT), "the small-frame traversal is an LDS-mask, non-counting path");
    for (int i = 0; i < W.max_outer; ++i) {
        if (t > tcap) return;
        if (BOX && i == 4) __builtin_amdgcn_s_setprio(1);
        if (BOX && i == 10) __builtin_amdgcn_s_setprio(2);
        if (BOX && i == 16) __builtin_amdgcn_s_setprio(3);
        const uint2 cw = W.cells[cell_index(W, bc[0], bc[1], bc[2])];
        uint32_t g = cw.x;
        const uint32_t meta = cw.y;
        if (COUNT) { c.outer++; c.bytes += 4; }
        if (sk == 5) {
            if (g == 5u) g = 0u;
            else if (g >= 256u) {
                if (COUNT) c.bytes += 1;
                if (((meta >> 8) & 255u) > 0u) g = 0u;
            }
        }
        if (g >= 256u) {
            if (brick_voxels<COUNT, LDSM, LS, BOX, PRE>(W, ro, sd, inv, st, up, faceA, bc, g, t, face, h, c, sk,
                                                   BOX ? gmin(gmin(tM8[0], tM8[1]), tM8[2]) : 0.0f))
                return;
        } else if (g != 0u && g != (uint32_t)sk) {
            const float e = t + 1e-3f;
            h.hit = 1; h.t = t; h.face = face; h.mat = g;
#pragma unroll
            for (int a = 0; a < 3; ++a) h.cell[a] = f2i(floorf(ro[a] + sd[a] * e));
            return;
        }
        bool stepped = false;
        if (g == 0u) {
            int32_t Cf = 0, cb[3] = {0, 0, 0};
            int32_t d = 0;
            if (W.use_dist) {
                d = (int32_t)(sk == 5 ? ((meta >> 8) & 255u) : (meta & 255u));
                if (COUNT) c.bytes += 1;
            }
            if (W.use_dist && d > 1) {
                Cf = 2 * d - 1;
#pragma unroll
                for (int a = 0; a < 3; ++a) cb[a] = bc[a] - (d - 1);
            } else {
                if (COUNT) c.bytes += 1;
                if (meta & (1u << 16)) Cf = 16;
                else {
                    if (COUNT) c.bytes += 1;
                    Cf = (meta & (1u << 17)) ? 4 : 0;
                }
                if (Cf > 0) {
#pragma unroll
                    for (int a = 0; a < 3; ++a) cb[a] = bc[a] - (bc[a] & (Cf - 1));
                }
            }
            if (Cf > 0) {
                float tFar[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const float cmin = (float)cb[a] * 8.0f;
                    const float cmax = cmin + (float)Cf * 8.0f;
                    tFar[a] = ((up[a] != 0.0f ? cmax : cmin) - ro[a]) * inv[a];
                }
                const bool ax0 = tFar[0] < tFar[1] && tFar[0] < tFar[2];
                const bool ax1 = !ax0 && tFar[1] < tFar[2];
                const bool ax2 = !ax0 && !ax1;
                t = ax0 ? tFar[0] : (ax1 ? tFar[1] : tFar[2]);
                face = ax0 ? faceA[0] : (ax1 ? faceA[1] : faceA[2]);
                const float e = t + 1e-3f;
                const bool axs[3] = {ax0, ax1, ax2};
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    int32_t v = iclamp(f2i(floorf((ro[a] + sd[a] * e) * 0.125f)), cb[a], cb[a] + Cf - 1);
                    if (axs[a]) v = st[a] > 0 ? cb[a] + Cf : cb[a] - 1;
                    bc[a] = v;
                    tM8[a] = (((float)bc[a] + up[a]) * 8.0f - ro[a]) * inv[a];
                }
                stepped = true;
            }
        }
        if (!stepped) {   // one brick step, the walk's axis rule (common.glsl:644-655)
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
        }
        // outside the loaded box on any axis (one unsigned comparison per axis, no short-circuit branches)
        if (((uint32_t)(bc[0] - W.minb[0]) >= (uint32_t)W.wb[0]) | ((uint32_t)(bc[1] - W.minb[1]) >= (uint32_t)W.wb[1]) |
            ((uint32_t)(bc[2] - W.minb[2]) >= (uint32_t)W.wb[2]))
            return;
    }
}

// ------------------------------------------------------------------------------------------------
// shared helpers of the passes
// ------------------------------------------------------------------------------------------------
// Workgroup id -> logical tile.  Workgroups are dealt round-robin to the 8 XCDs (each with its own L2);
// this gives each XCD runs of 16 consecutive tiles (a 256 x 16-pixel strip of a 16 x 16-tile pass), the
// runs of the 8 XCDs interleaved: neighbouring rays still share an L2, and every XCD gets a share of
// every part of the image.  (One contiguous band per XCD kept the L2s apart but left the XCDs with
// unequal work: up to 26 % longer passes.)  A bijection on [0, n); the last n mod 128 ids map to
// themselves.
__device__ __forceinline__ int32_t xcd_remap(int32_t i, int32_t n) {
    constexpr int32_t C = 16;
    const int32_t full = (n / (8 * C)) * (8 * C);
    if (i >= full) return i;
    const int32_t x = i & 7, k = i >> 3;
    return ((k / C) * 8 + x) * C + (k % C);
}

// xcd_remap for 64-thread blocks that each take a quarter of a 16x16 tile: runs of 64 blocks (16 tiles)
// per XCD, as xcd_remap deals 16 whole tiles.
__device__ __forceinline__ int32_t xcd_remap64(int32_t i, int32_t n) {
    constexpr int32_t C = 64;
    const int32_t full = (n / (8 * C)) * (8 * C);
    if (i >= full) return i;
    const int32_t x = i & 7, k = i >> 3;
    return ((k / C) * 8 + x) * C + (k % C);
}

__device__ __forceinline__ void store_hit(vf_hit *out, const Hit &h, bool body = false) {
    uint32_t *o = reinterpret_cast<uint32_t *>(out);
    if (h.hit) {
        o[0] = (uint32_t)h.cell[0]; o[1] = (uint32_t)h.cell[1]; o[2] = (uint32_t)h.cell[2];
        o[3] = __float_as_uint(h.t);
        o[4] = (h.mat & 255u) | ((uint32_t)h.face << 8) | (1u << 16) | ((uint32_t)body << 24);
    } else {
        o[0] = 0u; o[1] = 0u; o[2] = 0u;
        o[3] = __float_as_uint(-1.0f);
        o[4] = 0u;
    }
}

template <bool COUNT>
__device__ __forceinline__ void flush_count(const Cnt &c, unsigned long long *cnt) {
    if (COUNT) {
        atomicAdd(cnt + 0, c.rays);
        atomicAdd(cnt + 1, c.outer);
        atomicAdd(cnt + 2, c.voxel);
        atomicAdd(cnt + 3, c.bytes);
    }
}

__device__ __forceinline__ uint32_t unorm8(float v) { return (uint32_t)floorf(fclampg(v, 0.0f, 1.0f) * 255.0f + 0.5f); }
__device__ __forceinline__ float unorm8_dec(uint32_t b) { return ((float)b / 255.0f) * 2.0f - 1.0f; }

__device__ __forceinline__ void face_normal(int32_t face, float n[3]) {
    n[0] = n[1] = n[2] = 0.0f;
    if (face > 0) n[(face - 1) >> 1] = (face & 1) ? 1.0f : -1.0f;
}
// unorm8(n * 0.5 + 0.5) of face_normal(face)'s components, packed x | y << 8 | z << 16 (the G-buffer's
// RGBA8 normal without its alpha).  The components are -1, 0 or +1, whose bytes are exactly 0, 128 and 255
// (floor(0 * 255 + 0.5), floor(0.5 * 255 + 0.5), floor(255 + 0.5)), so the bytes come from the face code
// with integer operations instead of three float conversions.
__device__ __forceinline__ uint32_t face_normal_bytes(int32_t face) {
    const uint32_t sh = ((uint32_t)((face - 1) >> 1) & 3u) * 8u;   // (& 3: a valid shift for face 0 too)
    const uint32_t b = (face & 1) ? 0xFFu : 0u;
    return face > 0 ? ((0x808080u & ~(0xFFu << sh)) | (b << sh)) : 0x808080u;
}

// PCG hash and samplers (common.glsl:1233-1288)
__device__ __forceinline__ uint32_t pcg(uint32_t v) {
    v = v * 747796405u + 2891336453u;
    v = ((v >> ((v >> 28u) + 4u)) ^ v) * 277803737u;
    return (v >> 22u) ^ v;
}
__device__ __forceinline__ float fractg(float x) { return x - floorf(x); }

// sin / cos by Cody-Waite reduction with pi/2 = 1.5703125 + 4.838267948966e-4 and Taylor
// polynomials on [-pi/4, pi/4]; deterministic, so GI and probe directions are reproducible.
__device__ __forceinline__ void sincos_red(float x, float &s, float &c) {
    const float kf = floorf(x * 0.63661977f + 0.5f);
    const int32_t k = f2i(kf);
    const float r = (x - kf * 1.5703125f) - kf * 4.838267948966e-4f;
    const float r2 = r * r;
    const float ps = r + r * r2 * (-0.16666667f + r2 * (0.0083333333f + r2 * (-1.9841270e-4f + r2 * 2.7557319e-6f)));
    const float pc = 1.0f + r2 * (-0.5f + r2 * (0.041666667f + r2 * (-1.3888889e-3f + r2 * 2.4801587e-5f)));
    switch (k & 3) {
    case 0: s = ps; c = pc; break;
    case 1: s = pc; c = -ps; break;
    case 2: s = -ps; c = -pc; break;
    default: s = -pc; c = ps; break;
    }
}

__device__ __forceinline__ void cosine_hemisphere_u(const float n[3], float u0, float u1, float out[3]) {
    float t1[3], t2[3];
    const float upv[3] = {0.0f, 1.0f, 0.0f}, uxv[3] = {1.0f, 0.0f, 0.0f};
    cross3(n, fabsf(n[1]) < 0.9f ? upv : uxv, t1);
    normalize3(t1);
    cross3(n, t1, t2);
    const float r = sqrtf(u0);
    const float ph = 6.2831853f * u1;
    float s, c;
    sincos_red(ph, s, c);
    const float a = r * c, b = r * s, z = sqrtf(1.0f - u0);
#pragma unroll
    for (int i = 0; i < 3; ++i) out[i] = (t1[i] * a + t2[i] * b) + n[i] * z;
    normalize3(out);
}


// ------------------------------------------------------------------------------------------------
// traceVolumeS split into a resumable form for the persistent queue kernels: ray_begin is the
// set-up before the outer loop, ray_outer one iteration of it (including the voxel walk of a
// detailed brick).  The arithmetic is trace_world's, statement for statement, so a ray traced in
// pieces returns the same hit.
// ------------------------------------------------------------------------------------------------
struct RayState {
    float ro[3], sd[3], inv[3], tM8[3];
    float t, tcap;
    int32_t bc[3];
    int32_t face;
    int32_t iter;
};

enum : int32_t { RAY_RUNNING = 0, RAY_HIT = 1, RAY_MISS = 2 };

template <bool COUNT>
__device__ __forceinline__ int32_t ray_begin(const DevWorld &W, const float ro[3], const float rd[3], float tcap, RayState &r,
                                             Hit &h, Cnt &c) {
    h.hit = 0; h.t = -1.0f; h.face = 0; h.mat = 0; h.cell[0] = h.cell[1] = h.cell[2] = 0;
    if (COUNT) c.rays++;
    float tmin3[3], tmax3[3], up[3];
    int32_t st[3], faceA[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        r.ro[a] = ro[a];
        r.sd[a] = fabsf(rd[a]) < 1e-5f ? 1e-5f : rd[a];
        r.inv[a] = 1.0f / r.sd[a];
        const float lo = (float)(W.minb[a] * 8);
        const float dims = (float)(W.wb[a] * 8);
        const float t0 = (lo - ro[a]) * r.inv[a];
        const float t1 = ((lo + dims) - ro[a]) * r.inv[a];
        tmin3[a] = gmin(t0, t1);
        tmax3[a] = gmax(t0, t1);
        st[a] = r.sd[a] > 0.0f ? 1 : -1;
        up[a] = r.sd[a] < 0.0f ? 0.0f : 1.0f;
        faceA[a] = 2 * a + (st[a] > 0 ? 2 : 1);
    }
    const float tEnter = gmax(gmax(tmin3[0], tmin3[1]), tmin3[2]);
    const float tExit = gmin(gmin(tmax3[0], tmax3[1]), tmax3[2]);
    if (tExit < gmax(tEnter, 0.0f) || tEnter > tcap) return RAY_MISS;
    r.t = gmax(tEnter, 0.0f);
    r.tcap = tcap;
    const float e = r.t + 1e-3f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float entry = ro[a] + r.sd[a] * e;
        r.bc[a] = iclamp(f2i(floorf(entry * 0.125f)), W.minb[a], W.minb[a] + W.wb[a] - 1);
        r.tM8[a] = (((float)r.bc[a] + up[a]) * 8.0f - ro[a]) * r.inv[a];
    }
    r.face = 0;
    if (tEnter > 0.0f) {
        if (tEnter == tmin3[0]) r.face = faceA[0];
        else if (tEnter == tmin3[1]) r.face = faceA[1];
        else r.face = faceA[2];
    }
    r.iter = 0;
    return RAY_RUNNING;
}

template <bool COUNT, bool LDSM = false>
__device__ __forceinline__ int32_t ray_outer(const DevWorld &W, RayState &r, Hit &h, Cnt &c, const int skip) {
    if (r.iter >= W.max_outer) return RAY_MISS;
    r.iter++;
    if (r.t > r.tcap) return RAY_MISS;
    float up[3];
    int32_t st[3], faceA[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        st[a] = r.sd[a] > 0.0f ? 1 : -1;
        up[a] = r.sd[a] < 0.0f ? 0.0f : 1.0f;
        faceA[a] = 2 * a + (st[a] > 0 ? 2 : 1);
    }
    const uint2 cw = W.cells[cell_index(W, r.bc[0], r.bc[1], r.bc[2])];
    uint32_t g = cw.x;
    const uint32_t meta = cw.y;
    if (COUNT) { c.outer++; c.bytes += 4; }
    if (skip == 5) {
        if (g == 5u) g = 0u;
        else if (g >= 256u) {
            if (COUNT) c.bytes += 1;
            if (((meta >> 8) & 255u) > 0u) g = 0u;
        }
    }
    if (g >= 256u) {
        if (brick_voxels<COUNT, LDSM>(W, r.ro, r.sd, r.inv, st, up, faceA, r.bc, g, r.t, r.face, h, c, skip)) return RAY_HIT;
    } else if (g != 0u && g != (uint32_t)skip) {
        const float e = r.t + 1e-3f;
        h.hit = 1; h.t = r.t; h.face = r.face; h.mat = g;
#pragma unroll
        for (int a = 0; a < 3; ++a) h.cell[a] = f2i(floorf(r.ro[a] + r.sd[a] * e));
        return RAY_HIT;
    }
    bool stepped = false;
    if (g == 0u) {
        int32_t Cf = 0, cb[3] = {0, 0, 0};
        int32_t d = 0;
        if (W.use_dist) {
            d = (int32_t)(skip == 5 ? ((meta >> 8) & 255u) : (meta & 255u));
            if (COUNT) c.bytes += 1;
        }
        if (W.use_dist && d > 1) {
            Cf = 2 * d - 1;
#pragma unroll
            for (int a = 0; a < 3; ++a) cb[a] = r.bc[a] - (d - 1);
        } else {
            if (COUNT) c.bytes += 1;
            if (meta & (1u << 16)) Cf = 16;
            else {
                if (COUNT) c.bytes += 1;
                Cf = (meta & (1u << 17)) ? 4 : 0;
            }
            if (Cf > 0) {
#pragma unroll
                for (int a = 0; a < 3; ++a) cb[a] = r.bc[a] - (r.bc[a] & (Cf - 1));
            }
        }
        if (Cf > 0) {
            float tFar[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float cmin = (float)cb[a] * 8.0f;
                const float cmax = cmin + (float)Cf * 8.0f;
                tFar[a] = ((up[a] != 0.0f ? cmax : cmin) - r.ro[a]) * r.inv[a];
            }
            const bool ax0 = tFar[0] < tFar[1] && tFar[0] < tFar[2];
            const bool ax1 = !ax0 && tFar[1] < tFar[2];
            const bool ax2 = !ax0 && !ax1;
            r.t = ax0 ? tFar[0] : (ax1 ? tFar[1] : tFar[2]);
            r.face = ax0 ? faceA[0] : (ax1 ? faceA[1] : faceA[2]);
            const float e = r.t + 1e-3f;
            const bool axs[3] = {ax0, ax1, ax2};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                int32_t v = iclamp(f2i(floorf((r.ro[a] + r.sd[a] * e) * 0.125f)), cb[a], cb[a] + Cf - 1);
                if (axs[a]) v = st[a] > 0 ? cb[a] + Cf : cb[a] - 1;
                r.bc[a] = v;
                r.tM8[a] = (((float)r.bc[a] + up[a]) * 8.0f - r.ro[a]) * r.inv[a];
            }
            stepped = true;
        }
    }
    if (!stepped) {
        const bool a0 = r.tM8[0] < r.tM8[1] && r.tM8[0] < r.tM8[2];
        const bool a1 = !a0 && r.tM8[1] < r.tM8[2];
        const bool a2 = !a0 && !a1;
        r.t = a0 ? r.tM8[0] : (a1 ? r.tM8[1] : r.tM8[2]);
        r.face = a0 ? faceA[0] : (a1 ? faceA[1] : faceA[2]);
        const float n0 = r.tM8[0] + fabsf(r.inv[0]) * 8.0f, n1 = r.tM8[1] + fabsf(r.inv[1]) * 8.0f,
                    n2 = r.tM8[2] + fabsf(r.inv[2]) * 8.0f;
        r.tM8[0] = a0 ? n0 : r.tM8[0];
        r.tM8[1] = a1 ? n1 : r.tM8[1];
        r.tM8[2] = a2 ? n2 : r.tM8[2];
        r.bc[0] += a0 ? st[0] : 0;
        r.bc[1] += a1 ? st[1] : 0;
        r.bc[2] += a2 ? st[2] : 0;
    }
    if (r.bc[0] < W.minb[0] || r.bc[1] < W.minb[1] || r.bc[2] < W.minb[2] || r.bc[0] >= W.minb[0] + W.wb[0] ||
        r.bc[1] >= W.minb[1] + W.wb[1] || r.bc[2] >= W.minb[2] + W.wb[2])
        return RAY_MISS;
    return RAY_RUNNING;
}

// ------------------------------------------------------------------------------------------------
// camera inside solid terrain (gbuffer.comp:233-266)
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t voxel_mat_at(const DevWorld &W, const float p[3]) {
    const int32_t x = f2i(floorf(p[0])), y = f2i(floorf(p[1])), z = f2i(floorf(p[2]));
    if (x < W.minb[0] * 8 || y < W.minb[1] * 8 || z < W.minb[2] * 8 || x >= (W.minb[0] + W.wb[0]) * 8 ||
        y >= (W.minb[1] + W.wb[1]) * 8 || z >= (W.minb[2] + W.wb[2]) * 8)
        return 0u;
    const uint32_t g = W.cells[cell_index(W, x >> 3, y >> 3, z >> 3)].x;
    if (g == 0u) return 0u;
    if (g < 256u) return g;
    return W.atlas[(size_t)(g - 256u) * 512u + (uint32_t)(((z & 7) << 6) | ((y & 7) << 3) | (x & 7))];
}

// exitSolid (common.glsl:990-1016): the distance along rd to the first voxel that is not solid (air or
// water); -2 if the ray leaves the world first, -1 if it is still solid after max_steps.  Uniform
// solid bricks are crossed in one step; each step crosses exactly one face.
__device__ float exit_solid(const DevWorld &W, const float ro[3], const float rd[3], int32_t max_steps) {
    float sd[3], inv[3], up[3];
    int32_t st[3], v[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        sd[a] = fabsf(rd[a]) < 1e-5f ? 1e-5f : rd[a];
        inv[a] = 1.0f / sd[a];
        st[a] = sd[a] > 0.0f ? 1 : -1;
        up[a] = sd[a] < 0.0f ? 0.0f : 1.0f;
        v[a] = f2i(floorf(ro[a]));
    }
    float t = 0.0f;
    for (int32_t i = 0; i < max_steps; ++i) {
        bool out = false;
#pragma unroll
        for (int a = 0; a < 3; ++a) out = out || v[a] < W.minb[a] * 8 || v[a] >= (W.minb[a] + W.wb[a]) * 8;
        if (out) return -2.0f;
        const uint32_t g = W.cells[cell_index(W, v[0] >> 3, v[1] >> 3, v[2] >> 3)].x;
        if (g == 0u || g == 5u) return t;
        int32_t size = 1, lo[3] = {v[0], v[1], v[2]};
        if (g < 256u) {   // a uniform solid brick: its far side
            size = 8;
#pragma unroll
            for (int a = 0; a < 3; ++a) lo[a] = (v[a] >> 3) * 8;
        } else {
            const uint32_t m =
                W.atlas[(size_t)(g - 256u) * 512u + (uint32_t)(((v[2] & 7) << 6) | ((v[1] & 7) << 3) | (v[0] & 7))];
            if (m == 0u || m == 5u) return t;
        }
        float tf[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) tf[a] = (((float)lo[a] + up[a] * (float)size) - ro[a]) * inv[a];
        const bool ax = tf[0] < tf[1] && tf[0] < tf[2];
        const bool ay = !ax && tf[1] < tf[2];
        const bool az = !ax && !ay;
        t = ax ? tf[0] : (ay ? tf[1] : tf[2]);
        const bool sel[3] = {ax, ay, az};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int32_t nv = imin(imax(f2i(floorf(ro[a] + sd[a] * t)), lo[a]), lo[a] + (size - 1));
            v[a] = sel[a] ? (st[a] > 0 ? lo[a] + size : lo[a] - 1) : nv;
        }
    }
    return -1.0f;
}

// The primary ray of gbuffer.comp (233-266): a camera inside solid terrain (cam_mat = voxelMatAt(camPos)
// neither air nor water) starts each ray where it leaves that solid (tSkip, added to h.t), and a ray that
// only leaves it out of the world is a miss; a camera in water traces with skipMat 5 (WATER_CLEAR).
template <bool COUNT, bool LDSM = false>
__device__ __forceinline__ void primary_trace(const DevWorld &W, const float cam[3], uint32_t cam_mat,
                                              const float rd[3], Hit &h, Cnt &c) {
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
    h.t = h.t + tskip;
    if (void_ray) h.hit = 0;
}

