// The voxel walk's loop rules on the CPU: the one-step walk that tracks each voxel's entry time and face
// (the reference's brickVoxels order, common.glsl:474-539) against the packed mirrored-coordinate walk of
// csrc/vf_device.h (test after the step, tMax advanced only when the walk goes on, the hit's entry time
// and face from the last step's increment).  Random bricks and rays, half of them with exact tMax ties;
// exits non-zero on any difference in hit, voxel, entry time or face.  The float operations are IEEE
// single precision (built with -ffp-contract=off), as on the GPU.
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
#include <math.h>
#include <string.h>
static uint64_t rs = 88172645463325252ull;
static uint64_t rnd(void) { rs ^= rs << 13; rs ^= rs >> 7; rs ^= rs << 17; return rs; }
static float frand(void) { return (float)((rnd() >> 11) * (1.0 / 9007199254740992.0)); }
typedef struct { int hit, lx, ly, lz, f; float tv; } R;
static R walk_old(const uint64_t *layers, int lx, int ly, int lz, float tMx, float tMy, float tMz, float tDx, float tDy, float tDz,
                  int s0, int s1, int s2, float t, int face) {
    R r = {0}; float tv = t; int f2 = face;
    for (;;) {
        int hit = (layers[lz] >> (ly * 8 + lx)) & 1;
        int a0 = tMx < tMy && tMx < tMz, a1 = !a0 && tMy < tMz, a2 = !a0 && !a1;
        int nlx = lx + (a0 ? s0 : 0), nly = ly + (a1 ? s1 : 0), nlz = lz + (a2 ? s2 : 0);
        int out = ((unsigned)nlx | (unsigned)nly | (unsigned)nlz) > 7u;
        if (hit) { r.hit = 1; r.lx = lx; r.ly = ly; r.lz = lz; r.tv = tv; r.f = f2; return r; }
        if (out) return r;
        tv = a0 ? tMx : (a1 ? tMy : tMz); f2 = a0 ? 1 : (a1 ? 2 : 3);
        volatile float nx = tMx + tDx, ny = tMy + tDy, nz = tMz + tDz;
        tMx = a0 ? nx : tMx; tMy = a1 ? ny : tMy; tMz = a2 ? nz : tMz;
        lx = nlx; ly = nly; lz = nlz;
    }
}
static R walk_new(const uint64_t *layers, int lx, int ly, int lz, float tMx, float tMy, float tMz, float tDx, float tDy, float tDz,
                  int s0, int s1, int s2, float t, int face) {
    R r = {0};
    const uint32_t mir = (s0 > 0 ? 0u : 7u) | (s1 > 0 ? 0u : 0x70u) | (s2 > 0 ? 0u : 0x700u);
    uint32_t q = (uint32_t)lx | ((uint32_t)ly << 4) | ((uint32_t)lz << 8);
    uint32_t u = q ^ mir, d = 0;
    float tv = t; int f2 = face;
    int hit = (layers[lz] >> (((q >> 1) & 0x38u) | (q & 7u))) & 1;   /* the entry voxel */
    if (!hit) {
        uint32_t hb, ob;
        for (;;) {
            int a0 = tMx < tMy && tMx < tMz, a1 = !a0 && tMy < tMz, a2 = !a0 && !a1;
            d = a0 ? 1u : (a1 ? 16u : 256u);
            u += d;
            q = u ^ mir;
            hb = (uint32_t)(layers[(q >> 8) & 7u] >> (((q >> 1) & 0x38u) | (q & 7u))) & 1u;
            ob = u & 0x888u;
            int stop = (hb | ob) != 0u;
            /* tMax advanced only when the walk goes on (adding -0 keeps a value as it is) */
            volatile float nx = tMx + ((a0 && !stop) ? tDx : -0.0f);
            volatile float ny = tMy + ((a1 && !stop) ? tDy : -0.0f);
            volatile float nz = tMz + ((a2 && !stop) ? tDz : -0.0f);
            tMx = nx; tMy = ny; tMz = nz;
            if (stop) break;
        }
        hit = ob == 0u;
        if (hit) {   /* the step into the hit voxel: its axis from d, its tMax not yet advanced */
            tv = d == 1u ? tMx : (d == 16u ? tMy : tMz);
            f2 = d == 1u ? 1 : (d == 16u ? 2 : 3);
        }
    }
    if (!hit) return r;
    r.hit = 1; r.lx = q & 15; r.ly = (q >> 4) & 15; r.lz = q >> 8; r.tv = tv; r.f = f2;
    return r;
}
int main(int argc, char **argv) {
    long n = argc > 1 ? atol(argv[1]) : 10000000; long bad = 0, hits = 0, ties = 0;
    for (long i = 0; i < n; ++i) {
        uint64_t L[8]; int dens = rnd() % 4;
        for (int k = 0; k < 8; ++k) { L[k] = rnd(); for (int j = 0; j < dens; ++j) L[k] &= rnd(); if (rnd() % 3 == 0) L[k] = 0; }
        int lx = rnd() % 8, ly = rnd() % 8, lz = rnd() % 8;
        int s0 = rnd() & 1 ? 1 : -1, s1 = rnd() & 1 ? 1 : -1, s2 = rnd() & 1 ? 1 : -1;
        float tD[3], tM[3]; int mode = rnd() % 4;
        for (int a = 0; a < 3; ++a) {
            if (mode == 0) { tD[a] = 0.5f + 3.0f * frand(); tM[a] = 10.0f + tD[a] * frand(); }
            else if (mode == 1) { float v = (float)(1 + rnd() % 4) * 0.25f; tD[a] = v; tM[a] = 10.0f + (float)(rnd() % 4) * 0.25f; }   // exact ties
            else if (mode == 2) { tD[a] = (rnd() % 2) ? 1.0f : 1.0f + frand(); tM[a] = (rnd() % 2) ? 7.0f : 7.0f + frand(); }
            else { tD[a] = ldexpf(1.0f + frand(), (int)(rnd() % 6) - 2); tM[a] = 100.0f + tD[a] * frand(); }
        }
        R a = walk_old(L, lx, ly, lz, tM[0], tM[1], tM[2], tD[0], tD[1], tD[2], s0, s1, s2, 9.5f, 5);
        R b = walk_new(L, lx, ly, lz, tM[0], tM[1], tM[2], tD[0], tD[1], tD[2], s0, s1, s2, 9.5f, 5);
        if (mode == 1) ties++;
        hits += a.hit;
        if (memcmp(&a, &b, sizeof a) != 0 && (a.hit || b.hit)) { if (bad < 5) printf("mismatch %ld mode %d: %d %d%d%d %g f%d vs %d %d%d%d %g f%d\n", i, mode, a.hit, a.lx, a.ly, a.lz, a.tv, a.f, b.hit, b.lx, b.ly, b.lz, b.tv, b.f); bad++; }
    }
    printf("walks %ld hits %ld tie-mode %ld mismatches %ld\n", n, hits, ties, bad);
    return bad != 0;
}
