// Persistent work-queue traversal for incoherent and uneven ray batches (MI355X / gfx950).
//
// A grid of exactly as many workgroups as fit on the chip at once loops until its queue is empty.
// Each wavefront claims `chunk` ray ids at a time from a global counter and hands them to its
// lanes; a lane whose ray finishes takes the next id between two outer DDA iterations (a
// "while-while" loop with dynamic fetch), so lanes do not idle behind the wave's longest ray.  The
// arithmetic is ray_begin / ray_outer (vf_device.h), traceVolumeS statement for statement, so each
// ray's hit is exactly what a plain launch returns.
//
//   * vf_trace_gi_batch: 1 GI ray per G-buffer pixel + 1 sun ray per GI hit.  A generation kernel
//     compacts the non-sky pixels into a ray list (optionally bucket-sorted by direction octant and
//     origin cell), a persistent kernel traces it and compacts the sun rays of the hits into a
//     second list (wave ballot, one atomic per wave), a second persistent launch drains that.
//   * vf_trace_batch: camera views of any size + DDGI probe grids in ONE persistent launch.  Each
//     request owns a chunk-aligned id range (smallest views first); rays are generated on the fly;
//     the wave that retires a request's last ray stamps its completion time (100 MHz wall clock).
// With bodies in the scene (GI_OBJECTS) a ray meets them once its world walk is done, by the rules of the
// pass it stands for: a GI or probe ray the bodies nearer than min(its world hit or reach, 16), a sun ray
// that missed the world any body, a view ray the nearest body closer than its world hit.
// Queue counters are reset by a kernel (no memset), so a batch can be recorded into a graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/vf_trace.h"
#include "vf_bodies.h"
#include "vf_ctx.h"
#include "vf_device.h"
#include "vf_internal.h"

// Counters and stamps of vf_trace_batch_noalloc.  One copy per GPU, held until the process ends; host code
// reaches it only through the address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned long long scratch[64];

namespace {

constexpr int kMaxReq = VF_MAX_REQUESTS;
// counters: [0] ray-list length / batch head, [1] GI head, [2] sun tail, [3] sun head,
// [8 + r] remaining rays of request r
constexpr int kCtr = 8 + kMaxReq;
// The allocation-free batch's scratch: kCtr counters, then kMaxReq + 1 completion stamps.
constexpr int kScratchWords = 64;   // 512 B
static_assert(kCtr + kMaxReq + 1 <= kScratchWords, "batch scratch too small for the request table");

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint32_t rank_in(unsigned long long mask) {
    return (uint32_t)__popcll(mask & ((1ull << lane_id()) - 1ull));
}
__device__ __forceinline__ void miss(Hit &h) {
    h.hit = 0; h.t = -1.0f; h.face = 0; h.mat = 0; h.cell[0] = h.cell[1] = h.cell[2] = 0; h.obj = -1;
}

// The bodies after a ray's world walk (h: its world hit).  A bounce ray (gi.comp / probe_trace.comp with
// GI_OBJECTS, common.glsl:780, 2267-2271): the nearest body closer than min(the world hit or tcap, 16).
// A sun ray: if it missed the world, any body (common.glsl:2268-2270).
template <bool COUNT>
__device__ __forceinline__ void bounce_bodies(const DevWorld &W, const DevBodies &Bo, const float ro[3], const float rd[3],
                                              float tcap, Hit &h, Cnt &c) {
    h.obj = -1;
    trace_objects<COUNT>(W, Bo, ro, rd, h, gmin(h.hit ? h.t : tcap, 16.0f), false, c);
}
template <bool COUNT>
__device__ __forceinline__ void sun_bodies(const DevWorld &W, const DevBodies &Bo, const float ro[3], const float rd[3],
                                           Hit &h, Cnt &c) {
    h.obj = -1;
    if (!h.hit) trace_objects<COUNT>(W, Bo, ro, rd, h, 1e30f, true, c);
}
// The GI hit's sun ray origin (giPathImpl stage 1, common.glsl:2318-2330): a hit that is not water and
// whose normal (a body's rotated face normal for a body hit) faces the sun starts one at hit + N * 0.03.
template <bool BODIES>
__device__ __forceinline__ bool sun_origin(const DevBodies &Bo, const Hit &h, const float ro[3], const float rd[3],
                                           const float sun_dir[3], float so[3]) {
    if (!h.hit || h.mat == 5u) return false;
    float hn[3];
    hit_normal<BODIES>(&Bo, h, hn);
    if (!(gmax(dot3(hn, sun_dir), 0.0f) > 0.0f)) return false;
#pragma unroll
    for (int a = 0; a < 3; ++a) so[a] = (ro[a] + rd[a] * h.t) + hn[a] * 3e-2f;
    return true;
}

// Clears the GI-batch counters and the sort histogram.
__global__ void k_queue_reset(unsigned long long *ctr, uint32_t *buckets, int nbuckets) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < kCtr) ctr[i] = 0ull;
    if (buckets && i < nbuckets) buckets[i] = 0u;
}

// ---- GI batch ----------------------------------------------------------------------------------
struct GenArgs {
    int32_t w, h, tile_w, tile_h, tiles_x;
    uint32_t frame;
    float reach;
    int32_t sort;
    int32_t minb[3];
};

// Pixel -> GI ray of gi.comp at full resolution (fpix = pix): origin P + N*0.03, direction
// cosineHemisphereU(N, ldSample2(pix, frame >> 2)).  Sky pixels get miss records.  A workgroup covers
// kGenPer x 256 consecutive pixels of the tile order and reserves room for its rays with one atomic
// (atomics on the one list counter serialise, so fewer and larger reservations are faster); with
// `sort` it writes them grouped by direction octant (a counting sort in LDS), so the rays of a
// wavefront start close together AND share a direction octant.
constexpr int kGenPer = 4;
__device__ __forceinline__ void gi_ray(const GenArgs &A, const float4 *gbpos, const uint32_t *gbn, int32_t px, int32_t py,
                                       size_t fi, float o[3], float d[3]) {
    const float4 P4 = gbpos[fi];
    const uint32_t nb = gbn[fi];
    const float N[3] = {unorm8_dec(nb & 255u), unorm8_dec((nb >> 8) & 255u), unorm8_dec((nb >> 16) & 255u)};
    const uint32_t salt = 0x68e31da4u;
    const uint32_t hh = pcg((uint32_t)px + pcg((uint32_t)py + salt));
    const float rx = (float)(hh >> 8u) * (1.0f / 16777216.0f);
    const float ry = (float)(pcg(hh) >> 8u) * (1.0f / 16777216.0f);
    const float kf = (float)((A.frame >> 2u) & 0xFFFFu);
    const float u0 = fractg(rx + kf * 0.7548776662f), u1 = fractg(ry + kf * 0.5698402910f);
    cosine_hemisphere_u(N, u0, u1, d);
    o[0] = P4.x + N[0] * 3e-2f; o[1] = P4.y + N[1] * 3e-2f; o[2] = P4.z + N[2] * 3e-2f;
}

// The batch without a queue (refill -1): one lane per pixel of a 16x16 tile (8x8 per wave) traces its GI
// ray and then its sun ray, with the same rays and rules as the list kernels; for even work such as one
// ray per pixel the launch's own scheduling is enough.
template <bool COUNT, bool BODIES>
__global__ void __launch_bounds__(256, 7) k_gi_direct(DevWorld W, DevBodies Bo, GenArgs A, const float4 *gbpos,
                                                   const uint32_t *gbn, vf_hit *gi_hits, vf_hit *sun_hits, float sx,
                                                   float sy, float sz, unsigned long long *cnt, DirPre sp) {
    const int32_t L = xcd_remap((int32_t)blockIdx.x, (int32_t)gridDim.x);
    const int32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int32_t tiles_x = (A.w + 15) / 16;
    const int32_t px = (L % tiles_x) * 16 + (wv & 1) * 8 + (lane & 7);
    const int32_t py = (L / tiles_x) * 16 + (wv >> 1) * 8 + (lane >> 3);
    if (px >= A.w || py >= A.h) return;
    const size_t fi = (size_t)py * A.w + px;
    Cnt c = {0, 0, 0, 0};
    Hit h, sh;
    miss(sh);
    if (gbpos[fi].w < 0.0f) {
        miss(h);
    } else {
        float o[3], d[3];
        gi_ray(A, gbpos, gbn, px, py, fi, o, d);
        trace_world<0, COUNT, true>(W, o, d, A.reach, h, c);
        if (BODIES) bounce_bodies<COUNT>(W, Bo, o, d, A.reach, h, c);
        const float sun_dir[3] = {sx, sy, sz};
        float so[3];
        if (sun_origin<BODIES>(Bo, h, o, d, sun_dir, so)) {
            trace_world<0, COUNT, true, 256, true>(W, so, sun_dir, 1e30f, sh, c, 0, &sp);
            if (BODIES) sun_bodies<COUNT>(W, Bo, so, sun_dir, sh, c);
        }
    }
    store_hit(gi_hits + fi, h, BODIES && h.hit && h.obj >= 0);
    store_hit(sun_hits + fi, sh, BODIES && sh.hit && sh.obj >= 0);
    flush_count<COUNT>(c, cnt);
}
__global__ void __launch_bounds__(256) k_gi_gen(GenArgs A, const float4 *gbpos, const uint32_t *gbn, float4 *rays,
                                                unsigned long long *ctr, vf_hit *gi_hits, vf_hit *sun_hits) {
    __shared__ uint32_t oct_count[8], oct_base[8];
    __shared__ unsigned long long wg_base;
    const int t = threadIdx.x;
    if (t < 8) oct_count[t] = 0u;
    __syncthreads();
    const int32_t per = A.tile_w * A.tile_h;
    float o[kGenPer][3], d[kGenPer][3];
    uint32_t oct[kGenPer], local[kGenPer], tagv[kGenPer];
    bool valid[kGenPer];
#pragma unroll
    for (int k = 0; k < kGenPer; ++k) {
        const int64_t i = ((int64_t)blockIdx.x * kGenPer + k) * 256 + t;
        const int64_t tile = i / per;
        const int32_t in = (int32_t)(i % per);
        const int32_t px = (int32_t)(tile % A.tiles_x) * A.tile_w + in % A.tile_w;
        const int32_t py = (int32_t)(tile / A.tiles_x) * A.tile_h + in / A.tile_w;
        const bool inside = px < A.w && py < A.h;
        const size_t fi = inside ? (size_t)py * A.w + px : 0;
        valid[k] = inside && gbpos[fi].w >= 0.0f;
        tagv[k] = (uint32_t)fi;
        oct[k] = 0u;
        local[k] = 0u;
        o[k][0] = o[k][1] = o[k][2] = 0.0f;
        d[k][0] = d[k][1] = d[k][2] = 0.0f;
        if (valid[k]) {
            gi_ray(A, gbpos, gbn, px, py, fi, o[k], d[k]);
            oct[k] = A.sort ? ((d[k][0] < 0.0f ? 1u : 0u) | (d[k][1] < 0.0f ? 2u : 0u) | (d[k][2] < 0.0f ? 4u : 0u)) : 0u;
            local[k] = atomicAdd(&oct_count[oct[k]], 1u);
        } else if (inside) {
            Hit h;
            miss(h);
            store_hit(gi_hits + fi, h);
            store_hit(sun_hits + fi, h);
        }
    }
    __syncthreads();
    if (t == 0) {
        uint32_t run = 0u;
        for (int k = 0; k < 8; ++k) { oct_base[k] = run; run += oct_count[k]; }
        wg_base = run ? atomicAdd(ctr + 0, (unsigned long long)run) : 0ull;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kGenPer; ++k) {
        if (!valid[k]) continue;
        const uint64_t slot = wg_base + oct_base[oct[k]] + local[k];
        rays[slot * 2 + 0] = make_float4(o[k][0], o[k][1], o[k][2], __uint_as_float(tagv[k]));
        rays[slot * 2 + 1] = make_float4(d[k][0], d[k][1], d[k][2], A.reach);
    }
}

// Claims the next chunk of ray ids for the wave; false when the list is exhausted.
__device__ __forceinline__ bool claim(unsigned long long *head, uint64_t n, int32_t chunk, uint64_t &pool,
                                      uint64_t &pool_end) {
    unsigned long long b = 0;
    if (lane_id() == 0) b = atomicAdd(head, (unsigned long long)chunk);
    b = __shfl(b, 0);
    if (b >= n) return false;
    pool = b;
    pool_end = b + (uint64_t)chunk < n ? b + (uint64_t)chunk : n;
    return true;
}

// Traces a ray list (o.xyz, tag) (d.xyz, tcap) of length *len.  SUN: the list is the GI batch, and
// each GI hit that is not water and faces the sun appends its sun ray (giPathImpl stage 1,
// common.glsl:2318-2330) to sun_list; every GI ray writes one sun_hits record (miss when no ray).
// BODIES: the GI list's rays meet the bodies as bounce rays, the sun list's as sun rays.
template <bool SUN, bool COUNT, bool BODIES>
__global__ void __launch_bounds__(256) k_list_persist(DevWorld W, DevBodies Bo, const float4 *rays, const unsigned long long *len,
                                                      unsigned long long *head, vf_hit *hits, float4 *sun_list,
                                                      unsigned long long *sun_tail, vf_hit *sun_hits, float sx, float sy,
                                                      float sz, int32_t chunk, int32_t refill, unsigned long long *cnt) {
    const uint64_t n = *len;
    const float sun_dir[3] = {sx, sy, sz};
    Cnt c = {0, 0, 0, 0};
    RayState r;
    Hit h;
    float rd[3] = {0.0f, 0.0f, 0.0f};
    float tcap = 0.0f;
    bool active = false;
    uint32_t tag = 0;
    uint64_t pool = 0, pool_end = 0;
    bool drained = false;
    // sun rays waiting to be appended: at most one per lane, flushed with ONE atomic per wave once
    // enough lanes hold one (or a lane would need a second slot, or the wave is done)
    bool pend = false;
    float pso[3] = {0.0f, 0.0f, 0.0f};
    uint32_t ptag = 0;
    while (true) {
        bool finished = false;   // this lane finished a ray during this trip
        unsigned long long idle = __ballot(!active);
        // refill only when enough lanes are idle, so ray set-up runs for many lanes at once
        if (__popcll(idle) < (unsigned)refill && __any(active)) idle = 0ull;
        while (idle != 0ull && !drained) {
            if (pool >= pool_end && !claim(head, n, chunk, pool, pool_end)) { drained = true; break; }
            const uint64_t avail = pool_end - pool;
            const uint32_t need = (uint32_t)__popcll(idle);
            const uint32_t take = (uint32_t)(avail < need ? avail : need);
            const uint32_t rk = rank_in(idle);
            if (((idle >> lane_id()) & 1ull) != 0ull && rk < take) {
                const uint64_t id = pool + rk;
                const float4 a = rays[id * 2 + 0], b = rays[id * 2 + 1];
                const float ro[3] = {a.x, a.y, a.z};
                rd[0] = b.x; rd[1] = b.y; rd[2] = b.z;
                tcap = b.w;
                tag = __float_as_uint(a.w);
                if (ray_begin<COUNT>(W, ro, rd, b.w, r, h, c) == RAY_RUNNING) active = true;
                else finished = true;   // missed the world box: done at once
            }
            pool += take;
            idle = __ballot(!active && !finished);
        }
        if (active) {
            const int32_t st = ray_outer<COUNT, true>(W, r, h, c, 0);
            if (st != RAY_RUNNING) {
                if (st == RAY_MISS) miss(h);
                active = false;
                finished = true;
            }
        }
        bool emit = false;
        float so[3] = {0.0f, 0.0f, 0.0f};
        if (finished) {
            if (BODIES) {
                const float ro[3] = {r.ro[0], r.ro[1], r.ro[2]};
                if (SUN) bounce_bodies<COUNT>(W, Bo, ro, rd, tcap, h, c);
                else sun_bodies<COUNT>(W, Bo, ro, rd, h, c);
            }
            store_hit(hits + tag, h, BODIES && h.hit && h.obj >= 0);
            if (SUN) {
                const float ro[3] = {r.ro[0], r.ro[1], r.ro[2]};
                emit = sun_origin<BODIES>(Bo, h, ro, rd, sun_dir, so);
                if (!emit) {
                    Hit mh;
                    miss(mh);
                    store_hit(sun_hits + tag, mh);
                }
            }
        }
        const bool ending = drained && !__any(active);
        if (SUN) {
            const unsigned long long pm = __ballot(pend);
            const unsigned long long clash = __ballot(pend && emit);
            if (pm != 0ull && (clash != 0ull || __popcll(pm) >= 48 || ending)) {
                unsigned long long base = 0;
                if (lane_id() == 0) base = atomicAdd(sun_tail, (unsigned long long)__popcll(pm));
                base = __shfl(base, 0);
                if (pend) {
                    const uint64_t slot = base + rank_in(pm);
                    sun_list[slot * 2 + 0] = make_float4(pso[0], pso[1], pso[2], __uint_as_float(ptag));
                    sun_list[slot * 2 + 1] = make_float4(sx, sy, sz, 1e30f);
                }
                pend = false;
            }
            if (emit) {
                pend = true;
                pso[0] = so[0]; pso[1] = so[1]; pso[2] = so[2];
                ptag = tag;
            }
            if (ending) {
                const unsigned long long pm2 = __ballot(pend);
                if (pm2 != 0ull) {
                    unsigned long long base = 0;
                    if (lane_id() == 0) base = atomicAdd(sun_tail, (unsigned long long)__popcll(pm2));
                    base = __shfl(base, 0);
                    if (pend) {
                        const uint64_t slot = base + rank_in(pm2);
                        sun_list[slot * 2 + 0] = make_float4(pso[0], pso[1], pso[2], __uint_as_float(ptag));
                        sun_list[slot * 2 + 1] = make_float4(sx, sy, sz, 1e30f);
                    }
                    pend = false;
                }
            }
        }
        if (ending) break;
    }
    flush_count<COUNT>(c, cnt);
}

// ---- uneven batch: views + probe grids -----------------------------------------------------
struct ReqDev {
    int32_t kind;   // 0 view, 1 probe grid
    int32_t w, h, tiles_x, tile_w, tile_h;
    float v[12];    // view: pos, right, up, fwd;  probes: origin (3), spacing
    float thf, jit0, jit1, reach;
    float aspect;      // view: (float)w / (float)h, the IEEE quotient, computed on the host
    int32_t cam_air;   // view: the host's grid copy shows the camera in air (no per-ray camera-voxel lookup)
    int32_t dims[3], n_rays;
    uint32_t frame;
    uint64_t first, count;   // id range [first, first + count); first is chunk-aligned
    vf_hit *hits;
};

struct BatchArgs {
    int32_t n_req;
    uint64_t total;   // end of the last request's chunk-aligned range
    ReqDev req[kMaxReq];
};

// Batch counters from the request table itself (a kernel argument): recordable into a graph.
__global__ void k_batch_reset(BatchArgs B, unsigned long long *ctr, unsigned long long *stamps) {
    const int i = threadIdx.x;
    if (i < kCtr) ctr[i] = (i >= 8 && i - 8 < B.n_req) ? (unsigned long long)B.req[i - 8].count : 0ull;
    if (i < kMaxReq) stamps[i] = 0ull;
    if (i == kMaxReq) stamps[i] = ~0ull;
}

__device__ __forceinline__ int find_req(const BatchArgs &B, uint64_t id) {
    int r = 0;   // the request with the greatest first <= id (the id ranges need not follow request order)
    uint64_t best = 0;
    for (int k = 0; k < B.n_req; ++k)
        if (B.req[k].first <= id && B.req[k].first >= best) { best = B.req[k].first; r = k; }
    return r;
}

// Ray `local` of request q: a primary ray of a view (gbuffer.comp:194-202, tile order, with the
// camera-inside-solid pre-step of 233-266: tskip is added to the hit's t, void_ray makes it a miss), or a
// probe ray (probes.glsl:155-170).  out: the output index; skip: the ray's skipMat.
__device__ __forceinline__ void batch_ray(const DevWorld &W, const ReqDev &q, uint64_t local, float ro[3], float rd[3],
                                          float &tcap, uint32_t &out, int &skip, float &tskip, bool &void_ray) {
    tskip = 0.0f;
    void_ray = false;
    if (q.kind == 0) {
        const int32_t per = q.tile_w * q.tile_h;
        const int64_t tile = (int64_t)(local / (uint64_t)per);
        const int32_t in = (int32_t)(local % (uint64_t)per);
        const int32_t px = (int32_t)(tile % q.tiles_x) * q.tile_w + in % q.tile_w;
        const int32_t py = (int32_t)(tile / q.tiles_x) * q.tile_h + in / q.tile_w;
        const float sx = ((float)px + 0.5f) + q.jit0;
        const float sy = ((float)py + 0.5f) + q.jit1;
        const float nx = (sx / (float)q.w) * 2.0f - 1.0f;
        const float ny = (sy / (float)q.h) * 2.0f - 1.0f;
        const float a = (nx * q.aspect) * q.thf;
        const float b = ny * q.thf;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            ro[i] = q.v[i];
            rd[i] = (q.v[9 + i] + q.v[3 + i] * a) + q.v[6 + i] * b;
        }
        normalize3(rd);
        tcap = 1e30f;
        const uint32_t cam_mat = q.cam_air ? 0u : voxel_mat_at(W, ro);
        skip = cam_mat == 5u ? 5 : 0;
        if (cam_mat != 0u && cam_mat != 5u) {
            const float cam[3] = {ro[0], ro[1], ro[2]};
            tskip = exit_solid(W, cam, rd, 2048);
            void_ray = tskip < 0.0f;
            tskip = gmax(tskip, 0.0f);
#pragma unroll
            for (int i = 0; i < 3; ++i) ro[i] = cam[i] + rd[i] * tskip;
        }
        out = (px < q.w && py < q.h) ? (uint32_t)(py * q.w + px) : 0xFFFFFFFFu;
        return;
    }
    const int32_t i = (int32_t)(local % (uint64_t)q.n_rays);
    const int64_t p = (int64_t)(local / (uint64_t)q.n_rays);
    const int32_t pxi = (int32_t)(p % q.dims[0]), pyi = (int32_t)((p / q.dims[0]) % q.dims[1]),
                  pzi = (int32_t)(p / ((int64_t)q.dims[0] * q.dims[1]));
    ro[0] = q.v[0] + q.v[3] * (float)pxi;
    ro[1] = q.v[1] + q.v[3] * (float)pyi;
    ro[2] = q.v[2] + q.v[3] * (float)pzi;
    const float z = 1.0f - (2.0f * (float)i + 1.0f) / (float)q.n_rays;
    const float rr = sqrtf(gmax(0.0f, 1.0f - z * z));
    float s, c;
    sincos_red((float)i * 2.39996323f, s, c);
    const float d[3] = {c * rr, z, s * rr};
    const uint32_t h0 = pcg(q.frame * 747796405u + 1u);
    const float u1 = (float)(h0 & 0xFFFFu) / 65535.0f, u2 = (float)(h0 >> 16) / 65535.0f;
    const uint32_t h2 = pcg(h0);
    const float u3 = (float)(h2 & 0xFFFFu) / 65535.0f;
    float s2, c2, s3, c3;
    sincos_red(6.2831853f * u2, s2, c2);
    sincos_red(6.2831853f * u3, s3, c3);
    const float qa = sqrtf(1.0f - u1), qb = sqrtf(u1);
    const float qv[3] = {qa * s2, qa * c2, qb * s3};
    const float qw = qb * c3;
    float c1v[3], tmp[3], c2v[3];
    cross3(qv, d, c1v);
#pragma unroll
    for (int k = 0; k < 3; ++k) tmp[k] = c1v[k] + qw * d[k];
    cross3(qv, tmp, c2v);
#pragma unroll
    for (int k = 0; k < 3; ++k) rd[k] = d[k] + 2.0f * c2v[k];
    tcap = q.reach;
    skip = 5;   // probe_trace.comp: traceWorldProbe(..., kSkipMat)
    out = (uint32_t)local;
}

// Subtracts k retired rays from request rq; the wave whose subtraction reaches zero stamps the time.
__device__ __forceinline__ void retire(unsigned long long *ctr, unsigned long long *stamps, int rq, unsigned long long k) {
    if (k == 0ull) return;
    if (lane_id() == 0) {
        const unsigned long long old = atomicSub(ctr + 8 + rq, k);
        if (old == k) stamps[rq] = wall_clock64();
    }
}

// The bodies after a batch ray's world walk: a view ray (traceScenePrimary = traceSceneS, gbuffer.comp:233-266)
// the nearest body closer than its world hit, from the origin its world walk started at; a probe ray
// (probe_trace.comp:47) the bodies nearer than min(its world hit or reach, 16).
template <bool COUNT>
__device__ __forceinline__ void batch_bodies(const DevWorld &W, const DevBodies &Bo, const ReqDev &q, const float ro[3],
                                             const float rd[3], Hit &h, Cnt &c) {
    if (q.kind == 0) add_bodies<COUNT>(W, Bo, ro, rd, 1e30f, h, c);
    else bounce_bodies<COUNT>(W, Bo, ro, rd, q.reach, h, c);
}

template <bool COUNT, bool BODIES>
__global__ void __launch_bounds__(256) k_batch_persist(DevWorld W, DevBodies Bo, BatchArgs B, unsigned long long *ctr,
                                                       unsigned long long *stamps, int32_t chunk, int32_t refill,
                                                       unsigned long long *cnt) {
    if (threadIdx.x == 0) atomicMin(stamps + kMaxReq, (unsigned long long)wall_clock64());
    // every lane of the first trip waits for the start stamp of its own workgroup only; requests are
    // stamped relative to the earliest workgroup start
    Cnt c = {0, 0, 0, 0};
    RayState r;
    Hit h;
    bool active = false, vray = false;
    int32_t req = 0, skip = 0;
    uint32_t out = 0;
    float tskip = 0.0f;
    float rd[3] = {0.0f, 0.0f, 0.0f};   // the ray's direction as generated (the bodies need it unclamped)
    uint64_t pool = 0, pool_end = 0;
    int32_t pool_req = 0;
    bool drained = false;
    unsigned long long *head = ctr + 0;
    int acc_req = 0, trips = 0;
    unsigned long long acc_cnt = 0;
    while (true) {
        bool finished = false;
        bool claimed = false;
        unsigned long long idle = __ballot(!active);
        // refill only when enough lanes are idle, so ray set-up runs for many lanes at once
        if (__popcll(idle) < (unsigned)refill && __any(active)) idle = 0ull;
        while (idle != 0ull && !drained) {
            if (pool >= pool_end) {
                unsigned long long b = 0;
                if (lane_id() == 0) b = atomicAdd(head, (unsigned long long)chunk);
                b = __shfl(b, 0);
                if (b >= B.total) { drained = true; break; }
                claimed = true;
                pool_req = find_req(B, b);
                const uint64_t rend = B.req[pool_req].first + B.req[pool_req].count;
                pool = b;
                pool_end = b + (uint64_t)chunk < rend ? b + (uint64_t)chunk : rend;   // aligned: no straddling
                if (pool >= pool_end) continue;   // a chunk of the alignment padding: claim again
            }
            const uint64_t avail = pool_end - pool;
            const uint32_t need = (uint32_t)__popcll(idle);
            const uint32_t take = (uint32_t)(avail < need ? avail : need);
            const uint32_t rk = rank_in(idle);
            if (((idle >> lane_id()) & 1ull) != 0ull && rk < take) {
                const uint64_t id = pool + rk;
                req = pool_req;
                float ro[3], tcap;
                bool void_ray;
                batch_ray(W, B.req[req], id - B.req[req].first, ro, rd, tcap, out, skip, tskip, void_ray);
                vray = void_ray;
                if (out == 0xFFFFFFFFu) {
                    finished = true;   // tile padding outside the view: no ray
                } else if (void_ray) {
                    miss(h);
                    finished = true;
                } else if (ray_begin<COUNT>(W, ro, rd, tcap, r, h, c) == RAY_RUNNING) {
                    active = true;
                } else {
                    finished = true;
                }
            }
            pool += take;
            idle = __ballot(!active && !finished);
        }
        if (active) {
            const int32_t st = ray_outer<COUNT, true>(W, r, h, c, skip);
            if (st != RAY_RUNNING) {
                if (st == RAY_MISS) miss(h);
                active = false;
                finished = true;
            }
        }
        if (finished && out != 0xFFFFFFFFu) {
            if (BODIES && !vray) {   // a void ray is a miss whatever it meets
                const float ro[3] = {r.ro[0], r.ro[1], r.ro[2]};
                batch_bodies<COUNT>(W, Bo, B.req[req], ro, rd, h, c);
            }
            h.t = h.t + tskip;
            store_hit(B.req[req].hits + out, h, BODIES && h.hit && h.obj >= 0);
        }
        // retired rays are counted per wave (wave-uniform acc_req / acc_cnt) and subtracted from the
        // request's remaining count in one atomic when the request changes, the wave claimed a new
        // chunk, every 64 trips, or at the end; the atomic that reaches zero stamps the completion
        unsigned long long fm = __ballot(finished);
        while (fm != 0ull) {
            const int leader = __ffsll((long long)fm) - 1;
            const int rq = __shfl(req, leader);
            const unsigned long long same = __ballot(finished && req == rq);
            if (rq != acc_req) {
                retire(ctr, stamps, acc_req, acc_cnt);
                acc_req = rq;
                acc_cnt = 0;
            }
            acc_cnt += (unsigned long long)__popcll(same);
            fm &= ~same;
        }
        const bool ending = drained && !__any(active);
        if (++trips >= 64 || claimed || ending) {
            retire(ctr, stamps, acc_req, acc_cnt);
            acc_cnt = 0;
            trips = 0;
        }
        if (ending) break;
    }
    flush_count<COUNT>(c, cnt);
}

// Chunk mode (refill == 0): a wavefront claims `chunk` consecutive rays and traces them 64 at a time,
// each lane to completion with trace_world, then claims again.  No per-iteration bookkeeping, the
// register footprint of the plain kernels; the sun rays of 64 GI rays go out with one atomic.
template <bool SUN, bool COUNT, bool BODIES>
__global__ void __launch_bounds__(256) k_list_chunked(DevWorld W, DevBodies Bo, const float4 *rays, const unsigned long long *len,
                                                      unsigned long long *head, vf_hit *hits, float4 *sun_list,
                                                      unsigned long long *sun_tail, vf_hit *sun_hits, float sx, float sy,
                                                      float sz, int32_t chunk, unsigned long long *cnt) {
    const uint64_t n = *len;
    const float sun_dir[3] = {sx, sy, sz};
    Cnt c = {0, 0, 0, 0};
    uint64_t pool = 0, pool_end = 0;
    while (claim(head, n, chunk, pool, pool_end)) {
        for (uint64_t base = pool; base < pool_end; base += 64) {
            const uint64_t id = base + lane_id();
            bool emit = false;
            float so[3] = {0.0f, 0.0f, 0.0f};
            uint32_t tag = 0;
            if (id < pool_end) {
                const float4 a = rays[id * 2 + 0], b = rays[id * 2 + 1];
                const float ro[3] = {a.x, a.y, a.z}, rd[3] = {b.x, b.y, b.z};
                tag = __float_as_uint(a.w);
                Hit h;
                trace_world<0, COUNT, true>(W, ro, rd, b.w, h, c);
                if (BODIES) {
                    if (SUN) bounce_bodies<COUNT>(W, Bo, ro, rd, b.w, h, c);
                    else sun_bodies<COUNT>(W, Bo, ro, rd, h, c);
                }
                store_hit(hits + tag, h, BODIES && h.hit && h.obj >= 0);
                if (SUN) {
                    emit = sun_origin<BODIES>(Bo, h, ro, rd, sun_dir, so);
                    if (!emit) {
                        Hit mh;
                        miss(mh);
                        store_hit(sun_hits + tag, mh);
                    }
                }
            }
            if (SUN) {
                const unsigned long long em = __ballot(emit);
                if (em != 0ull) {
                    unsigned long long sb = 0;
                    if (lane_id() == 0) sb = atomicAdd(sun_tail, (unsigned long long)__popcll(em));
                    sb = __shfl(sb, 0);
                    if (emit) {
                        const uint64_t slot = sb + rank_in(em);
                        sun_list[slot * 2 + 0] = make_float4(so[0], so[1], so[2], __uint_as_float(tag));
                        sun_list[slot * 2 + 1] = make_float4(sx, sy, sz, 1e30f);
                    }
                }
            }
        }
    }
    flush_count<COUNT>(c, cnt);
}

// Chunk mode of the uneven batch: a claimed chunk belongs to one request (wave-uniform), so the
// request's fields are scalar loads and views (skipMat 0) and probes (skipMat 5) never diverge.
template <bool COUNT, bool BODIES>
__global__ void __launch_bounds__(256) k_batch_chunked(DevWorld W, DevBodies Bo, BatchArgs B, unsigned long long *ctr,
                                                       unsigned long long *stamps, int32_t chunk,
                                                       unsigned long long *cnt) {
    if (threadIdx.x == 0) atomicMin(stamps + kMaxReq, (unsigned long long)wall_clock64());
    Cnt c = {0, 0, 0, 0};
    while (true) {
        unsigned long long b = 0;
        if (lane_id() == 0) b = atomicAdd(ctr + 0, (unsigned long long)chunk);
        b = __shfl(b, 0);
        if (b >= B.total) break;
        const int rq = find_req(B, b);
        const ReqDev &q = B.req[rq];
        const uint64_t rend = q.first + q.count;
        const uint64_t end = b + (uint64_t)chunk < rend ? b + (uint64_t)chunk : rend;
        for (uint64_t base = b; base < end; base += 64) {
            const uint64_t id = base + lane_id();
            if (id < end) {
                float ro[3], rd[3], tcap, tskip;
                uint32_t out;
                int skip;
                bool void_ray;
                batch_ray(W, q, id - q.first, ro, rd, tcap, out, skip, tskip, void_ray);
                if (out != 0xFFFFFFFFu) {
                    Hit h;
                    if (skip == 0) trace_world<0, COUNT, true>(W, ro, rd, tcap, h, c);
                    else trace_world<5, COUNT, true>(W, ro, rd, tcap, h, c);
                    if (BODIES) batch_bodies<COUNT>(W, Bo, q, ro, rd, h, c);
                    h.t = h.t + tskip;
                    if (void_ray) h.hit = 0;
                    store_hit(q.hits + out, h, BODIES && h.hit && h.obj >= 0);
                }
            }
        }
        if (end > b) retire(ctr, stamps, rq, (unsigned long long)(end - b));
    }
    flush_count<COUNT>(c, cnt);
}

// Flat mode of the uneven batch (refill -1): no queue.  One launch with a workgroup per 256 consecutive ray
// ids of the batch (requests in their service order, each 256-aligned, so a workgroup serves one request),
// dealt to the XCDs in runs like the plain passes; each lane traces one id to completion, and the workgroup
// subtracts its ids from the request's count in one atomic, so the request's last workgroup stamps its
// completion as in the queue modes.  (A per-wave atomic serialised on the few counters: 3x slower.)  The
// start stamp is the first workgroup's.  BOX: the small-frame traversal (trace_world), held at 5 waves per
// SIMD (its box test would spill registers at 7).
template <bool COUNT, bool BODIES, bool BOX = false>
__global__ void __launch_bounds__(256, BODIES ? 1 : (BOX ? 5 : 7)) k_batch_flat(DevWorld W, DevBodies Bo, BatchArgs B, unsigned long long *ctr,
                                                    unsigned long long *stamps, unsigned long long *cnt) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicMin(stamps + kMaxReq, (unsigned long long)wall_clock64());
    const int32_t L = xcd_remap((int32_t)blockIdx.x, (int32_t)gridDim.x);
    const uint64_t b = (uint64_t)L * 256u;   // the workgroup's first id
    if (b >= B.total) return;
    const int rq = find_req(B, b);
    const ReqDev &q = B.req[rq];
    const uint64_t rend = q.first + q.count;
    if (b >= rend) return;   // the alignment padding after a request
    Cnt c = {0, 0, 0, 0};
    const uint64_t id = b + threadIdx.x;
    if (id < rend) {
        float ro[3], rd[3], tcap, tskip;
        uint32_t out;
        int skip;
        bool void_ray;
        batch_ray(W, q, id - q.first, ro, rd, tcap, out, skip, tskip, void_ray);
        if (out != 0xFFFFFFFFu) {
            Hit h;
            trace_world<-1, COUNT, true, 256, false, BOX && !COUNT>(W, ro, rd, tcap, h, c, skip);
            if (BODIES) batch_bodies<COUNT>(W, Bo, q, ro, rd, h, c);
            h.t = h.t + tskip;
            if (void_ray) h.hit = 0;
            store_hit(q.hits + out, h, BODIES && h.hit && h.obj >= 0);
        }
    }
    flush_count<COUNT>(c, cnt);
    __syncthreads();
    if (threadIdx.x == 0) retire(ctr, stamps, rq, (unsigned long long)((rend - b) < 256u ? (rend - b) : 256u));
}

}  // namespace

// ------------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------------
namespace vf {

// The device address of `scratch` on the context's device, taken once per device (with that device
// current) and kept for every later context on it.
int queue_scratch_init(vf_ctx *ctx) {
    static_assert(sizeof(scratch) == kScratchWords * sizeof(unsigned long long), "scratch size");
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static unsigned long long *per_device[kMaxDevices] = {};
    if (ctx->device < 0 || ctx->device >= kMaxDevices) return VF_OK;   // the allocation-free batch then refuses
    std::lock_guard<std::mutex> lock(mu);
    if (!per_device[ctx->device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return vf::fail(ctx, VF_E_DEVICE, "batch scratch: hipGetSymbolAddress failed: %s", hipGetErrorString(e));
        per_device[ctx->device] = static_cast<unsigned long long *>(p);
    }
    ctx->batch_scratch = per_device[ctx->device];
    return VF_OK;
}

void free_queue(vf_ctx *ctx) {
    if (ctx->batch_stamps == ctx->q_stamps) ctx->batch_stamps = nullptr;
    void *ptrs[] = {ctx->q_ctr, ctx->q_stamps, ctx->q_rays, ctx->q_sun};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    ctx->q_ctr = ctx->q_stamps = nullptr;
    ctx->q_rays = ctx->q_sun = nullptr;
    ctx->q_cap = 0;
}

}  // namespace vf

namespace {

int reserve_queue(vf_ctx *ctx, int64_t rays) {
    if (ctx->q_ctr && ctx->q_cap >= rays) return VF_OK;
    if (ctx->capturing)
        return vf::fail(ctx, VF_E_INVALID, "queue buffers for %lld rays are not allocated: run the batch once before recording it",
                        (long long)rays);
    VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    vf::free_queue(ctx);
    const int64_t cap = std::max<int64_t>(rays, 1);
#define VF_QALLOC(ptr, bytes)                                                                        \
    do {                                                                                             \
        hipError_t e_ = hipMalloc((void **)&(ptr), (bytes));                                         \
        if (e_ != hipSuccess) {                                                                      \
            vf::free_queue(ctx);                                                                     \
            return vf::fail(ctx, VF_E_OOM, "queue buffers: hipMalloc(%zu): %s", (size_t)(bytes), hipGetErrorString(e_)); \
        }                                                                                            \
    } while (0)
    VF_QALLOC(ctx->q_ctr, kCtr * 8);
    VF_QALLOC(ctx->q_stamps, (kMaxReq + 1) * 8);
    VF_QALLOC(ctx->q_rays, (size_t)cap * 32);
    VF_QALLOC(ctx->q_sun, (size_t)cap * 32);
#undef VF_QALLOC
    ctx->q_cap = cap;
    return VF_OK;
}

int persistent_groups(vf_ctx *ctx, const void *kernel, const vf_queue_cfg *cfg) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess || per_cu <= 0) per_cu = 4;
    if (cfg && cfg->groups_per_cu > 0) per_cu = std::min(per_cu, (int)cfg->groups_per_cu);
    return std::max(1, per_cu * ctx->n_cu);
}

int check_cfg(vf_ctx *ctx, const vf_queue_cfg *cfg, const char *what) {
    if (!cfg) return VF_OK;
    const bool tiles = (cfg->tile_w == 8 && cfg->tile_h == 8) || (cfg->tile_w == 16 && cfg->tile_h == 8) ||
                       (cfg->tile_w == 16 && cfg->tile_h == 16) || (cfg->tile_w == 8 && cfg->tile_h == 16) ||
                       (cfg->tile_w == 32 && cfg->tile_h == 8);
    if (!tiles) return vf::fail(ctx, VF_E_INVALID, "%s: tile %dx%d not supported", what, cfg->tile_w, cfg->tile_h);
    if (cfg->chunk < 64 || cfg->chunk > 4096 || (cfg->chunk & 63))
        return vf::fail(ctx, VF_E_INVALID, "%s: chunk %d must be a multiple of 64 in [64, 4096]", what, cfg->chunk);
    if (cfg->refill < -1 || cfg->refill > 64)
        return vf::fail(ctx, VF_E_INVALID, "%s: refill %d must be in [0, 64], or -1 (no queue)", what, cfg->refill);
    return VF_OK;
}

vf_queue_cfg default_cfg() {
    vf_queue_cfg c;
    c.tile_w = 8; c.tile_h = 8; c.chunk = 256; c.sort = 1; c.groups_per_cu = 0; c.refill = 32;
    c.traversal = 0;
    return c;
}

}  // namespace

extern "C" {

int vf_queue_defaults(vf_queue_cfg *cfg) {
    if (!cfg) return vf::fail(nullptr, VF_E_INVALID, "vf_queue_defaults: null cfg");
    *cfg = default_cfg();
    return VF_OK;
}

int vf_trace_gi_batch(vf_ctx *ctx, const vf_pass_ubo *pass, int32_t w, int32_t h, float reach, const float *gb_pos,
                      const uint32_t *gb_normal, vf_hit *gi_hits, vf_hit *sun_hits, const vf_queue_cfg *cfg_in) {
    int rc = vf::check_ready(ctx, "vf_trace_gi_batch");
    if (rc) return rc;
    if ((rc = vf::check_bodies(ctx, "vf_trace_gi_batch"))) return rc;
    if (!ctx->has_scene) return vf::fail(ctx, VF_E_INVALID, "vf_trace_gi_batch: no scene (vf_set_scene)");
    if (!pass || !gb_pos || !gb_normal || !gi_hits || !sun_hits) return vf::fail(ctx, VF_E_INVALID, "vf_trace_gi_batch: null buffer");
    if (w <= 0 || h <= 0 || (int64_t)w * h > (1ll << 28)) return vf::fail(ctx, VF_E_INVALID, "vf_trace_gi_batch: bad size %dx%d", w, h);
    if (!(reach > 0.0f)) return vf::fail(ctx, VF_E_INVALID, "vf_trace_gi_batch: reach must be > 0");
    if ((rc = check_cfg(ctx, cfg_in, "vf_trace_gi_batch"))) return rc;
    const vf_queue_cfg cfg = cfg_in ? *cfg_in : default_cfg();
    const int64_t n = (int64_t)w * h;
    if ((rc = reserve_queue(ctx, n))) return rc;
    GenArgs A;
    A.w = w; A.h = h; A.tile_w = cfg.tile_w; A.tile_h = cfg.tile_h;
    A.tiles_x = (w + cfg.tile_w - 1) / cfg.tile_w;
    const int64_t tiles_y = (h + cfg.tile_h - 1) / cfg.tile_h;
    A.frame = pass->frame_idx;
    A.reach = reach;
    A.sort = cfg.sort ? 1 : 0;
    for (int a = 0; a < 3; ++a) A.minb[a] = ctx->W.minb[a];
    const int64_t ngen = (int64_t)A.tiles_x * tiles_y * cfg.tile_w * cfg.tile_h;
    const float *s = ctx->S.sun_dir;
    const DevBodies Bo = vf::dev_bodies(ctx);
    const bool bodies = Bo.n > 0, counting = ctx->counting != 0;
    unsigned long long *ctr = ctx->q_ctr;
    if ((rc = vf::begin_timed(ctx))) return rc;
    if (cfg.refill < 0) {   // no queue
        const int32_t nt = ((w + 15) / 16) * ((h + 15) / 16);
        const DirPre sp = dir_pre(s);   // the sun ray's set-up, as trace_world computes it
        using DirectFn = void (*)(DevWorld, DevBodies, GenArgs, const float4 *, const uint32_t *, vf_hit *, vf_hit *, float,
                                  float, float, unsigned long long *, DirPre);
        const DirectFn fn = bodies ? (counting ? k_gi_direct<true, true> : k_gi_direct<false, true>)
                                   : (counting ? k_gi_direct<true, false> : k_gi_direct<false, false>);
        hipLaunchKernelGGL(fn, dim3(nt), dim3(256), 0, ctx->stream, ctx->W, Bo, A, reinterpret_cast<const float4 *>(gb_pos),
                           gb_normal, gi_hits, sun_hits, s[0], s[1], s[2], ctx->d_cnt, sp);
        if ((rc = vf::launch_check(ctx, "vf_trace_gi_batch"))) return rc;
        return vf::end_timed(ctx);
    }
    hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(256), 0, ctx->stream, ctr, nullptr, 0);
    hipLaunchKernelGGL(k_gi_gen, dim3(vf::blocks_for(ngen, 256 * kGenPer)), dim3(256), 0, ctx->stream, A,
                       reinterpret_cast<const float4 *>(gb_pos), gb_normal, ctx->q_rays, ctr, gi_hits, sun_hits);
    unsigned long long *C0 = ctx->d_cnt;
    if (cfg.refill == 0) {
        using ListFn = void (*)(DevWorld, DevBodies, const float4 *, const unsigned long long *, unsigned long long *, vf_hit *,
                                float4 *, unsigned long long *, vf_hit *, float, float, float, int32_t, unsigned long long *);
        ListFn gi_fn, sun_fn;
        if (bodies) {
            gi_fn = counting ? k_list_chunked<true, true, true> : k_list_chunked<true, false, true>;
            sun_fn = counting ? k_list_chunked<false, true, true> : k_list_chunked<false, false, true>;
        } else {
            gi_fn = counting ? k_list_chunked<true, true, false> : k_list_chunked<true, false, false>;
            sun_fn = counting ? k_list_chunked<false, true, false> : k_list_chunked<false, false, false>;
        }
        const int g = persistent_groups(ctx, reinterpret_cast<const void *>(gi_fn), &cfg);
        hipLaunchKernelGGL(gi_fn, dim3(g), dim3(256), 0, ctx->stream, ctx->W, Bo, ctx->q_rays, ctr, ctr + 1, gi_hits,
                           ctx->q_sun, ctr + 2, sun_hits, s[0], s[1], s[2], cfg.chunk, C0);
        hipLaunchKernelGGL(sun_fn, dim3(g), dim3(256), 0, ctx->stream, ctx->W, Bo, ctx->q_sun, ctr + 2, ctr + 3, sun_hits,
                           nullptr, nullptr, nullptr, s[0], s[1], s[2], cfg.chunk, C0);
    } else {
        using ListFn = void (*)(DevWorld, DevBodies, const float4 *, const unsigned long long *, unsigned long long *, vf_hit *,
                                float4 *, unsigned long long *, vf_hit *, float, float, float, int32_t, int32_t,
                                unsigned long long *);
        ListFn gi_fn, sun_fn;
        if (bodies) {
            gi_fn = counting ? k_list_persist<true, true, true> : k_list_persist<true, false, true>;
            sun_fn = counting ? k_list_persist<false, true, true> : k_list_persist<false, false, true>;
        } else {
            gi_fn = counting ? k_list_persist<true, true, false> : k_list_persist<true, false, false>;
            sun_fn = counting ? k_list_persist<false, true, false> : k_list_persist<false, false, false>;
        }
        const int g = persistent_groups(ctx, reinterpret_cast<const void *>(gi_fn), &cfg);
        hipLaunchKernelGGL(gi_fn, dim3(g), dim3(256), 0, ctx->stream, ctx->W, Bo, ctx->q_rays, ctr, ctr + 1, gi_hits,
                           ctx->q_sun, ctr + 2, sun_hits, s[0], s[1], s[2], cfg.chunk, cfg.refill, C0);
        hipLaunchKernelGGL(sun_fn, dim3(g), dim3(256), 0, ctx->stream, ctx->W, Bo, ctx->q_sun, ctr + 2, ctr + 3, sun_hits,
                           nullptr, nullptr, nullptr, s[0], s[1], s[2], cfg.chunk, cfg.refill, C0);
    }
    if ((rc = vf::launch_check(ctx, "vf_trace_gi_batch"))) return rc;
    return vf::end_timed(ctx);
}

}  // extern "C"

namespace {

// vf_trace_batch's work with its counters and stamps either in the context's queue buffers (allocated on
// first use) or, with noalloc, in this device's `scratch` (nothing allocated).
int trace_batch(vf_ctx *ctx, const vf_request *reqs, int32_t n_req, const vf_queue_cfg *cfg_in, bool noalloc) {
    const char *what = noalloc ? "vf_trace_batch_noalloc" : "vf_trace_batch";
    int rc = vf::check_ready(ctx, what);
    if (rc) return rc;
    if ((rc = vf::check_bodies(ctx, what))) return rc;
    if (!reqs || n_req <= 0 || n_req > kMaxReq)
        return vf::fail(ctx, VF_E_INVALID, "%s: 1..%d requests required (the batch's counters and stamps are sized for %d)",
                        what, kMaxReq, kMaxReq);
    if (noalloc && !ctx->batch_scratch)
        return vf::fail(ctx, VF_E_UNSUPPORTED, "%s: no batch scratch on device %d", what, ctx->device);
    if ((rc = check_cfg(ctx, cfg_in, what))) return rc;
    const vf_queue_cfg cfg = cfg_in ? *cfg_in : default_cfg();
    BatchArgs B;
    std::memset(&B, 0, sizeof(B));
    B.n_req = n_req;
    uint64_t at = 0;
    uint64_t counts[kMaxReq] = {};
    for (int k = 0; k < n_req; ++k) {
        const vf_request &q = reqs[k];
        ReqDev &d = B.req[k];
        if (!q.d_hits) return vf::fail(ctx, VF_E_INVALID, "%s: request %d has no output", what, k);
        d.kind = q.kind;
        d.hits = q.d_hits;
        if (q.kind == VF_REQ_VIEW) {
            if (q.width <= 0 || q.height <= 0 || (int64_t)q.width * q.height > (1ll << 28))
                return vf::fail(ctx, VF_E_INVALID, "%s: request %d: bad view size", what, k);
            d.w = q.width; d.h = q.height;
            d.tile_w = cfg.tile_w; d.tile_h = cfg.tile_h;
            d.tiles_x = (q.width + cfg.tile_w - 1) / cfg.tile_w;
            const int64_t ty = (q.height + cfg.tile_h - 1) / cfg.tile_h;
            for (int a = 0; a < 3; ++a) {
                d.v[a] = q.cam_pos[a]; d.v[3 + a] = q.cam_right[a]; d.v[6 + a] = q.cam_up[a]; d.v[9 + a] = q.cam_fwd[a];
            }
            d.thf = q.tan_half_fov;
            d.aspect = (float)q.width / (float)q.height;
            d.cam_air = vf::camera_in_air(ctx, q.cam_pos) ? 1 : 0;
            d.jit0 = q.jitter[0];
            d.jit1 = q.jitter[1];
            d.count = (uint64_t)d.tiles_x * ty * cfg.tile_w * cfg.tile_h;   // tile padding: no output
            counts[k] = d.count;
        } else if (q.kind == VF_REQ_PROBES) {
            if (q.n_rays <= 0 || q.n_rays > 65536 || q.dims[0] <= 0 || q.dims[1] <= 0 || q.dims[2] <= 0)
                return vf::fail(ctx, VF_E_INVALID, "%s: request %d: bad probe grid", what, k);
            for (int a = 0; a < 3; ++a) { d.v[a] = q.origin[a]; d.dims[a] = q.dims[a]; }
            d.v[3] = q.spacing;
            d.n_rays = q.n_rays;
            d.frame = q.frame;
            d.reach = q.reach;
            d.count = (uint64_t)q.dims[0] * q.dims[1] * q.dims[2] * (uint64_t)q.n_rays;
            counts[k] = d.count;
        } else {
            return vf::fail(ctx, VF_E_INVALID, "%s: request %d: unknown kind %d", what, k, q.kind);
        }
        d.first = at;
        const uint64_t align = cfg.refill < 0 ? 256u : (uint64_t)cfg.chunk;   // flat: one workgroup, one request
        at += (d.count + align - 1) / align * align;
    }
    B.total = at;
    (void)counts;
    unsigned long long *ctr, *stamps;
    if (noalloc) {
        ctr = ctx->batch_scratch;
        stamps = ctx->batch_scratch + kCtr;
    } else {
        if ((rc = reserve_queue(ctx, 1))) return rc;
        ctr = ctx->q_ctr;
        stamps = ctx->q_stamps;
    }
    const DevBodies Bo = vf::dev_bodies(ctx);
    const bool bodies = Bo.n > 0, counting = ctx->counting != 0;
    if ((rc = vf::begin_timed(ctx))) return rc;
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(256), 0, ctx->stream, B, ctr, stamps);
    if (cfg.refill < 0) {   // flat: one workgroup per 256 ids, no queue
        using FlatFn = void (*)(DevWorld, DevBodies, BatchArgs, unsigned long long *, unsigned long long *,
                                unsigned long long *);
        const bool box = cfg.traversal == 1 && !counting;
        const FlatFn fn = bodies ? (counting ? k_batch_flat<true, true> : box ? k_batch_flat<false, true, true> : k_batch_flat<false, true>)
                                 : (counting ? k_batch_flat<true, false> : box ? k_batch_flat<false, false, true> : k_batch_flat<false, false>);
        hipLaunchKernelGGL(fn, dim3((unsigned)(B.total / 256u)), dim3(256), 0, ctx->stream, ctx->W, Bo, B, ctr, stamps,
                           ctx->d_cnt);
    } else if (cfg.refill == 0) {
        using ChunkFn = void (*)(DevWorld, DevBodies, BatchArgs, unsigned long long *, unsigned long long *, int32_t,
                                 unsigned long long *);
        const ChunkFn fn = bodies ? (counting ? k_batch_chunked<true, true> : k_batch_chunked<false, true>)
                                  : (counting ? k_batch_chunked<true, false> : k_batch_chunked<false, false>);
        const int g = persistent_groups(ctx, reinterpret_cast<const void *>(fn), &cfg);
        hipLaunchKernelGGL(fn, dim3(g), dim3(256), 0, ctx->stream, ctx->W, Bo, B, ctr, stamps, cfg.chunk, ctx->d_cnt);
    } else {
        using PersistFn = void (*)(DevWorld, DevBodies, BatchArgs, unsigned long long *, unsigned long long *, int32_t, int32_t,
                                   unsigned long long *);
        const PersistFn fn = bodies ? (counting ? k_batch_persist<true, true> : k_batch_persist<false, true>)
                                    : (counting ? k_batch_persist<true, false> : k_batch_persist<false, false>);
        const int g = persistent_groups(ctx, reinterpret_cast<const void *>(fn), &cfg);
        hipLaunchKernelGGL(fn, dim3(g), dim3(256), 0, ctx->stream, ctx->W, Bo, B, ctr, stamps, cfg.chunk, cfg.refill,
                           ctx->d_cnt);
    }
    if ((rc = vf::launch_check(ctx, what))) return rc;
    ctx->batch_stamps = stamps;
    return vf::end_timed(ctx);
}

}  // namespace

extern "C" {

int vf_trace_batch(vf_ctx *ctx, const vf_request *reqs, int32_t n_req, const vf_queue_cfg *cfg_in) {
    return trace_batch(ctx, reqs, n_req, cfg_in, false);
}

int vf_trace_batch_noalloc(vf_ctx *ctx, const vf_request *reqs, int32_t n_req, const vf_queue_cfg *cfg_in) {
    return trace_batch(ctx, reqs, n_req, cfg_in, true);
}

int vf_read_batch_times(vf_ctx *ctx, int32_t n_req, double *ms) {
    if (!ctx || !ms || n_req <= 0 || n_req > kMaxReq) return vf::fail(ctx, VF_E_INVALID, "vf_read_batch_times: bad argument");
    if (int rc = vf::no_capture(ctx, "vf_read_batch_times")) return rc;
    if (!ctx->batch_stamps) return vf::fail(ctx, VF_E_INVALID, "vf_read_batch_times: no batch has run");
    unsigned long long st[kMaxReq + 1];
    VF_HIP(ctx, hipMemcpyAsync(st, ctx->batch_stamps, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
    VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const double tick_ms = 1.0 / (double)ctx->wall_khz;
    for (int k = 0; k < n_req; ++k) ms[k] = st[k] ? (double)(st[k] - st[kMaxReq]) * tick_ms : -1.0;
    return VF_OK;
}

int vf_autotune(vf_ctx *ctx, int32_t workload, const void *args, const vf_queue_cfg *base_cfg, vf_queue_cfg *best,
                vf_tune_result *results, int32_t max_results, int32_t *n_results) {
    if (!ctx || !args || !best) return vf::fail(ctx, VF_E_INVALID, "vf_autotune: null argument");
    if (int rc = vf::no_capture(ctx, "vf_autotune")) return rc;
    static const int tiles[3][2] = {{8, 8}, {16, 8}, {16, 16}};
    static const int chunks[5] = {64, 128, 256, 512, 1024};
    const int sorts = workload == VF_TUNE_GI_BATCH ? 2 : 1;
    static const int refills[3] = {-1, 0, 32};   // -1: no queue (GI batch: one configuration; batch: per tile)
    const int travs = workload == VF_TUNE_BATCH ? 2 : 1;   // the batch's flat mode with either traversal
    int k = 0;
    float best_ms = 1e30f;
    hipEvent_t e0, e1;
    VF_HIP(ctx, hipEventCreate(&e0));
    VF_HIP(ctx, hipEventCreate(&e1));
    int rc = VF_OK;
    for (int rf : refills)
    for (int tv = 0; tv < (rf < 0 ? travs : 1); ++tv)
    for (int so = 0; so < sorts && rc == VF_OK; ++so)
        for (const auto &t : tiles)
            for (int ch : chunks) {
                if (rf < 0 && (ch != 64 || (workload == VF_TUNE_GI_BATCH && (so != 0 || t[0] != 16 || t[1] != 16)))) continue;
                vf_queue_cfg c = default_cfg();
                c.tile_w = t[0]; c.tile_h = t[1]; c.chunk = ch; c.sort = so;
                c.refill = rf;
                c.traversal = tv;
                if (base_cfg) c.groups_per_cu = base_cfg->groups_per_cu;
                auto run = [&]() -> int {
                    if (workload == VF_TUNE_GI_BATCH) {
                        const vf_tune_gi_args *a = static_cast<const vf_tune_gi_args *>(args);
                        return vf_trace_gi_batch(ctx, a->pass, a->width, a->height, a->reach, a->d_gb_pos, a->d_gb_normal,
                                                 a->d_gi_hits, a->d_sun_hits, &c);
                    }
                    const vf_tune_batch_args *a = static_cast<const vf_tune_batch_args *>(args);
                    return vf_trace_batch(ctx, a->reqs, a->n_req, &c);
                };
                if ((rc = run())) break;   // warm-up (and buffer allocation)
                float ms_med[3];
                for (int rep = 0; rep < 3 && rc == VF_OK; ++rep) {
                    if (hipEventRecord(e0, ctx->stream) != hipSuccess) { rc = VF_E_DEVICE; break; }
                    if ((rc = run())) break;
                    if (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                        hipEventElapsedTime(&ms_med[rep], e0, e1) != hipSuccess) {
                        rc = VF_E_DEVICE;
                        break;
                    }
                }
                if (rc) break;
                std::sort(ms_med, ms_med + 3);
                if (results && k < max_results) { results[k].cfg = c; results[k].ms = ms_med[1]; }
                ++k;
                if (ms_med[1] < best_ms) { best_ms = ms_med[1]; *best = c; }
            }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc == VF_E_DEVICE ? vf::fail(ctx, rc, "vf_autotune: timing failed") : rc;
    if (n_results) *n_results = k;
    return VF_OK;
}

}  // extern "C"
