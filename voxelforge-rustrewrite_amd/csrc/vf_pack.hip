// Device-side world pack: GpuWorld::upload's CPU pack (reference src/gfx/gpu_world.rs:228-330) done by
// kernels.  The reference walks the grid on the CPU, gives each detailed brick the next atlas slot
// (encode_brick, :191-214), fills a staged atlas layer by layer and computes every brick's masks, then
// copies it all to the GPU.  Here the source grid and bricks go to the device as they are and:
//   * k_first_cell / k_first_flags: the first cell (in grid order) that names each brick;
//   * an exclusive scan of those flags: slot = the number of bricks first named before it, exactly the
//     reference's sequential allocation on a fresh atlas (a brick named again reuses its slot);
//   * k_majority: past the atlas capacity a brick's cells take its majority material instead;
//   * k_encode / k_scatter / k_brick_mask: the encoded grid, the atlas slots (TORCH_FLAME as air,
//     gpu_voxels :58-67) in 16-byte pieces, the reference's brickMasks of each placed brick;
// then the shared tail of the world upload (occupancy masks, distance fields, cell words, sky-top).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "../../include/vf_trace.h"
#include "vf_ctx.h"

namespace {

constexpr uint8_t kTorchFlame = 254;   // wm::TORCH_FLAME

__global__ void k_fill_i32(int32_t *p, int32_t v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_first_cell(const uint32_t *grid, int64_t n, uint32_t *first) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t v = grid[i];
    if (v >= 256u) atomicMin(&first[v - 256u], (uint32_t)i);
}

__global__ void k_first_flags(const uint32_t *grid, int64_t n, const uint32_t *first, uint32_t *flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t v = grid[i];
    flag[i] = (v >= 256u && first[v - 256u] == (uint32_t)i) ? 1u : 0u;
}

// slot of each brick at its first cell: the rank among first cells, or -2 past the capacity
__global__ void k_slots(const uint32_t *grid, int64_t n, const uint32_t *flag, const uint32_t *rank, uint32_t cap,
                        int32_t *slot_of) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    slot_of[grid[i] - 256u] = rank[i] < cap ? (int32_t)rank[i] : -2;
}

// encode_brick's full-atlas branch (gpu_world.rs:202-213): the most frequent material of the brick other
// than air and TORCH_FLAME, ties to the lower value, starting from 1; used only if it fills half the brick.
__global__ void __launch_bounds__(256) k_majority(const uint8_t *bricks, const int32_t *slot_of, uint32_t *major) {
    const uint32_t b = blockIdx.x;
    if (slot_of[b] != -2) return;
    __shared__ uint32_t cnt[256];
    cnt[threadIdx.x] = 0u;
    __syncthreads();
    const uint8_t *v = bricks + (size_t)b * 512u;
    atomicAdd(&cnt[v[threadIdx.x]], 1u);
    atomicAdd(&cnt[v[threadIdx.x + 256]], 1u);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t best = 1u;
        for (uint32_t m = 2u; m < 256u; ++m)
            if (cnt[m] > cnt[best] && m != kTorchFlame) best = m;
        major[b] = cnt[best] * 2u >= 512u ? best : 0u;
    }
}

__global__ void k_encode(const uint32_t *src, int64_t n, const int32_t *slot_of, const uint32_t *major, uint32_t *grid) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t v = src[i];
    if (v < 256u) {
        grid[i] = v;
        return;
    }
    const int32_t s = slot_of[v - 256u];
    grid[i] = s >= 0 ? 256u + (uint32_t)s : major[v - 256u];
}

__device__ __forceinline__ uint32_t no_flame(uint32_t w) {
    // bytes equal to TORCH_FLAME -> 0
    uint32_t o = w;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (((w >> (8 * k)) & 255u) == kTorchFlame) o &= ~(255u << (8 * k));
    return o;
}

// one thread per 16-byte piece of a brick (32 per brick)
__global__ void k_scatter(const uint4 *bricks, int64_t n_pieces, const int32_t *slot_of, uint4 *atlas) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_pieces) return;
    const int64_t b = i >> 5;
    const int32_t s = slot_of[b];
    if (s < 0) return;
    const uint4 v = bricks[i];
    atlas[(size_t)s * 32u + (size_t)(i & 31)] = make_uint4(no_flame(v.x), no_flame(v.y), no_flame(v.z), no_flame(v.w));
}

// brick_mask (sparse_volume.rs:48-63) of each placed brick's own voxels: word 2*octant + h holds voxel
// (x, y, z) of octant (x>>2, y>>2, z>>2) with ((z&3)>>1) == h at bit (x&3) | (y&3)<<2 | ((z&3)&1)<<4.
// One thread per word (16 per brick).
__global__ void k_brick_mask(const uint8_t *bricks, int64_t n_words, const int32_t *slot_of, uint32_t *masks) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_words) return;
    const int64_t b = i >> 4;
    const int32_t s = slot_of[b];
    if (s < 0) return;
    const uint32_t word = (uint32_t)(i & 15), oct = word >> 1, h = word & 1u;
    const uint32_t ox = (oct & 1u) * 4u, oy = ((oct >> 1) & 1u) * 4u, oz = ((oct >> 2) & 1u) * 4u + h * 2u;
    const uint8_t *v = bricks + (size_t)b * 512u;
    uint32_t m = 0u;
#pragma unroll
    for (uint32_t dz = 0; dz < 2; ++dz)
#pragma unroll
        for (uint32_t y = 0; y < 4; ++y)
#pragma unroll
            for (uint32_t x = 0; x < 4; ++x)
                if (v[(oz + dz) * 64u + (oy + y) * 8u + ox + x] != 0u) m |= 1u << (x | (y << 2) | (dz << 4));
    masks[(size_t)s * 16u + word] = m;
}

struct DevTemp {
    std::vector<void *> p;
    ~DevTemp() {
        for (void *q : p) (void)hipFree(q);
    }
    template <typename T>
    hipError_t alloc(T *&ptr, size_t count) {
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, count * sizeof(T) > 0 ? count * sizeof(T) : 1);
        if (e == hipSuccess) p.push_back(q);
        ptr = (T *)q;
        return e;
    }
};

}  // namespace

extern "C" int vf_world_pack(vf_ctx *ctx, const vf_world_src *src, int32_t *slot_of_brick) {
    if (!ctx || !src) return vf::fail(ctx, VF_E_INVALID, "vf_world_pack: null argument");
    if (int rc = vf::no_capture(ctx, "vf_world_pack")) return rc;
    if (!src->grid || !src->palette || (src->n_bricks && !src->bricks))
        return vf::fail(ctx, VF_E_INVALID, "vf_world_pack: grid, palette and bricks are required");
    for (int a = 0; a < 3; ++a)
        if (src->world_bricks[a] <= 0 || src->world_bricks[a] > 4096)
            return vf::fail(ctx, VF_E_INVALID, "vf_world_pack: world_bricks[%d] = %d out of range", a, src->world_bricks[a]);
    if (src->n_bricks > (1u << 26)) return vf::fail(ctx, VF_E_INVALID, "vf_world_pack: %u bricks", src->n_bricks);
    const int64_t n = (int64_t)src->world_bricks[0] * src->world_bricks[1] * src->world_bricks[2];
    if (n >= (int64_t)1 << 31) return vf::fail(ctx, VF_E_INVALID, "vf_world_pack: %lld grid cells", (long long)n);
    for (int64_t i = 0; i < n; ++i)
        if (src->grid[i] >= 256u && src->grid[i] - 256u >= src->n_bricks)
            return vf::fail(ctx, VF_E_INVALID, "vf_world_pack: grid[%lld] = %u names brick %u of %u", (long long)i, src->grid[i],
                            src->grid[i] - 256u, src->n_bricks);
    int rc = vf::set_device(ctx);
    if (rc) return rc;
    const int64_t nb = std::max<int64_t>(src->n_bricks, 1);
    DevTemp t;
    uint32_t *d_src = nullptr, *first = nullptr, *flag = nullptr, *rank = nullptr, *major = nullptr;
    int32_t *slot_of = nullptr;
    uint8_t *d_bricks = nullptr;
    hipError_t e;
    if ((e = t.alloc(d_src, n)) != hipSuccess || (e = t.alloc(first, nb)) != hipSuccess || (e = t.alloc(flag, n)) != hipSuccess ||
        (e = t.alloc(rank, n)) != hipSuccess || (e = t.alloc(slot_of, nb)) != hipSuccess || (e = t.alloc(major, nb)) != hipSuccess ||
        (e = t.alloc(d_bricks, (size_t)nb * 512u)) != hipSuccess)
        return vf::fail(ctx, VF_E_OOM, "vf_world_pack: hipMalloc: %s", hipGetErrorString(e));
    hipStream_t st = ctx->stream;
    VF_HIP(ctx, hipMemcpyAsync(d_src, src->grid, (size_t)n * 4u, hipMemcpyHostToDevice, st));
    if (src->n_bricks)
        VF_HIP(ctx, hipMemcpyAsync(d_bricks, src->bricks, (size_t)src->n_bricks * 512u, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_fill_i32, dim3(vf::blocks_for(nb)), dim3(256), 0, st, (int32_t *)first, -1, nb);
    hipLaunchKernelGGL(k_fill_i32, dim3(vf::blocks_for(nb)), dim3(256), 0, st, slot_of, -1, nb);
    hipLaunchKernelGGL(k_first_cell, dim3(vf::blocks_for(n)), dim3(256), 0, st, (const uint32_t *)d_src, n, first);
    hipLaunchKernelGGL(k_first_flags, dim3(vf::blocks_for(n)), dim3(256), 0, st, (const uint32_t *)d_src, n,
                       (const uint32_t *)first, flag);
    if ((rc = vf::launch_check(ctx, "world pack"))) return rc;
    size_t scan_bytes = 0;
    VF_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, flag, rank, (int)n, st));
    uint8_t *scan_tmp = nullptr;
    if ((e = t.alloc(scan_tmp, scan_bytes)) != hipSuccess)
        return vf::fail(ctx, VF_E_OOM, "vf_world_pack: hipMalloc: %s", hipGetErrorString(e));
    VF_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, flag, rank, (int)n, st));
    uint32_t last[2] = {0u, 0u};
    VF_HIP(ctx, hipMemcpyAsync(&last[0], rank + (n - 1), 4, hipMemcpyDeviceToHost, st));
    VF_HIP(ctx, hipMemcpyAsync(&last[1], flag + (n - 1), 4, hipMemcpyDeviceToHost, st));
    VF_HIP(ctx, hipStreamSynchronize(st));
    const uint32_t total = last[0] + last[1];
    const uint32_t cap = src->atlas_slots ? src->atlas_slots : total;
    if ((rc = vf::world_prepare(ctx, src->world_bricks, src->world_win, cap, "vf_world_pack"))) return rc;
    hipLaunchKernelGGL(k_slots, dim3(vf::blocks_for(n)), dim3(256), 0, st, (const uint32_t *)d_src, n, (const uint32_t *)flag,
                       (const uint32_t *)rank, cap, slot_of);
    if (total > cap)
        hipLaunchKernelGGL(k_majority, dim3((unsigned)src->n_bricks), dim3(256), 0, st, (const uint8_t *)d_bricks,
                           (const int32_t *)slot_of, major);
    hipLaunchKernelGGL(k_encode, dim3(vf::blocks_for(n)), dim3(256), 0, st, (const uint32_t *)d_src, n, (const int32_t *)slot_of,
                       (const uint32_t *)major, ctx->grid);
    if (src->n_bricks) {
        const int64_t pieces = (int64_t)src->n_bricks * 32, words = (int64_t)src->n_bricks * 16;
        hipLaunchKernelGGL(k_scatter, dim3(vf::blocks_for(pieces)), dim3(256), 0, st, (const uint4 *)d_bricks, pieces,
                           (const int32_t *)slot_of, (uint4 *)ctx->atlas);
        hipLaunchKernelGGL(k_brick_mask, dim3(vf::blocks_for(words)), dim3(256), 0, st, (const uint8_t *)d_bricks, words,
                           (const int32_t *)slot_of, ctx->masks_ref);
    }
    if ((rc = vf::launch_check(ctx, "world pack"))) return rc;
    VF_HIP(ctx, hipMemcpyAsync(ctx->palette, src->palette, 256 * 48, hipMemcpyHostToDevice, st));
    if ((rc = vf::build_coarse(ctx, d_src))) return rc;   // from the source grid, as the reference's pack
    if ((rc = vf::world_finish(ctx, true))) return rc;
    ctx->atlas_used = std::min(total, cap);
    ctx->free_slots.clear();
    ctx->world_slot.assign(cap, 0);
    std::fill(ctx->world_slot.begin(), ctx->world_slot.begin() + ctx->atlas_used, (uint8_t)1);
    // the host copies the camera test reads (the encoded grid, each slot's occupancy bits)
    ctx->host_grid.resize((size_t)n);
    ctx->host_occ.resize((size_t)cap * 8u);
    VF_HIP(ctx, hipMemcpyAsync(ctx->host_grid.data(), ctx->grid, (size_t)n * 4u, hipMemcpyDeviceToHost, st));
    if (cap) VF_HIP(ctx, hipMemcpyAsync(ctx->host_occ.data(), ctx->occ, (size_t)cap * 64u, hipMemcpyDeviceToHost, st));
    if (slot_of_brick && src->n_bricks)
        VF_HIP(ctx, hipMemcpyAsync(slot_of_brick, slot_of, (size_t)src->n_bricks * 4u, hipMemcpyDeviceToHost, st));
    VF_HIP(ctx, hipStreamSynchronize(st));
    if (slot_of_brick)
        for (uint32_t b = 0; b < src->n_bricks; ++b)
            if (slot_of_brick[b] < 0) slot_of_brick[b] = -1;
    return VF_OK;
}

extern "C" int vf_world_read(vf_ctx *ctx, uint32_t *grid, uint8_t *atlas, uint32_t n_slots, uint32_t *masks) {
    int rc = vf::check_ready(ctx, "vf_world_read");
    if (rc) return rc;
    if ((rc = vf::no_capture(ctx, "vf_world_read"))) return rc;
    if (n_slots > ctx->n_slots) return vf::fail(ctx, VF_E_INVALID, "vf_world_read: %u slots of %u", n_slots, ctx->n_slots);
    VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (grid) VF_HIP(ctx, hipMemcpy(grid, ctx->grid, ctx->n_cells * 4u, hipMemcpyDeviceToHost));
    if (atlas && n_slots) VF_HIP(ctx, hipMemcpy(atlas, ctx->atlas, (size_t)n_slots * 512u, hipMemcpyDeviceToHost));
    if (masks && n_slots) VF_HIP(ctx, hipMemcpy(masks, ctx->masks_ref, (size_t)n_slots * 64u, hipMemcpyDeviceToHost));
    return VF_OK;
}

extern "C" int vf_read_coarse(vf_ctx *ctx, uint8_t *coarse1, uint8_t *coarse2) {
    int rc = vf::check_ready(ctx, "vf_read_coarse");
    if (rc) return rc;
    if ((rc = vf::no_capture(ctx, "vf_read_coarse"))) return rc;
    VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const size_t n1 = (size_t)ctx->c1d[0] * ctx->c1d[1] * ctx->c1d[2], n2 = (size_t)ctx->c2d[0] * ctx->c2d[1] * ctx->c2d[2];
    if (coarse1) VF_HIP(ctx, hipMemcpy(coarse1, ctx->c1, n1, hipMemcpyDeviceToHost));
    if (coarse2) VF_HIP(ctx, hipMemcpy(coarse2, ctx->c2, n2, hipMemcpyDeviceToHost));
    return VF_OK;
}

// ------------------------------------------------------------------------------------------------
// Chunk sync (GpuWorld::sync_chunks, gpu_world.rs:1194-1376)
// ------------------------------------------------------------------------------------------------
namespace {

constexpr int32_t kChunkB = 16;   // RING_CHUNK_B

struct SyncArgs {
    int32_t wbx, wby, wbz;
    int32_t n_loaded;             // loaded chunks in `cells` order
    int64_t cells_per_chunk;      // 16 * wby * 16
    uint32_t n_free, used, cap;
};

// grid position of cell k of the concatenated loaded chunks
__device__ __forceinline__ size_t sync_cell_pos(const SyncArgs &A, const int2 *org, int64_t k) {
    const int64_t c = k / A.cells_per_chunk, r = k % A.cells_per_chunk;
    const int32_t x = (int32_t)(r % kChunkB), y = (int32_t)((r / kChunkB) % A.wby), z = (int32_t)(r / ((int64_t)kChunkB * A.wby));
    const int2 o = org[c];   // the chunk's wrapped brick origin (tx, tz)
    return ((size_t)(o.y + z) * A.wby + y) * A.wbx + (size_t)(o.x + x);
}

__global__ void k_sync_first(const uint32_t *cells, int64_t n, const int32_t *gslot, uint32_t *first, uint8_t *ref) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t v = cells[k];
    if (v < 256u) return;
    ref[v - 256u] = 1u;
    if (gslot[v - 256u] < 0) atomicMin(&first[v - 256u], (uint32_t)k);
}

__global__ void k_sync_flags(const uint32_t *cells, int64_t n, const int32_t *gslot, const uint32_t *first, uint32_t *flag) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t v = cells[k];
    flag[k] = (v >= 256u && gslot[v - 256u] < 0 && first[v - 256u] == (uint32_t)k) ? 1u : 0u;
}

// alloc_slot (gpu_world.rs:178-186) for the k-th new brick: the free list's back first, then the next
// unused slot, else -2 (no slot: majority material)
__global__ void k_sync_assign(const SyncArgs A, const uint32_t *cells, int64_t n, const uint32_t *flag, const uint32_t *rank,
                              const uint32_t *free_list, int32_t *slot) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n || !flag[k]) return;
    const uint32_t r = rank[k];
    int32_t s;
    if (r < A.n_free) s = (int32_t)free_list[A.n_free - 1u - r];
    else s = (uint64_t)A.used + (r - A.n_free) < (uint64_t)A.cap ? (int32_t)(A.used + (r - A.n_free)) : -2;
    slot[cells[k] - 256u] = s;
}

__global__ void k_sync_encode(const SyncArgs A, const int2 *org, const uint32_t *cells, int64_t n, const int32_t *slot,
                              const uint32_t *major, uint32_t *grid) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t v = cells[k];
    uint32_t e = v;
    if (v >= 256u) {
        const int32_t s = slot[v - 256u];
        e = s >= 0 ? 256u + (uint32_t)s : major[v - 256u];
    }
    grid[sync_cell_pos(A, org, k)] = e;
}

// the slots whose bricks are written (named by the loaded chunks and placed), or -1
__global__ void k_sync_written(const int32_t *slot, const uint8_t *ref, int64_t nb, int32_t *written) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) written[b] = (ref[b] && slot[b] >= 0) ? slot[b] : -1;
}

// an unloaded chunk's grid cells -> 0
__global__ void k_sync_clear(const SyncArgs A, int2 o, uint32_t *grid) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= A.cells_per_chunk) return;
    const int32_t x = (int32_t)(k % kChunkB), y = (int32_t)((k / kChunkB) % A.wby), z = (int32_t)(k / ((int64_t)kChunkB * A.wby));
    grid[((size_t)(o.y + z) * A.wby + y) * A.wbx + (size_t)(o.x + x)] = 0u;
}

// coarse_cell1 (gpu_world.rs:944-957) of the chunk's 4^3-brick cells from its own cells (cells == NULL:
// an unloaded chunk, all 0)
__global__ void k_sync_coarse1(const SyncArgs A, int2 o, const uint32_t *cells, int32_t c1x, int32_t c1y, uint8_t *c1) {
    const int32_t per = (kChunkB / 4) * c1y * (kChunkB / 4);
    const int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= per) return;
    const int32_t cx = k % 4, cy = (k / 4) % c1y, cz = k / (4 * c1y);
    uint8_t occ = 0u;
    if (cells) {
        for (int32_t z = cz * 4; z < cz * 4 + 4 && !occ; ++z)
            for (int32_t y = cy * 4; y < min(cy * 4 + 4, A.wby) && !occ; ++y)
                for (int32_t x = cx * 4; x < cx * 4 + 4; ++x)
                    if (cells[((size_t)z * A.wby + y) * kChunkB + x] != 0u) { occ = 1u; break; }
    }
    c1[((size_t)(o.y / 4 + cz) * c1y + cy) * c1x + (size_t)(o.x / 4 + cx)] = occ;
}

// coarse_cell2 (gpu_world.rs:959-973) of the chunk's column of 16^3-brick cells from coarse1
__global__ void k_sync_coarse2(int2 o, int32_t c1x, int32_t c1y, int32_t c1z, int32_t c2x, int32_t c2y, const uint8_t *c1,
                               uint8_t *c2) {
    const int32_t cy = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (cy >= c2y) return;
    const int32_t cx = o.x / 16, cz = o.y / 16;
    uint8_t occ = 0u;
    for (int32_t z = cz * 4; z < min(cz * 4 + 4, c1z) && !occ; ++z)
        for (int32_t y = cy * 4; y < min(cy * 4 + 4, c1y) && !occ; ++y)
            for (int32_t x = cx * 4; x < min(cx * 4 + 4, c1x); ++x)
                if (c1[((size_t)z * c1y + y) * c1x + x] != 0u) { occ = 1u; break; }
    c2[((size_t)cz * c2y + cy) * c2x + cx] = occ;
}

}  // namespace

extern "C" int vf_world_atlas_state(vf_ctx *ctx, uint32_t *used, uint32_t *n_free, uint32_t *capacity) {
    int rc = vf::check_ready(ctx, "vf_world_atlas_state");
    if (rc) return rc;
    if (used) *used = ctx->atlas_used;
    if (n_free) *n_free = (uint32_t)ctx->free_slots.size();
    if (capacity) *capacity = ctx->n_slots;
    return VF_OK;
}

extern "C" int vf_world_sync_chunks(vf_ctx *ctx, const vf_chunk *chunks, int32_t n_chunks, const uint8_t *bricks, uint32_t n_bricks,
                                    int32_t *gpu_slot, const int32_t *released, int32_t n_released) {
    int rc = vf::check_ready(ctx, "vf_world_sync_chunks");
    if (rc) return rc;
    if ((rc = vf::no_capture(ctx, "vf_world_sync_chunks"))) return rc;
    const DevWorld &W = ctx->W;
    const auto pow2 = [](int32_t v) { return v > 0 && (v & (v - 1)) == 0; };
    if (!W.ring || W.wb[0] != W.wb[2] || !pow2(W.wb[0]) || W.wb[0] % kChunkB)
        return vf::fail(ctx, VF_E_INVALID, "vf_world_sync_chunks: needs a square ring world of whole chunks (%d x %d bricks)",
                        W.wb[0], W.wb[2]);
    if (n_chunks < 0 || n_released < 0 || (n_chunks && !chunks) || (n_released && !released) || (n_bricks && (!bricks || !gpu_slot)))
        return vf::fail(ctx, VF_E_INVALID, "vf_world_sync_chunks: null or negative argument");
    const int32_t ring_w = W.wb[0] / kChunkB, ring_d = W.wb[2] / kChunkB;
    const int64_t per = (int64_t)kChunkB * W.wb[1] * kChunkB;
    // chunks once per ring slot, in order; loaded ones concatenated
    std::vector<uint8_t> done((size_t)ring_w * ring_d, 0);
    std::vector<const vf_chunk *> list;
    std::vector<int2> org_loaded;
    std::vector<uint32_t> cat;
    for (int32_t i = 0; i < n_chunks; ++i) {
        const vf_chunk &c = chunks[i];
        if (c.cx < -(1 << 26) || c.cx >= (1 << 26) || c.cz < -(1 << 26) || c.cz >= (1 << 26))
            return vf::fail(ctx, VF_E_INVALID, "vf_world_sync_chunks: chunk (%d, %d) out of range", c.cx, c.cz);
        const size_t slot = (size_t)(c.cz & (ring_d - 1)) * ring_w + (size_t)(c.cx & (ring_w - 1));
        if (done[slot]) continue;
        done[slot] = 1;
        list.push_back(&c);
        if (c.loaded) {
            if (!c.cells) return vf::fail(ctx, VF_E_INVALID, "vf_world_sync_chunks: loaded chunk (%d, %d) without cells", c.cx, c.cz);
            for (int64_t k = 0; k < per; ++k)
                if (c.cells[k] >= 256u && c.cells[k] - 256u >= n_bricks)
                    return vf::fail(ctx, VF_E_INVALID, "vf_world_sync_chunks: chunk (%d, %d) names brick %u of %u", c.cx, c.cz,
                                    c.cells[k] - 256u, n_bricks);
            org_loaded.push_back(make_int2((c.cx * kChunkB) & (W.wb[0] - 1), (c.cz * kChunkB) & (W.wb[2] - 1)));
            cat.insert(cat.end(), c.cells, c.cells + per);
        }
    }
    const int64_t n = (int64_t)cat.size();
    if (n > (int64_t)INT32_MAX) return vf::fail(ctx, VF_E_INVALID, "vf_world_sync_chunks: %lld cells in one call", (long long)n);
    const int64_t nb = std::max<int64_t>(n_bricks, 1);
    // released slots that hold world bricks, highest first (gpu_world.rs:1199-1207); taken below, once
    // nothing can fail before the world changes
    std::vector<int32_t> rel(released, released + n_released);
    std::sort(rel.begin(), rel.end(), [](int32_t a, int32_t b) { return a > b; });
    rel.erase(std::unique(rel.begin(), rel.end()), rel.end());
    rel.erase(std::remove_if(rel.begin(), rel.end(),
                             [&](int32_t sl) { return sl < 0 || (size_t)sl >= ctx->world_slot.size() || !ctx->world_slot[(size_t)sl]; }),
              rel.end());
    hipStream_t st = ctx->stream;
    DevTemp t;
    uint32_t *d_cells = nullptr, *first = nullptr, *flag = nullptr, *rank = nullptr, *major = nullptr, *d_free = nullptr;
    int32_t *slot = nullptr, *written = nullptr;
    uint8_t *ref = nullptr, *d_bricks = nullptr;
    int2 *d_org = nullptr;
    hipError_t e;
    const size_t nf = ctx->free_slots.size() + rel.size();
    size_t scan_bytes = 0;
    uint8_t *scan_tmp = nullptr;
    if (n) VF_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, flag, rank, (int)n, st));
    if ((e = t.alloc(scan_tmp, std::max<size_t>(scan_bytes, 1))) != hipSuccess ||
        (e = t.alloc(d_cells, std::max<int64_t>(n, 1))) != hipSuccess || (e = t.alloc(first, nb)) != hipSuccess ||
        (e = t.alloc(flag, std::max<int64_t>(n, 1))) != hipSuccess || (e = t.alloc(rank, std::max<int64_t>(n, 1))) != hipSuccess ||
        (e = t.alloc(major, nb)) != hipSuccess || (e = t.alloc(d_free, std::max<size_t>(nf, 1))) != hipSuccess ||
        (e = t.alloc(slot, nb)) != hipSuccess || (e = t.alloc(written, nb)) != hipSuccess || (e = t.alloc(ref, nb)) != hipSuccess ||
        (e = t.alloc(d_bricks, (size_t)nb * 512u)) != hipSuccess || (e = t.alloc(d_org, std::max<size_t>(org_loaded.size(), 1))) != hipSuccess)
        return vf::fail(ctx, VF_E_OOM, "vf_world_sync_chunks: hipMalloc: %s", hipGetErrorString(e));
    for (int32_t sl : rel) {
        ctx->world_slot[(size_t)sl] = 0;
        ctx->free_slots.push_back((uint32_t)sl);
    }
    SyncArgs A{};
    A.wbx = W.wb[0]; A.wby = W.wb[1]; A.wbz = W.wb[2];
    A.n_loaded = (int32_t)org_loaded.size();
    A.cells_per_chunk = per;
    A.n_free = (uint32_t)nf;
    A.used = ctx->atlas_used;
    A.cap = ctx->n_slots;
    uint32_t total_new = 0;
    if (n) {
        VF_HIP(ctx, hipMemcpyAsync(d_cells, cat.data(), (size_t)n * 4u, hipMemcpyHostToDevice, st));
        VF_HIP(ctx, hipMemcpyAsync(d_org, org_loaded.data(), org_loaded.size() * sizeof(int2), hipMemcpyHostToDevice, st));
        if (n_bricks) {
            VF_HIP(ctx, hipMemcpyAsync(d_bricks, bricks, (size_t)n_bricks * 512u, hipMemcpyHostToDevice, st));
            VF_HIP(ctx, hipMemcpyAsync(slot, gpu_slot, (size_t)n_bricks * 4u, hipMemcpyHostToDevice, st));
        }
        if (nf) VF_HIP(ctx, hipMemcpyAsync(d_free, ctx->free_slots.data(), nf * 4u, hipMemcpyHostToDevice, st));
        VF_HIP(ctx, hipMemsetAsync(ref, 0, (size_t)nb, st));
        hipLaunchKernelGGL(k_fill_i32, dim3(vf::blocks_for(nb)), dim3(256), 0, st, (int32_t *)first, -1, nb);
        hipLaunchKernelGGL(k_sync_first, dim3(vf::blocks_for(n)), dim3(256), 0, st, (const uint32_t *)d_cells, n,
                           (const int32_t *)slot, first, ref);
        hipLaunchKernelGGL(k_sync_flags, dim3(vf::blocks_for(n)), dim3(256), 0, st, (const uint32_t *)d_cells, n,
                           (const int32_t *)slot, (const uint32_t *)first, flag);
        if ((rc = vf::launch_check(ctx, "chunk sync"))) return rc;
        VF_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, flag, rank, (int)n, st));
        uint32_t last[2] = {0u, 0u};
        VF_HIP(ctx, hipMemcpyAsync(&last[0], rank + (n - 1), 4, hipMemcpyDeviceToHost, st));
        VF_HIP(ctx, hipMemcpyAsync(&last[1], flag + (n - 1), 4, hipMemcpyDeviceToHost, st));
        VF_HIP(ctx, hipStreamSynchronize(st));
        total_new = last[0] + last[1];
        hipLaunchKernelGGL(k_sync_assign, dim3(vf::blocks_for(n)), dim3(256), 0, st, A, (const uint32_t *)d_cells, n,
                           (const uint32_t *)flag, (const uint32_t *)rank, (const uint32_t *)d_free, slot);
        if (n_bricks && total_new > nf && (uint64_t)A.used + (total_new - nf) > A.cap)
            hipLaunchKernelGGL(k_majority, dim3(n_bricks), dim3(256), 0, st, (const uint8_t *)d_bricks, (const int32_t *)slot, major);
        hipLaunchKernelGGL(k_sync_encode, dim3(vf::blocks_for(n)), dim3(256), 0, st, A, (const int2 *)d_org, (const uint32_t *)d_cells,
                           n, (const int32_t *)slot, (const uint32_t *)major, ctx->grid);
        if (n_bricks) {
            hipLaunchKernelGGL(k_sync_written, dim3(vf::blocks_for(nb)), dim3(256), 0, st, (const int32_t *)slot, (const uint8_t *)ref,
                               nb, written);
            const int64_t pieces = (int64_t)n_bricks * 32, words = (int64_t)n_bricks * 16;
            hipLaunchKernelGGL(k_scatter, dim3(vf::blocks_for(pieces)), dim3(256), 0, st, (const uint4 *)d_bricks, pieces,
                               (const int32_t *)written, (uint4 *)ctx->atlas);
            hipLaunchKernelGGL(k_brick_mask, dim3(vf::blocks_for(words)), dim3(256), 0, st, (const uint8_t *)d_bricks, words,
                               (const int32_t *)written, ctx->masks_ref);
            if ((rc = vf::masks_for(ctx, (const uint32_t *)written, n_bricks))) return rc;   // -1 entries are skipped
        }
    }
    // cleared chunks, coarse cells of every synced chunk
    int64_t li = 0;
    for (const vf_chunk *c : list) {
        const int2 o = make_int2((c->cx * kChunkB) & (W.wb[0] - 1), (c->cz * kChunkB) & (W.wb[2] - 1));
        const uint32_t *cells = nullptr;
        if (c->loaded) cells = d_cells + (li++) * per;
        else hipLaunchKernelGGL(k_sync_clear, dim3(vf::blocks_for(per)), dim3(256), 0, st, A, o, ctx->grid);
        const int32_t c1per = (kChunkB / 4) * ctx->c1d[1] * (kChunkB / 4);
        hipLaunchKernelGGL(k_sync_coarse1, dim3(vf::blocks_for(c1per)), dim3(256), 0, st, A, o, cells, ctx->c1d[0], ctx->c1d[1], ctx->c1);
        hipLaunchKernelGGL(k_sync_coarse2, dim3(vf::blocks_for(ctx->c2d[1])), dim3(256), 0, st, o, ctx->c1d[0], ctx->c1d[1], ctx->c1d[2],
                           ctx->c2d[0], ctx->c2d[1], (const uint8_t *)ctx->c1, ctx->c2);
    }
    if ((rc = vf::launch_check(ctx, "chunk sync"))) return rc;
    // the distance field over each chunk's box +-16 bricks (gpu_world.rs:1355-1361), then cell words and sky-top
    for (const vf_chunk *c : list) {
        const int32_t lo[3] = {c->cx * kChunkB, 0, c->cz * kChunkB};
        const int32_t hi[3] = {c->cx * kChunkB + kChunkB - 1, W.wb[1] - 1, c->cz * kChunkB + kChunkB - 1};
        const int32_t flo[3] = {lo[0] - 16, lo[1] - 16, lo[2] - 16}, fhi[3] = {hi[0] + 16, hi[1] + 16, hi[2] + 16};
        if ((rc = vf::distance_box(ctx, flo, fhi, lo, hi))) return rc;
    }
    if ((rc = vf::refresh_world(ctx))) return rc;
    // bookkeeping: the slots popped and handed out, world-owned slots, the callers' gpu_slot
    std::vector<int32_t> out((size_t)nb, -1);
    if (n && n_bricks) {
        VF_HIP(ctx, hipMemcpyAsync(out.data(), slot, (size_t)n_bricks * 4u, hipMemcpyDeviceToHost, st));
        std::vector<int32_t> wr((size_t)n_bricks);
        VF_HIP(ctx, hipMemcpyAsync(wr.data(), written, (size_t)n_bricks * 4u, hipMemcpyDeviceToHost, st));
        VF_HIP(ctx, hipMemcpyAsync(ctx->host_grid.data(), ctx->grid, ctx->n_cells * 4u, hipMemcpyDeviceToHost, st));
        VF_HIP(ctx, hipStreamSynchronize(st));
        const uint32_t pops = std::min<uint32_t>(total_new, (uint32_t)nf);
        ctx->free_slots.resize(nf - pops);
        const uint32_t fresh = total_new > nf ? std::min<uint32_t>(total_new - (uint32_t)nf, A.cap - A.used) : 0u;
        ctx->atlas_used += fresh;
        for (uint32_t b = 0; b < n_bricks; ++b) {
            gpu_slot[b] = out[b] >= 0 ? out[b] : -1;
            if (wr[b] >= 0) {
                ctx->world_slot[(size_t)wr[b]] = 1;
                uint8_t vox[512];
                for (int k = 0; k < 512; ++k) vox[k] = bricks[(size_t)b * 512u + k] == kTorchFlame ? 0 : bricks[(size_t)b * 512u + k];
                vf::slot_occupancy(vox, &ctx->host_occ[(size_t)wr[b] * 8u]);
            }
        }
    } else {
        VF_HIP(ctx, hipMemcpyAsync(ctx->host_grid.data(), ctx->grid, ctx->n_cells * 4u, hipMemcpyDeviceToHost, st));
        VF_HIP(ctx, hipStreamSynchronize(st));
    }
    return VF_OK;
}
