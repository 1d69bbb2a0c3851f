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
#include <cstring>
#include <thread>
#include <unordered_map>
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

// The caller's bricks: voxels and slots at byte strides (packed arrays, or the fields of brick records).
struct BrickView {
    const uint8_t *vox;
    size_t vstride;
    int32_t *slot;
    size_t sstride;
    const uint8_t *at(size_t b) const { return vox + b * vstride; }
    int32_t get_slot(size_t b) const {
        int32_t v;
        std::memcpy(&v, reinterpret_cast<const uint8_t *>(slot) + b * sstride, 4);
        return v;
    }
    void set_slot(size_t b, int32_t v) const { std::memcpy(reinterpret_cast<uint8_t *>(slot) + b * sstride, &v, 4); }
};

// Checks the strides (0 = packed) and fills the view.
int brick_view(vf_ctx *ctx, const char *what, const uint8_t *vox, uint32_t vstride, int32_t *slot, uint32_t sstride,
               BrickView &v) {
    if (vstride && vstride < 512u) return vf::fail(ctx, VF_E_INVALID, "%s: brick stride %u < 512", what, vstride);
    if (sstride && (sstride % 4u)) return vf::fail(ctx, VF_E_INVALID, "%s: slot stride %u is not a multiple of 4", what, sstride);
    v.vox = vox;
    v.vstride = vstride ? vstride : 512u;
    v.slot = slot;
    v.sstride = sstride ? sstride : 4u;
    return VF_OK;
}

// Copies n bricks' voxels into a packed array (threads for large counts); idx == NULL: bricks 0..n-1.
void gather_bricks(const BrickView &v, const uint32_t *idx, size_t n, uint8_t *out) {
    const unsigned nt = n >= 4096u ? std::max(1u, std::min(8u, std::thread::hardware_concurrency())) : 1u;
    const auto work = [&](unsigned t) {
        for (size_t l = t; l < n; l += nt) std::memcpy(out + l * 512u, v.at(idx ? idx[l] : l), 512);
    };
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; ++t) th.emplace_back(work, t);
    work(0);
    for (auto &x : th) x.join();
}

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
    BrickView bv;
    int rc = brick_view(ctx, "vf_world_pack", src->bricks, src->brick_stride, slot_of_brick, src->slot_stride, bv);
    if (rc) return rc;
    if ((rc = vf::set_device(ctx))) return rc;
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
    std::vector<uint8_t> packed;   // bricks given as records: their voxels gathered first
    if (src->n_bricks && bv.vstride != 512u) {
        packed.resize((size_t)src->n_bricks * 512u);
        gather_bricks(bv, nullptr, src->n_bricks, packed.data());
    }
    if (src->n_bricks)
        VF_HIP(ctx, hipMemcpyAsync(d_bricks, packed.empty() ? src->bricks : packed.data(), (size_t)src->n_bricks * 512u,
                                   hipMemcpyHostToDevice, st));
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
    if ((rc = vf::world_finish(ctx, true, true))) return rc;
    ctx->atlas_used = std::min(total, cap);
    ctx->free_slots.clear();
    ctx->world_slot.assign(cap, 0);
    std::fill(ctx->world_slot.begin(), ctx->world_slot.begin() + ctx->atlas_used, (uint8_t)1);
    // the host copies the camera test reads (the encoded grid, each slot's occupancy bits)
    ctx->host_grid.resize((size_t)n);
    ctx->host_occ.resize((size_t)cap * 8u);
    VF_HIP(ctx, hipMemcpyAsync(ctx->host_grid.data(), ctx->grid, (size_t)n * 4u, hipMemcpyDeviceToHost, st));
    if (cap) VF_HIP(ctx, hipMemcpyAsync(ctx->host_occ.data(), ctx->occ, (size_t)cap * 64u, hipMemcpyDeviceToHost, st));
    std::vector<int32_t> slots(slot_of_brick ? (size_t)src->n_bricks : 0u);
    if (!slots.empty()) VF_HIP(ctx, hipMemcpyAsync(slots.data(), slot_of, slots.size() * 4u, hipMemcpyDeviceToHost, st));
    VF_HIP(ctx, hipStreamSynchronize(st));
    for (size_t b = 0; b < slots.size(); ++b) bv.set_slot(b, slots[b] < 0 ? -1 : slots[b]);
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

__global__ void k_sync_first(const uint32_t *cells, int64_t n, const int32_t *gslot, uint32_t *first) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t v = cells[k];
    if (v >= 256u && gslot[v - 256u] < 0) atomicMin(&first[v - 256u], (uint32_t)k);
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
                              const uint32_t *major, uint32_t *grid, uint8_t *srcnz, uint32_t *enc) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t v = cells[k];
    uint32_t e = v;
    if (v >= 256u) {
        const int32_t s = slot[v - 256u];
        e = s >= 0 ? 256u + (uint32_t)s : major[v - 256u];
    }
    const size_t pos = sync_cell_pos(A, org, k);
    grid[pos] = e;
    srcnz[pos] = v != 0u ? 1u : 0u;
    enc[k] = e;
}

// the slots the named bricks are written to, or -1 (no slot: majority material)
__global__ void k_sync_written(const int32_t *slot, int64_t nb, int32_t *written) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) written[b] = slot[b] >= 0 ? slot[b] : -1;
}

// the occupancy rows of the slots the named bricks went to (-1: none), for the host copy
__global__ void k_sync_occ(const int32_t *written, int64_t n_words, const uint32_t *occ, uint32_t *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_words) return;
    const int32_t s = written[i >> 4];
    out[i] = s >= 0 ? occ[(size_t)s * 16u + (size_t)(i & 15)] : 0u;
}

// Up to 64 chunks of one sync, one per blockIdx.y: their grid origins (bricks, wrapped into the ring) and the
// offset of their cells in the staged cell block (-1: an unloaded chunk). A sync's chunks sit in distinct ring
// slots, so their grid cells and coarse cells are disjoint.
struct SyncChunks {
    int32_t n;
    int2 o[64];
    int64_t cells[64];
};

// an unloaded chunk's grid cells -> 0
__global__ void k_sync_clear(const SyncArgs A, const SyncChunks C, uint32_t *grid, uint8_t *srcnz) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t w = (int32_t)blockIdx.y;
    if (k >= A.cells_per_chunk || w >= C.n || C.cells[w] >= 0) return;
    const int2 o = C.o[w];
    const int32_t x = (int32_t)(k % kChunkB), y = (int32_t)((k / kChunkB) % A.wby), z = (int32_t)(k / ((int64_t)kChunkB * A.wby));
    const size_t pos = ((size_t)(o.y + z) * A.wby + y) * A.wbx + (size_t)(o.x + x);
    grid[pos] = 0u;
    srcnz[pos] = 0u;
}

// coarse_cell1 (gpu_world.rs:944-957) of each chunk's 4^3-brick cells from its own cells (an unloaded chunk: all 0)
__global__ void k_sync_coarse1(const SyncArgs A, const SyncChunks C, const uint32_t *staged, int32_t c1x, int32_t c1y,
                               uint8_t *c1) {
    const int32_t per = (kChunkB / 4) * c1y * (kChunkB / 4);
    const int32_t k = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    const int32_t w = (int32_t)blockIdx.y;
    if (k >= per || w >= C.n) return;
    const int2 o = C.o[w];
    const int32_t cx = k % 4, cy = (k / 4) % c1y, cz = k / (4 * c1y);
    uint8_t occ = 0u;
    if (C.cells[w] >= 0) {
        const uint32_t *cells = staged + C.cells[w];
        for (int32_t z = cz * 4; z < cz * 4 + 4 && !occ; ++z)
            for (int32_t y = cy * 4; y < min(cy * 4 + 4, A.wby) && !occ; ++y)
                for (int32_t x = cx * 4; x < cx * 4 + 4; ++x)
                    if (cells[((size_t)z * A.wby + y) * kChunkB + x] != 0u) { occ = 1u; break; }
    }
    c1[((size_t)(o.y / 4 + cz) * c1y + cy) * c1x + (size_t)(o.x / 4 + cx)] = occ;
}

// coarse_cell2 (gpu_world.rs:959-973) of each chunk's column of 16^3-brick cells from coarse1
__global__ void k_sync_coarse2(const SyncChunks C, int32_t c1x, int32_t c1y, int32_t c1z, int32_t c2x, int32_t c2y,
                               const uint8_t *c1, uint8_t *c2) {
    const int32_t cy = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    const int32_t w = (int32_t)blockIdx.y;
    if (cy >= c2y || w >= C.n) return;
    const int32_t cx = C.o[w].x / 16, cz = C.o[w].y / 16;
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
                                    uint32_t brick_stride, int32_t *gpu_slot, uint32_t slot_stride, const int32_t *released,
                                    int32_t n_released) {
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
    BrickView bv;
    if ((rc = brick_view(ctx, "vf_world_sync_chunks", bricks, brick_stride, gpu_slot, slot_stride, bv))) return rc;
    const int32_t ring_w = W.wb[0] / kChunkB, ring_d = W.wb[2] / kChunkB;
    const int64_t per = (int64_t)kChunkB * W.wb[1] * kChunkB;
    // chunks once per ring slot, in order; loaded ones concatenated with their bricks renumbered by first
    // mention (only the bricks the chunks name travel to the device)
    std::vector<uint8_t> done((size_t)ring_w * ring_d, 0);
    std::vector<const vf_chunk *> list;
    std::vector<int2> org_loaded;
    std::vector<uint32_t> cat;
    std::vector<int32_t> loc(n_bricks, -1);
    std::vector<uint32_t> named;                     // local brick -> caller's brick
    for (int32_t i = 0; i < n_chunks; ++i) {
        const vf_chunk &c = chunks[i];
        if (c.cx < -(1 << 26) || c.cx >= (1 << 26) || c.cz < -(1 << 26) || c.cz >= (1 << 26))
            return vf::fail(ctx, VF_E_INVALID, "vf_world_sync_chunks: chunk (%d, %d) out of range", c.cx, c.cz);
        const size_t slot = (size_t)(c.cz & (ring_d - 1)) * ring_w + (size_t)(c.cx & (ring_w - 1));
        if (done[slot]) continue;
        done[slot] = 1;
        list.push_back(&c);
        if (!c.loaded) continue;
        if (!c.cells) return vf::fail(ctx, VF_E_INVALID, "vf_world_sync_chunks: loaded chunk (%d, %d) without cells", c.cx, c.cz);
        if ((int64_t)cat.size() + per > (int64_t)INT32_MAX)
            return vf::fail(ctx, VF_E_INVALID, "vf_world_sync_chunks: more than 2^31 cells in one call");
        org_loaded.push_back(make_int2((c.cx * kChunkB) & (W.wb[0] - 1), (c.cz * kChunkB) & (W.wb[2] - 1)));
        const size_t base = cat.size();
        cat.insert(cat.end(), c.cells, c.cells + per);
        for (size_t k = base; k < cat.size(); ++k) {
            const uint32_t v = cat[k];
            if (v < 256u) continue;
            if (v - 256u >= n_bricks)
                return vf::fail(ctx, VF_E_INVALID, "vf_world_sync_chunks: chunk (%d, %d) names brick %u of %u", c.cx, c.cz, v - 256u,
                                n_bricks);
            int32_t &l = loc[v - 256u];
            if (l < 0) {
                l = (int32_t)named.size();
                named.push_back(v - 256u);
            }
            cat[k] = 256u + (uint32_t)l;
        }
    }
    const int64_t n = (int64_t)cat.size();
    const uint32_t nl = (uint32_t)named.size();
    const int64_t nb = std::max<int64_t>(nl, 1);
    // one pinned staging block: the named bricks, their slots, the cells (one DMA each way)
    const size_t off_slot = (size_t)nl * 512u, off_cells = off_slot + (size_t)nb * 4u, stage = off_cells + (size_t)n * 4u;
    if (stage > ctx->pin_cap) {
        if (ctx->pin) (void)hipHostFree(ctx->pin);
        ctx->pin = nullptr;
        ctx->pin_cap = 0;
        const size_t want = std::max<size_t>(stage + stage / 4, (size_t)1 << 20);
        if (hipHostMalloc((void **)&ctx->pin, want, hipHostMallocDefault) != hipSuccess)
            return vf::fail(ctx, VF_E_OOM, "vf_world_sync_chunks: %zu bytes of pinned staging", want);
        ctx->pin_cap = want;
    }
    uint8_t *pb = ctx->pin;
    int32_t *pslot = reinterpret_cast<int32_t *>(pb + off_slot);
    uint32_t *pcells = reinterpret_cast<uint32_t *>(pb + off_cells);
    gather_bricks(bv, named.data(), nl, pb);
    for (uint32_t l = 0; l < nl; ++l) pslot[l] = bv.get_slot(named[l]);
    if (n) std::memcpy(pcells, cat.data(), (size_t)n * 4u);
    // released slots that hold world bricks, highest first (gpu_world.rs:1199-1207); taken below, once
    // nothing can fail before the world changes
    std::vector<int32_t> rel(released, released + n_released);
    std::sort(rel.begin(), rel.end(), [](int32_t a, int32_t b) { return a > b; });
    rel.erase(std::unique(rel.begin(), rel.end()), rel.end());
    rel.erase(std::remove_if(rel.begin(), rel.end(),
                             [&](int32_t sl) { return sl < 0 || (size_t)sl >= ctx->world_slot.size() || !ctx->world_slot[(size_t)sl]; }),
              rel.end());
    hipStream_t st = ctx->stream;
    const size_t nf = ctx->free_slots.size() + rel.size();
    const int64_t nc = std::max<int64_t>(n, 1);
    size_t scan_bytes = 0;
    if (n) VF_HIP(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)n, st));
    // device workspace, carved in 256-byte aligned pieces; kept by the context and grown on demand
    size_t ws = 0;
    const auto take = [&ws](size_t bytes) { const size_t o = ws; ws += (std::max<size_t>(bytes, 1) + 255u) & ~(size_t)255u; return o; };
    const size_t o_stage = take(stage), o_scan = take(scan_bytes), o_enc = take((size_t)nc * 4u), o_flag = take((size_t)nc * 4u),
                 o_rank = take((size_t)nc * 4u), o_first = take((size_t)nb * 4u), o_major = take((size_t)nb * 4u),
                 o_free = take(std::max<size_t>(nf, 1) * 4u), o_written = take((size_t)nb * 4u), o_gocc = take((size_t)nb * 64u),
                 o_org = take(std::max<size_t>(org_loaded.size(), 1) * sizeof(int2));
    if (ws > ctx->sync_ws_cap) {
        VF_HIP(ctx, hipStreamSynchronize(st));
        if (ctx->sync_ws) (void)hipFree(ctx->sync_ws);
        ctx->sync_ws = nullptr;
        ctx->sync_ws_cap = 0;
        const size_t want = ws + ws / 4;
        hipError_t e = hipMalloc((void **)&ctx->sync_ws, want);
        if (e != hipSuccess) return vf::fail(ctx, VF_E_OOM, "vf_world_sync_chunks: hipMalloc(%zu): %s", want, hipGetErrorString(e));
        ctx->sync_ws_cap = want;
    }
    uint8_t *wsb = ctx->sync_ws;
    uint8_t *d_stage = wsb + o_stage, *scan_tmp = wsb + o_scan, *d_bricks = d_stage;
    uint32_t *enc = (uint32_t *)(wsb + o_enc), *flag = (uint32_t *)(wsb + o_flag), *rank = (uint32_t *)(wsb + o_rank),
             *first = (uint32_t *)(wsb + o_first), *major = (uint32_t *)(wsb + o_major), *d_free = (uint32_t *)(wsb + o_free),
             *gocc = (uint32_t *)(wsb + o_gocc);
    int32_t *written = (int32_t *)(wsb + o_written);
    int2 *d_org = (int2 *)(wsb + o_org);
    int32_t *slot = reinterpret_cast<int32_t *>(d_stage + off_slot);
    uint32_t *d_cells = reinterpret_cast<uint32_t *>(d_stage + off_cells);
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
        VF_HIP(ctx, hipMemcpyAsync(d_stage, pb, stage, hipMemcpyHostToDevice, st));
        VF_HIP(ctx, hipMemcpyAsync(d_org, org_loaded.data(), org_loaded.size() * sizeof(int2), hipMemcpyHostToDevice, st));
        if (nf) VF_HIP(ctx, hipMemcpyAsync(d_free, ctx->free_slots.data(), nf * 4u, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_fill_i32, dim3(vf::blocks_for(nb)), dim3(256), 0, st, (int32_t *)first, -1, nb);
        hipLaunchKernelGGL(k_sync_first, dim3(vf::blocks_for(n)), dim3(256), 0, st, (const uint32_t *)d_cells, n,
                           (const int32_t *)slot, first);
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
        if (nl && total_new > nf && (uint64_t)A.used + (total_new - nf) > A.cap)
            hipLaunchKernelGGL(k_majority, dim3(nl), dim3(256), 0, st, (const uint8_t *)d_bricks, (const int32_t *)slot, major);
        hipLaunchKernelGGL(k_sync_encode, dim3(vf::blocks_for(n)), dim3(256), 0, st, A, (const int2 *)d_org, (const uint32_t *)d_cells,
                           n, (const int32_t *)slot, (const uint32_t *)major, ctx->grid, ctx->srcnz, enc);
        if (nl) {
            // every named brick with a slot is written (encode_brick's `needs upload`)
            hipLaunchKernelGGL(k_sync_written, dim3(vf::blocks_for(nb)), dim3(256), 0, st, (const int32_t *)slot, nb, written);
            const int64_t pieces = (int64_t)nl * 32, words = (int64_t)nl * 16;
            hipLaunchKernelGGL(k_scatter, dim3(vf::blocks_for(pieces)), dim3(256), 0, st, (const uint4 *)d_bricks, pieces,
                               (const int32_t *)written, (uint4 *)ctx->atlas);
            hipLaunchKernelGGL(k_brick_mask, dim3(vf::blocks_for(words)), dim3(256), 0, st, (const uint8_t *)d_bricks, words,
                               (const int32_t *)written, ctx->masks_ref);
            if ((rc = vf::masks_for(ctx, (const uint32_t *)written, nl))) return rc;   // -1 entries are skipped
            hipLaunchKernelGGL(k_sync_occ, dim3(vf::blocks_for(words)), dim3(256), 0, st, (const int32_t *)written, words,
                               (const uint32_t *)ctx->occ, gocc);
        }
    }
    // cleared chunks, coarse cells of every synced chunk: 64 chunks per launch
    int64_t li = 0;
    const int32_t c1per = (kChunkB / 4) * ctx->c1d[1] * (kChunkB / 4);
    for (size_t c0 = 0; c0 < list.size(); c0 += 64) {
        SyncChunks C;
        C.n = (int32_t)std::min<size_t>(64, list.size() - c0);
        bool any_clear = false;
        for (int32_t w = 0; w < C.n; ++w) {
            const vf_chunk *c = list[c0 + (size_t)w];
            C.o[w] = make_int2((c->cx * kChunkB) & (W.wb[0] - 1), (c->cz * kChunkB) & (W.wb[2] - 1));
            C.cells[w] = c->loaded ? (li++) * per : -1;
            any_clear |= !c->loaded;
        }
        const dim3 g1(vf::blocks_for(per), (unsigned)C.n), gc1(vf::blocks_for(c1per), (unsigned)C.n),
            gc2(vf::blocks_for(ctx->c2d[1]), (unsigned)C.n);
        if (any_clear) hipLaunchKernelGGL(k_sync_clear, g1, dim3(256), 0, st, A, C, ctx->grid, ctx->srcnz);
        hipLaunchKernelGGL(k_sync_coarse1, gc1, dim3(256), 0, st, A, C, (const uint32_t *)d_cells, ctx->c1d[0], ctx->c1d[1], ctx->c1);
        hipLaunchKernelGGL(k_sync_coarse2, gc2, dim3(256), 0, st, C, ctx->c1d[0], ctx->c1d[1], ctx->c1d[2], ctx->c2d[0],
                           ctx->c2d[1], (const uint8_t *)ctx->c1, ctx->c2);
    }
    if ((rc = vf::launch_check(ctx, "chunk sync"))) return rc;
    // the distance field over each chunk's box +-16 bricks (gpu_world.rs:1355-1361), then cell words and sky-top
    for (const vf_chunk *c : list) {
        const int32_t lo[3] = {c->cx * kChunkB, 0, c->cz * kChunkB};
        const int32_t hi[3] = {c->cx * kChunkB + kChunkB - 1, W.wb[1] - 1, c->cz * kChunkB + kChunkB - 1};
        const int32_t flo[3] = {lo[0] - 16, lo[1] - 16, lo[2] - 16}, fhi[3] = {hi[0] + 16, hi[1] + 16, hi[2] + 16};
        if ((rc = vf::distance_box(ctx, flo, fhi, lo, hi))) return rc;
    }
    if ((rc = vf::cell_words(ctx))) return rc;
    {   // record_grass_top over each chunk's columns (gpu_world.rs:1358): one ring slot each, so the windows are
        // disjoint and go in one launch
        std::vector<int32_t> wx, wz;
        for (const vf_chunk *c : list) {
            wx.push_back(((c->cx * kChunkB) & (W.wb[0] - 1)) * 8);
            wz.push_back(((c->cz * kChunkB) & (W.wb[2] - 1)) * 8);
        }
        if ((rc = vf::skytop_windows(ctx, wx.data(), wz.data(), (int32_t)wx.size(), kChunkB * 8, kChunkB * 8))) return rc;
    }
    // bookkeeping: the host grid copy, the slots popped and handed out, world-owned slots, the caller's gpu_slot
    // back through the staging block: slots over the brick voxels' place, occupancy rows after them
    int32_t *out = pslot;
    uint32_t *hocc = reinterpret_cast<uint32_t *>(pb);
    if (n) {
        VF_HIP(ctx, hipMemcpyAsync(pcells, enc, (size_t)n * 4u, hipMemcpyDeviceToHost, st));
        if (nl) {
            VF_HIP(ctx, hipMemcpyAsync(pslot, slot, (size_t)nl * 4u, hipMemcpyDeviceToHost, st));
            VF_HIP(ctx, hipMemcpyAsync(hocc, gocc, (size_t)nl * 64u, hipMemcpyDeviceToHost, st));
        }
    }
    VF_HIP(ctx, hipStreamSynchronize(st));
    li = 0;
    for (const vf_chunk *c : list) {
        const int2 o = make_int2((c->cx * kChunkB) & (W.wb[0] - 1), (c->cz * kChunkB) & (W.wb[2] - 1));
        const uint32_t *src = c->loaded ? pcells + (li++) * per : nullptr;
        for (int32_t z = 0; z < kChunkB; ++z)
            for (int32_t y = 0; y < W.wb[1]; ++y) {
                uint32_t *row = &ctx->host_grid[((size_t)(o.y + z) * W.wb[1] + y) * W.wb[0] + o.x];
                if (src) std::memcpy(row, src + ((size_t)z * W.wb[1] + y) * kChunkB, kChunkB * 4u);
                else std::memset(row, 0, kChunkB * 4u);
            }
    }
    const uint32_t pops = std::min<uint32_t>(total_new, (uint32_t)nf);
    ctx->free_slots.resize(nf - pops);
    ctx->atlas_used += total_new > nf ? std::min<uint32_t>(total_new - (uint32_t)nf, A.cap - A.used) : 0u;
    for (uint32_t l = 0; l < nl; ++l) {
        bv.set_slot(named[l], out[l] >= 0 ? out[l] : -1);
        if (out[l] < 0) continue;
        ctx->world_slot[(size_t)out[l]] = 1;
        std::memcpy(&ctx->host_occ[(size_t)out[l] * 8u], hocc + (size_t)l * 16u, 64);
    }
    return VF_OK;
}

// ------------------------------------------------------------------------------------------------
// Edit sync (GpuWorld::sync, gpu_world.rs:985-1142)
// ------------------------------------------------------------------------------------------------
namespace {

constexpr int32_t kDistCap = 16;   // DIST_CAP (gpu_world.rs:31)

// the dirty cells' encoded and source values
__global__ void k_edit_cells(const uint32_t *idx, const uint32_t *enc, const uint32_t *src, int64_t n, uint32_t *grid,
                             uint8_t *srcnz) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    grid[idx[i]] = enc[i];
    srcnz[idx[i]] = src[i] != 0u ? 1u : 0u;
}

// coarse_cell1 (gpu_world.rs:944-957) of the listed cells, over the source cells
__global__ void k_coarse1_list(const uint32_t *list, int64_t n, const uint8_t *srcnz, int32_t wbx, int32_t wby, int32_t wbz,
                               int32_t c1x, int32_t c1y, uint8_t *c1) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t ci = list[i];
    const int32_t cx = (int32_t)(ci % (uint32_t)c1x), cy = (int32_t)((ci / (uint32_t)c1x) % (uint32_t)c1y),
                  cz = (int32_t)(ci / ((uint32_t)c1x * (uint32_t)c1y));
    uint8_t occ = 0u;
    for (int32_t z = cz * 4; z < min(cz * 4 + 4, wbz) && !occ; ++z)
        for (int32_t y = cy * 4; y < min(cy * 4 + 4, wby) && !occ; ++y)
            for (int32_t x = cx * 4; x < min(cx * 4 + 4, wbx); ++x)
                if (srcnz[((size_t)z * wby + y) * wbx + x]) { occ = 1u; break; }
    c1[ci] = occ;
}

// coarse_cell2 (gpu_world.rs:959-973) of the listed cells, over coarse1
__global__ void k_coarse2_list(const uint32_t *list, int64_t n, const uint8_t *c1, int32_t c1x, int32_t c1y, int32_t c1z,
                               int32_t c2x, int32_t c2y, uint8_t *c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t ci = list[i];
    const int32_t cx = (int32_t)(ci % (uint32_t)c2x), cy = (int32_t)((ci / (uint32_t)c2x) % (uint32_t)c2y),
                  cz = (int32_t)(ci / ((uint32_t)c2x * (uint32_t)c2y));
    uint8_t occ = 0u;
    for (int32_t z = cz * 4; z < min(cz * 4 + 4, c1z) && !occ; ++z)
        for (int32_t y = cy * 4; y < min(cy * 4 + 4, c1y) && !occ; ++y)
            for (int32_t x = cx * 4; x < min(cx * 4 + 4, c1x); ++x)
                if (c1[((size_t)z * c1y + y) * c1x + x]) { occ = 1u; break; }
    c2[ci] = occ;
}

// encode_brick's full-atlas value (gpu_world.rs:202-213) on the host
uint32_t majority_host(const uint8_t *vox) {
    uint32_t count[256] = {0};
    for (int k = 0; k < 512; ++k) count[vox[k]]++;
    uint32_t best = 1;
    for (uint32_t m = 2; m < 256; ++m)
        if (count[m] > count[best] && m != kTorchFlame) best = m;
    return count[best] * 2u >= 512u ? best : 0u;
}

struct BrickBox {
    int32_t lo[3], hi[3];
};

}  // namespace

extern "C" int vf_world_sync_edits(vf_ctx *ctx, const uint32_t *cells, uint32_t *values, uint32_t n, const uint8_t *bricks,
                                   uint32_t n_bricks, uint32_t brick_stride, int32_t *gpu_slot, uint32_t slot_stride) {
    int rc = vf::check_ready(ctx, "vf_world_sync_edits");
    if (rc) return rc;
    if ((rc = vf::no_capture(ctx, "vf_world_sync_edits"))) return rc;
    if (!n) return VF_OK;
    if (!cells || !values || (n_bricks && (!bricks || !gpu_slot)))
        return vf::fail(ctx, VF_E_INVALID, "vf_world_sync_edits: null argument");
    BrickView bv;
    if ((rc = brick_view(ctx, "vf_world_sync_edits", bricks, brick_stride, gpu_slot, slot_stride, bv))) return rc;
    const DevWorld &W = ctx->W;
    for (uint32_t i = 0; i < n; ++i) {
        if (cells[i] >= ctx->n_cells) return vf::fail(ctx, VF_E_INVALID, "vf_world_sync_edits: cell %u of %zu", cells[i], ctx->n_cells);
        if (values[i] >= 256u && values[i] - 256u >= n_bricks)
            return vf::fail(ctx, VF_E_INVALID, "vf_world_sync_edits: cell %u names brick %u of %u", cells[i], values[i] - 256u, n_bricks);
    }
    // the dirty cells sorted, once each (a cell listed twice takes its last value)
    std::vector<uint32_t> ord(n);
    for (uint32_t i = 0; i < n; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return cells[a] < cells[b]; });
    std::vector<uint32_t> dirty, val;
    std::vector<uint32_t> pos_of(n);   // input entry -> dirty index
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t i = ord[k];
        if (dirty.empty() || dirty.back() != cells[i]) {
            dirty.push_back(cells[i]);
            val.push_back(values[i]);
        } else {
            val.back() = values[i];
        }
        pos_of[i] = (uint32_t)dirty.size() - 1u;
    }
    const size_t nd = dirty.size();
    // the sequential part (gpu_world.rs:997-1021), on copies of the bookkeeping: compaction of bricks
    // that became uniform (SparseVolume::compact_brick), then encode_brick with alloc_slot
    std::vector<uint32_t> fl = ctx->free_slots;
    uint32_t used = ctx->atlas_used;
    std::unordered_map<uint32_t, int32_t> slot_of;   // bricks whose slot changes
    const auto cur_slot = [&](uint32_t b) {
        auto it = slot_of.find(b);
        return it != slot_of.end() ? it->second : bv.get_slot(b);
    };
    std::vector<std::pair<uint32_t, uint8_t>> ws_ops;   // world_slot writes in order
    std::vector<uint32_t> enc(nd), up;
    for (size_t j = 0; j < nd; ++j) {
        if (val[j] >= 256u) {
            const uint32_t b = val[j] - 256u;
            const uint8_t *v = bv.at(b);
            bool uni = true;
            for (int k = 1; k < 512 && uni; ++k) uni = v[k] == v[0];
            if (uni) {
                const int32_t s = cur_slot(b);
                slot_of[b] = -1;
                val[j] = v[0];
                if (s >= 0) {
                    ws_ops.emplace_back((uint32_t)s, 0);
                    fl.push_back((uint32_t)s);
                }
            }
        }
        if (val[j] < 256u) {
            enc[j] = val[j];
            continue;
        }
        const uint32_t b = val[j] - 256u;
        int32_t s = cur_slot(b);
        if (s < 0) {
            if (!fl.empty()) {
                s = (int32_t)fl.back();
                fl.pop_back();
            } else if (used < ctx->n_slots) {
                s = (int32_t)used++;
            } else {
                s = -1;
            }
            slot_of[b] = s;
        }
        if (s >= 0) {
            ws_ops.emplace_back((uint32_t)s, 1);
            enc[j] = 256u + (uint32_t)s;
            up.push_back(b);
        } else {
            enc[j] = majority_host(bv.at(b));
        }
    }
    std::sort(up.begin(), up.end());
    up.erase(std::unique(up.begin(), up.end()), up.end());
    // coarse cells touched (gpu_world.rs:1033-1065)
    std::vector<uint32_t> c1l(nd), c2l(nd);
    const int32_t wbx = W.wb[0], wby = W.wb[1], wbz = W.wb[2];
    for (size_t j = 0; j < nd; ++j) {
        const int32_t g = (int32_t)dirty[j];
        const int32_t bx = g % wbx, by = (g / wbx) % wby, bz = g / (wbx * wby);
        c1l[j] = (uint32_t)(((bz / 4) * ctx->c1d[1] + by / 4) * ctx->c1d[0] + bx / 4);
        c2l[j] = (uint32_t)(((bz / 16) * ctx->c2d[1] + by / 16) * ctx->c2d[0] + bx / 16);
    }
    for (auto *l : {&c1l, &c2l}) {
        std::sort(l->begin(), l->end());
        l->erase(std::unique(l->begin(), l->end()), l->end());
    }
    // clusters of dirty bricks, absolute coordinates (gpu_world.rs:1103-1130)
    std::vector<BrickBox> boxes;
    for (size_t j = 0; j < nd; ++j) {
        const int32_t g = (int32_t)dirty[j];
        int32_t b[3] = {g % wbx, (g / wbx) % wby, g / (wbx * wby)};
        if (W.ring) {   // World::brick_coords: the ring slot's chunk inside the window
            b[0] = W.minb[0] + (((b[0] - W.minb[0]) % wbx) + wbx) % wbx;
            b[2] = W.minb[2] + (((b[2] - W.minb[2]) % wbz) + wbz) % wbz;
        }
        BrickBox *hit = nullptr;
        for (auto &x : boxes) {
            bool in = true;
            for (int k = 0; k < 3; ++k) in = in && b[k] >= x.lo[k] - kDistCap && b[k] <= x.hi[k] + kDistCap;
            if (in) { hit = &x; break; }
        }
        if (hit) {
            for (int k = 0; k < 3; ++k) { hit->lo[k] = std::min(hit->lo[k], b[k]); hit->hi[k] = std::max(hit->hi[k], b[k]); }
        } else {
            boxes.push_back(BrickBox{{b[0], b[1], b[2]}, {b[0], b[1], b[2]}});
        }
    }
    for (bool merged = true; merged;) {
        merged = false;
        for (size_t i = 0; i < boxes.size() && !merged; ++i)
            for (size_t j = i + 1; j < boxes.size(); ++j) {
                bool near = true;
                for (int k = 0; k < 3; ++k)
                    near = near && boxes[j].lo[k] <= boxes[i].hi[k] + 2 * kDistCap && boxes[j].hi[k] >= boxes[i].lo[k] - 2 * kDistCap;
                if (near) {
                    const BrickBox bj = boxes[j];
                    boxes.erase(boxes.begin() + (ptrdiff_t)j);
                    for (int k = 0; k < 3; ++k) { boxes[i].lo[k] = std::min(boxes[i].lo[k], bj.lo[k]); boxes[i].hi[k] = std::max(boxes[i].hi[k], bj.hi[k]); }
                    merged = true;
                    break;
                }
            }
    }
    // staging: dirty cells, their encoded and source values, the uploaded bricks and their slots, coarse lists
    const size_t nu = up.size(), nbk = std::max<size_t>(nu, 1);
    size_t hs = 0;
    const auto htake = [&hs](size_t bytes) { const size_t o = hs; hs += (std::max<size_t>(bytes, 1) + 255u) & ~(size_t)255u; return o; };
    const size_t h_br = htake(nu * 512u), h_sl = htake(nbk * 4u), h_idx = htake(nd * 4u), h_enc = htake(nd * 4u), h_src = htake(nd * 4u),
                 h_c1 = htake(c1l.size() * 4u), h_c2 = htake(c2l.size() * 4u), h_occ = htake(nbk * 64u);
    if (hs > ctx->pin_cap) {
        if (ctx->pin) (void)hipHostFree(ctx->pin);
        ctx->pin = nullptr;
        ctx->pin_cap = 0;
        const size_t want = std::max<size_t>(hs + hs / 4, (size_t)1 << 20);
        if (hipHostMalloc((void **)&ctx->pin, want, hipHostMallocDefault) != hipSuccess)
            return vf::fail(ctx, VF_E_OOM, "vf_world_sync_edits: %zu bytes of pinned staging", want);
        ctx->pin_cap = want;
    }
    if (hs > ctx->sync_ws_cap) {
        VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (ctx->sync_ws) (void)hipFree(ctx->sync_ws);
        ctx->sync_ws = nullptr;
        ctx->sync_ws_cap = 0;
        const size_t want = hs + hs / 4;
        hipError_t e = hipMalloc((void **)&ctx->sync_ws, want);
        if (e != hipSuccess) return vf::fail(ctx, VF_E_OOM, "vf_world_sync_edits: hipMalloc(%zu): %s", want, hipGetErrorString(e));
        ctx->sync_ws_cap = want;
    }
    uint8_t *pb = ctx->pin, *db = ctx->sync_ws;
    int32_t *pslot = (int32_t *)(pb + h_sl);
    gather_bricks(bv, up.data(), nu, pb + h_br);
    for (size_t u = 0; u < nu; ++u) pslot[u] = cur_slot(up[u]);
    std::memcpy(pb + h_idx, dirty.data(), nd * 4u);
    std::memcpy(pb + h_enc, enc.data(), nd * 4u);
    std::memcpy(pb + h_src, val.data(), nd * 4u);
    std::memcpy(pb + h_c1, c1l.data(), c1l.size() * 4u);
    std::memcpy(pb + h_c2, c2l.data(), c2l.size() * 4u);
    hipStream_t st = ctx->stream;
    VF_HIP(ctx, hipMemcpyAsync(db, pb, h_occ, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_edit_cells, dim3(vf::blocks_for((int64_t)nd)), dim3(256), 0, st, (const uint32_t *)(db + h_idx),
                       (const uint32_t *)(db + h_enc), (const uint32_t *)(db + h_src), (int64_t)nd, ctx->grid, ctx->srcnz);
    if (nu) {
        const int32_t *dsl = (const int32_t *)(db + h_sl);
        hipLaunchKernelGGL(k_scatter, dim3(vf::blocks_for((int64_t)nu * 32)), dim3(256), 0, st, (const uint4 *)(db + h_br),
                           (int64_t)nu * 32, dsl, (uint4 *)ctx->atlas);
        hipLaunchKernelGGL(k_brick_mask, dim3(vf::blocks_for((int64_t)nu * 16)), dim3(256), 0, st, (const uint8_t *)(db + h_br),
                           (int64_t)nu * 16, dsl, ctx->masks_ref);
        if ((rc = vf::masks_for(ctx, (const uint32_t *)dsl, (uint32_t)nu))) return rc;
        hipLaunchKernelGGL(k_sync_occ, dim3(vf::blocks_for((int64_t)nu * 16)), dim3(256), 0, st, dsl, (int64_t)nu * 16,
                           (const uint32_t *)ctx->occ, (uint32_t *)(db + h_occ));
    }
    hipLaunchKernelGGL(k_coarse1_list, dim3(vf::blocks_for((int64_t)c1l.size())), dim3(256), 0, st, (const uint32_t *)(db + h_c1),
                       (int64_t)c1l.size(), (const uint8_t *)ctx->srcnz, wbx, wby, wbz, ctx->c1d[0], ctx->c1d[1], ctx->c1);
    hipLaunchKernelGGL(k_coarse2_list, dim3(vf::blocks_for((int64_t)c2l.size())), dim3(256), 0, st, (const uint32_t *)(db + h_c2),
                       (int64_t)c2l.size(), (const uint8_t *)ctx->c1, ctx->c1d[0], ctx->c1d[1], ctx->c1d[2], ctx->c2d[0], ctx->c2d[1],
                       ctx->c2);
    if ((rc = vf::launch_check(ctx, "edit sync"))) return rc;
    for (const BrickBox &x : boxes) {   // grass tops, then the distance field, per cluster (gpu_world.rs:1133-1141)
        if ((rc = vf::skytop_box(ctx, x.lo[0] * 8, x.lo[2] * 8, x.hi[0] * 8 + 7, x.hi[2] * 8 + 7))) return rc;
        const int32_t lo[3] = {x.lo[0] - kDistCap, x.lo[1] - kDistCap, x.lo[2] - kDistCap};
        const int32_t hi[3] = {x.hi[0] + kDistCap, x.hi[1] + kDistCap, x.hi[2] + kDistCap};
        if ((rc = vf::distance_box(ctx, lo, hi, x.lo, x.hi))) return rc;
    }
    if ((rc = vf::cell_words(ctx))) return rc;
    if (nu) VF_HIP(ctx, hipMemcpyAsync(pb + h_occ, db + h_occ, nu * 64u, hipMemcpyDeviceToHost, st));
    VF_HIP(ctx, hipStreamSynchronize(st));
    // commit the bookkeeping
    ctx->free_slots.swap(fl);
    ctx->atlas_used = used;
    for (const auto &op : ws_ops) ctx->world_slot[op.first] = op.second;
    for (size_t j = 0; j < nd; ++j) ctx->host_grid[dirty[j]] = enc[j];
    for (size_t u = 0; u < nu; ++u) std::memcpy(&ctx->host_occ[(size_t)pslot[u] * 8u], pb + h_occ + u * 64u, 64);
    for (const auto &kv : slot_of) bv.set_slot(kv.first, kv.second);
    for (uint32_t i = 0; i < n; ++i) values[i] = val[pos_of[i]];
    return VF_OK;
}
