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
