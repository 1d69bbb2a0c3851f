// The context object behind the C ABI's opaque vf_ctx (host side, shared by the library's .hip files).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "vf_bodies.h"
#include "vf_device.h"
#include "vf_internal.h"

#define VF_HIP(ctx, call)                                                                                         \
    do {                                                                                                          \
        hipError_t e_ = (call);                                                                                   \
        if (e_ != hipSuccess)                                                                                     \
            return vf::fail((ctx), e_ == hipErrorOutOfMemory ? VF_E_OOM : VF_E_DEVICE, "%s failed: %s (%s:%d)", #call, \
                            hipGetErrorString(e_), __FILE__, __LINE__);                                           \
    } while (0)

struct vf_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    bool has_world = false;
    bool has_scene = false;
    DevWorld W{};
    SkyParams S{};
    vf_scene_ubo scene{};
    // owned device buffers
    uint32_t *grid = nullptr;
    uint2 *cells = nullptr;
    uint8_t *atlas = nullptr;
    uint32_t *occ = nullptr, *occdry = nullptr, *masks_ref = nullptr;
    uint8_t *dist[4] = {nullptr, nullptr, nullptr, nullptr};
    uint8_t *dist_tmp = nullptr;   // the normal field's x+y values of a two-launch build (not a reference field)
    uint8_t *dry_any = nullptr;    // per atlas slot: 1 if the brick holds a voxel that is neither air nor water
    uint8_t *c1 = nullptr, *c2 = nullptr;
    uint8_t *srcnz = nullptr;   // per cell: the caller's (source) cell is not air; the coarse pyramid's input
    uint16_t *skytop = nullptr;
    float4 *palette = nullptr;
    int32_t c1d[3] = {0, 0, 0}, c2d[3] = {0, 0, 0};
    uint32_t n_slots = 0;
    std::vector<uint32_t> host_grid;   // host copies of the grid and of the slots' occupancy bits
    std::vector<uint64_t> host_occ;    // (8 words per slot): classify the camera voxel without a device read
    size_t n_cells = 0;
    // the atlas bookkeeping of GpuWorld (gpu_world.rs:146-186): slots handed out so far, the free list
    // (popped from the back), which slots hold world bricks (only those are freed by chunk syncs)
    uint32_t atlas_used = 0;
    std::vector<uint32_t> free_slots;
    std::vector<uint8_t> world_slot;
    uint8_t *pin = nullptr;     // pinned host staging for chunk syncs (hipHostMalloc), grown on demand
    size_t pin_cap = 0;
    uint8_t *sync_ws = nullptr;  // device workspace of chunk syncs, grown on demand
    size_t sync_ws_cap = 0;
    bool counting = false;
    unsigned long long *d_cnt = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
    bool timing = false;   // vf_set_timing: record ev0 / ev1 around each trace entry point
    bool capturing = false;   // inside vf_graph_begin / vf_graph_end: no syncs, no timing events
    int n_cu = 256;           // compute units (persistent grids)
    int32_t traversal = 0;    // vf_set_traversal: 0 by launch size, 1 the plain traversal, 2 the small-frame one
    long long wall_khz = 100000;   // wall clock rate (batch completion stamps)
    // persistent-queue buffers (vf_queue.hip), grown on first use outside a capture
    unsigned long long *q_ctr = nullptr, *q_stamps = nullptr;
    float4 *q_rays = nullptr, *q_sun = nullptr;
    int64_t q_cap = 0;
    // the allocation-free batch's counters and stamps: this device's copy of a file-scope __device__ array
    // (vf_queue.hip), its address taken once per device at context creation; never freed
    unsigned long long *batch_scratch = nullptr;
    const unsigned long long *batch_stamps = nullptr;   // where the last batch on this context stamped
    // bodies (vf_bodies.hip): repacked headers, grids, BVH; grown, never shrunk, so frame graphs that
    // recorded them stay valid while the body count is unchanged
    int32_t n_bodies = 0;
    float4 *b_hdr = nullptr, *b_nodes = nullptr;
    int32_t *b_esc = nullptr;
    uint32_t *b_grid = nullptr, *b_prims = nullptr;
    size_t b_hdr_cap = 0, b_grid_cap = 0, b_node_cap = 0, b_esc_cap = 0, b_prim_cap = 0;   // elements
    float4 *b_prev = nullptr;                 // 3 float4 per body: rows of prevObjToWorld (motion vectors)
    size_t b_prev_cap = 0;
    uint32_t b_slot_end = 0;                  // one past the highest atlas slot the body grids name
    std::vector<uint32_t> b_host_grid;        // host copy of the body grids
    std::vector<vf_bvh_node> b_host_nodes;    // the BVH in use
    std::vector<uint32_t> b_host_prims;
};

struct vf_graph {
    int device = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace vf {
// helpers defined in vf_trace.hip
int set_device(vf_ctx *ctx);
int launch_check(vf_ctx *ctx, const char *what);
int begin_timed(vf_ctx *ctx);
int end_timed(vf_ctx *ctx);
int check_ready(vf_ctx *ctx, const char *what);
// the host's grid copy shows world position pos in air (or outside the loaded box): a camera there needs
// no per-ray camera-voxel lookup (vf_trace.hip)
bool camera_in_air(const vf_ctx *ctx, const float pos[3]);
int no_capture(vf_ctx *ctx, const char *what);
void free_queue(vf_ctx *ctx);
int queue_scratch_init(vf_ctx *ctx);   // vf_queue.hip: the device address of the batch scratch (device current)
inline unsigned blocks_for(int64_t n, int per = 256) { return (unsigned)((n + per - 1) / per); }
// The small-frame traversal for a launch of `waves` wavefronts: by the context's setting, or, by default,
// for launches of at most 16 waves per SIMD (4 SIMDs per compute unit), whose time is that of their
// slowest waves rather than the chip's issue rate.
inline bool small_frame(const vf_ctx *ctx, int64_t waves) {
    if (ctx->counting || ctx->traversal == 1) return false;
    return ctx->traversal == 2 || waves <= (int64_t)ctx->n_cu * 4 * 16;
}
// bodies (vf_bodies.hip): the view the kernels get (n = 0 unless the scene's num_objects is set), and the
// launch-time check that the scene's num_objects matches the uploaded bodies and their bricks exist
DevBodies dev_bodies(const vf_ctx *ctx);
int check_bodies(vf_ctx *ctx, const char *what);
void free_bodies(vf_ctx *ctx);
void slot_occupancy(const uint8_t *vox, uint64_t out[8]);
void host_occupancy(std::vector<uint64_t> &occ, const uint8_t *atlas, uint32_t n_slots);
// the world's buffers and derived data (vf_trace.hip), shared by vf_world_upload and vf_world_pack
int world_prepare(vf_ctx *ctx, const int32_t wb[3], const int32_t win[4], uint32_t n_slots, const char *what);
int world_finish(vf_ctx *ctx, bool coarse_given, bool source_given = false);
int build_coarse(vf_ctx *ctx, const uint32_t *d_grid);
int masks_for(vf_ctx *ctx, const uint32_t *d_slots, uint32_t n);
int distance_box(vf_ctx *ctx, const int32_t lo[3], const int32_t hi[3], const int32_t dlo[3], const int32_t dhi[3]);
int refresh_world(vf_ctx *ctx);
int skytop_columns(vf_ctx *ctx, int32_t x0, int32_t z0, int32_t nx, int32_t nz);
int skytop_windows(vf_ctx *ctx, const int32_t *x0, const int32_t *z0, int32_t n, int32_t nx, int32_t nz);
int cell_words(vf_ctx *ctx);
int skytop_box(vf_ctx *ctx, int32_t x0, int32_t z0, int32_t x1, int32_t z1);
}  // namespace vf
