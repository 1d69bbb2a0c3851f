// Bodies (dynamic voxel objects) on the host side of the C ABI: their headers repacked for the
// kernels, their brick grids, and the BVH over their world boxes (build_bvh, reference
// src/gfx/gpu_world.rs:550-600) with the per-node continuation links that let the kernels walk it in
// the reference's order without a stack (shaders/common.glsl:693-716).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/vf_trace.h"
#include "vf_ctx.h"
#include "vf_internal.h"

namespace {

constexpr int32_t kMaxBodies = 1 << 20;
constexpr int32_t kMaxBodyBricks = 1024;   // per axis
constexpr int32_t kStackDepth = 24;        // the reference walk's stack (common.glsl:697)

float bits_f(int32_t v) {
    float f;
    std::memcpy(&f, &v, 4);
    return f;
}
float bits_f(uint32_t v) {
    float f;
    std::memcpy(&f, &v, 4);
    return f;
}

// The kernels' 128-byte header: worldToObj rows (with the translation in .w), objToWorld columns (with
// the brick dims in .w), the world box (with the grid offset in .w).  Matrices are column-major:
// m[c][r] = a[4c + r].
void body_header(const vf_obj_header &h, float4 out[kBodyWords]) {
    const float *w = h.world_to_obj, *o = h.obj_to_world;
    for (int r = 0; r < 3; ++r) out[r] = make_float4(w[r], w[4 + r], w[8 + r], w[12 + r]);
    for (int c = 0; c < 3; ++c) out[3 + c] = make_float4(o[4 * c], o[4 * c + 1], o[4 * c + 2], bits_f(h.dims_bricks[c]));
    out[6] = make_float4(h.aabb_min[0], h.aabb_min[1], h.aabb_min[2], bits_f(h.dims_bricks[3]));
    out[7] = make_float4(h.aabb_max[0], h.aabb_max[1], h.aabb_max[2], 0.0f);
}

int check_headers(vf_ctx *ctx, const char *what, const vf_obj_header *hd, int32_t n, int64_t grid_len) {
    for (int32_t i = 0; i < n; ++i) {
        const int32_t *d = hd[i].dims_bricks;
        for (int a = 0; a < 3; ++a)
            if (d[a] <= 0 || d[a] > kMaxBodyBricks)
                return vf::fail(ctx, VF_E_INVALID, "%s: body %d has %d bricks along axis %d (1..%d)", what, i, d[a], a,
                                kMaxBodyBricks);
        const int64_t cells = (int64_t)d[0] * d[1] * d[2];
        if (d[3] < 0 || (int64_t)d[3] + cells > grid_len)
            return vf::fail(ctx, VF_E_INVALID, "%s: body %d's grid [%d, +%lld) is outside the %lld grid entries", what, i, d[3],
                            (long long)cells, (long long)grid_len);
    }
    return VF_OK;
}

// build_bvh (gpu_world.rs:550-600): node = box of its bodies; at most two bodies make a leaf; otherwise
// split at the median of the box centroids (min + max) along the axis of their widest extent (ties to
// the lower axis), the two children appended as a pair and the left subtree built before the right.
// The median selection orders a range by (centroid, body index): one of the partitions the reference's
// unstable selection may produce, and the same every time.
void build_node(const std::vector<std::array<float, 6>> &bb, uint32_t *prims, uint32_t count, std::vector<vf_bvh_node> &nodes,
                size_t n, uint32_t first) {
    float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
    for (uint32_t i = 0; i < count; ++i)
        for (int k = 0; k < 3; ++k) {
            mn[k] = std::min(mn[k], bb[prims[i]][k]);
            mx[k] = std::max(mx[k], bb[prims[i]][3 + k]);
        }
    vf_bvh_node &nd = nodes[n];
    for (int k = 0; k < 3; ++k) { nd.min[k] = mn[k]; nd.max[k] = mx[k]; }
    if (count <= 2) {
        nd.first = first;
        nd.count = count;
        return;
    }
    auto cen = [&](uint32_t p, int k) { return bb[p][k] + bb[p][3 + k]; };
    float cmn[3] = {1e30f, 1e30f, 1e30f}, cmx[3] = {-1e30f, -1e30f, -1e30f};
    for (uint32_t i = 0; i < count; ++i)
        for (int k = 0; k < 3; ++k) {
            cmn[k] = std::min(cmn[k], cen(prims[i], k));
            cmx[k] = std::max(cmx[k], cen(prims[i], k));
        }
    int ax = 0;
    for (int k = 1; k < 3; ++k)
        if (cmx[k] - cmn[k] > cmx[ax] - cmn[ax]) ax = k;
    const uint32_t mid = count / 2;
    std::sort(prims, prims + count, [&](uint32_t a, uint32_t b) {
        const float ca = cen(a, ax), cb = cen(b, ax);
        return ca < cb || (!(cb < ca) && a < b);
    });
    const size_t left = nodes.size();
    nodes.push_back(vf_bvh_node{});
    nodes.push_back(vf_bvh_node{});
    nodes[n].first = (uint32_t)left;
    nodes[n].count = 0;
    build_node(bb, prims, mid, nodes, left, first);
    build_node(bb, prims + mid, count - mid, nodes, left + 1, first + mid);
}

void build_bvh(const vf_obj_header *hd, int32_t n, std::vector<vf_bvh_node> &nodes, std::vector<uint32_t> &prims) {
    nodes.clear();
    prims.clear();
    if (n <= 0) return;
    std::vector<std::array<float, 6>> bb((size_t)n);
    for (int32_t i = 0; i < n; ++i)
        bb[i] = {hd[i].aabb_min[0], hd[i].aabb_min[1], hd[i].aabb_min[2], hd[i].aabb_max[0], hd[i].aabb_max[1], hd[i].aabb_max[2]};
    prims.resize((size_t)n);
    for (int32_t i = 0; i < n; ++i) prims[i] = (uint32_t)i;
    nodes.push_back(vf_bvh_node{});
    build_node(bb, prims.data(), (uint32_t)n, nodes, 0, 0);
}

// The continuation of every node reachable from the root: the node the reference's stack holds on top
// once the node's subtree is done (-1: empty, the walk ends).  A first child continues with its
// sibling, a second child with its parent's continuation.  Refuses what is not a tree over the node
// array, prim ranges or body indices out of range, and trees whose walk would overflow the reference's
// 24-entry stack (the reference then drops subtrees).
int link_bvh(vf_ctx *ctx, const char *what, const std::vector<vf_bvh_node> &nodes, const std::vector<uint32_t> &prims,
             int32_t n_bodies, std::vector<int32_t> &esc) {
    esc.assign(nodes.size(), -1);
    if (n_bodies == 0) return VF_OK;
    if (nodes.empty()) return vf::fail(ctx, VF_E_INVALID, "%s: no BVH nodes for %d bodies", what, n_bodies);
    std::vector<uint8_t> seen(nodes.size(), 0);
    struct Item { uint32_t node; int32_t cont; int32_t depth; };
    std::vector<Item> todo{{0u, -1, 0}};
    while (!todo.empty()) {
        const Item it = todo.back();
        todo.pop_back();
        if (seen[it.node]) return vf::fail(ctx, VF_E_INVALID, "%s: BVH node %u is reached twice", what, it.node);
        seen[it.node] = 1;
        esc[it.node] = it.cont;
        const vf_bvh_node &nd = nodes[it.node];
        if (nd.count == 0u) {
            if ((uint64_t)nd.first + 1u >= nodes.size())
                return vf::fail(ctx, VF_E_INVALID, "%s: BVH node %u's children %u, %u are out of range", what, it.node, nd.first,
                                nd.first + 1u);
            if (it.depth >= kStackDepth)
                return vf::fail(ctx, VF_E_INVALID, "%s: BVH deeper than the traversal's %d-entry stack", what, kStackDepth);
            todo.push_back({nd.first + 1u, it.cont, it.depth});
            todo.push_back({nd.first, (int32_t)(nd.first + 1u), it.depth + 1});
        } else {
            if ((uint64_t)nd.first + nd.count > prims.size())
                return vf::fail(ctx, VF_E_INVALID, "%s: BVH leaf %u's prims [%u, +%u) are out of range", what, it.node, nd.first,
                                nd.count);
            for (uint32_t k = 0; k < nd.count; ++k)
                if (prims[nd.first + k] >= (uint32_t)n_bodies)
                    return vf::fail(ctx, VF_E_INVALID, "%s: BVH prim %u names body %u of %d", what, nd.first + k,
                                    prims[nd.first + k], n_bodies);
        }
    }
    return VF_OK;
}

template <typename T>
int grow(vf_ctx *ctx, const char *what, T *&ptr, size_t &cap, size_t need) {
    if (need <= cap && ptr) return VF_OK;
    need = std::max<size_t>(need, 1);
    VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
    hipError_t e = hipMalloc((void **)&ptr, need * sizeof(T));
    if (e != hipSuccess) {
        ptr = nullptr;
        return vf::fail(ctx, VF_E_OOM, "%s: hipMalloc(%zu): %s", what, need * sizeof(T), hipGetErrorString(e));
    }
    cap = need;
    return VF_OK;
}

// One past the highest atlas slot the body grids name (0: none).
uint32_t slot_end(const std::vector<uint32_t> &grid) {
    uint32_t e = 0;
    for (uint32_t g : grid)
        if (g >= 256u) e = std::max(e, g - 255u);
    return e;
}

// The BVH and its links to the device.
int write_bvh(vf_ctx *ctx, const char *what, std::vector<vf_bvh_node> nodes, std::vector<uint32_t> prims,
              const std::vector<int32_t> &esc) {
    int rc;
    if ((rc = grow(ctx, what, ctx->b_esc, ctx->b_esc_cap, esc.size()))) return rc;
    if ((rc = grow(ctx, what, ctx->b_nodes, ctx->b_node_cap, nodes.size() * 2))) return rc;
    if ((rc = grow(ctx, what, ctx->b_prims, ctx->b_prim_cap, prims.size()))) return rc;
    std::vector<float4> packed(nodes.size() * 2);
    for (size_t i = 0; i < nodes.size(); ++i) {
        const vf_bvh_node &nd = nodes[i];
        packed[2 * i] = make_float4(nd.min[0], nd.min[1], nd.min[2], bits_f(nd.first));
        packed[2 * i + 1] = make_float4(nd.max[0], nd.max[1], nd.max[2], bits_f(nd.count));
    }
    if (!packed.empty())
        VF_HIP(ctx, hipMemcpyAsync(ctx->b_nodes, packed.data(), packed.size() * sizeof(float4), hipMemcpyHostToDevice, ctx->stream));
    if (!esc.empty())
        VF_HIP(ctx, hipMemcpyAsync(ctx->b_esc, esc.data(), esc.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    if (!prims.empty())
        VF_HIP(ctx, hipMemcpyAsync(ctx->b_prims, prims.data(), prims.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->b_host_nodes = std::move(nodes);
    ctx->b_host_prims = std::move(prims);
    return VF_OK;
}

int write_headers(vf_ctx *ctx, const char *what, const vf_obj_header *hd, int32_t n) {
    std::vector<float4> packed((size_t)n * kBodyWords);
    for (int32_t i = 0; i < n; ++i) body_header(hd[i], &packed[(size_t)i * kBodyWords]);
    // rows of prevObjToWorld's top 3x4 (the motion vectors' previous pose, gbuffer.comp:84-88)
    std::vector<float4> prev((size_t)n * 3);
    for (int32_t i = 0; i < n; ++i)
        for (int r = 0; r < 3; ++r) {
            const float *m = hd[i].prev_obj_to_world;
            prev[(size_t)i * 3 + r] = make_float4(m[r], m[4 + r], m[8 + r], m[12 + r]);
        }
    int rc;
    if ((rc = grow(ctx, what, ctx->b_hdr, ctx->b_hdr_cap, packed.size()))) return rc;
    if ((rc = grow(ctx, what, ctx->b_prev, ctx->b_prev_cap, prev.size()))) return rc;
    if (!packed.empty())
        VF_HIP(ctx, hipMemcpyAsync(ctx->b_hdr, packed.data(), packed.size() * sizeof(float4), hipMemcpyHostToDevice, ctx->stream));
    if (!prev.empty())
        VF_HIP(ctx, hipMemcpyAsync(ctx->b_prev, prev.data(), prev.size() * sizeof(float4), hipMemcpyHostToDevice, ctx->stream));
    VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return VF_OK;
}

int bodies_call(vf_ctx *ctx, const char *what) {
    if (!ctx) return vf::fail(nullptr, VF_E_INVALID, "%s: null context", what);
    if (int rc = vf::no_capture(ctx, what)) return rc;
    return vf::set_device(ctx);
}

}  // namespace

namespace vf {

DevBodies dev_bodies(const vf_ctx *ctx) {
    DevBodies B{};
    B.n = ctx->scene.num_objects > 0 ? ctx->n_bodies : 0;
    B.hdr = ctx->b_hdr;
    B.grid = ctx->b_grid;
    B.nodes = ctx->b_nodes;
    B.esc = ctx->b_esc;
    B.prims = ctx->b_prims;
    return B;
}

int check_bodies(vf_ctx *ctx, const char *what) {
    const int32_t k = ctx->scene.num_objects;
    if (k == 0) return VF_OK;
    if (k != ctx->n_bodies)
        return fail(ctx, VF_E_INVALID, "%s: the scene has num_objects = %d but %d bodies are uploaded", what, k, ctx->n_bodies);
    if (ctx->b_slot_end > ctx->n_slots)
        return fail(ctx, VF_E_INVALID, "%s: a body brick names atlas slot %u of %u", what, ctx->b_slot_end - 1u, ctx->n_slots);
    return VF_OK;
}

void free_bodies(vf_ctx *ctx) {
    for (void *p : {(void *)ctx->b_hdr, (void *)ctx->b_nodes, (void *)ctx->b_esc, (void *)ctx->b_grid, (void *)ctx->b_prims,
                    (void *)ctx->b_prev})
        if (p) (void)hipFree(p);
    ctx->b_hdr = ctx->b_nodes = ctx->b_prev = nullptr;
    ctx->b_prev_cap = 0;
    ctx->b_esc = nullptr;
    ctx->b_grid = ctx->b_prims = nullptr;
    ctx->b_hdr_cap = ctx->b_grid_cap = ctx->b_node_cap = ctx->b_esc_cap = ctx->b_prim_cap = 0;
    ctx->n_bodies = 0;
    ctx->b_slot_end = 0;
    ctx->b_host_grid.clear();
    ctx->b_host_nodes.clear();
    ctx->b_host_prims.clear();
}

}  // namespace vf

extern "C" {

int vf_objects_upload(vf_ctx *ctx, const vf_obj_header *hd, int32_t n, const uint32_t *grid, int64_t grid_len) {
    const char *what = "vf_objects_upload";
    int rc = bodies_call(ctx, what);
    if (rc) return rc;
    if (n < 0 || n > kMaxBodies) return vf::fail(ctx, VF_E_INVALID, "%s: %d bodies (0..%d)", what, n, kMaxBodies);
    if (n == 0) {
        VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
        vf::free_bodies(ctx);
        return VF_OK;
    }
    if (!hd || !grid) return vf::fail(ctx, VF_E_INVALID, "%s: null headers or grid", what);
    if (grid_len <= 0 || grid_len > (1ll << 31)) return vf::fail(ctx, VF_E_INVALID, "%s: grid length %lld", what, (long long)grid_len);
    if ((rc = check_headers(ctx, what, hd, n, grid_len))) return rc;
    std::vector<vf_bvh_node> nodes;
    std::vector<uint32_t> prims;
    std::vector<int32_t> esc;
    build_bvh(hd, n, nodes, prims);
    if ((rc = link_bvh(ctx, what, nodes, prims, n, esc))) return rc;
    VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->n_bodies = 0;   // until everything is on the device
    std::vector<uint32_t> g(grid, grid + grid_len);
    if ((rc = grow(ctx, what, ctx->b_grid, ctx->b_grid_cap, g.size()))) return rc;
    VF_HIP(ctx, hipMemcpyAsync(ctx->b_grid, g.data(), g.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    if ((rc = write_headers(ctx, what, hd, n))) return rc;
    if ((rc = write_bvh(ctx, what, std::move(nodes), std::move(prims), esc))) return rc;
    ctx->b_slot_end = slot_end(g);
    ctx->b_host_grid = std::move(g);
    ctx->n_bodies = n;
    return VF_OK;
}

int vf_objects_set_headers(vf_ctx *ctx, const vf_obj_header *hd, int32_t n) {
    const char *what = "vf_objects_set_headers";
    int rc = bodies_call(ctx, what);
    if (rc) return rc;
    if (n != ctx->n_bodies) return vf::fail(ctx, VF_E_INVALID, "%s: %d headers for %d bodies", what, n, ctx->n_bodies);
    if (n == 0) return VF_OK;
    if (!hd) return vf::fail(ctx, VF_E_INVALID, "%s: null headers", what);
    if ((rc = check_headers(ctx, what, hd, n, (int64_t)ctx->b_host_grid.size()))) return rc;
    std::vector<vf_bvh_node> nodes;
    std::vector<uint32_t> prims;
    std::vector<int32_t> esc;
    build_bvh(hd, n, nodes, prims);
    if ((rc = link_bvh(ctx, what, nodes, prims, n, esc))) return rc;
    if ((rc = write_headers(ctx, what, hd, n))) return rc;
    return write_bvh(ctx, what, std::move(nodes), std::move(prims), esc);
}

int vf_objects_set_bvh(vf_ctx *ctx, const vf_bvh_node *nodes, int32_t n_nodes, const uint32_t *prims, int32_t n_prims) {
    const char *what = "vf_objects_set_bvh";
    int rc = bodies_call(ctx, what);
    if (rc) return rc;
    if (ctx->n_bodies == 0) return vf::fail(ctx, VF_E_INVALID, "%s: no bodies uploaded", what);
    if (n_nodes <= 0 || n_prims <= 0 || n_nodes > 4 * kMaxBodies || n_prims > 4 * kMaxBodies || !nodes || !prims)
        return vf::fail(ctx, VF_E_INVALID, "%s: %d nodes, %d prims", what, n_nodes, n_prims);
    std::vector<vf_bvh_node> nv(nodes, nodes + n_nodes);
    std::vector<uint32_t> pv(prims, prims + n_prims);
    std::vector<int32_t> esc;
    if ((rc = link_bvh(ctx, what, nv, pv, ctx->n_bodies, esc))) return rc;
    return write_bvh(ctx, what, std::move(nv), std::move(pv), esc);
}

int vf_objects_update_grid(vf_ctx *ctx, const uint32_t *idx, const uint32_t *val, int32_t n) {
    const char *what = "vf_objects_update_grid";
    int rc = bodies_call(ctx, what);
    if (rc) return rc;
    if (n < 0 || (n && (!idx || !val))) return vf::fail(ctx, VF_E_INVALID, "%s: bad arrays", what);
    for (int32_t i = 0; i < n; ++i)
        if (idx[i] >= ctx->b_host_grid.size())
            return vf::fail(ctx, VF_E_INVALID, "%s: entry %u of %zu", what, idx[i], ctx->b_host_grid.size());
    if (n == 0) return VF_OK;
    for (int32_t i = 0; i < n; ++i) ctx->b_host_grid[idx[i]] = val[i];
    VF_HIP(ctx, hipMemcpyAsync(ctx->b_grid, ctx->b_host_grid.data(), ctx->b_host_grid.size() * 4, hipMemcpyHostToDevice,
                               ctx->stream));
    VF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ctx->b_slot_end = slot_end(ctx->b_host_grid);
    return VF_OK;
}

int vf_objects_read_bvh(vf_ctx *ctx, vf_bvh_node *nodes, int32_t max_nodes, int32_t *n_nodes, uint32_t *prims, int32_t max_prims,
                        int32_t *n_prims) {
    if (!ctx || !n_nodes || !n_prims) return vf::fail(ctx, VF_E_INVALID, "vf_objects_read_bvh: null argument");
    *n_nodes = (int32_t)ctx->b_host_nodes.size();
    *n_prims = (int32_t)ctx->b_host_prims.size();
    if ((nodes && max_nodes < *n_nodes) || (prims && max_prims < *n_prims))
        return vf::fail(ctx, VF_E_INVALID, "vf_objects_read_bvh: arrays too small (%d nodes, %d prims)", *n_nodes, *n_prims);
    if (nodes && *n_nodes) std::memcpy(nodes, ctx->b_host_nodes.data(), ctx->b_host_nodes.size() * sizeof(vf_bvh_node));
    if (prims && *n_prims) std::memcpy(prims, ctx->b_host_prims.data(), ctx->b_host_prims.size() * 4);
    return VF_OK;
}

}  // extern "C"
