// Synthetic worlds for the traversal workloads, packed the way GpuWorld::upload packs the reference
// world on the CPU (src/gfx/gpu_world.rs:243-313, 364-382): grid encoding 0 / 1-255 / 256 + slot,
// slots in grid-index order, atlas in slot order (Brick::vox order lz*64 + ly*8 + lx), 64-byte
// brick masks (src/voxel/sparse_volume.rs:48-63) and the C1 = 4 / C2 = 16 coarse pyramid.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <new>
#include <vector>

#include <omp.h>

#include "../../include/vf_worldgen.h"
#include "vf_internal.h"

namespace {

// Palette indices of the generated world (src/voxel/world.rs:40-56 wm).
enum : uint8_t { AIR = 0, GRASS = 1, DIRT = 2, STONE = 3, SAND = 4, WATER = 5, WOOD = 6, LEAVES = 7, LAMP = 8, BRICKM = 9 };

uint32_t hash32(uint32_t v) {   // integer finaliser for the generator's lattice values
    v ^= v >> 16; v *= 0x7feb352dU; v ^= v >> 15; v *= 0x846ca68bU; v ^= v >> 16;
    return v;
}
uint32_t hash3(int32_t a, int32_t b, int32_t c, uint32_t seed) {
    return hash32((uint32_t)a * 0x8da6b343U ^ hash32((uint32_t)b * 0xd8163841U ^ hash32((uint32_t)c * 0xcb1ab31fU ^ seed)));
}
float lattice(int32_t x, int32_t z, uint32_t seed, int32_t oct) {
    return (float)(hash3(x, z, oct, seed) >> 8) * (1.0f / 16777216.0f);
}
// 2-D value noise, smoothstep-interpolated lattice values; octave o has period period/2^o and
// weight 1/2^o; result in [0, 1).
float value_noise(float x, float z, uint32_t seed, int octaves, float period) {
    float sum = 0.0f, wsum = 0.0f, p = period, wgt = 1.0f;
    for (int o = 0; o < octaves; ++o) {
        float fx = x / p, fz = z / p;
        int32_t ix = (int32_t)std::floor(fx), iz = (int32_t)std::floor(fz);
        float tx = fx - (float)ix, tz = fz - (float)iz;
        tx = tx * tx * (3.0f - 2.0f * tx);
        tz = tz * tz * (3.0f - 2.0f * tz);
        float a = lattice(ix, iz, seed, o), b = lattice(ix + 1, iz, seed, o);
        float c = lattice(ix, iz + 1, seed, o), d = lattice(ix + 1, iz + 1, seed, o);
        float v = (a + (b - a) * tx) + ((c + (d - c) * tx) - (a + (b - a) * tx)) * tz;
        sum += v * wgt;
        wsum += wgt;
        p *= 0.5f;
        wgt *= 0.5f;
    }
    return sum / wsum;
}

void default_palette(vf_material pal[256]) {
    std::memset(pal, 0, sizeof(vf_material) * 256);
    auto set = [&](int i, float r, float g, float b, float e, float kind) {
        pal[i].r = r; pal[i].g = g; pal[i].b = b; pal[i].emissive = e;
        pal[i].props[0] = 1.0f; pal[i].props[2] = kind;
        pal[i].surf[0] = 0.9f; pal[i].surf[2] = 0.04f;
    };
    set(GRASS, 0.30f, 0.55f, 0.20f, 0.0f, 2.0f);
    set(DIRT, 0.45f, 0.32f, 0.20f, 0.0f, 0.0f);
    set(STONE, 0.50f, 0.50f, 0.52f, 0.0f, 0.0f);
    set(SAND, 0.80f, 0.74f, 0.55f, 0.0f, 0.0f);
    set(WATER, 0.10f, 0.25f, 0.35f, 0.0f, 1.0f);
    set(WOOD, 0.40f, 0.28f, 0.16f, 0.0f, 0.0f);
    set(LEAVES, 0.22f, 0.45f, 0.15f, 0.0f, 3.0f);
    set(LAMP, 1.00f, 0.85f, 0.60f, 1.0f, 0.0f);
    set(BRICKM, 0.55f, 0.25f, 0.20f, 0.0f, 0.0f);
    for (int i = 10; i < 256; ++i) set(i, 0.5f, 0.5f, 0.5f, 0.0f, 0.0f);
}

}  // namespace

struct vf_gen_world {
    int32_t wb[3] = {0, 0, 0};
    int32_t win[4] = {0, 0, 0, 0};
    std::vector<uint32_t> grid;
    std::vector<uint8_t> atlas;
    std::vector<uint32_t> masks;
    std::vector<uint8_t> c1, c2;
    vf_material palette[256];
    // terrain model for ground queries
    int32_t kind = 0;
    vf_gen_params params{};
    uint64_t stats[4] = {0, 0, 0, 0};
};

namespace {

void brick_mask(const uint8_t *vox, uint32_t *m) {
    std::memset(m, 0, 64);
    for (int z = 0; z < 8; ++z)
        for (int y = 0; y < 8; ++y)
            for (int x = 0; x < 8; ++x) {
                if (vox[z * 64 + y * 8 + x] == 0) continue;
                int oct = (x >> 2) | ((y >> 2) << 1) | ((z >> 2) << 2);
                int bit = (x & 3) | ((y & 3) << 2) | ((z & 3) << 4);
                m[oct * 2 + (bit >> 5)] |= 1u << (bit & 31);
            }
}

// Per-brick classification result: 0 empty, 1-255 uniform, or detailed (index into a brick store).
struct BrickOut {
    uint32_t value;   // < 256 uniform/empty; kDetailed = detailed
    uint32_t local;   // detailed: index in the producing block's store
};
constexpr uint32_t kDetailed = 0xFFFFFFFFu;

// Classify one 8^3 brick given a fetch(lx, ly, lz) functor; TORCH_FLAME (254) becomes air as the
// GPU sees it (gpu_world.rs:58-67).
template <class F>
uint32_t classify(F fetch, uint8_t *vox) {
    bool uniform = true;
    uint8_t first = 0;
    for (int z = 0; z < 8; ++z)
        for (int y = 0; y < 8; ++y)
            for (int x = 0; x < 8; ++x) {
                uint8_t v = fetch(x, y, z);
                if (v == 254) v = 0;
                int i = z * 64 + y * 8 + x;
                vox[i] = v;
                if (i == 0) first = v;
                else if (v != first) uniform = false;
            }
    return uniform ? (uint32_t)first : kDetailed;
}

// Assign slots in grid-index order (GpuWorld::upload's loop over world.grid) and build the atlas,
// masks and coarse pyramid.  stores[b][k] = detailed brick k of producer b; owner(gi) gives the
// producer and local index of a detailed cell.
void finish(vf_gen_world *w, std::vector<BrickOut> &cells, const std::vector<std::vector<uint8_t>> &stores,
            const std::vector<uint32_t> &producer) {
    const size_t n = cells.size();
    w->grid.assign(n, 0);
    uint32_t slots = 0;
    for (size_t i = 0; i < n; ++i)
        if (cells[i].value == kDetailed) slots++;
    w->atlas.assign((size_t)slots * 512, 0);
    w->masks.assign((size_t)slots * 16, 0);
    std::vector<size_t> slot_of(n, 0);
    uint32_t s = 0;
    for (size_t i = 0; i < n; ++i) {
        if (cells[i].value == kDetailed) {
            slot_of[i] = s;
            w->grid[i] = 256u + s;
            s++;
        } else {
            w->grid[i] = cells[i].value;
        }
    }
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        if (cells[i].value != kDetailed) continue;
        const uint8_t *src = stores[producer[i]].data() + (size_t)cells[i].local * 512;
        uint8_t *dst = w->atlas.data() + slot_of[i] * 512;
        std::memcpy(dst, src, 512);
        brick_mask(dst, w->masks.data() + slot_of[i] * 16);
    }
    int32_t d1[3], d2[3];
    for (int a = 0; a < 3; ++a) { d1[a] = (w->wb[a] + 3) / 4; d2[a] = (w->wb[a] + 15) / 16; }
    w->c1.assign((size_t)d1[0] * d1[1] * d1[2], 0);
    w->c2.assign((size_t)d2[0] * d2[1] * d2[2], 0);
    uint64_t nonempty = 0, uniform = 0;
    for (int32_t z = 0; z < w->wb[2]; ++z)
        for (int32_t y = 0; y < w->wb[1]; ++y)
            for (int32_t x = 0; x < w->wb[0]; ++x) {
                uint32_t g = w->grid[((size_t)z * w->wb[1] + y) * w->wb[0] + x];
                if (g == 0) continue;
                nonempty++;
                if (g < 256) uniform++;
                w->c1[((size_t)(z / 4) * d1[1] + y / 4) * d1[0] + x / 4] = 1;
                w->c2[((size_t)(z / 16) * d2[1] + y / 16) * d2[0] + x / 16] = 1;
            }
    w->stats[0] = nonempty;
    w->stats[1] = slots;
    w->stats[2] = uniform;
    uint64_t solid = 0;
    for (size_t i = 0; i < w->atlas.size(); ++i) solid += w->atlas[i] != 0;
    w->stats[3] = solid;
}

int32_t terrain_height(const vf_gen_params &p, int32_t x, int32_t z) {
    float n = value_noise((float)x + 0.5f, (float)z + 0.5f, p.seed, p.octaves, p.period);
    return (int32_t)std::floor(p.surface_mean + p.amplitude * (2.0f * n - 1.0f));
}

// Kind 1: value-noise heightfield, materials by depth band, water below the water level, sparse
// surface lamps.
uint8_t kind1_voxel(const vf_gen_params &p, int32_t x, int32_t y, int32_t z, int32_t h) {
    if (y < h) {
        int32_t d = h - 1 - y;
        if (d == 0) {
            if (h <= p.water_level + 1) return SAND;
            return (hash3(x, y, z, p.seed ^ 0x1a77u) & 511u) == 0 ? LAMP : GRASS;
        }
        if (d < 4) return DIRT;
        if (d < 12) return STONE;
        if (d < 16) return BRICKM;
        if (d < 20) return WOOD;
        return STONE;
    }
    if (y < p.water_level) return WATER;
    return AIR;
}

void gen_kind1(vf_gen_world *w) {
    const vf_gen_params &p = w->params;
    for (int a = 0; a < 3; ++a) w->wb[a] = p.world_bricks[a];
    const int32_t nx = w->wb[0] * 8, nz = w->wb[2] * 8;
    std::vector<int32_t> heights((size_t)nx * nz);
#pragma omp parallel for
    for (int32_t z = 0; z < nz; ++z)
        for (int32_t x = 0; x < nx; ++x) heights[(size_t)z * nx + x] = terrain_height(p, x, z);
    const size_t ncell = (size_t)w->wb[0] * w->wb[1] * w->wb[2];
    std::vector<BrickOut> cells(ncell);
    // one producer per brick z-slab
    std::vector<std::vector<uint8_t>> stores(w->wb[2]);
    std::vector<uint32_t> producer(ncell);
#pragma omp parallel for schedule(dynamic, 1)
    for (int32_t bz = 0; bz < w->wb[2]; ++bz) {
        uint8_t vox[512];
        auto &store = stores[bz];
        for (int32_t by = 0; by < w->wb[1]; ++by)
            for (int32_t bx = 0; bx < w->wb[0]; ++bx) {
                size_t gi = ((size_t)bz * w->wb[1] + by) * w->wb[0] + bx;
                uint32_t v = classify(
                    [&](int lx, int ly, int lz) {
                        int32_t x = bx * 8 + lx, y = by * 8 + ly, z = bz * 8 + lz;
                        return kind1_voxel(p, x, y, z, heights[(size_t)z * nx + x]);
                    },
                    vox);
                producer[gi] = (uint32_t)bz;
                cells[gi].value = v;
                if (v == kDetailed) {
                    cells[gi].local = (uint32_t)(store.size() / 512);
                    store.insert(store.end(), vox, vox + 512);
                }
            }
    }
    finish(w, cells, stores, producer);
}

// Kind 2: 128 x 256 x 128-voxel chunks (16 x 32 x 16 bricks) of 4-octave terrain plus trees (a
// cylinder trunk of radius 2-5 and height 40-120 under a sphere crown of radius 10-30), held in a
// ring grid whose window is centred on chunk (0, 0) like the reference's world_win
// (frame.rs:583-590).
struct Tree { int32_t x, z, base, r, h, R; };

void chunk_trees(const vf_gen_params &p, int32_t cx, int32_t cz, std::vector<Tree> &out) {
    out.clear();
    for (int32_t i = 0; i < p.trees_per_chunk; ++i) {
        uint32_t a = hash3(cx, cz, i, p.seed ^ 0x7e3u), b = hash32(a ^ 0x9e3779b9u), c = hash32(b ^ 0x85ebca6bu);
        Tree t;
        t.x = cx * 128 + (int32_t)(a & 127u);
        t.z = cz * 128 + (int32_t)((a >> 7) & 127u);
        t.r = 2 + (int32_t)(b % 4u);
        t.h = 40 + (int32_t)((b >> 8) % 81u);
        t.R = 10 + (int32_t)(c % 21u);
        t.base = terrain_height(p, t.x, t.z);
        out.push_back(t);
    }
}

void gen_kind2(vf_gen_world *w) {
    const vf_gen_params &p = w->params;
    for (int a = 0; a < 3; ++a) w->wb[a] = p.world_bricks[a];
    const int32_t ringc = w->wb[0] / 16;   // chunk slots per side
    w->win[0] = (0 - ringc / 2) * 16;
    w->win[1] = 0;
    w->win[2] = (0 - ringc / 2) * 16;
    w->win[3] = 1;
    const int32_t half = p.chunks / 2;
    const int32_t nchunk = p.chunks * p.chunks;
    const int32_t H = w->wb[1] * 8;
    const size_t ncell = (size_t)w->wb[0] * w->wb[1] * w->wb[2];
    std::vector<BrickOut> cells(ncell, BrickOut{0, 0});
    std::vector<uint32_t> producer(ncell, 0);
    std::vector<std::vector<uint8_t>> stores(nchunk);
#pragma omp parallel
    {
        std::vector<uint8_t> dense((size_t)128 * 128 * H);   // (z*128 + x)*H + y
        std::vector<int32_t> ground(128 * 128);
        std::vector<Tree> trees;
        uint8_t vox[512];
#pragma omp for schedule(dynamic, 1)
        for (int32_t ci = 0; ci < nchunk; ++ci) {
            const int32_t cx = ci % p.chunks - half, cz = ci / p.chunks - half;
            const int32_t x0 = cx * 128, z0 = cz * 128;
            std::fill(dense.begin(), dense.end(), 0);
            for (int32_t lz = 0; lz < 128; ++lz)
                for (int32_t lx = 0; lx < 128; ++lx) {
                    int32_t h = std::min(terrain_height(p, x0 + lx, z0 + lz), H - 1);
                    ground[lz * 128 + lx] = h;
                    uint8_t *col = &dense[((size_t)lz * 128 + lx) * H];
                    for (int32_t y = 0; y < h; ++y) {
                        int32_t d = h - 1 - y;
                        col[y] = d == 0 ? GRASS : (d < 4 ? DIRT : STONE);
                    }
                }
            // crowns (leaves fill air), then trunks (wood overwrites), from this chunk and its neighbours
            for (int pass = 0; pass < 2; ++pass)
                for (int32_t nz = cz - 1; nz <= cz + 1; ++nz)
                    for (int32_t nx = cx - 1; nx <= cx + 1; ++nx) {
                        if (nx < -half || nx > half || nz < -half || nz > half) continue;
                        chunk_trees(p, nx, nz, trees);
                        for (const Tree &t : trees) {
                            if (pass == 0) {
                                const int32_t cyc = t.base + t.h, R = t.R;
                                for (int32_t z = std::max(t.z - R, z0); z <= std::min(t.z + R, z0 + 127); ++z)
                                    for (int32_t x = std::max(t.x - R, x0); x <= std::min(t.x + R, x0 + 127); ++x) {
                                        int32_t dd = (x - t.x) * (x - t.x) + (z - t.z) * (z - t.z);
                                        if (dd > R * R) continue;
                                        int32_t span = (int32_t)std::floor(std::sqrt((double)(R * R - dd)));
                                        uint8_t *col = &dense[((size_t)(z - z0) * 128 + (x - x0)) * H];
                                        for (int32_t y = std::max(cyc - span, 0); y <= std::min(cyc + span, H - 1); ++y)
                                            if (col[y] == AIR) col[y] = LEAVES;
                                    }
                            } else {
                                const int32_t r = t.r;
                                for (int32_t z = std::max(t.z - r, z0); z <= std::min(t.z + r, z0 + 127); ++z)
                                    for (int32_t x = std::max(t.x - r, x0); x <= std::min(t.x + r, x0 + 127); ++x) {
                                        if ((x - t.x) * (x - t.x) + (z - t.z) * (z - t.z) > r * r) continue;
                                        uint8_t *col = &dense[((size_t)(z - z0) * 128 + (x - x0)) * H];
                                        for (int32_t y = std::max(t.base, 0); y < std::min(t.base + t.h, H); ++y) col[y] = WOOD;
                                    }
                            }
                        }
                    }
            auto &store = stores[ci];
            for (int32_t bz = 0; bz < 16; ++bz)
                for (int32_t by = 0; by < w->wb[1]; ++by)
                    for (int32_t bx = 0; bx < 16; ++bx) {
                        uint32_t v = classify(
                            [&](int lx, int ly, int lz) {
                                return dense[((size_t)(bz * 8 + lz) * 128 + (bx * 8 + lx)) * H + by * 8 + ly];
                            },
                            vox);
                        int32_t abx = cx * 16 + bx, abz = cz * 16 + bz;
                        size_t gi = ((size_t)(abz & (w->wb[2] - 1)) * w->wb[1] + by) * w->wb[0] + (abx & (w->wb[0] - 1));
                        producer[gi] = (uint32_t)ci;
                        cells[gi].value = v;
                        if (v == kDetailed) {
                            cells[gi].local = (uint32_t)(store.size() / 512);
                            store.insert(store.end(), vox, vox + 512);
                        }
                    }
        }
    }
    finish(w, cells, stores, producer);
}

}  // namespace

extern "C" {

int vf_gen_defaults(int32_t kind, vf_gen_params *p) {
    if (!p || (kind != 1 && kind != 2)) return vf::fail(nullptr, VF_E_INVALID, "vf_gen_defaults: kind must be 1 or 2");
    std::memset(p, 0, sizeof(*p));
    p->kind = kind;
    p->seed = 1337;
    if (kind == 1) {
        p->world_bricks[0] = 32; p->world_bricks[1] = 16; p->world_bricks[2] = 32;
        p->octaves = 2; p->surface_mean = 48.0f; p->amplitude = 24.0f; p->period = 64.0f;
        p->water_level = 40;
    } else {
        p->world_bricks[0] = 256; p->world_bricks[1] = 32; p->world_bricks[2] = 256;
        p->chunks = 13; p->octaves = 4; p->surface_mean = 64.0f; p->amplitude = 32.0f; p->period = 256.0f;
        p->trees_per_chunk = 400;
    }
    return VF_OK;
}

int vf_gen_world_create(const vf_gen_params *p, vf_gen_world **out) {
    if (!p || !out) return vf::fail(nullptr, VF_E_INVALID, "vf_gen_world_create: null argument");
    *out = nullptr;
    if (p->kind != 1 && p->kind != 2) return vf::fail(nullptr, VF_E_INVALID, "vf_gen_world_create: kind must be 1 or 2");
    for (int a = 0; a < 3; ++a)
        if (p->world_bricks[a] <= 0 || p->world_bricks[a] > 4096)
            return vf::fail(nullptr, VF_E_INVALID, "vf_gen_world_create: world_bricks out of range");
    if (p->octaves < 1 || p->octaves > 8 || !(p->period > 0.0f))
        return vf::fail(nullptr, VF_E_INVALID, "vf_gen_world_create: octaves 1-8 and period > 0 required");
    if (p->kind == 2) {
        const int32_t x = p->world_bricks[0], z = p->world_bricks[2];
        if ((x & (x - 1)) || (z & (z - 1)) || x != z || x % 16)
            return vf::fail(nullptr, VF_E_INVALID, "vf_gen_world_create: ring grid x/z must be equal powers of two, multiples of 16");
        if (p->chunks < 1 || p->chunks > x / 16 - 1 || p->world_bricks[1] * 8 > 65535)
            return vf::fail(nullptr, VF_E_INVALID, "vf_gen_world_create: chunks must fit inside the ring window");
        if (p->trees_per_chunk < 0 || p->trees_per_chunk > 100000)
            return vf::fail(nullptr, VF_E_INVALID, "vf_gen_world_create: trees_per_chunk out of range");
    }
    std::unique_ptr<vf_gen_world> w(new (std::nothrow) vf_gen_world());
    if (!w) return vf::fail(nullptr, VF_E_OOM, "vf_gen_world_create: out of host memory");
    w->params = *p;
    w->kind = p->kind;
    default_palette(w->palette);
    int prev = omp_get_max_threads();
    if (p->threads > 0) omp_set_num_threads(p->threads);
    try {
        if (p->kind == 1) gen_kind1(w.get());
        else gen_kind2(w.get());
    } catch (const std::bad_alloc &) {
        if (p->threads > 0) omp_set_num_threads(prev);
        return vf::fail(nullptr, VF_E_OOM, "vf_gen_world_create: out of host memory");
    }
    if (p->threads > 0) omp_set_num_threads(prev);
    *out = w.release();
    return VF_OK;
}

int vf_gen_world_from_dense(const uint8_t *vox, const int32_t wb[3], const int32_t win[4], const vf_material *pal,
                            vf_gen_world **out) {
    if (!vox || !wb || !out) return vf::fail(nullptr, VF_E_INVALID, "vf_gen_world_from_dense: null argument");
    *out = nullptr;
    for (int a = 0; a < 3; ++a)
        if (wb[a] <= 0 || wb[a] > 1024) return vf::fail(nullptr, VF_E_INVALID, "vf_gen_world_from_dense: bad dims");
    std::unique_ptr<vf_gen_world> w(new (std::nothrow) vf_gen_world());
    if (!w) return vf::fail(nullptr, VF_E_OOM, "vf_gen_world_from_dense: out of host memory");
    for (int a = 0; a < 3; ++a) w->wb[a] = wb[a];
    if (win)
        for (int a = 0; a < 4; ++a) w->win[a] = win[a];
    if (w->win[3] != 0) {
        if ((wb[0] & (wb[0] - 1)) || (wb[2] & (wb[2] - 1)))
            return vf::fail(nullptr, VF_E_INVALID, "vf_gen_world_from_dense: ring mode needs power-of-two x/z");
    }
    if (pal) std::memcpy(w->palette, pal, sizeof(w->palette));
    else default_palette(w->palette);
    const int32_t nx = wb[0] * 8, ny = wb[1] * 8;
    const size_t ncell = (size_t)wb[0] * wb[1] * wb[2];
    std::vector<BrickOut> cells(ncell);
    std::vector<std::vector<uint8_t>> stores(wb[2]);
    std::vector<uint32_t> producer(ncell);
#pragma omp parallel for schedule(dynamic, 1)
    for (int32_t bz = 0; bz < wb[2]; ++bz) {
        uint8_t b[512];
        for (int32_t by = 0; by < wb[1]; ++by)
            for (int32_t bx = 0; bx < wb[0]; ++bx) {
                size_t gi = ((size_t)bz * wb[1] + by) * wb[0] + bx;
                uint32_t v = classify(
                    [&](int lx, int ly, int lz) {
                        return vox[((size_t)(bz * 8 + lz) * ny + (by * 8 + ly)) * nx + (bx * 8 + lx)];
                    },
                    b);
                producer[gi] = (uint32_t)bz;
                cells[gi].value = v;
                if (v == kDetailed) {
                    cells[gi].local = (uint32_t)(stores[bz].size() / 512);
                    stores[bz].insert(stores[bz].end(), b, b + 512);
                }
            }
    }
    finish(w.get(), cells, stores, producer);
    *out = w.release();
    return VF_OK;
}

void vf_gen_world_destroy(vf_gen_world *w) { delete w; }

int vf_gen_world_desc(const vf_gen_world *w, vf_world_desc *d) {
    if (!w || !d) return vf::fail(nullptr, VF_E_INVALID, "vf_gen_world_desc: null argument");
    std::memset(d, 0, sizeof(*d));
    for (int a = 0; a < 3; ++a) d->world_bricks[a] = w->wb[a];
    for (int a = 0; a < 4; ++a) d->world_win[a] = w->win[a];
    d->grid = w->grid.data();
    d->atlas = w->atlas.empty() ? nullptr : w->atlas.data();
    d->n_slots = (uint32_t)(w->atlas.size() / 512);
    d->masks = w->masks.empty() ? nullptr : w->masks.data();
    d->coarse1 = w->c1.data();
    d->coarse2 = w->c2.data();
    d->palette = w->palette;
    return VF_OK;
}

int vf_gen_world_stats(const vf_gen_world *w, uint64_t out[4]) {
    if (!w || !out) return vf::fail(nullptr, VF_E_INVALID, "vf_gen_world_stats: null argument");
    for (int i = 0; i < 4; ++i) out[i] = w->stats[i];
    return VF_OK;
}

uint32_t vf_gen_world_voxel(const vf_gen_world *w, int32_t x, int32_t y, int32_t z) {
    if (!w) return 0;
    const int ring = w->win[3] != 0;
    const int32_t mn[3] = {ring ? w->win[0] : 0, 0, ring ? w->win[2] : 0};
    const int32_t v[3] = {x, y, z};
    for (int a = 0; a < 3; ++a)
        if (v[a] < mn[a] * 8 || v[a] >= (mn[a] + w->wb[a]) * 8) return 0;
    int32_t bx = x >> 3, by = y >> 3, bz = z >> 3;
    if (ring) { bx &= w->wb[0] - 1; bz &= w->wb[2] - 1; }
    uint32_t g = w->grid[((size_t)bz * w->wb[1] + by) * w->wb[0] + bx];
    if (g < 256) return g;
    return w->atlas[(size_t)(g - 256) * 512 + (size_t)((z & 7) * 64 + (y & 7) * 8 + (x & 7))];
}

int32_t vf_gen_world_ground(const vf_gen_world *w, int32_t x, int32_t z) {
    if (!w) return 0;
    if (w->kind == 1 || w->kind == 2) return terrain_height(w->params, x, z);
    for (int32_t y = w->wb[1] * 8 - 1; y >= 0; --y)
        if (vf_gen_world_voxel(w, x, y, z) != 0) return y + 1;
    return 0;
}

}  // extern "C"
