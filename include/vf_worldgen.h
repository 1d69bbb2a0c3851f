/*
 * vf_worldgen.h — synthetic worlds for the traversal workloads (part of libvftrace.so).
 *
 * Not a reference entry point: the reference generates its world in src/worldgen/ (World::generate,
 * src/voxel/world.rs:390), which is outside this library.  These helpers build seeded worlds of the
 * benchmark shapes and pack them exactly as GpuWorld::upload does on the CPU (gpu_world.rs:243-313,
 * 364-382): grid encoding 0 / 1-255 / 256+slot with slots assigned in grid-index order, the atlas in
 * slot order, 64-byte brick masks and the coarse pyramid.  The result is a vf_world_desc.
 */
#ifndef VF_WORLDGEN_H
#define VF_WORLDGEN_H

#include "vf_trace.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vf_gen_params {
    int32_t kind;            /* 1: value-noise terrain with water (fixed world)
                                2: streamed-chunk forest in a ring grid (endless shape) */
    uint32_t seed;
    int32_t world_bricks[3]; /* kind 1: grid dims (e.g. 32,16,32).  kind 2: ring grid dims (e.g. 256,32,256) */
    int32_t chunks;          /* kind 2: loaded chunks per side (e.g. 13), centred on chunk (0,0) */
    int32_t octaves;         /* value-noise octaves (kind 1: 2, kind 2: 4) */
    float surface_mean;      /* voxels (kind 1: 48) */
    float amplitude;         /* voxels (kind 1: 24) */
    float period;            /* first octave period, voxels */
    int32_t water_level;     /* kind 1: water fills y < water_level above the ground (0 = none) */
    int32_t trees_per_chunk; /* kind 2 (e.g. 400) */
    int32_t threads;         /* 0 = all */
} vf_gen_params;

typedef struct vf_gen_world vf_gen_world;

/* Fills params with the defaults of `kind` (1 or 2). */
int vf_gen_defaults(int32_t kind, vf_gen_params *p);
int vf_gen_world_create(const vf_gen_params *p, vf_gen_world **out);
/* From a dense voxel array (x fastest, then y, then z) of dims = world_bricks*8 in grid (wrapped)
 * coordinates; palette NULL = the default palette.  For tests and small scenes. */
int vf_gen_world_from_dense(const uint8_t *vox, const int32_t world_bricks[3], const int32_t world_win[4],
                            const vf_material *palette, vf_gen_world **out);
void vf_gen_world_destroy(vf_gen_world *w);
/* Pointers stay valid until vf_gen_world_destroy. */
int vf_gen_world_desc(const vf_gen_world *w, vf_world_desc *out);
/* Stats: {non-empty bricks, detailed bricks (slots), uniform bricks, solid voxels (approx. excl. uniform)}. */
int vf_gen_world_stats(const vf_gen_world *w, uint64_t out[4]);
/* The material of absolute voxel (x, y, z); 0 outside the window. */
uint32_t vf_gen_world_voxel(const vf_gen_world *w, int32_t x, int32_t y, int32_t z);
/* Terrain surface height (first air voxel above the ground) of absolute column (x, z). */
int32_t vf_gen_world_ground(const vf_gen_world *w, int32_t x, int32_t z);

#ifdef __cplusplus
}
#endif
#endif
