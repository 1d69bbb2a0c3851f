/*
 * Two-bounce GI paths and the learned GI radiance cache (C ABI).
 *
 * The reference renders indirect light with giPath (shaders/common.glsl:2250-2366; two bounces with
 * r.gi.bounces 2, giPathImpl(..., lastStage 3, traceOn)) and, with r.gi.cache, reads bounce 2 from a
 * spatial-hash radiance cache (common.glsl:2150-2230, gi_cache_update.comp, gi_cache_resolve.comp,
 * record_gi_cache in src/gfx/renderer/record.rs:243-315).  This library traces the same two-bounce path
 * (vf_trace_gi_path) and replaces the hash table by a small MLP trained online on the tracer's own
 * samples (vf_gicache_*).  The MLP has no counterpart in the reference; its entry points take the
 * place of record_gi_cache's two dispatches and of giCacheLookup / giCacheAdd.
 *
 * Network: input = 12 frequency bands (sin, cos) of each position coordinate (72) + surface normal (3)
 * + view direction (3) = 78 features, zero-padded to 96; five hidden layers of 64 (ReLU); 3 outputs
 * (RGB radiance, stored as FP16).  23,043 parameters.  Operands are BF16 (MFMA), accumulation FP32,
 * master weights and Adam moments FP32.
 *
 * Parameter order (vf_gicache_get_params / set_params / the gradient of vf_gicache_grad): for layer
 * l = 0..5 with (in, out) = (96,64), (64,64) x 4, (64,3): W_l as in x out row-major, then b_l (out).
 *
 * All device pointers are caller-owned device memory of the context's device; calls are asynchronous
 * on the context's stream unless stated otherwise.  Status codes are those of vf_trace.h.
 */
#ifndef VF_GICACHE_H
#define VF_GICACHE_H
#include "vf_trace.h"

#ifdef __cplusplus
extern "C" {
#endif

#define VF_GICACHE_PARAMS 23043
#define VF_GICACHE_INPUTS 96

typedef struct vf_gicache vf_gicache;

/* A point query / training sample (64 B).  pos in voxels (absolute), n = unit surface normal, view =
 * unit direction from the point back along the arriving ray.  For a training sample tag = 1 (valid) or
 * 0 (ignored) and rgb = the traced radiance target; for a render query tag = the output slot and rgb =
 * the path throughput the cache's answer is weighted by. */
typedef struct vf_gi_record {
    float pos[3];
    uint32_t tag;
    float n[3];
    float pad0;
    float view[3];
    float pad1;
    float rgb[3];
    float pad2;
} vf_gi_record;

typedef struct vf_gicache_params {
    int32_t terminate;       /* 1: render traces bounce 1 (+ its sun ray) and takes bounce 2 onward from the
                              *    cache at the bounce-1 hit; 0: the cache answers at the G-buffer point */
    int32_t train_samples;   /* traced samples per frame (65536) */
    int32_t train_steps;     /* optimizer steps the frame's samples are split into (4) */
    int32_t spp;             /* bounce-1 samples per pixel when rendering with terminate = 1 (1..8) */
    float lr, beta1, beta2, eps;   /* Adam (1e-2, 0.9, 0.9, 1e-8) */
    float loss_eps;          /* relative L2 loss (y - t)^2 / (sg(mean(y))^2 + loss_eps) (0.01) */
    float fallback_ratio;    /* render the traced path while the running error > ratio x baseline (2) */
    int32_t warmup_frames;   /* frames rendered with the traced path while the cache first trains (16) */
    int32_t enc_shift;       /* band k of coordinate x is sin/cos(pi * x * 2^(k - enc_shift)), k = 0..11 */
    uint32_t seed;           /* weight initialisation */
    uint32_t salt;           /* training-sample and render random streams */
    float gi_reach;          /* GI_REACH (48) */
    int32_t lights_nee;      /* LIGHTS_NEE: emitters with emission >= 0.5 add nothing along the path (1) */
    int32_t bounces;         /* bounces of the GI the cache stands in for (2): 2 = the reference's giPath; 3 = the
                              * three-bounce path of vf_trace_gi_path_n.  Training targets, the fallback path and
                              * (terminate 1) the tail's share follow it: with 3 the tail after bounce 1 has two
                              * bounces, so a terminate-1 sample is a traced three-bounce path */
    float weight_ema;        /* the render's network weights: an exponential mean of the trained weights updated
                              * at every Adam step, we = ema * we + (1 - ema) * w (0: the trained weights) */
} vf_gicache_params;

typedef struct vf_gicache_stats {
    int64_t frames;          /* frames run */
    int64_t steps;           /* optimizer steps taken */
    float last_loss;         /* mean loss of the last frame's fresh samples before its first update */
    float running_error;     /* exponential mean (0.1) of last_loss */
    float baseline;          /* lowest running error since the cache (re)started trusting itself */
    int32_t fallback;        /* 1: the last frame rendered the traced path */
    int32_t n_queries;       /* cache queries of the last frame */
    int32_t n_valid_samples; /* valid training samples of the last frame */
} vf_gicache_stats;

/* Defaults: terminate 0, 65536 samples in 4 steps, spp 1, Adam (lr 1e-2, 0.9, 0.9, 1e-8), loss_eps 0.01,
 * fallback 2x, warmup 16, enc_shift 8, seed 1337, salt 0x2f6b9a31, gi_reach 48, lights_nee 1, bounces 2,
 * weight_ema 0.9.  (The fast learning rate tracks a moving camera; the weights' exponential mean removes
 * most of the step noise it adds to the render.) */
int vf_gicache_defaults(vf_gicache_params *p);
int vf_gicache_create(vf_ctx *ctx, const vf_gicache_params *p, vf_gicache **out);
void vf_gicache_destroy(vf_gicache *c);
/* Re-initialises weights, Adam state and statistics (synchronous). */
int vf_gicache_reset(vf_ctx *ctx, vf_gicache *c);

/* Plain traced GI, the reference's two-bounce giPath (common.glsl:2359-2364 with GI_DOUBLE): for each
 * non-sky G-buffer pixel, spp samples; sample s uses index k = frame * spp + s: bounce-1 direction
 * cosineHemisphereU(N, ldSample2(pix, k, salt)) (gi.comp:69-72), path seed pixelSeed(pix, k, salt)
 * (common.glsl:1242).  gi_reach and lights_nee as vf_trace_gi.  out_rgba (RGBA32F per pixel) = mean
 * radiance, a = 1 (0 for sky). */
int vf_trace_gi_path(vf_ctx *ctx, const vf_pass_ubo *pass, int32_t width, int32_t height, float gi_reach,
                     int32_t lights_nee, const float *gb_pos, const uint32_t *gb_normal, uint32_t frame, int32_t spp,
                     uint32_t salt, float *out_rgba);
/* vf_trace_gi_path with 2 or 3 bounces.  The three-bounce path has no counterpart in the reference (its giPath
 * stops after two): it is the two-bounce path up to where bounce 2 starts (at the bounce-1 hit, throughput
 * thr), then a two-bounce path from that point (its first direction the next cosine draw of the same seed,
 * never traced on past the reach): L = L_head + thr * L_rest.  bounces 2 = vf_trace_gi_path. */
int vf_trace_gi_path_n(vf_ctx *ctx, const vf_pass_ubo *pass, int32_t width, int32_t height, float gi_reach,
                       int32_t lights_nee, const float *gb_pos, const uint32_t *gb_normal, uint32_t frame, int32_t spp,
                       uint32_t salt, int32_t bounces, float *out_rgba);

/* One frame of cached GI: trace params.train_samples fresh samples (paths from stratified G-buffer
 * pixels; target = the traced radiance at the cache's query point), take train_steps Adam steps on
 * them, update the running error, then render: the path with its tail from the cache, or the traced
 * path (vf_trace_gi_path with spp, salt, gi_reach and lights_nee from params) while warming up or when the
 * running error exceeds fallback_ratio x baseline.  out_rgba as vf_trace_gi_path. */
int vf_gicache_frame(vf_ctx *ctx, vf_gicache *c, const vf_pass_ubo *pass, int32_t width, int32_t height,
                     const float *gb_pos, const uint32_t *gb_normal, uint32_t frame, float *out_rgba);
/* Synchronous. */
int vf_gicache_read_stats(vf_ctx *ctx, vf_gicache *c, vf_gicache_stats *out);

/* The training samples vf_gicache_frame traces for this frame (n = params.train_samples records). */
int vf_gicache_samples(vf_ctx *ctx, vf_gicache *c, const vf_pass_ubo *pass, int32_t width, int32_t height,
                       const float *gb_pos, const uint32_t *gb_normal, uint32_t frame, vf_gi_record *out);

/* --- the network, for checks and for callers with their own samples --- */
/* FP32 master weights in parameter order (synchronous). set_params also clears the Adam moments. */
int vf_gicache_get_params(vf_ctx *ctx, vf_gicache *c, float *host_out);
/* The weights the render and vf_gicache_infer use (parameter order, synchronous): the exponential mean of the
 * trained weights with weight_ema > 0, else the trained weights; set_params and reset set them equal. */
int vf_gicache_get_render_params(vf_ctx *ctx, vf_gicache *c, float *host_out);
int vf_gicache_set_params(vf_ctx *ctx, vf_gicache *c, const float *host_in);
/* The 96 input features of n records (FP32, n x 96). */
int vf_gicache_encode(vf_ctx *ctx, vf_gicache *c, const vf_gi_record *recs, int64_t n, float *enc);
/* Network output for n records with the render's weights (the trained ones, or their exponential mean with
 * weight_ema > 0): RGB as FP16 in 4 halves per record (the 4th is 0). */
int vf_gicache_infer(vf_ctx *ctx, vf_gicache *c, const vf_gi_record *recs, int64_t n, uint16_t *out_rgbh);
/* Gradient (parameter order) of the mean loss over n samples at the current weights, and that loss
 * (sum over valid samples of the per-sample loss, divided by n); no update. */
int vf_gicache_grad(vf_ctx *ctx, vf_gicache *c, const vf_gi_record *samples, int64_t n, float *grad, float *loss);
/* One Adam step with the given gradient (parameter order, device memory). */
int vf_gicache_adam(vf_ctx *ctx, vf_gicache *c, const float *grad);
/* grad + Adam on n samples. */
int vf_gicache_train(vf_ctx *ctx, vf_gicache *c, const vf_gi_record *samples, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
