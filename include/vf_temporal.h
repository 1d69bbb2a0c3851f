/*
 * vf_temporal.h — the temporal consumers of the traced output, on MI355X (gfx950).
 *
 * These are the reference's per-frame passes that turn one noisy traced sample per pixel into a
 * stable signal: the GI pass's reprojected accumulation (shaders/gi.comp), the lighting pass's
 * sun-shadow history (shaders/lighting.comp) and the SVGF a-trous filter over the half-res GI
 * (shaders/atrous.comp).  They read the traced outputs of vf_trace.h (vf_trace_gi's radiance,
 * vf_trace_sun_cone's shadow sample) and the G-buffers of this frame and the previous one.
 *
 * Image layouts (row-major, one record per texel, device pointers):
 *   gb_pos       RGBA32F  float[4]   full res  (xyz world position, w = hit distance; w < 0 sky)
 *   gb_normal    RGBA8    uint32     full res  (xyz = n*0.5+0.5 in bytes 0..2)
 *   gb_motion    RG16F    uint32     full res  (prevUV - curUV; x in the low 16 bits)
 *   gb_delta     RGBA16F  uint16[4]  full res  (P - Pprev of the surface, e.g. a moving body)
 *   GI history   RGBA16F  uint16[4]  half res  (indirect rgb, history length)
 *   GI moments   RGBA32F  float[4]   half res  (luma m1, m2, fast luma, 0)
 *   GI guide     RGBA32UI uint32[4]  half res  (position bits, RGBA8 normal or 0 for sky)
 *   shadow hist  RGBA32F  float[4]   full res  (shadow, 2nd moment, history length, fast)
 *   a-trous out  RGBA16F  uint16[4]  half res  (filtered rgb, variance)
 * Half res = ((width+1)/2) x ((height+1)/2) (renderer/mod.rs:525).  Half-float stores round to
 * nearest even.  The previous-frame images of a ping-pong pair must not alias the current ones.
 * Calls are stream-ordered on the context's stream and return VF_OK or a VF_E_* code.
 */
#ifndef VF_TEMPORAL_H
#define VF_TEMPORAL_H

#include <stdint.h>

#include "vf_trace.h"

#ifdef __cplusplus
extern "C" {
#endif

/* gbuffer.comp's motion outputs (gbuffer.comp:84-104 prevWorldPos / reprojMotion / STORE_MOTION,
 * 405-410 sky), from the G-buffer vf_trace_gbuffer / vf_trace_primary just wrote and the pass's
 * current and previous camera (cam_*, prev_cam_*, tan_half_fov, prev_tan_half_fov, jitter):
 * gbMotion = prevUV - curUV of the surface point where it was last frame (a body's voxel carried by
 * the body's prev_obj_to_world), 0 when that point is behind the previous camera; gbDelta = P - Pprev
 * with .a = body + 1 (0 for the world).  Sky pixels: the view direction's motion at infinity
 * (prevCamPos + rd*1e6), delta 0.  d_body (nullable): vf_trace_gbuffer's per-pixel body index;
 * NULL or a scene without bodies: every hit is the static world (motion from the camera only). */
int vf_gbuffer_motion(vf_ctx *ctx, const vf_pass_ubo *pass, int32_t width, int32_t height, const float *d_gb_pos,
                      const int32_t *d_body, uint32_t *d_gb_motion, uint16_t *d_gb_delta);

/* P::Gi's accumulation half (gi.comp:37-63 guide and all-sky texels, 93-184): per half-res texel,
 * the representative full-res texel becomes the filter guide; this frame's sample (the rotating
 * sub-texel of pass->frame_idx, or the representative when that one is sky) is reprojected through
 * gb_motion / gb_delta, the 4 bilinear history taps are validated against the previous G-buffer
 * (plane distance < 0.5 + t*0.004, normal dot > 0.85, pass->shift_p added to old positions), and
 * the sample is blended in with the firefly cap, the anti-lag clamp and 48 frames of history.
 * bicubic != 0: r.gi.bicubic (the reference's default), Catmull-Rom history when all 4 taps are
 * valid.  d_indirect: the traced radiance (vf_trace_gi with half_res, RGBA32F; rgb used). */
int vf_gi_accumulate(vf_ctx *ctx, const vf_pass_ubo *pass, int32_t width, int32_t height, int32_t bicubic,
                     const float *d_gb_pos, const uint32_t *d_gb_normal, const float *d_gb_pos_prev,
                     const uint32_t *d_gb_normal_prev, const uint32_t *d_gb_motion, const uint16_t *d_gb_delta,
                     const float *d_indirect, const uint16_t *d_hist_prev, const float *d_mom_prev,
                     uint16_t *d_hist_cur, float *d_mom_cur, uint32_t *d_guide);

/* P::Light's history (lighting.comp:36-42 sky, 92-139): per full-res pixel the raw sun sample
 * (vf_trace_sun_cone's d_shadow) is blended into the reprojected shadow history with the same
 * tap validation, the fast history and the anti-lag clamp.  Sky pixels get (1, 1, 0, 0). */
int vf_light_accumulate(vf_ctx *ctx, const vf_pass_ubo *pass, int32_t width, int32_t height, const float *d_gb_pos,
                        const uint32_t *d_gb_normal, const float *d_gb_pos_prev, const uint32_t *d_gb_normal_prev,
                        const uint32_t *d_gb_motion, const uint16_t *d_gb_delta, const float *d_shadow,
                        const float *d_hist_prev, float *d_hist_cur);

/* P::Atrous x levels (record.rs:372-376; atrous.comp): level i (tap spacing 2^i) writes
 * d_filt[i & 1] and reads d_filt[(i + 1) & 1]; level 0 reads the GI history and moments, so the
 * result of an odd level count is in d_filt0 (levels 3 = r.gi.filter "normal", 5 = "smooth").
 * histfix != 0: r.gi.histfix (default), freshly revealed texels borrow converged neighbours' GI.
 * Edge stops: normal^64, exp(-plane distance / (1 + t*0.01)), exp(-|dluma| / (4 sigma + 1e-4)). */
int vf_gi_atrous(vf_ctx *ctx, int32_t width, int32_t height, int32_t levels, int32_t histfix, const float *d_gb_pos,
                 const uint32_t *d_gb_normal, const float *d_gi_mom, const uint16_t *d_gi_hist, const uint32_t *d_guide,
                 uint16_t *d_filt0, uint16_t *d_filt1);

#ifdef __cplusplus
}
#endif
#endif
