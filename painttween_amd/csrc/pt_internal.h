/*
 * pt_internal.h -- types shared by the HIP kernels (pt_kernels.hip) and the
 * host runtime (pt_runtime.hip).  Not part of the public C ABI.
 */
#ifndef PT_INTERNAL_H
#define PT_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pt_api.h"

/* Nearest-neighbour field as structure-of-arrays planes: {x, y, theta,
 * hysteresis} and {scaleU, scaleV}.  Style layer and luminance shift are
 * fixed at 0 on this path (PatchXF, types.h:98-114). */
struct PtNnf {
    float4 *xyth;
    float2 *sc;
};

/* AdvectedBuffersDevice (advectedBuffers.h:37-52). */
struct PtAdv {
    float4 *xyth;
    float2 *sc;
    float4 *base;
    float4 *out;
    float *res;
    int32_t *ts;
};

#define PT_MAX_R 7

/* Everything a kernel needs, passed by value.  Level-0-sized buffers have
 * pitch W (elements); level-sized textures are tightly packed. */
struct PtArgs {
    int W, H;          /* level-0 image size (buffer pitch)            */
    int level, lw, lh; /* active level rectangle                       */
    int hw, hh;        /* offsets-histogram size at this level (wrap)  */
    int ew, eh;        /* exemplar texture size (this level)           */
    int ow, oh;        /* image orientation texture size               */
    int scw, sch;      /* image scale texture size                     */
    int rw, rh;        /* random texture size                          */
    int cur;           /* gCurrentBuffer                               */

    int r;             /* residual window radius                       */
    float gauss[2 * PT_MAX_R + 1];
    float alpha_w, in_w, canvas_w, dt_w, su, sv;
    int cull, transp, orient_en, direction;
    float tdi, tdo, tcw, hyst_w, max_dist, jitter;
    int samples, rng_mode;
    uint32_t key0, key1, call;
    int time_active;   /* some advected offset may be valid */
    int early_exit;    /* all energy terms non-negative: exact early rejection allowed */
    int unit_parallel; /* candidate units spread over the lanes of a workgroup */

    PtNnf base, work;
    float *res_base, *res_work;
    float4 *canvas;
    PtAdv adv[2];
    const float4 *color[2];
    const float *dt[2];
    const float2 *velb[2];
    const float2 *velf[2];
    const float *orient;
    const float2 *scale;
    const float4 *random;
    const float4 *ex; /* interleaved exemplar: ex[2i] base, ex[2i+1] output */
    const float *ex_orient;
    const float *ex_dt;
};

/* Host-side launchers (pt_kernels.hip).  All run on `stream`. */
hipError_t pt_launch_update_residual(const PtArgs &a, hipStream_t s);
hipError_t pt_launch_canvas(const PtArgs &a, hipStream_t s);
hipError_t pt_launch_sync_canvas(const PtArgs &a, int dw, int dh, hipStream_t s);
hipError_t pt_launch_copy_random(const PtArgs &a, hipStream_t s);
hipError_t pt_launch_propagate(const PtArgs &a, int jump, hipStream_t s);
hipError_t pt_launch_scatter(const PtArgs &a, hipStream_t s);
hipError_t pt_launch_upsample(const PtArgs &a, hipStream_t s);
hipError_t pt_launch_advect(const PtArgs &a, int forwards, int time_step, int flags, hipStream_t s);
hipError_t pt_launch_merge_randomly(const PtArgs &a, int forwards, float weight, int force, hipStream_t s);
hipError_t pt_launch_merge_residual(const PtArgs &a, int forwards, hipStream_t s);
hipError_t pt_launch_downsample(const float *src, int pw, float *dst, int lw, int lh, int ch,
                                int scale_mode, hipStream_t s);
hipError_t pt_launch_fill(float *dst, size_t n, float v, hipStream_t s);
hipError_t pt_launch_tensor_angle(const float *tensor, float *angle, size_t n, hipStream_t s);
hipError_t pt_launch_interleave(const float4 *src, int w, int h, float4 *ex, int ew, int slot, hipStream_t s);
hipError_t pt_launch_random_pyramid(const float4 *color, const float *iorient, int w, int h,
                                    const float4 *eout, const float *eorient, int ew, int eh,
                                    int transp, int orient_en, uint32_t key0, uint32_t key1,
                                    float4 *out, hipStream_t s);
hipError_t pt_launch_pack_nnf(const PtNnf &src, int W, int lw, int lh, void *dst_patchxf, hipStream_t s);
hipError_t pt_launch_unpack_nnf(const void *src_patchxf, int lw, int lh, PtNnf dst, int W, hipStream_t s);

#endif
