/*
 * pt_oracle.c -- TEST INFRASTRUCTURE ONLY: CPU restatement of the
 * PaintTween per-frame synthesis hot path, the correctness oracle for the HIP
 * kernels.  Only tests, the smoke check and the benchmark's CPU-baseline leg
 * load it.
 *
 * Reference files followed (paths relative to src/PaintTweenCUDA/ unless
 * prefixed G/ for src/PaintTweenGUI/):
 *   texSynth_kernel.cu  -- every kernel and launcher (line numbers per function)
 *   texSynth_util.h     -- clamp, SSD, basis, new exemplar coordinate
 *   cudaImageBuffer.h   -- level-0-sized buffers, pixelWrap, inBounds
 *   cudaImagePyramid.h  -- PYRAMID_FETCH: +0.5 texel centre, clamp addressing
 *   convolution.cu      -- Gaussian tap weights
 *   downsample.cpp, imagePyramid.cpp -- box pyramid
 *   synthesisFrame.cpp, synthesisProcessor.cpp -- schedule
 *   G/nlDataAccess.cpp, imageIO.h -- orientation tensor to angle, defaults
 *
 * Deliberate, documented definitions where the reference depends on hardware
 * or libraries that cannot run here (applied identically on the GPU):
 *   - texture filtering: bilinear weights in FP32 from the exact texel-space
 *     coordinate (the reference's texture unit uses 8-bit weights); point
 *     sampling picks floor(x + 0.5f);
 *   - cos/sin/atan2 come from pt_math.h (Cody-Waite + Taylor), not CUDA libm;
 *   - mul+add pairs are fused exactly where written with fmaf (the reference's
 *     nvcc build fuses at the compiler's discretion); nothing else is fused;
 *   - glibc rand() in the random pyramid is replaced by Philox4x32-10;
 *   - residuals kept between passes stay FP32 (the reference round-trips them
 *     through half-float EXR files).
 *
 * Pixel loops are Jacobi (read base, write working), so the result does not
 * depend on the OpenMP thread count.
 */
#include "pt_oracle.h"
#include "../painttween_amd/csrc/pt_math.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdarg.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { float x, y, z, w; } f4;
typedef struct { float x, y; } f2;

/* PatchXF restricted to one style layer and no luminance shift: layer is
 * always 0 and the shift always (0,0) on this path (multi-style and
 * luminance shift are rejected by setup). */
/* PatchXF: ly is luminanceShift.y (its .a is always 0 here), la the style
 * layer (0 except with keyframe interpolation) */
typedef struct { float x, y, th, hy, su, sv, ly; int la; } nnf;

/* RibbonP (types.h:116-126): offset to a keyframe, time step, style layer
 * (-1: invalid).  Packed in textures as Color4(time_step, y, x, layer). */
typedef struct { float x, y; int32_t ts, layer; } ribbon;

typedef struct {
    nnf *off;
    f4 *base;
    f4 *out;
    float *res;
    int32_t *ts;
} adv_t;

struct ptref_pyramid {
    int elem, w, h, levels;
    int lw[32], lh[32];
    float *data[32];
};

typedef struct {
    ptref_pyramid *color, *orient, *dt, *scale, *velb, *velf, *id;
} frame_pyr;

typedef struct {
    pt_patchxf *off;
    float *res;
    float *canvas;
} snap_t;

struct ptref_ctx {
    char err[512];
    int nthreads;
    pt_params p;
    int have_params;
    float gauss[64];
    int gauss_r;
    /* WorkingBuffers, level-0 sized (workingBuffers.cpp:48-95) */
    int W, H, SW, SH;
    nnf *off_base, *off_work;
    float *res_base, *res_work;
    f4 *canvas;
    adv_t adv[2]; /* 0 = advectedF, 1 = advectedB */
    /* two-layer image textures at level-0 size (texSynth_kernel.cu:2111-2117) */
    f4 *img_color[2];
    float *img_dt[2];
    f2 *img_velb[2], *img_velf[2];
    int cur; /* gCurrentBuffer */
    /* non-layered image textures, reallocated at each upload's size */
    float *img_orient; int ow, oh;
    f2 *img_scale; int scw, sch;
    f4 *random; int rw, rh;
    int32_t *img_id; int idw, idh; /* gImageId, a single 0 texel until uploaded */
    int lum_live; /* some luminance shift may be nonzero: the canvas adds it */
    /* exemplar textures, one level, nl style layers of ew x eh texels each
     * (texSynth_kernel.cu:2119-2126) */
    f4 *ex_base, *ex_out;
    float *ex_orient, *ex_dt;
    int32_t *ex_id;                 /* gExemplarObjectIDs */
    int ew, eh, nl;
    /* keyframe interpolation: cumulative key-to-key fields (layered, cl
     * layers of cw x ch, linear filter), frame-to-key ribbons (level sized,
     * one 0 texel until uploaded) and the ribbon field buffer pair */
    f4 *cumul[2]; int cw, ch, cl;   /* [0] = gCumulFieldB, [1] = gCumulFieldF */
    f4 *ftk[2]; int fw[2], fh[2];   /* [0] = gFrameToKeyRibbonB, [1] = ...F */
    ribbon *rib_base, *rib_work;
    uint32_t scatter_calls;
    uint64_t pde;
    /* sequence state of ptref_synthesize_next_frame */
    int seq_frame, seq_levels;
    snap_t *seq_snap[2];       /* [frame parity][level] */
    frame_pyr seq_pyr[2];      /* [frame parity] */
    ptref_pyramid *seq_style[5]; /* base, output, orientation angle, distance transform, object IDs */
};

static int fail(ptref_ctx *c, int code, const char *fmt, ...)
{
    if (c) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(c->err, sizeof(c->err), fmt, ap);
        va_end(ap);
    }
    return code;
}

/* ------------------------------------------------------------------ */
/* math entry points                                                   */
/* ------------------------------------------------------------------ */
void ptref_sincosf(float x, float *s, float *c) { pt_sincosf(x, s, c); }
float ptref_atan2f(float y, float x) { return pt_atan2f(y, x); }
float ptref_logf(float x) { return pt_logf(x); }
float ptref_acosf(float x) { return pt_acosf(x); }
double ptref_log_d(double x) { return pt_log_d(x); }
double ptref_exp_d(double x) { return pt_exp_d(x); }
void ptref_rgb_to_lab(float r, float g, float b, float *out) { pt_rgb_to_lab(r, g, b, &out[0], &out[1], &out[2]); }
float ptref_luminance_delta(float cr, float cg, float cb, float pr, float pg, float pb)
{
    return pt_luminance_delta(cr, cg, cb, pr, pg, pb);
}
float ptref_coherence_term(float cx, float cy, float tcx, float tcy, float tx, float ty, int i, int j, int level,
                           float ratio_range, float angle_range)
{
    return pt_coherence_term(cx, cy, tcx, tcy, tx, ty, i, j, level, ratio_range, angle_range);
}
float ptref_tensor_to_angle(float xx, float xy, float yy) { return pt_tensor_to_angle(xx, xy, yy); }
void ptref_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    pt_philox4x32_10(ctr, key, out);
}

/* ------------------------------------------------------------------ */
/* texture sampling (cudaImagePyramid.h:99-108, clamp addressing,      */
/* unnormalised coordinates, fetch at (X+0.5, Y+0.5))                  */
/* ------------------------------------------------------------------ */
static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

static inline f4 tex4_lin(const f4 *t, int pitch, int w, int h, float x, float y)
{
    float fx = floorf(x), fy = floorf(y);
    float a = x - fx, b = y - fy;
    int x0 = (int)fx, y0 = (int)fy;
    int xa = clampi(x0, 0, w - 1), xb = clampi(x0 + 1, 0, w - 1);
    int ya = clampi(y0, 0, h - 1), yb = clampi(y0 + 1, 0, h - 1);
    float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b);
    float w01 = (1.0f - a) * b, w11 = a * b;
    f4 t00 = t[ya * pitch + xa], t10 = t[ya * pitch + xb];
    f4 t01 = t[yb * pitch + xa], t11 = t[yb * pitch + xb];
    f4 r;
    r.x = fmaf(w11, t11.x, fmaf(w01, t01.x, fmaf(w10, t10.x, w00 * t00.x)));
    r.y = fmaf(w11, t11.y, fmaf(w01, t01.y, fmaf(w10, t10.y, w00 * t00.y)));
    r.z = fmaf(w11, t11.z, fmaf(w01, t01.z, fmaf(w10, t10.z, w00 * t00.z)));
    r.w = fmaf(w11, t11.w, fmaf(w01, t01.w, fmaf(w10, t10.w, w00 * t00.w)));
    return r;
}

static inline f2 tex2_lin(const f2 *t, int pitch, int w, int h, float x, float y)
{
    float fx = floorf(x), fy = floorf(y);
    float a = x - fx, b = y - fy;
    int x0 = (int)fx, y0 = (int)fy;
    int xa = clampi(x0, 0, w - 1), xb = clampi(x0 + 1, 0, w - 1);
    int ya = clampi(y0, 0, h - 1), yb = clampi(y0 + 1, 0, h - 1);
    float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b);
    float w01 = (1.0f - a) * b, w11 = a * b;
    f2 t00 = t[ya * pitch + xa], t10 = t[ya * pitch + xb];
    f2 t01 = t[yb * pitch + xa], t11 = t[yb * pitch + xb];
    f2 r;
    r.x = fmaf(w11, t11.x, fmaf(w01, t01.x, fmaf(w10, t10.x, w00 * t00.x)));
    r.y = fmaf(w11, t11.y, fmaf(w01, t01.y, fmaf(w10, t10.y, w00 * t00.y)));
    return r;
}

static inline float tex1_lin(const float *t, int pitch, int w, int h, float x, float y)
{
    float fx = floorf(x), fy = floorf(y);
    float a = x - fx, b = y - fy;
    int x0 = (int)fx, y0 = (int)fy;
    int xa = clampi(x0, 0, w - 1), xb = clampi(x0 + 1, 0, w - 1);
    int ya = clampi(y0, 0, h - 1), yb = clampi(y0 + 1, 0, h - 1);
    float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b);
    float w01 = (1.0f - a) * b, w11 = a * b;
    return fmaf(w11, t[yb * pitch + xb], fmaf(w01, t[yb * pitch + xa],
                fmaf(w10, t[ya * pitch + xb], w00 * t[ya * pitch + xa])));
}

static inline float tex1_point(const float *t, int w, int h, float x, float y)
{
    int xi = clampi((int)floorf(x + 0.5f), 0, w - 1);
    int yi = clampi((int)floorf(y + 0.5f), 0, h - 1);
    return t[yi * w + xi];
}

static inline int point_index(int w, int h, float x, float y)
{
    int xi = clampi((int)floorf(x + 0.5f), 0, w - 1);
    int yi = clampi((int)floorf(y + 0.5f), 0, h - 1);
    return yi * w + xi;
}

static inline f4 tex4_point(const f4 *t, int w, int h, float x, float y)
{
    int xi = clampi((int)floorf(x + 0.5f), 0, w - 1);
    int yi = clampi((int)floorf(y + 0.5f), 0, h - 1);
    return t[yi * w + xi];
}

/* Exact texel at integer coordinates (a linear fetch at an integer position
 * has weights (1,0,0,0)). */
static inline f4 img_color_at(const ptref_ctx *c, int layer, int x, int y)
{
    return c->img_color[layer][clampi(y, 0, c->H - 1) * c->W + clampi(x, 0, c->W - 1)];
}
static inline float img_dt_at(const ptref_ctx *c, int layer, int x, int y)
{
    return c->img_dt[layer][clampi(y, 0, c->H - 1) * c->W + clampi(x, 0, c->W - 1)];
}
static inline f2 img_scale_at(const ptref_ctx *c, float x, float y)
{
    return tex2_lin(c->img_scale, c->scw, c->scw, c->sch, x, y);
}

/* Style layer `l` of the exemplar textures: a layered texture clamps the
 * layer index into its range. */
static inline size_t ex_layer(const ptref_ctx *c, int l)
{
    l = l < 0 ? 0 : (l >= c->nl ? c->nl - 1 : l);
    return (size_t)l * c->ew * c->eh;
}
#define EXB(c, l) ((c)->ex_base + ex_layer((c), (l)))
#define EXO(c, l) ((c)->ex_out + ex_layer((c), (l)))
#define EXOR(c, l) ((c)->ex_orient + ex_layer((c), (l)))
#define EXDT(c, l) ((c)->ex_dt + ex_layer((c), (l)))
#define EXID(c, l) ((c)->ex_id + ex_layer((c), (l)))

/* ------------------------------------------------------------------ */
/* helpers (texSynth_util.h)                                           */
/* ------------------------------------------------------------------ */
static inline int in_bounds(int x, int y, int w, int h, int border)
{
    return x >= border && y >= border && x < w - border && y < h - border;
}

/* clampToExemplar, float version (texSynth_util.h:50-57) */
static inline void clamp_ex(float *x, float *y, int w, int h)
{
    if (*x < 0) *x = 0;
    if (*y < 0) *y = 0;
    if (*x >= (float)w) *x = (float)(w - 1);
    if (*y >= (float)h) *y = (float)(h - 1);
}

/* sumSquaredDiff with alpha weight (texSynth_util.h:69-76) */
static inline float ssd_alpha(f4 a, f4 b, float alpha)
{
    float d0 = a.x - b.x, d1 = a.y - b.y, d2 = a.z - b.z, d3 = a.w - b.w;
    float s = d0 * d0;
    s = fmaf(d1, d1, s);
    s = fmaf(d2, d2, s);
    s = fmaf(alpha * d3, d3, s);
    return s;
}

/* sumSquaredDiff without alpha weight (texSynth_util.h:78-85) */
static inline float ssd4(f4 a, f4 b)
{
    float d0 = a.x - b.x, d1 = a.y - b.y, d2 = a.z - b.z, d3 = a.w - b.w;
    float s = d0 * d0;
    s = fmaf(d1, d1, s);
    s = fmaf(d2, d2, s);
    s = fmaf(d3, d3, s);
    return s;
}

static inline f4 sub4(f4 a, f4 b)
{
    f4 r = {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w};
    return r;
}

/* screenToExemplarBasis (texSynth_util.h:95-104) */
static inline void basis(float angle, float su, float sv, float *b0x, float *b0y, float *b1x, float *b1y)
{
    float s, co;
    pt_sincosf(angle, &s, &co);
    *b1x = -s * sv;
    *b1y = co * sv;
    *b0x = co * su;
    *b0y = s * su;
}

/* calculateNewExemplarCoord (texSynth_util.h:106-118) */
static inline void new_coord(float i, float j, float ex, float ey, float b0x, float b0y,
                             float b1x, float b1y, float *ox, float *oy)
{
    float cx = i * b0x + j * b1x;
    float cy = i * b0y + j * b1y;
    *ox = ex + cx;
    *oy = ey + cy;
}

/* screenToExemplarAngle (texSynth_kernel.cu:425-436), exemplar orientation
 * of style layer `layer` */
static inline float screen_to_ex_angle_l(const ptref_ctx *c, float x, float y, float ex, float ey, int layer)
{
    if (c->p.orientation_enabled) {
        float theta = tex1_point(EXOR(c, layer), c->ew, c->eh, ex, ey) -
                      tex1_point(c->img_orient, c->ow, c->oh, x, y);
        if ((double)theta > 1.57079632679489661923)
            return (float)((double)theta - 3.14159265358979323846);
        if ((double)theta < -1.57079632679489661923)
            return (float)(3.14159265358979323846 + (double)theta);
        return theta;
    }
    return 0.0f;
}

static inline float screen_to_ex_angle(const ptref_ctx *c, float x, float y, float ex, float ey)
{
    return screen_to_ex_angle_l(c, x, y, ex, ey, 0);
}

/* isImagePixelCulled (texSynth_kernel.cu:176-186) */
static inline int image_culled(const ptref_ctx *c, int x, int y)
{
    float d = img_dt_at(c, c->cur, x, y);
    return d > ((float)c->p.dist_trans_cull - 0.0f);
}

/* isExemplarPixelCulled (texSynth_kernel.cu:200-205): int parameters, so
 * the position is truncated toward zero and the texel there read exactly
 * (clamp addressing). */
static inline int exemplar_culled(const ptref_ctx *c, float x, float y, int layer)
{
    float d = EXDT(c, layer)[clampi((int)y, 0, c->eh - 1) * c->ew + clampi((int)x, 0, c->ew - 1)];
    return (double)d > (double)c->p.dist_trans_cull * 1.25;
}

/* hysteresisTerm (texSynth_kernel.cu:506-535) */
static float hysteresis_term(const ptref_ctx *c, int x, int y, int level)
{
    /* powf(2, level) of the reference is exactly 2^level */
    float maxD = c->p.max_distortion / (float)(1 << level);
    const f2 *v = (c->p.direction == PT_FORWARD) ? c->img_velb[c->cur] : c->img_velf[c->cur];
    f2 d00 = tex2_lin(v, c->W, c->W, c->H, (float)x, (float)y);
    f2 d10 = tex2_lin(v, c->W, c->W, c->H, (float)x + 1.f, (float)y);
    f2 d01 = tex2_lin(v, c->W, c->W, c->H, (float)x, (float)y + 1.f);
    f2 d11 = tex2_lin(v, c->W, c->W, c->H, (float)x + 1.f, (float)y + 1.f);
    float maxX = fmaxf(d00.x, fmaxf(d01.x, fmaxf(d10.x, d11.x)));
    float maxY = fmaxf(d00.y, fmaxf(d01.y, fmaxf(d10.y, d11.y)));
    float minX = fminf(d00.x, fminf(d01.x, fminf(d10.x, d11.x)));
    float minY = fminf(d00.y, fminf(d01.y, fminf(d10.y, d11.y)));
    if ((maxX - minX) < maxD && (maxY - minY) < maxD)
        return c->p.hysteresis_weight;
    return 0.f;
}

/* timeDerivativeTerms (texSynth_kernel.cu:438-469) */
/* RGBtoLAB (texSynth_util.h:220-249), alpha unchanged */
static f4 lab4(f4 v)
{
    f4 o;
    pt_rgb_to_lab(v.x, v.y, v.z, &o.x, &o.y, &o.z);
    o.w = v.w;
    return o;
}

static float time_terms(const ptref_ctx *c, const adv_t *a, nnf prev, float pb0x, float pb0y,
                        float pb1x, float pb1y, f4 base, f4 test_base, f4 canvas, f4 test_out,
                        int x, int y, int i, int j)
{
    int o = (y + j) * c->W + (x + i);
    f4 prev_base = a->base[o];
    f4 prev_out = a->out[o];
    float ptx, pty;
    new_coord((float)i, (float)j, prev.x, prev.y, pb0x, pb0y, pb1x, pb1y, &ptx, &pty);
    f4 peb = tex4_lin(EXB(c, prev.la), c->ew, c->ew, c->eh, ptx, pty);
    f4 peo = tex4_lin(EXO(c, prev.la), c->ew, c->ew, c->eh, ptx, pty);
    if (c->p.colorspace == 1) { /* texSynth_kernel.cu:454-459 */
        prev_base = lab4(prev_base);
        prev_out = lab4(prev_out);
        peb = lab4(peb);
        peo = lab4(peo);
    }
    float r = ssd4(sub4(base, prev_base), sub4(test_base, peb)) * c->p.time_derivative_input_weight;
    r = fmaf(ssd4(sub4(canvas, prev_out), sub4(test_out, peo)), c->p.time_derivative_output_weight, r);
    r = fmaf(ssd4(prev_out, test_out), c->p.temporal_coherence_weight, r);
    return r;
}

/* calculateResidual (texSynth_kernel.cu:537-659): one patch distance. */
static float calc_residual(const ptref_ctx *c, int x, int y, int level, nnf t)
{
    const pt_params *p = &c->p;
    int lw = c->W >> level, lh = c->H >> level;
    int hw = c->SW >> level, hh = (c->SH * p->num_styles) >> level;
    int o = y * c->W + x;
    nnf pf = c->adv[0].off[o], pb = c->adv[1].off[o];
    int cpv = pf.hy > 0, cnv = pb.hy > 0;
    int any_time_w = p->time_derivative_input_weight > 0.f || p->time_derivative_output_weight > 0.f ||
                     p->temporal_coherence_weight > 0.f;
    int fwd_dir = p->direction == PT_FORWARD || p->direction == PT_BIDIRECTIONAL;
    int bwd_dir = p->direction == PT_BACKWARD || p->direction == PT_BIDIRECTIONAL;
    float fb0x = 0, fb0y = 0, fb1x = 0, fb1y = 0, bb0x = 0, bb0y = 0, bb1x = 0, bb1y = 0;
    if (cpv) {
        float ang = screen_to_ex_angle_l(c, (float)x, (float)y, pf.x, pf.y, pf.la);
        basis(ang, p->residual_window_scale_u * pf.su, p->residual_window_scale_v * pf.sv,
              &fb0x, &fb0y, &fb1x, &fb1y);
    }
    if (cnv) {
        float ang = screen_to_ex_angle_l(c, (float)x, (float)y, pb.x, pb.y, pb.la);
        basis(ang, p->residual_window_scale_u * pb.su, p->residual_window_scale_v * pb.sv,
              &bb0x, &bb0y, &bb1x, &bb1y);
    }
    float sum_w = 0.f, sum_r = 0.f;
    if (cpv)
        sum_r -= hysteresis_term(c, x, y, level);
    float b0x, b0y, b1x, b1y;
    basis(t.th, p->residual_window_scale_u * t.su, p->residual_window_scale_v * t.sv, &b0x, &b0y, &b1x, &b1y);
    int r = p->residual_window_size;
    for (int i = -r; i <= r; i++) {
        for (int j = -r; j <= r; j++) {
            if (!in_bounds(x + i, y + j, lw, lh, 0))
                continue;
            int oo = (y + j) * c->W + (x + i);
            f4 base = img_color_at(c, c->cur, x + i, y + j);
            f4 canvas = c->canvas[oo];
            if (p->colorspace == 1) { /* texSynth_kernel.cu:575-578 */
                base = lab4(base);
                canvas = lab4(canvas);
            }
            int pv = cpv && c->adv[0].off[oo].hy > 0;
            int nv = cnv && c->adv[1].off[oo].hy > 0;
            float tx, ty;
            new_coord((float)i, (float)j, t.x, t.y, b0x, b0y, b1x, b1y, &tx, &ty);
            if (tx >= (float)hw) tx -= (float)hw;
            else if (tx < 0) tx += (float)hw;
            if (ty >= (float)hh) ty -= (float)hh;
            else if (ty < 0) ty += (float)hh;
            f4 tb = tex4_lin(EXB(c, t.la), c->ew, c->ew, c->eh, tx, ty);
            f4 to = tex4_lin(EXO(c, t.la), c->ew, c->ew, c->eh, tx, ty);
            if (p->colorspace == 1) { /* texSynth_kernel.cu:603-606 */
                tb = lab4(tb);
                to = lab4(to);
            }
            float res = ssd_alpha(base, tb, p->alpha_weight) * p->input_analogy_weight;
            if (fwd_dir && pv && any_time_w)
                res = res + time_terms(c, &c->adv[0], pf, fb0x, fb0y, fb1x, fb1y, base, tb, canvas, to, x, y, i, j);
            if (bwd_dir && nv && any_time_w)
                res = res + time_terms(c, &c->adv[1], pb, bb0x, bb0y, bb1x, bb1y, base, tb, canvas, to, x, y, i, j);
            if (i == 0 && j == 0) {
                if (p->dist_trans_weight > 0.f) {
                    /* distTransformTerm (texSynth_kernel.cu:471-485) */
                    float de = tex1_lin(EXDT(c, t.la), c->ew, c->ew, c->eh, tx, ty);
                    de /= t.sv;
                    float di = img_dt_at(c, c->cur, x, y);
                    float diff = fabsf(de - di);
                    diff = fminf(fmaxf(diff, 0.0f), 10.0f);
                    res = res + diff * p->dist_trans_weight;
                }
                /* objectIdTerm (texSynth_kernel.cu:497-504, 637-638): point
                 * fetches, the image ID at the pixel, the style's at the tap */
                if (p->style_obj_id_weight > 0.f) {
                    int iid = c->img_id[point_index(c->idw, c->idh, (float)x, (float)y)];
                    int sid = EXID(c, t.la)[point_index(c->ew, c->eh, tx, ty)];
                    if (iid != sid) res = res + p->style_obj_id_weight;
                }
            } else {
                res = fmaf(ssd_alpha(canvas, to, p->alpha_weight), p->canvas_output_weight, res);
                /* coherence term (texSynth_kernel.cu:642-646) */
                if (p->coherence_weight > 0.f) {
                    const nnf *cn = &c->off_base[oo];
                    float ct = pt_coherence_term(cn->x, cn->y, t.x, t.y, tx, ty, i, j, level,
                                                 p->coherence_ratio_range, p->coherence_angle_range);
                    res = res + ct * p->coherence_weight;
                }
            }
            float wgt = c->gauss[i + r] * c->gauss[j + r];
            sum_r = fmaf(res, wgt, sum_r);
            sum_w += wgt;
        }
    }
    if (sum_w < 0.000001f)
        return 1e31f;
    return sum_r / sum_w;
}

/* test_offset (texSynth_kernel.cu:695-721) */
static float test_offset(const ptref_ctx *c, int x, int y, nnf t, int level, float image_alpha)
{
    if (!exemplar_culled(c, t.x, t.y, t.la)) {
        if (!c->p.transparency_ok) {
            float ea = tex4_lin(EXO(c, t.la), c->ew, c->ew, c->eh, t.x, t.y).w;
            if (ea < 0.99f && image_alpha > 0.99f)
                return 1e31f;
        }
        return calc_residual(c, x, y, level, t);
    }
    return 1e31f;
}

static void ensure_threads(const ptref_ctx *c)
{
#ifdef _OPENMP
    if (c->nthreads > 0) omp_set_num_threads(c->nthreads);
#else
    (void)c;
#endif
}

/* ------------------------------------------------------------------ */
/* context, setup, buffers                                             */
/* ------------------------------------------------------------------ */
/* NLParameters::getTsDefaultParams (G/nlParameters.cpp:63-101), with
 * distTransCull 50 (synthesisProcessor.cpp:220) and direction FORWARD. */
void ptref_params_default(pt_params *p)
{
    memset(p, 0, sizeof(*p));
    p->residual_window_size = 4;
    p->coherence_weight = 0.0f;
    p->dist_trans_weight = 0.0f;
    p->lum_shift_weight = 0.0f;
    p->lum_shift_enable = 0;
    p->input_analogy_weight = 1.0f;
    p->canvas_output_weight = 1.0f;
    p->orientation_enabled = 1;
    p->num_styles = 1;
    p->max_distortion = 4.0f;
    p->alpha_weight = 1.0f;
    p->transparency_ok = 0;
    p->coherence_ratio_range = 2.0f;
    p->coherence_angle_range = 0.25f;
    p->levels = 4;
    p->scatter_samples_count = 25;
    p->advection_jitter = 0.0f;
    p->time_derivative_input_weight = 1.f;
    p->time_derivative_output_weight = 1.f;
    p->temporal_coherence_weight = 0.2f;
    p->hysteresis_weight = 0.f;
    p->interpolate_key_frame = 0;
    p->use_guide = 0;
    p->dist_trans_cull = 50;
    p->residual_window_scale_u = 1;
    p->residual_window_scale_v = 1;
    p->first_key_index = -1;
    p->last_key_index = -1;
    p->style_obj_id_weight = 0.f;
    p->direction = PT_FORWARD;
    p->colorspace = 0;
    p->rng_mode = PT_RNG_REFERENCE;
    p->seed = 0;
    p->frame_number = 0;
}

int ptref_create(ptref_ctx **out, int num_threads)
{
    if (!out) return PT_ERR_INVALID_ARG;
    ptref_ctx *c = (ptref_ctx *)calloc(1, sizeof(ptref_ctx));
    if (!c) return PT_ERR_OUT_OF_MEMORY;
    c->nthreads = num_threads;
    *out = c;
    return PT_OK;
}

static void free_buffers(ptref_ctx *c)
{
    free(c->off_base); free(c->off_work); free(c->res_base); free(c->res_work); free(c->canvas);
    for (int k = 0; k < 2; k++) {
        free(c->adv[k].off); free(c->adv[k].base); free(c->adv[k].out); free(c->adv[k].res); free(c->adv[k].ts);
        free(c->img_color[k]); free(c->img_dt[k]); free(c->img_velb[k]); free(c->img_velf[k]);
        memset(&c->adv[k], 0, sizeof(adv_t));
        c->img_color[k] = NULL; c->img_dt[k] = NULL; c->img_velb[k] = NULL; c->img_velf[k] = NULL;
    }
    c->off_base = c->off_work = NULL; c->res_base = c->res_work = NULL; c->canvas = NULL;
    free(c->rib_base); free(c->rib_work);
    c->rib_base = c->rib_work = NULL;
}

static void free_sequence(ptref_ctx *c);

int ptref_destroy(ptref_ctx *c)
{
    if (!c) return PT_OK;
    free_sequence(c);
    free_buffers(c);
    free(c->img_orient); free(c->img_scale); free(c->random); free(c->img_id);
    free(c->ex_base); free(c->ex_out); free(c->ex_orient); free(c->ex_dt); free(c->ex_id);
    for (int k = 0; k < 2; k++) { free(c->cumul[k]); free(c->ftk[k]); }
    free(c);
    return PT_OK;
}

const char *ptref_last_error(ptref_ctx *c) { return c ? c->err : ""; }

static int check_supported(ptref_ctx *c, const pt_params *p)
{
    if (p->residual_window_size < 1 || p->residual_window_size > 15)
        return fail(c, PT_ERR_INVALID_ARG, "residual_window_size must be in [1, 15]");
    if (p->num_styles < 1 || p->num_styles > 4)
        return fail(c, PT_ERR_INVALID_ARG, "num_styles must be in [1, 4]");
    if (p->coherence_weight > 0.f && !(p->coherence_ratio_range > 1.f && p->coherence_angle_range > 0.f))
        return fail(c, PT_ERR_INVALID_ARG, "coherence needs ratio range > 1 and angle range > 0");
    if (p->colorspace != 0 && p->colorspace != 1)
        return fail(c, PT_ERR_UNSUPPORTED, "colourspace must be RGB (0) or LAB (1)");
    if (p->interpolate_key_frame &&
        (p->first_key_index < 0 || p->last_key_index < p->first_key_index || p->last_key_index >= PT_MAX_STYLE_LAYERS))
        return fail(c, PT_ERR_INVALID_ARG, "keyframe interpolation needs 0 <= first_key_index <= last_key_index < %d",
                    PT_MAX_STYLE_LAYERS);
    /* The offsets-histogram term (histogramTerm, texSynth_kernel.cu:486-494)
     * reads a histogram that stays all zero: its accumulation is commented out
     * (:258), the buffer is reset before it and the two blur passes of zeros
     * give zeros (:1250-1276).  The term is therefore the constant
     * (0 > threshold) ? (0 - threshold) * slope : 0 of the candidate's style
     * layer, which is layer 0 without keyframes; it is exactly 0 for every
     * slope when that threshold is not negative (the GUI's range is
     * [0, 1000], nlDialsAndKnobs.cpp:279), and adding 0 to a non-negative
     * float sum changes no bit.  A negative layer-0 threshold with a nonzero
     * slope is not implemented. */
    {
        float thr = p->offsets_histogram_threshold[0], slope = p->offsets_histogram_slope[0];
        float term = (0.f > thr) ? (0.f - thr) * slope : 0.f;
        if (term != 0.f)
            return fail(c, PT_ERR_UNSUPPORTED, "offsets histogram with a negative threshold is not supported");
    }
    if (p->rng_mode != PT_RNG_REFERENCE && p->rng_mode != PT_RNG_COUNTER)
        return fail(c, PT_ERR_INVALID_ARG, "unknown rng_mode");
    if (p->first_pass_direction != 0 && p->first_pass_direction != 1)
        return fail(c, PT_ERR_INVALID_ARG, "first_pass_direction must be 0 or 1");
    if (p->max_level < 0 || p->max_level >= p->levels)
        return fail(c, PT_ERR_INVALID_ARG, "max_level must be in [0, levels)");
    return PT_OK;
}

/* allocateGaussianKernel(r, normalize=false) (convolution.cu:37-60) */
static void gaussian(float *k, int r)
{
    float sigma = (float)(r) / 3.0;
    float sigma2 = sigma * sigma;
    for (int i = 0; i <= r; i++) {
        float dist2 = (float)(i * i);
        float w = expf(-dist2 / (2.f * sigma2));
        k[r + i] = w;
        k[r - i] = w;
    }
}

int ptref_upload_parameters(ptref_ctx *c, const pt_params *p)
{
    if (!c || !p) return PT_ERR_INVALID_ARG;
    int rc = check_supported(c, p);
    if (rc) return rc;
    c->p = *p;
    c->have_params = 1;
    if (p->lum_shift_enable) c->lum_live = 1;
    return PT_OK;
}

int ptref_setup(ptref_ctx *c, const pt_params *p)
{
    int rc = ptref_upload_parameters(c, p);
    if (rc) return rc;
    if (c->gauss_r != p->residual_window_size) {
        gaussian(c->gauss, p->residual_window_size);
        c->gauss_r = p->residual_window_size;
    }
    return PT_OK;
}

int ptref_gaussian_weights(ptref_ctx *c, float *out, int n)
{
    if (!c || !out || n != 2 * c->gauss_r + 1) return PT_ERR_INVALID_ARG;
    memcpy(out, c->gauss, sizeof(float) * n);
    return PT_OK;
}

int ptref_initialize(ptref_ctx *c, int iw, int ih, int sw, int sh)
{
    if (!c || iw < 2 || ih < 2 || sw < 2 || sh < 2) return PT_ERR_INVALID_ARG;
    if (c->off_base && c->W == iw && c->H == ih && c->SW == sw && c->SH == sh)
        return PT_OK;
    free_buffers(c);
    size_t n = (size_t)iw * ih;
    c->W = iw; c->H = ih; c->SW = sw; c->SH = sh;
    c->off_base = (nnf *)calloc(n, sizeof(nnf));
    c->off_work = (nnf *)calloc(n, sizeof(nnf));
    c->res_base = (float *)calloc(n, sizeof(float));
    c->res_work = (float *)calloc(n, sizeof(float));
    c->canvas = (f4 *)calloc(n, sizeof(f4));
    for (int k = 0; k < 2; k++) {
        c->adv[k].off = (nnf *)calloc(n, sizeof(nnf));
        c->adv[k].base = (f4 *)calloc(n, sizeof(f4));
        c->adv[k].out = (f4 *)calloc(n, sizeof(f4));
        c->adv[k].res = (float *)calloc(n, sizeof(float));
        c->adv[k].ts = (int32_t *)calloc(n, sizeof(int32_t));
        c->img_color[k] = (f4 *)calloc(n, sizeof(f4));
        c->img_dt[k] = (float *)calloc(n, sizeof(float));
        c->img_velb[k] = (f2 *)calloc(n, sizeof(f2));
        c->img_velf[k] = (f2 *)calloc(n, sizeof(f2));
        if (!c->adv[k].ts || !c->img_velf[k]) return fail(c, PT_ERR_OUT_OF_MEMORY, "out of memory");
    }
    if (!c->canvas) return fail(c, PT_ERR_OUT_OF_MEMORY, "out of memory");
    c->rib_base = (ribbon *)malloc(n * sizeof(ribbon));
    c->rib_work = (ribbon *)malloc(n * sizeof(ribbon));
    if (!c->rib_base || !c->rib_work) return fail(c, PT_ERR_OUT_OF_MEMORY, "out of memory");
    for (size_t i = 0; i < n; i++) {
        ribbon r0 = {0.f, 0.f, 0, -1}; /* RibbonP(): invalid */
        c->rib_base[i] = r0;
        c->rib_work[i] = r0;
    }
    for (int k = 0; k < 2; k++) { /* unbound frame-to-key ribbons read as 0 */
        free(c->ftk[k]);
        c->ftk[k] = (f4 *)calloc(1, sizeof(f4));
        c->fw[k] = c->fh[k] = 1;
    }
    c->lum_live = c->have_params && c->p.lum_shift_enable;
    free(c->img_id); /* a fresh image-ID texture: one 0 texel until uploaded */
    c->img_id = (int32_t *)calloc(1, sizeof(int32_t));
    c->idw = c->idh = 1;
    c->cur = 0;
    return PT_OK;
}

int ptref_initialize_exemplars(ptref_ctx *c, int w, int h, int num_layers)
{
    if (!c || w < 1 || h < 1 || num_layers < 1 || num_layers > PT_MAX_STYLE_LAYERS) return PT_ERR_INVALID_ARG;
    if (c->ex_base && c->ew == w && c->eh == h && c->nl == num_layers) return PT_OK;
    free(c->ex_base); free(c->ex_out); free(c->ex_orient); free(c->ex_dt); free(c->ex_id);
    c->ew = w; c->eh = h; c->nl = num_layers;
    size_t n = (size_t)w * h * num_layers;
    c->ex_id = (int32_t *)calloc(n, sizeof(int32_t));
    c->ex_base = (f4 *)calloc(n, sizeof(f4));
    c->ex_out = (f4 *)calloc(n, sizeof(f4));
    c->ex_orient = (float *)calloc(n, sizeof(float));
    c->ex_dt = (float *)calloc(n, sizeof(float));
    return PT_OK;
}

/* TexSynth::initialize_cumulativeFields (texSynth_kernel.cu:2136-2140):
 * zero-filled layered fields. */
int ptref_initialize_cumulative_fields(ptref_ctx *c, int w, int h, int num_layers)
{
    if (!c || w < 1 || h < 1 || num_layers < 1 || num_layers > PT_MAX_STYLE_LAYERS) return PT_ERR_INVALID_ARG;
    for (int k = 0; k < 2; k++) {
        free(c->cumul[k]);
        c->cumul[k] = (f4 *)calloc((size_t)w * h * num_layers, sizeof(f4));
        if (!c->cumul[k]) return fail(c, PT_ERR_OUT_OF_MEMORY, "out of memory");
    }
    c->cw = w; c->ch = h; c->cl = num_layers;
    return PT_OK;
}

int ptref_reset_advected_buffers(ptref_ctx *c)
{
    if (!c || !c->off_base) return PT_ERR_STATE;
    size_t n = (size_t)c->W * c->H;
    for (int k = 0; k < 2; k++) {
        memset(c->adv[k].off, 0, n * sizeof(nnf));
        memset(c->adv[k].base, 0, n * sizeof(f4));
        memset(c->adv[k].out, 0, n * sizeof(f4));
        memset(c->adv[k].res, 0, n * sizeof(float));
        memset(c->adv[k].ts, 0, n * sizeof(int32_t));
    }
    return PT_OK;
}

int ptref_flip_anim_buffers(ptref_ctx *c)
{
    if (!c) return PT_ERR_INVALID_ARG;
    c->cur = (c->cur + 1) % 2;
    return PT_OK;
}

/* ------------------------------------------------------------------ */
/* pyramids (imagePyramid.cpp:81-106, downsample.cpp:41-160)           */
/* ------------------------------------------------------------------ */
static int elem_channels(int elem) { return elem == PT_ELEM_I32 ? 1 : (elem >= 1 && elem <= 4) ? elem : 0; }

/* downsampleVote (downsample.cpp:162-210): the most frequent of the four
 * IDs; a tie goes to the first of them in scan order (the reference's
 * winner among ties follows QHash iteration order). */
static void downsample_vote(const int32_t *prev, int pw, int32_t *dst, int lw, int lh)
{
    for (int y = 0; y < lh; y++)
        for (int x = 0; x < lw; x++) {
            int32_t v[4] = {prev[(2 * y) * pw + 2 * x], prev[(2 * y) * pw + 2 * x + 1],
                            prev[(2 * y + 1) * pw + 2 * x], prev[(2 * y + 1) * pw + 2 * x + 1]};
            int best = 0, best_n = 0;
            for (int k = 0; k < 4; k++) {
                int n = 0;
                for (int m = 0; m < 4; m++) n += v[m] == v[k];
                if (n > best_n) { best = k; best_n = n; }
            }
            dst[y * lw + x] = v[best];
        }
}

int ptref_pyramid_create(ptref_ctx *c, ptref_pyramid **out, int elem, int w, int h,
                         int scale_mode, const void *host0, float defv)
{
    int ch = elem_channels(elem);
    if (!out || !ch || w < 2 || h < 2) return fail(c, PT_ERR_INVALID_ARG, "bad pyramid arguments");
    ptref_pyramid *p = (ptref_pyramid *)calloc(1, sizeof(ptref_pyramid));
    p->elem = elem; p->w = w; p->h = h;
    int lv = (int)fmin(floor(log2f((float)w)), floor(log2f((float)h)));
    p->levels = lv;
    for (int l = 0; l < lv; l++) {
        p->lw[l] = w >> l;
        p->lh[l] = h >> l;
        p->data[l] = (float *)malloc(sizeof(float) * ch * p->lw[l] * p->lh[l]);
    }
    size_t n0 = (size_t)w * h * ch;
    if (host0) memcpy(p->data[0], host0, n0 * sizeof(float));
    else if (elem == PT_ELEM_I32) for (size_t i = 0; i < n0; i++) ((int32_t *)p->data[0])[i] = (int32_t)defv;
    else for (size_t i = 0; i < n0; i++) p->data[0][i] = defv;
    if (elem == PT_ELEM_I32) {
        for (int l = 1; l < lv; l++)
            downsample_vote((const int32_t *)p->data[l - 1], p->lw[l - 1], (int32_t *)p->data[l], p->lw[l], p->lh[l]);
        *out = p;
        return PT_OK;
    }
    if (scale_mode == PT_DOWNSAMPLE_RIBBON) {
        /* initializeRibbonPyramid (imagePyramid.cpp:211-235): the top-left
         * texel of each 2x2 block as RibbonP -- time step and layer through
         * int, x and y halved -- packed back as Color4(time_step, y, x, layer) */
        if (elem != PT_ELEM_F32X4) {
            ptref_pyramid_destroy(p);
            return fail(c, PT_ERR_INVALID_ARG, "ribbon pyramids are F32X4");
        }
        for (int l = 1; l < lv; l++) {
            int lw = p->lw[l], lh = p->lh[l], pw = p->lw[l - 1];
            const f4 *prev = (const f4 *)p->data[l - 1];
            f4 *dst = (f4 *)p->data[l];
            for (int y = 0; y < lh; y++)
                for (int x = 0; x < lw; x++) {
                    f4 q = prev[(y * 2) * pw + x * 2];
                    f4 r = {(float)(int32_t)q.x, q.y * 0.5f, q.z * 0.5f, (float)(int32_t)q.w};
                    dst[y * lw + x] = r;
                }
        }
        *out = p;
        return PT_OK;
    }
    for (int l = 1; l < lv; l++) {
        int lw = p->lw[l], lh = p->lh[l], pw = p->lw[l - 1];
        const float *prev = p->data[l - 1];
        float *dst = p->data[l];
        for (int y = 0; y < lh; y++)
            for (int x = 0; x < lw; x++)
                for (int k = 0; k < ch; k++) {
                    float s = 0.f;
                    s += prev[((y * 2) * pw + x * 2) * ch + k];
                    s += prev[((y * 2) * pw + x * 2 + 1) * ch + k];
                    s += prev[((y * 2 + 1) * pw + x * 2) * ch + k];
                    s += prev[((y * 2 + 1) * pw + x * 2 + 1) * ch + k];
                    s *= 0.25f;
                    if (scale_mode == PT_DOWNSAMPLE_SCALE_WITH_LEVELS)
                        s *= 0.5f;
                    dst[(y * lw + x) * ch + k] = s;
                }
    }
    *out = p;
    return PT_OK;
}

/* NLDataAccess::toOrientationPyramid (G/nlDataAccess.cpp:82-93) */
int ptref_pyramid_orientation(ptref_ctx *c, ptref_pyramid **out, const ptref_pyramid *t)
{
    if (!out || !t || t->elem != PT_ELEM_F32X3) return fail(c, PT_ERR_INVALID_ARG, "need a tensor pyramid");
    ptref_pyramid *p = (ptref_pyramid *)calloc(1, sizeof(ptref_pyramid));
    *p = *t;
    p->elem = PT_ELEM_F32;
    for (int l = 0; l < t->levels; l++) {
        size_t n = (size_t)t->lw[l] * t->lh[l];
        p->data[l] = (float *)malloc(sizeof(float) * n);
        for (size_t i = 0; i < n; i++)
            p->data[l][i] = pt_tensor_to_angle(t->data[l][3 * i], t->data[l][3 * i + 1], t->data[l][3 * i + 2]);
    }
    *out = p;
    return PT_OK;
}

int ptref_pyramid_destroy(ptref_pyramid *p)
{
    if (!p) return PT_OK;
    for (int l = 0; l < p->levels; l++) free(p->data[l]);
    free(p);
    return PT_OK;
}

int ptref_pyramid_levels(const ptref_pyramid *p) { return p ? p->levels : 0; }

int ptref_pyramid_dims(const ptref_pyramid *p, int level, int *w, int *h)
{
    if (!p || level < 0 || level >= p->levels) return PT_ERR_INVALID_ARG;
    *w = p->lw[level];
    *h = p->lh[level];
    return PT_OK;
}

int ptref_pyramid_download(const ptref_pyramid *p, int level, void *host)
{
    if (!p || level < 0 || level >= p->levels || !host) return PT_ERR_INVALID_ARG;
    memcpy(host, p->data[level], sizeof(float) * elem_channels(p->elem) * p->lw[level] * p->lh[level]);
    return PT_OK;
}

/* copy a level into the top-left rectangle of a level-0-sized layer */
static void copy_rect(float *dst, int dst_pitch_elems, const float *src, int w, int h, int ch)
{
    for (int y = 0; y < h; y++)
        memcpy(dst + (size_t)y * dst_pitch_elems * ch, src + (size_t)y * w * ch, sizeof(float) * w * ch);
}

static float *realloc_tex(float *old, size_t n)
{
    free(old);
    return (float *)calloc(n, sizeof(float));
}

/* TexSynth::uploadImage_* (texSynth_kernel.cu:2128-2253) */
int ptref_upload_image(ptref_ctx *c, int level, int layer, const ptref_pyramid *pyr, int style_layer)
{
    if (!c || !pyr || level < 0 || level >= pyr->levels) return fail(c, PT_ERR_INVALID_ARG, "bad upload");
    if (style_layer < 0 || style_layer >= PT_MAX_STYLE_LAYERS)
        return fail(c, PT_ERR_INVALID_ARG, "style layer must be in [0, %d]", PT_MAX_STYLE_LAYERS - 1);
    int w = pyr->lw[level], h = pyr->lh[level];
    const float *src = pyr->data[level];
    int prev = (c->cur + 1) % 2;
    switch (layer) {
    case PT_LAYER_INPUT_COLOR:
        if (pyr->elem != 4 || w > c->W || h > c->H) break;
        copy_rect((float *)c->img_color[c->cur], c->W, src, w, h, 4);
        return PT_OK;
    case PT_LAYER_INPUT_DIST_TRANS:
        if (pyr->elem != 1 || w > c->W || h > c->H) break;
        copy_rect(c->img_dt[c->cur], c->W, src, w, h, 1);
        return PT_OK;
    case PT_LAYER_INPUT_VEL_B:
    case PT_LAYER_INPUT_VEL_F:
    case PT_LAYER_INPUT_VEL_B_PREVIOUS:
    case PT_LAYER_INPUT_VEL_F_PREVIOUS: {
        if (pyr->elem != 2 || w > c->W || h > c->H) break;
        f2 *dst = layer == PT_LAYER_INPUT_VEL_B ? c->img_velb[c->cur]
                  : layer == PT_LAYER_INPUT_VEL_F ? c->img_velf[c->cur]
                  : layer == PT_LAYER_INPUT_VEL_B_PREVIOUS ? c->img_velb[prev] : c->img_velf[prev];
        copy_rect((float *)dst, c->W, src, w, h, 2);
        return PT_OK;
    }
    case PT_LAYER_INPUT_ORIENTATION:
        if (pyr->elem != 1) break;
        c->img_orient = realloc_tex(c->img_orient, (size_t)w * h);
        memcpy(c->img_orient, src, sizeof(float) * w * h);
        c->ow = w; c->oh = h;
        return PT_OK;
    case PT_LAYER_INPUT_SCALE:
        if (pyr->elem != 2) break;
        c->img_scale = (f2 *)realloc_tex((float *)c->img_scale, (size_t)w * h * 2);
        memcpy(c->img_scale, src, sizeof(float) * 2 * w * h);
        c->scw = w; c->sch = h;
        return PT_OK;
    case PT_LAYER_RANDOM:
        if (pyr->elem != 4) break;
        c->random = (f4 *)realloc_tex((float *)c->random, (size_t)w * h * 4);
        memcpy(c->random, src, sizeof(float) * 4 * w * h);
        c->rw = w; c->rh = h;
        return PT_OK;
    case PT_LAYER_INPUT_ID:
        if (pyr->elem != PT_ELEM_I32) break;
        c->img_id = (int32_t *)realloc_tex((float *)c->img_id, (size_t)w * h);
        memcpy(c->img_id, src, sizeof(int32_t) * w * h);
        c->idw = w; c->idh = h;
        return PT_OK;
    case PT_LAYER_EXEMPLAR_OBJECT_IDS:
        if (pyr->elem != PT_ELEM_I32 || !c->ex_id || w > c->ew || h > c->eh || style_layer >= c->nl) break;
        copy_rect((float *)EXID(c, style_layer), c->ew, src, w, h, 1);
        return PT_OK;
    case PT_LAYER_CUMUL_FIELD_B:
    case PT_LAYER_CUMUL_FIELD_F: {
        int k = layer == PT_LAYER_CUMUL_FIELD_B ? 0 : 1;
        if (pyr->elem != 4 || !c->cumul[k] || w > c->cw || h > c->ch || style_layer >= c->cl) break;
        copy_rect((float *)(c->cumul[k] + (size_t)style_layer * c->cw * c->ch), c->cw, src, w, h, 4);
        return PT_OK;
    }
    case PT_LAYER_FRAME_TO_KEY_RIBBON_B:
    case PT_LAYER_FRAME_TO_KEY_RIBBON_F: {
        int k = layer == PT_LAYER_FRAME_TO_KEY_RIBBON_B ? 0 : 1;
        if (pyr->elem != 4) break;
        c->ftk[k] = (f4 *)realloc_tex((float *)c->ftk[k], (size_t)w * h * 4);
        memcpy(c->ftk[k], src, sizeof(float) * 4 * w * h);
        c->fw[k] = w; c->fh[k] = h;
        return PT_OK;
    }
    case PT_LAYER_EXEMPLAR_BASE:
    case PT_LAYER_EXEMPLAR_OUTPUT:
    case PT_LAYER_EXEMPLAR_ORIENTATION:
    case PT_LAYER_EXEMPLAR_DIST_TRANS: {
        if (!c->ex_base || w > c->ew || h > c->eh || style_layer >= c->nl) break;
        int ch = (layer == PT_LAYER_EXEMPLAR_BASE || layer == PT_LAYER_EXEMPLAR_OUTPUT) ? 4 : 1;
        if (pyr->elem != ch) break;
        float *dst = layer == PT_LAYER_EXEMPLAR_BASE ? (float *)EXB(c, style_layer)
                     : layer == PT_LAYER_EXEMPLAR_OUTPUT ? (float *)EXO(c, style_layer)
                     : layer == PT_LAYER_EXEMPLAR_ORIENTATION ? EXOR(c, style_layer) : EXDT(c, style_layer);
        copy_rect(dst, c->ew, src, w, h, ch);
        return PT_OK;
    }
    default:
        return fail(c, PT_ERR_UNSUPPORTED, "layer %d not supported", layer);
    }
    return fail(c, PT_ERR_INVALID_ARG, "pyramid kind or size does not fit layer %d", layer);
}

/* SynthesisFrame::createRandomPyramid (synthesisFrame.cpp:66-150), with
 * Philox4x32-10 draws in place of glibc rand(): key (seed, frame*level),
 * counter (x, y, trial, 0) for trial 0..49, (x, y, 50, 0) for the
 * fractional parts and, with keys, (x, y, 51, 0) for the style
 * (thisStyle = rand() % keys + first key, :107-110). */
static int random_pyramid(ptref_ctx *c, int level, int frame, const ptref_pyramid *icol, const ptref_pyramid *iori,
                          int ns, const ptref_pyramid *const *eouts, const ptref_pyramid *const *eoris, int keys)
{
    if (!c || !icol || !iori || ns < 1 || !eouts || !eoris) return PT_ERR_INVALID_ARG;
    if (level >= icol->levels || level >= iori->levels) return fail(c, PT_ERR_INVALID_ARG, "level out of range");
    for (int k = 0; k < ns; k++) {
        if (!eouts[k] || !eoris[k] || level >= eouts[k]->levels || level >= eoris[k]->levels)
            return fail(c, PT_ERR_INVALID_ARG, "style %d: level out of range", k);
        if (eouts[k]->lw[level] < 2 || eouts[k]->lh[level] < 2)
            return fail(c, PT_ERR_INVALID_ARG, "exemplar level too small");
    }
    int nk = c->p.last_key_index - c->p.first_key_index + 1, k0 = c->p.first_key_index;
    if (keys && (nk < 1 || k0 < 0 || k0 + nk > ns)) return fail(c, PT_ERR_INVALID_ARG, "keys outside the styles");
    int w = icol->lw[level], h = icol->lh[level];
    const f4 *col = (const f4 *)icol->data[level];
    const float *ior = iori->data[level];
    c->random = (f4 *)realloc_tex((float *)c->random, (size_t)w * h * 4);
    c->rw = w; c->rh = h;
    uint32_t key[2] = {c->p.seed, (uint32_t)(frame * level)};
    int transp = c->p.transparency_ok, orient = c->p.orientation_enabled;
    ensure_threads(c);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            uint32_t ctr[4] = {(uint32_t)x, (uint32_t)y, 0u, 0u}, o[4];
            int st = 0;
            if (keys) {
                ctr[2] = 51u;
                pt_philox4x32_10(ctr, key, o);
                st = pt_rand31(o[0]) % nk + k0;
            }
            int ew = eouts[st]->lw[level], eh = eouts[st]->lh[level];
            const f4 *eo = (const f4 *)eouts[st]->data[level];
            const float *eor = eoris[st]->data[level];
            int ox, oy, trial = 0;
            do {
                ctr[2] = (uint32_t)trial;
                pt_philox4x32_10(ctr, key, o);
                ox = pt_rand31(o[0]) % (ew - 1);
                oy = pt_rand31(o[1]) % (eh - 1);
                trial++;
            } while (!transp && col[y * w + x].w > 0.99f && eo[oy * ew + ox].w < 0.99f && trial < 50);
            float orientation = orient ? ior[y * w + x] - eor[oy * ew + ox] : 0.f;
            ctr[2] = 50u;
            pt_philox4x32_10(ctr, key, o);
            float rx = pt_rand_unit(o[0]);
            float ry = pt_rand_unit(o[1]);
            f4 v = {(float)ox + rx, (float)oy + ry, orientation, (float)st};
            c->random[y * w + x] = v;
        }
    }
    return PT_OK;
}

int ptref_create_random_pyramid(ptref_ctx *c, int level, int frame, const ptref_pyramid *icol,
                                const ptref_pyramid *iori, const ptref_pyramid *eout,
                                const ptref_pyramid *eori)
{
    return random_pyramid(c, level, frame, icol, iori, 1, &eout, &eori, 0);
}

int ptref_create_random_pyramid_keys(ptref_ctx *c, int level, int frame, const ptref_pyramid *icol,
                                     const ptref_pyramid *iori, int num_styles,
                                     const ptref_pyramid *const *eouts, const ptref_pyramid *const *eoris)
{
    return random_pyramid(c, level, frame, icol, iori, num_styles, eouts, eoris, 1);
}

/* ------------------------------------------------------------------ */
/* kernels                                                             */
/* ------------------------------------------------------------------ */
static int need_level(ptref_ctx *c, int level)
{
    if (!c || !c->off_base) return fail(c, PT_ERR_STATE, "not initialized");
    if (!c->have_params) return fail(c, PT_ERR_STATE, "no parameters");
    if (!c->ex_base) return fail(c, PT_ERR_STATE, "no exemplar");
    if (level < 0 || (c->W >> level) < 1 || (c->H >> level) < 1) return fail(c, PT_ERR_INVALID_ARG, "bad level");
    if (!c->img_orient || !c->img_scale) return fail(c, PT_ERR_STATE, "orientation/scale not uploaded");
    return PT_OK;
}

/* updateResidualCache_kernel (texSynth_kernel.cu:661-693) */
int ptref_update_residual_cache(ptref_ctx *c, int level)
{
    int rc = need_level(c, level);
    if (rc) return rc;
    int lw = c->W >> level, lh = c->H >> level;
    ensure_threads(c);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < lh; y++)
        for (int x = 0; x < lw; x++) {
            int o = y * c->W + x;
            if (image_culled(c, x, y)) { c->res_base[o] = 0.0f; continue; }
            c->res_base[o] = calc_residual(c, x, y, level, c->off_base[o]);
        }
    c->pde += (uint64_t)lw * lh;
    return PT_OK;
}

/* offsetsToOutput (texSynth_kernel.cu:263-274) */
static void offsets_to_output(ptref_ctx *c, int level)
{
    int lw = c->W >> level, lh = c->H >> level;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < lh; y++)
        for (int x = 0; x < lw; x++) {
            nnf n = c->off_base[y * c->W + x];
            f4 v = tex4_lin(EXO(c, n.la), c->ew, c->ew, c->eh, n.x, n.y);
            if (c->lum_live) { /* toColor4(output, luminanceShift) (texSynth_kernel.cu:272) */
                v.x = v.x + n.ly;
                v.y = v.y + n.ly;
                v.z = v.z + n.ly;
                v.w = v.w + 0.0f;
            }
            c->canvas[y * c->W + x] = v;
        }
}

/* TexSynth::syncBuffers (texSynth_kernel.cu:1240-1303) */
int ptref_sync_buffers(ptref_ctx *c, int level, int flags)
{
    int rc = need_level(c, level);
    if (rc) return rc;
    if (!(flags & PT_DONT_COPY_OFFSETS))
        memcpy(c->off_base, c->off_work, sizeof(nnf) * c->W * c->H);
    ensure_threads(c);
    offsets_to_output(c, level);
    if (!(flags & PT_DONT_UPDATE_RESIDUAL))
        return ptref_update_residual_cache(c, level);
    return PT_OK;
}

/* copyRandomKernel (texSynth_kernel.cu:2044-2059) + randomizeOffsets (:2100-2109) */
int ptref_randomize_offsets(ptref_ctx *c, int level)
{
    int rc = need_level(c, level);
    if (rc) return rc;
    if (!c->random) return fail(c, PT_ERR_STATE, "no random texture");
    int lw = c->W >> level, lh = c->H >> level;
    for (int y = 0; y < lh; y++)
        for (int x = 0; x < lw; x++) {
            /* toPatchXF(random): the style layer is the texel's w (texSynth_util.h:212-214) */
            f4 r = tex4_point(c->random, c->rw, c->rh, (float)x, (float)y);
            nnf n = {r.x, r.y, r.z, 0.f, 1.f, 1.f, 0.f, (int)r.w};
            c->off_work[y * c->W + x] = n;
        }
    return ptref_sync_buffers(c, level, 0);
}

/* pixelWrap (cudaImageBuffer.h:81-87): modulo the level-0 dimensions */
static inline nnf pixel_wrap(const ptref_ctx *c, const nnf *buf, int x, int y)
{
    int wx = (int)((float)x - (float)c->W * floorf((float)x / (float)c->W));
    int wy = (int)((float)y - (float)c->H * floorf((float)y / (float)c->H));
    return buf[wy * c->W + wx];
}

/* calculatePropagateOffset (texSynth_kernel.cu:956-982): the neighbour's
 * layer is kept */
static nnf propagate_offset(const ptref_ctx *c, int tx, int ty, float x, float y)
{
    nnf t = pixel_wrap(c, c->off_base, tx, ty);
    f2 sc = img_scale_at(c, (float)tx, (float)ty);
    float b0x, b0y, b1x, b1y;
    basis(t.th, c->p.residual_window_scale_u * sc.x, c->p.residual_window_scale_v * sc.y, &b0x, &b0y, &b1x, &b1y);
    float ex, ey;
    new_coord(x - (float)tx, y - (float)ty, t.x, t.y, b0x, b0y, b1x, b1y, &ex, &ey);
    clamp_ex(&ex, &ey, c->ew, c->eh);
    float ang = screen_to_ex_angle_l(c, x, y, ex, ey, t.la);
    f2 s2 = img_scale_at(c, x, y);
    nnf r = {ex, ey, ang, 0.f, s2.x, s2.y, 0.f, t.la};
    return r;
}

/* The keymask of a keyframe (texSynth_kernel.cu:1137-1142, 887-892): the
 * exemplar output of style layer key_index at the pixel's own position
 * (integer fetch: the texel itself, clamped); alpha > 0 marks a painted
 * keyframe pixel. */
static inline int keymask(const ptref_ctx *c, int key_index, int x, int y)
{
    if (key_index < 0) return 0;
    f4 m = EXO(c, key_index)[clampi(y, 0, c->eh - 1) * c->ew + clampi(x, 0, c->ew - 1)];
    return m.w > 0.f;
}

/* A cumulative key-to-key field texel at exemplar (ex, ey) of field k
 * (0 = gCumulFieldB, 1 = gCumulFieldF), layer `layer`: bilinear, clamped in
 * x, y and layer (unbound fields read 0). */
static inline f4 cumul_at(const ptref_ctx *c, int k, float ex, float ey, int layer)
{
    if (!c->cumul[k]) {
        f4 z = {0.f, 0.f, 0.f, 0.f};
        return z;
    }
    int l = layer < 0 ? 0 : (layer >= c->cl ? c->cl - 1 : layer);
    return tex4_lin(c->cumul[k] + (size_t)l * c->cw * c->ch, c->cw, c->cw, c->ch, ex, ey);
}

/* toRibbonP (texSynth_util.h:216-218): Color4(time_step, y, x, layer) */
static inline ribbon to_ribbon(f4 v)
{
    ribbon r = {v.z, v.y, (int32_t)v.x, (int32_t)v.w};
    return r;
}

/* The candidates of propagate_kernel_keys / scatter_kernel_keys that follow
 * candidate (ex, ey) of layer `elayer` through the cumulative fields to the
 * earlier keys (field B, layers elayer-1 down to firstKeyIndex) and to the
 * later keys (field F, layers elayer+1 up to lastKeyIndex), each tested in
 * turn (texSynth_kernel.cu:1061-1107, 792-834).  Propagation stops a walk
 * on a texel whose time-step channel is below -EPSILON, scatter on an
 * invalid ribbon (layer < 0). */
static void key_walks(const ptref_ctx *c, int x, int y, float ex0, float ey0, int elayer, int level,
                      float image_alpha, int scatter_rule, nnf *ow, float *best)
{
    for (int dir = 0; dir < 2; dir++) {
        float ex = ex0, ey = ey0;
        int j = dir == 0 ? elayer - 1 : elayer + 1;
        for (; dir == 0 ? j >= c->p.first_key_index : j <= c->p.last_key_index; j += dir == 0 ? -1 : 1) {
            f4 m = cumul_at(c, dir, ex, ey, dir == 0 ? j + 1 : j - 1);
            if (scatter_rule) {
                ribbon rb = to_ribbon(m);
                if (rb.layer < 0) break;
                ex += rb.x;
                ey += rb.y;
            } else {
                if (m.x < -0.000001f) break;
                ex += m.z;
                ey += m.y;
            }
            float ang = screen_to_ex_angle_l(c, (float)x, (float)y, ex, ey, j);
            nnf t = {ex, ey, ang, 0.f, 1.f, 1.f, 0.f, j};
            float r = test_offset(c, x, y, t, level, image_alpha);
            if (r < *best) { *ow = t; *best = r; }
        }
    }
}

/* propagate_kernel and propagate_kernel_keys (texSynth_kernel.cu:984-1219)
 * + TexSynth::propagate (:1306-1357) */
int ptref_propagate(ptref_ctx *c, int level, int jump, int key_index)
{
    int rc = need_level(c, level);
    if (rc) return rc;
    int lw = c->W >> level, lh = c->H >> level;
    const int keys = c->p.interpolate_key_frame != 0;
    uint64_t walks = 0;
    ensure_threads(c);
#pragma omp parallel for schedule(static) reduction(+ : walks)
    for (int y = 0; y < lh; y++)
        for (int x = 0; x < lw; x++) {
            int o = y * c->W + x;
            if (image_culled(c, x, y)) {
                nnf z = {0.f, 0.f, 0.f, 0.f, 1.f, 1.f, 0.f, 0};
                c->off_work[o] = z;
                continue;
            }
            if (keymask(c, key_index, x, y)) {
                nnf k = {(float)x, (float)y, 0.f, 1.f, 1.f, 1.f, 0.f, key_index};
                c->off_work[o] = k;
                continue;
            }
            float best = c->res_base[o];
            float ia = img_color_at(c, c->cur, x, y).w;
            nnf ow = c->off_base[o];
            for (int i = 0; i < 4; i++) {
                int tx = x, ty = y;
                if (i == 0) tx = x + jump;
                else if (i == 1) tx = x - jump;
                else if (i == 2) ty = y + jump;
                else ty = y - jump;
                nnf cand = propagate_offset(c, tx, ty, (float)x, (float)y);
                float r = test_offset(c, x, y, cand, level, ia);
                if (r < best) { ow = cand; best = r; }
                if (keys) {
                    key_walks(c, x, y, cand.x, cand.y, cand.la, level, ia, 0, &ow, &best);
                    walks++;
                }
            }
            /* The two advected candidates of propagate_kernel (:1184-1216) are
             * built with hysteresis 0 by calculatePropagateOffset and so never
             * tested; propagate_kernel_keys has none. */
            c->off_work[o] = ow;
        }
    (void)walks;
    c->pde += (uint64_t)lw * lh * 4;
    return ptref_sync_buffers(c, level, 0);
}

/* averageOffset (texSynth_kernel.cu:841-862): the mean position of the
 * in-bounds 8-neighbours of style layer `layer` */
static nnf average_offset(const ptref_ctx *c, int x, int y, int lw, int lh, int layer)
{
    nnf a = {0.f, 0.f, 0.f, 1.f, 1.f, 1.f, 0.f, layer};
    int sum = 0;
    for (int i = -1; i <= 1; i++)
        for (int j = -1; j <= 1; j++) {
            if (!in_bounds(x + i, y + j, lw, lh, 0) || (i == 0 && j == 0)) continue;
            nnf n = c->off_base[(y + j) * c->W + (x + i)];
            if (n.la != layer) continue;
            a.x += n.x;
            a.y += n.y;
            sum++;
        }
    a.x /= (float)sum;
    a.y /= (float)sum;
    return a;
}

static inline float u24(uint32_t w) { return (float)(w >> 8) * (1.0f / 16777216.0f); }

/* scatter_kernel and scatter_kernel_keys (texSynth_kernel.cu:723-839,
 * 865-952) + TexSynth::scatter (:1360-1401) */
int ptref_scatter(ptref_ctx *c, int level, int key_index)
{
    int rc = need_level(c, level);
    if (rc) return rc;
    int lw = c->W >> level, lh = c->H >> level;
    int ew = c->ew, eh = c->eh;
    const int keys = c->p.interpolate_key_frame != 0;
    int samples = c->p.scatter_samples_count;
    if (keys) {
        int nk = c->p.last_key_index - c->p.first_key_index + 1;
        if (nk < 1) return fail(c, PT_ERR_INVALID_ARG, "last_key_index below first_key_index");
        samples = samples / nk; /* the samples are shared by the keys (:765) */
    }
    uint32_t call = c->scatter_calls++;
    uint32_t key[2] = {c->p.seed ^ 0x5CA77E12u, (uint32_t)(c->p.frame_number * level)};
    int counter_rng = c->p.rng_mode == PT_RNG_COUNTER;
    ensure_threads(c);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < lh; y++)
        for (int x = 0; x < lw; x++) {
            int o = y * c->W + x;
            if (image_culled(c, x, y)) continue;
            nnf off = c->off_base[o];
            if (keymask(c, key_index, x, y)) { /* keep the current offset */
                c->off_work[o] = off;
                continue;
            }
            float cached = c->res_base[o];
            float best = (float)((double)cached + 10e-7);
            nnf ow = off;
            float ia = img_color_at(c, c->cur, x, y).w;
            float dist = (float)(ew > eh ? ew : eh);
            const float closest = 10.f, factor = 0.9f;
            f2 sc = img_scale_at(c, (float)x, (float)y);
            for (int i = 0; (i < samples) && (dist * factor >= closest); i++) {
                float r = 0.0f, theta = 0.0f;
                if (counter_rng) {
                    uint32_t ctr[4] = {(uint32_t)x, (uint32_t)y, call, (uint32_t)i}, rnd[4];
                    pt_philox4x32_10(ctr, key, rnd);
                    r = u24(rnd[0]);
                    theta = (float)((double)u24(rnd[1]) * 0.63661977236758134308);
                }
                float dmax = dist, dmin = dist * factor;
                float s, co;
                pt_sincosf(theta, &s, &co);
                float len = r * (dmax - dmin) + dmin;
                float ex = ow.x + co * len;
                float ey = ow.y + s * len;
                clamp_ex(&ex, &ey, ew, eh);
                if (keys) {
                    /* the sample stays in the winner's layer, scale (1, 1) */
                    int el = ow.la;
                    float ang = screen_to_ex_angle_l(c, (float)x, (float)y, ex, ey, el);
                    nnf cand = {ex, ey, ang, 0.f, 1.f, 1.f, 0.f, el};
                    float res = test_offset(c, x, y, cand, level, ia);
                    if (res < best) { ow = cand; best = res; }
                    key_walks(c, x, y, ex, ey, el, level, ia, 1, &ow, &best);
                } else {
                    float ang = screen_to_ex_angle(c, (float)x, (float)y, ex, ey);
                    nnf cand = {ex, ey, ang, 0.f, sc.x, sc.y, 0.f, 0};
                    float res = test_offset(c, x, y, cand, level, ia);
                    if (res < best) { ow = cand; best = res; }
                }
                dist *= factor;
            }
            if (!keys) { /* the averaged-neighbour candidate of scatter_kernel (:938-947) */
                nnf avg = average_offset(c, x, y, lw, lh, 0);
                avg.th = screen_to_ex_angle(c, (float)x, (float)y, avg.x, avg.y);
                float res = test_offset(c, x, y, avg, level, ia);
                if (res < best) { ow = avg; best = res; }
            }
            if (best < cached) c->off_work[o] = ow;
        }
    {
        /* nominal count: every sample tested plus the average candidate */
        float dist = (float)(ew > eh ? ew : eh);
        int n = 0;
        for (int i = 0; (i < samples) && (dist * 0.9f >= 10.f); i++) { n++; dist *= 0.9f; }
        c->pde += (uint64_t)lw * lh * (uint64_t)(n + (keys ? 0 : 1));
    }
    return ptref_sync_buffers(c, level, 0);
}

/* offsetUpsampleNearest (texSynth_kernel.cu:1403-1417) */
static void upsample_nearest(const ptref_ctx *c, nnf cx, float dx, float dy, float *ux, float *uy)
{
    float b0x, b0y, b1x, b1y;
    basis(cx.th, 1, 1, &b0x, &b0y, &b1x, &b1y);
    float ix, iy;
    new_coord(dx, dy, cx.x, cx.y, b0x, b0y, b1x, b1y, &ix, &iy);
    clamp_ex(&ix, &iy, c->ew, c->eh);
    *ux = (float)(((double)ix * 2.0) + 0.5);
    *uy = (float)(((double)iy * 2.0) + 0.5);
}

/* upsampleBilinearKernel (texSynth_kernel.cu:1453-1510) + TexSynth::upsample (:1513-1532) */
int ptref_upsample(ptref_ctx *c, int fine_level)
{
    int rc = need_level(c, fine_level);
    if (rc) return rc;
    int lw = c->W >> fine_level, lh = c->H >> fine_level;
    ensure_threads(c);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < lh; y++)
        for (int x = 0; x < lw; x++) {
            if (!in_bounds(x, y, lw, lh, 1)) continue;
            float cxf = (float)((double)x - 0.5) / 2.0f;
            float cyf = (float)((double)y - 0.5) / 2.0f;
            int lcx = (int)floorf(cxf), lcy = (int)floorf(cyf);
            nnf o00 = c->off_base[lcy * c->W + lcx];
            nnf o10 = c->off_base[lcy * c->W + lcx + 1];
            nnf o01 = c->off_base[(lcy + 1) * c->W + lcx];
            nnf o11 = c->off_base[(lcy + 1) * c->W + lcx + 1];
            float a = cxf - (float)lcx, b = cyf - (float)lcy;
            float w00 = (1 - a) * (1 - b), w10 = a * (1 - b), w01 = (1 - a) * b, w11 = a * b;
            float ix = ((w00 * o00.x + w10 * o10.x) + w01 * o01.x) + w11 * o11.x;
            float iy = ((w00 * o00.y + w10 * o10.y) + w01 * o01.y) + w11 * o11.y;
            float ux = ix * 2.0f + 0.5f, uy = iy * 2.0f + 0.5f;
            const float M = 4.0f * 4.0f;
#define D2(o) (((ix - (o).x) * (ix - (o).x)) + ((iy - (o).y) * (iy - (o).y)))
            if (D2(o00) > M || D2(o10) > M || D2(o01) > M || D2(o11) > M) {
                int rcx = (int)roundf(cxf), rcy = (int)roundf(cyf);
                nnf cx = c->off_base[rcy * c->W + rcx];
                upsample_nearest(c, cx, cxf - (float)rcx, cyf - (float)rcy, &ux, &uy);
            }
#undef D2
            float rw = c->res_base[lcy * c->W + lcx];
            float ang = screen_to_ex_angle_l(c, (float)x, (float)y, ux, uy, o00.la);
            f2 sc = img_scale_at(c, (float)x, (float)y);
            nnf n = {ux, uy, ang, 0.f, sc.x, sc.y, 0.f, o00.la};
            c->off_work[y * c->W + x] = n;
            c->res_work[y * c->W + x] = rw;
        }
    memcpy(c->off_base, c->off_work, sizeof(nnf) * c->W * c->H);
    memcpy(c->res_base, c->res_work, sizeof(float) * c->W * c->H);
    return PT_OK;
}

/* calculateTranslatedPosition (texSynth_kernel.cu:1541-1575) */
static int translated_position(const ptref_ctx *c, int fwd, int x, int y, float *tx, float *ty, int lw, int lh)
{
    int prev = (c->cur + 1) % 2;
    f2 forward, back;
    if (fwd) {
        back = tex2_lin(c->img_velb[c->cur], c->W, c->W, c->H, (float)x, (float)y);
        *tx = (float)x + back.x;
        *ty = (float)y + back.y;
        forward = tex2_lin(c->img_velf[prev], c->W, c->W, c->H, *tx, *ty);
    } else {
        forward = tex2_lin(c->img_velf[c->cur], c->W, c->W, c->H, (float)x, (float)y);
        *tx = (float)x + forward.x;
        *ty = (float)y + forward.y;
        back = tex2_lin(c->img_velb[prev], c->W, c->W, c->H, *tx, *ty);
    }
    const float valid_dist = 1.414 * 2;
    forward.x = -forward.x;
    forward.y = -forward.y;
    float dist = sqrtf((forward.x - back.x) * (forward.x - back.x) + (forward.y - back.y) * (forward.y - back.y));
    return in_bounds((int)ceilf(*tx), (int)ceilf(*ty), lw, lh, 0) &&
           in_bounds((int)floorf(*tx), (int)floorf(*ty), lw, lh, 0) && (dist < valid_dist);
}

/* A frame-to-key ribbon texel (k = 0: gFrameToKeyRibbonB, 1: ...F) at the
 * integer pixel (x, y): the texel itself, clamped (one 0 texel, a valid
 * layer-0 ribbon, until uploaded). */
static inline ribbon ftk_at(const ptref_ctx *c, int k, int x, int y)
{
    return to_ribbon(c->ftk[k][clampi(y, 0, c->fh[k] - 1) * c->fw[k] + clampi(x, 0, c->fw[k] - 1)]);
}

/* advectKernel (texSynth_kernel.cu:1577-1693) + TexSynth::advectOffsets (:1695-1726) */
int ptref_advect_offsets(ptref_ctx *c, int level, int time_step, int forwards, int flags)
{
    int rc = need_level(c, level);
    if (rc) return rc;
    if (!c->random) return fail(c, PT_ERR_STATE, "no random texture");
    int lw = c->W >> level, lh = c->H >> level;
    adv_t *d = &c->adv[forwards ? 0 : 1];
    int prev = (c->cur + 1) % 2;
    ensure_threads(c);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < lh; y++)
        for (int x = 0; x < lw; x++) {
            int o = y * c->W + x;
            if (image_culled(c, x, y)) {
                f4 z = {0, 0, 0, 0};
                nnf dn = {0.f, 0.f, 0.f, 0.f, 1.f, 1.f, 0.f};
                d->base[o] = z; d->out[o] = z; d->off[o] = dn;
                continue;
            }
            f4 rnd = tex4_point(c->random, c->rw, c->rh, (float)x, (float)y);
            f4 frac = {rnd.x - floorf(rnd.x), rnd.y - floorf(rnd.y), rnd.z - floorf(rnd.z), rnd.w - floorf(rnd.w)};
            nnf out = {0.f, 0.f, 0.f, 0.f, 1.f, 1.f, 0.f};
            float tx, ty;
            /* isImagePixelCulledPreviousFrame (texSynth_kernel.cu:188-198)
             * has int parameters: (tx, ty) is truncated (it is in bounds,
             * so non-negative) and the texel read exactly. */
            if (!(flags & PT_ADVECT_DO_NOT_CHECK_PREVIOUS) &&
                translated_position(c, forwards, x, y, &tx, &ty, lw, lh) &&
                !(c->img_dt[prev][(int)ty * c->W + (int)tx] > ((float)c->p.dist_trans_cull - 8.0f))) {
                int rtx = (int)roundf(tx), rty = (int)roundf(ty);
                nnf nr = c->off_base[rty * c->W + rtx];
                float b0x, b0y, b1x, b1y;
                basis(nr.th, 1, 1, &b0x, &b0y, &b1x, &b1y);
                float ix, iy;
                new_coord(tx - (float)rtx, ty - (float)rty, nr.x, nr.y, b0x, b0y, b1x, b1y, &ix, &iy);
                clamp_ex(&ix, &iy, c->ew, c->eh);
                f4 prev_color = tex4_lin(c->img_color[prev], c->W, c->W, c->H, tx, ty);
                out.x = ix;
                out.y = iy;
                out.th = nr.th;
                out.la = nr.la;
                out.hy = 1.f;
                /* luminanceShift = nearest's + (color - prev_color).toLuminance(), .a = 0 (:1626-1631) */
                if (c->p.lum_shift_enable) {
                    f4 col = c->img_color[c->cur][o];
                    out.ly = nr.ly + pt_luminance_delta(col.x, col.y, col.z, prev_color.x, prev_color.y, prev_color.z);
                }
                d->base[o] = prev_color;
                d->out[o] = tex4_lin(EXO(c, out.la), c->ew, c->ew, c->eh, out.x, out.y);
                d->ts[o] = time_step;
            }
            /* ADVECT_CHECK_FRAME_TO_KEY (:1643-1668): a long step, or a failed
             * advection, may take the pixel's ribbon to a keyframe instead */
            if ((flags & PT_ADVECT_CHECK_FRAME_TO_KEY) && (time_step > 1 || out.hy == 0.f)) {
                ribbon fk = ftk_at(c, forwards ? 0 : 1, x, y);
                if (fk.layer >= 0 && (abs(fk.ts) < time_step || out.hy == 0.f)) {
                    out.x = fk.x + (float)x;
                    out.y = fk.y + (float)y;
                    out.la = fk.layer;
                    out.hy = 1.f;
                    out.ly = 0.f;
                }
                d->base[o] = tex4_lin(EXB(c, out.la), c->ew, c->ew, c->eh, out.x, out.y);
                d->out[o] = tex4_lin(EXO(c, out.la), c->ew, c->ew, c->eh, out.x, out.y);
                d->ts[o] = abs(fk.ts);
            }
            if (out.hy == 0.f) {
                nnf rn = {rnd.x, rnd.y, rnd.z, 0.f, 1.f, 1.f, 0.f, (int)rnd.w};
                f4 z = {0, 0, 0, 0};
                out = rn;
                d->base[o] = z;
                d->out[o] = z;
                d->ts[o] = -1;
            }
            float jit = c->p.advection_jitter;
            out.x = out.x * (1 - jit) + frac.x * (jit);
            out.y = out.y * (1 - jit) + frac.y * (jit);
            out.th = screen_to_ex_angle_l(c, (float)x, (float)y, out.x, out.y, out.la);
            d->off[o] = out;
        }
    size_t n = (size_t)c->W * c->H;
    memcpy(c->off_base, d->off, sizeof(nnf) * n);
    rc = ptref_sync_buffers(c, level, PT_DONT_COPY_OFFSETS | PT_DONT_UPDATE_RESIDUAL);
    if (rc) return rc;
    if (flags & PT_ADVECT_UPDATE_RESIDUAL) {
        rc = ptref_update_residual_cache(c, level);
        if (rc) return rc;
    }
    memcpy(d->res, c->res_base, sizeof(float) * n);
    return PT_OK;
}

/* mergeRandomlyKernel (texSynth_kernel.cu:1891-1936) + TexSynth::mergeRandomly (:2003-2024).
 * The advected offset wins where its direction's frame-to-key ribbon is
 * valid and the other one is not; unbound ribbon textures read as zero,
 * "valid" (layer 0), which leaves that clause false. */
int ptref_merge_randomly(ptref_ctx *c, int level, int forwards, float weight, int force_valid)
{
    int rc = need_level(c, level);
    if (rc) return rc;
    if (!c->random) return fail(c, PT_ERR_STATE, "no random texture");
    int lw = c->W >> level, lh = c->H >> level;
    const adv_t *d = &c->adv[forwards ? 0 : 1];
    for (int y = 0; y < lh; y++)
        for (int x = 0; x < lw; x++) {
            int o = y * c->W + x;
            nnf a = d->off[o], u = c->off_base[o];
            int av = a.hy > 0;
            f4 rnd = tex4_point(c->random, c->rw, c->rh, (float)x, (float)y);
            float r = forwards ? rnd.x - floorf(rnd.x) : rnd.y - floorf(rnd.y);
            ribbon ar = ftk_at(c, forwards ? 0 : 1, x, y), ur = ftk_at(c, forwards ? 1 : 0, x, y);
            int rib = ar.layer >= 0 && !(ur.layer >= 0);
            c->off_work[o] = (rib || ((av || force_valid) && r < weight)) ? a : u;
        }
    return ptref_sync_buffers(c, level, 0);
}

/* mergeUsingResidualKernel (texSynth_kernel.cu:1856-1889) + launcher (:1978-2001) */
int ptref_merge_using_residual(ptref_ctx *c, int level, int forwards)
{
    int rc = need_level(c, level);
    if (rc) return rc;
    int lw = c->W >> level, lh = c->H >> level;
    const adv_t *d = &c->adv[forwards ? 0 : 1];
    for (int y = 0; y < lh; y++)
        for (int x = 0; x < lw; x++) {
            int o = y * c->W + x;
            nnf a = d->off[o], u = c->off_base[o], r;
            if (d->res[o] < c->res_base[o]) { r = a; r.hy = 1.f; }
            else { r = u; r.hy = 0.f; }
            c->off_work[o] = r;
        }
    return ptref_sync_buffers(c, level, 0);
}

/* mergeAdvectedKernel (texSynth_kernel.cu:1938-1976) + TexSynth::
 * mergeAdvectedBuffers (:2026-2042): both valid advected offsets -> one of
 * them by the fractional random x, one valid -> it, none -> the fractional
 * random texel as an offset (MERGE_FALLBACK_RANDOM) or the current base. */
int ptref_merge_advected_buffers(ptref_ctx *c, int level, int fallback)
{
    int rc = need_level(c, level);
    if (rc) return rc;
    if (!c->random) return fail(c, PT_ERR_STATE, "no random texture");
    int lw = c->W >> level, lh = c->H >> level;
    for (int y = 0; y < lh; y++)
        for (int x = 0; x < lw; x++) {
            int o = y * c->W + x;
            nnf f = c->adv[0].off[o], b = c->adv[1].off[o], r;
            f4 rnd = tex4_point(c->random, c->rw, c->rh, (float)x, (float)y);
            /* randomFloat4's out parameter: the fractional parts (:207-213) */
            f4 fr = {rnd.x - floorf(rnd.x), rnd.y - floorf(rnd.y), rnd.z - floorf(rnd.z), rnd.w - floorf(rnd.w)};
            if (f.hy > 0 && b.hy > 0) r = (fr.x < 0.5f) ? f : b;
            else if (f.hy > 0) r = f;
            else if (b.hy > 0) r = b;
            else if (fallback == PT_MERGE_FALLBACK_RANDOM) {
                nnf t = {fr.x, fr.y, fr.z, 0.f, 1.f, 1.f, 0.f, (int)fr.w};
                r = t;
            } else r = c->off_base[o];
            c->off_work[o] = r;
        }
    return ptref_sync_buffers(c, level, 0);
}

/* copyFrameToKeyRibbonKernel (texSynth_kernel.cu:2061-2097) + TexSynth::
 * offsetsFromFrameToKeyRibbon (:2111-2118): each pixel starts at the
 * keyframe texel its ribbon reaches (hysteresis 1), or at its random
 * texel where the ribbon is invalid. */
int ptref_offsets_from_frame_to_key_ribbon(ptref_ctx *c, int level, int look_forwards)
{
    int rc = need_level(c, level);
    if (rc) return rc;
    if (!c->random) return fail(c, PT_ERR_STATE, "no random texture");
    int lw = c->W >> level, lh = c->H >> level;
    for (int y = 0; y < lh; y++)
        for (int x = 0; x < lw; x++) {
            ribbon rb = ftk_at(c, look_forwards ? 1 : 0, x, y);
            nnf out;
            if (rb.layer >= 0) {
                nnf t = {rb.x + (float)x, rb.y + (float)y, 0.f, 1.f, 1.f, 1.f, 0.f, rb.layer};
                t.th = screen_to_ex_angle_l(c, (float)x, (float)y, t.x, t.y, t.la);
                out = t;
            } else {
                f4 rnd = tex4_point(c->random, c->rw, c->rh, (float)x, (float)y);
                nnf t = {rnd.x, rnd.y, rnd.z, 0.f, 1.f, 1.f, 0.f, (int)rnd.w};
                out = t;
            }
            c->off_work[y * c->W + x] = out;
        }
    return ptref_sync_buffers(c, level, 0);
}

/* accumulateAdvectionKernel (texSynth_kernel.cu:1728-1826) and
 * TexSynth::accumulateRibbonField (:1828-1854): the ribbon field (base)
 * advected one step of `time_step` frames along the image velocities into
 * the working buffer, which is then copied to the base whole;
 * cumulativeAdvectionFieldToOutput (:290-305) then writes the canvas texel
 * each valid ribbon reaches into the forward advected output colours (the
 * reference reads that texel unchecked; here it is clamped into the
 * level-0-sized canvas). */
int ptref_accumulate_ribbon_field(ptref_ctx *c, int level, int forwards, int time_step, int overwrite)
{
    if (!c || !c->off_base) return fail(c, PT_ERR_STATE, "not initialized");
    if (level < 0 || (c->W >> level) < 1 || (c->H >> level) < 1) return fail(c, PT_ERR_INVALID_ARG, "bad level");
    int lw = c->W >> level, lh = c->H >> level;
    const ribbon *src = c->rib_base;
    ribbon *dst = c->rib_work;
    ensure_threads(c);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < lh; y++)
        for (int x = 0; x < lw; x++) {
            int o = y * c->W + x;
            ribbon out = {0.f, 0.f, 0, -1};
            f4 col = img_color_at(c, c->cur, x, y);
            float tx, ty;
            if (col.w == 0.f || !translated_position(c, forwards, x, y, &tx, &ty, lw, lh)) {
                if (overwrite) dst[o] = out;
                continue;
            }
            int ftx = (int)floorf(tx), fty = (int)floorf(ty), ctx = (int)ceilf(tx), cty = (int)ceilf(ty);
            ribbon s00 = src[fty * c->W + ftx], s10 = src[fty * c->W + ctx];
            ribbon s01 = src[cty * c->W + ftx], s11 = src[cty * c->W + ctx];
            int v00 = s00.layer >= 0, v10 = s10.layer >= 0, v01 = s01.layer >= 0, v11 = s11.layer >= 0;
            if (!(v00 || v10 || v01 || v11)) {
                if (overwrite) dst[o] = out;
                continue;
            }
            int all_valid = v00 && v10 && v01 && v11;
            /* (s00.layer == s10.layer == s01.layer == s11.layer), evaluated
             * left to right as C++ does: a bool compared with the next int */
            int same = ((((s00.layer == s10.layer) ? 1 : 0) == s01.layer) ? 1 : 0) == s11.layer;
            const float th = 5.f;
            int diverges = fabsf(s00.x - s10.x) > th || fabsf(s00.x - s01.x) > th || fabsf(s00.x - s11.x) > th ||
                           fabsf(s00.y - s10.y) > th || fabsf(s00.y - s01.y) > th || fabsf(s00.y - s11.y) > th;
            if (!all_valid || !same || diverges) {
                int rtx = (int)roundf(tx), rty = (int)roundf(ty);
                if (ftx == rtx && fty == rty) out = s00;
                else if (ftx + 1 == rtx && fty == rty) out = s10;
                else if (ftx == rtx && fty + 1 == rty) out = s01;
                else out = s11;
            } else {
                float wx = tx - (float)ftx, wy = ty - (float)fty;
                float rx0 = s00.x * (1.f - wx) + s10.x * wx;
                float ry0 = s00.y * (1.f - wx) + s10.y * wx;
                float rx1 = s01.x * (1.f - wx) + s11.x * wx;
                float ry1 = s01.y * (1.f - wx) + s11.y * wx;
                out.x = rx0 * (1.f - wy) + rx1 * wy;
                out.y = ry0 * (1.f - wy) + ry1 * wy;
                out.ts = s00.ts;
                out.layer = s00.layer;
            }
            out.x += (tx - (float)x);
            out.y += (ty - (float)y);
            out.ts += time_step;
            ribbon cur = dst[o];
            if (overwrite || !(cur.layer >= 0) || fabsf((float)cur.ts) >= fabsf((float)out.ts)) dst[o] = out;
        }
    memcpy(c->rib_base, c->rib_work, sizeof(ribbon) * (size_t)c->W * c->H);
    for (int y = 0; y < lh; y++)
        for (int x = 0; x < lw; x++) {
            int o = y * c->W + x;
            ribbon r = c->rib_base[o];
            f4 v = {0.f, 0.f, 0.f, 0.f};
            if (r.layer >= 0) {
                int sx = clampi((int)((float)x + r.x), 0, c->W - 1), sy = clampi((int)((float)y + r.y), 0, c->H - 1);
                v = c->canvas[sy * c->W + sx];
            }
            c->adv[0].out[o] = v;
        }
    return PT_OK;
}

int ptref_load_ribbon_field(ptref_ctx *c, int level, const pt_ribbonp *in)
{
    if (!c || !c->off_base || !in || level < 0) return PT_ERR_INVALID_ARG;
    int lw = c->W >> level, lh = c->H >> level;
    for (int y = 0; y < lh; y++)
        for (int x = 0; x < lw; x++) {
            const pt_ribbonp *p = &in[y * lw + x];
            ribbon r = {p->x, p->y, p->time_step, p->layer};
            c->rib_base[y * c->W + x] = r;
        }
    return PT_OK;
}

int ptref_get_ribbon_field(ptref_ctx *c, int level, pt_ribbonp *out)
{
    if (!c || !c->off_base || !out || level < 0) return PT_ERR_INVALID_ARG;
    int lw = c->W >> level, lh = c->H >> level;
    for (int y = 0; y < lh; y++)
        for (int x = 0; x < lw; x++) {
            ribbon r = c->rib_base[y * c->W + x];
            pt_ribbonp *p = &out[y * lw + x];
            p->x = r.x; p->y = r.y; p->time_step = r.ts; p->layer = r.layer;
        }
    return PT_OK;
}

/* ------------------------------------------------------------------ */
/* host readback / load                                                */
/* ------------------------------------------------------------------ */
static pt_patchxf to_patchxf(nnf n)
{
    pt_patchxf p;
    p.x = n.x; p.y = n.y; p.theta = n.th; p.hysteresis = n.hy; p.layer = n.la;
    p.scale_u = n.su; p.scale_v = n.sv; p.lum_y = n.ly; p.lum_a = 0.f;
    return p;
}

int ptref_get_output(ptref_ctx *c, int level, pt_patchxf *offsets, float *canvas, float *residual)
{
    if (!c || !c->off_base || level < 0) return PT_ERR_INVALID_ARG;
    int lw = c->W >> level, lh = c->H >> level;
    for (int y = 0; y < lh; y++)
        for (int x = 0; x < lw; x++) {
            int o = y * c->W + x, d = y * lw + x;
            if (offsets) offsets[d] = to_patchxf(c->off_base[o]);
            if (canvas) memcpy(canvas + 4 * d, &c->canvas[o], sizeof(f4));
            if (residual) residual[d] = c->res_base[o];
        }
    return PT_OK;
}

int ptref_get_advected(ptref_ctx *c, int level, int which, pt_patchxf *offsets, float *base,
                       float *output, float *residual, int32_t *ts)
{
    if (!c || !c->off_base || level < 0 || which < 0 || which > 1) return PT_ERR_INVALID_ARG;
    int lw = c->W >> level, lh = c->H >> level;
    const adv_t *d = &c->adv[which];
    for (int y = 0; y < lh; y++)
        for (int x = 0; x < lw; x++) {
            int o = y * c->W + x, k = y * lw + x;
            if (offsets) offsets[k] = to_patchxf(d->off[o]);
            if (base) memcpy(base + 4 * k, &d->base[o], sizeof(f4));
            if (output) memcpy(output + 4 * k, &d->out[o], sizeof(f4));
            if (residual) residual[k] = d->res[o];
            if (ts) ts[k] = d->ts[o];
        }
    return PT_OK;
}

int ptref_load_offsets(ptref_ctx *c, int level, const pt_patchxf *in)
{
    if (!c || !c->off_base || !in || level < 0) return PT_ERR_INVALID_ARG;
    int lw = c->W >> level, lh = c->H >> level;
    for (int y = 0; y < lh; y++)
        for (int x = 0; x < lw; x++) {
            const pt_patchxf *p = &in[y * lw + x];
            if (p->lum_a != 0.f)
                return fail(c, PT_ERR_UNSUPPORTED, "offsets with an alpha luminance shift");
            if (p->lum_y != 0.f) c->lum_live = 1;
            nnf n = {p->x, p->y, p->theta, p->hysteresis, p->scale_u, p->scale_v, p->lum_y, p->layer};
            c->off_base[y * c->W + x] = n;
        }
    return PT_OK;
}

int ptref_load_residual(ptref_ctx *c, int level, const float *in)
{
    if (!c || !c->off_base || !in || level < 0) return PT_ERR_INVALID_ARG;
    int lw = c->W >> level, lh = c->H >> level;
    for (int y = 0; y < lh; y++)
        memcpy(c->res_base + (size_t)y * c->W, in + (size_t)y * lw, sizeof(float) * lw);
    return PT_OK;
}

int ptref_pde_count(ptref_ctx *c, uint64_t *out)
{
    if (!c || !out) return PT_ERR_INVALID_ARG;
    *out = c->pde;
    return PT_OK;
}

float ptref_patch_distance(ptref_ctx *c, int x, int y, int level, const pt_patchxf *cand)
{
    nnf n = {cand->x, cand->y, cand->theta, cand->hysteresis, cand->scale_u, cand->scale_v, 0.f, cand->layer};
    return calc_residual(c, x, y, level, n);
}

float ptref_screen_to_exemplar_angle(ptref_ctx *c, float x, float y, float ex, float ey)
{
    return screen_to_ex_angle(c, x, y, ex, ey);
}

/* ------------------------------------------------------------------ */
/* schedule                                                            */
/* ------------------------------------------------------------------ */

static void snap_save(ptref_ctx *c, int level, snap_t *s)
{
    int n = (c->W >> level) * (c->H >> level);
    if (!s->off) {
        s->off = (pt_patchxf *)malloc(sizeof(pt_patchxf) * n);
        s->res = (float *)malloc(sizeof(float) * n);
        s->canvas = (float *)malloc(sizeof(float) * 4 * n);
    }
    ptref_get_output(c, level, s->off, s->canvas, s->res);
}

/* Object-ID pyramids are built only when the term is on (style_obj_id_weight
 * > 0); an absent ID image is all 0 (NLDataAccess::getInputElement). */
static int build_frame_pyramids(ptref_ctx *c, int w, int h, const pt_frame_guides *g, frame_pyr *fp, int ids)
{
    ptref_pyramid *ten = NULL;
    int rc = ptref_pyramid_create(c, &fp->color, PT_ELEM_F32X4, w, h, PT_DOWNSAMPLE_SCALE_CONSTANT, g->color_rgba, 0.f);
    if (!rc) rc = ptref_pyramid_create(c, &ten, PT_ELEM_F32X3, w, h, PT_DOWNSAMPLE_SCALE_CONSTANT, g->orientation_xx_xy_yy, 0.f);
    if (!rc) rc = ptref_pyramid_orientation(c, &fp->orient, ten);
    ptref_pyramid_destroy(ten);
    if (!rc) rc = ptref_pyramid_create(c, &fp->dt, PT_ELEM_F32, w, h, PT_DOWNSAMPLE_SCALE_CONSTANT, g->dist_trans, 1.0f);
    if (!rc) rc = ptref_pyramid_create(c, &fp->scale, PT_ELEM_F32X2, w, h, PT_DOWNSAMPLE_SCALE_CONSTANT, g->scale_uv, 1.0f);
    if (!rc) rc = ptref_pyramid_create(c, &fp->velb, PT_ELEM_F32X2, w, h, PT_DOWNSAMPLE_SCALE_WITH_LEVELS, g->vel_b, 0.f);
    if (!rc) rc = ptref_pyramid_create(c, &fp->velf, PT_ELEM_F32X2, w, h, PT_DOWNSAMPLE_SCALE_WITH_LEVELS, g->vel_f, 0.f);
    if (!rc && ids) rc = ptref_pyramid_create(c, &fp->id, PT_ELEM_I32, w, h, 0, g->object_id, 0.f);
    return rc;
}

static void free_frame_pyramids(frame_pyr *fp)
{
    ptref_pyramid_destroy(fp->color); ptref_pyramid_destroy(fp->orient); ptref_pyramid_destroy(fp->dt);
    ptref_pyramid_destroy(fp->scale); ptref_pyramid_destroy(fp->velb); ptref_pyramid_destroy(fp->velf);
    ptref_pyramid_destroy(fp->id);
}

/* propagateAndScatter (synthesisFrame.cpp:483-508) */
static int propagate_and_scatter(ptref_ctx *c, int level, int iterations)
{
    int rc = PT_OK;
    for (int itr = 0; itr < iterations && !rc; itr++) {
        for (int i = 128; i > 1 && !rc; i /= 2) rc = ptref_propagate(c, level, i, -1);
        if (!rc) rc = ptref_propagate(c, level, 1, -1);
        if (!rc) rc = ptref_propagate(c, level, 1, -1);
        if (!rc && itr < iterations - 1) rc = ptref_scatter(c, level, -1);
    }
    return rc;
}

int ptref_synthesize(ptref_ctx *c, const pt_params *params, int scheme, int iterations, int fixup_passes,
                     int W, int H, int F, const pt_frame_guides *frames, int SW, int SH,
                     const pt_style_guides *style, pt_patchxf *out_off, float *out_canvas, float *out_res)
{
    if (!c || !params || !frames || !style || F < 1 || iterations < 1 || fixup_passes < 0)
        return fail(c, PT_ERR_INVALID_ARG, "bad synthesize arguments");
    if (scheme < PT_S_FMBM || scheme > PT_S_IND) return fail(c, PT_ERR_UNSUPPORTED, "scheme");
    if (scheme == PT_S_TCTF && params->first_pass_direction != 0)
        return fail(c, PT_ERR_UNSUPPORTED, "TCTF with a backward first pass");
    if (params->max_level != 0 && scheme != PT_S_FMBM && scheme != PT_S_FMBM_RM)
        return fail(c, PT_ERR_UNSUPPORTED, "max_level above 0 with FMBM and FMBM_RM only");
    int levels = params->levels;
    int img_lv = (int)fmin(floor(log2f((float)W)), floor(log2f((float)H)));
    int sty_lv = (int)fmin(floor(log2f((float)SW)), floor(log2f((float)SH)));
    if (levels < 1 || levels > img_lv || levels > sty_lv) return fail(c, PT_ERR_INVALID_ARG, "levels");
    int rc = ptref_setup(c, params);
    if (rc) return rc;
    /* a one-shot run starts from freshly allocated, zeroed WorkingBuffers */
    free_buffers(c);
    rc = ptref_initialize(c, W, H, SW, SH);
    if (rc) return rc;
    c->scatter_calls = 0;
    /* style pyramids (Style::load, style.cpp:62-140) */
    ptref_pyramid *sb = NULL, *so = NULL, *st = NULL, *sor = NULL, *sdt = NULL, *sid = NULL;
    const int ids = params->style_obj_id_weight > 0.f;
    rc = ptref_pyramid_create(c, &sb, PT_ELEM_F32X4, SW, SH, PT_DOWNSAMPLE_SCALE_CONSTANT, style->base_rgba, 0.f);
    if (!rc) rc = ptref_pyramid_create(c, &so, PT_ELEM_F32X4, SW, SH, PT_DOWNSAMPLE_SCALE_CONSTANT, style->output_rgba, 0.f);
    if (!rc) rc = ptref_pyramid_create(c, &st, PT_ELEM_F32X3, SW, SH, PT_DOWNSAMPLE_SCALE_CONSTANT, style->orientation_xx_xy_yy, 0.f);
    if (!rc) rc = ptref_pyramid_orientation(c, &sor, st);
    if (!rc) rc = ptref_pyramid_create(c, &sdt, PT_ELEM_F32, SW, SH, PT_DOWNSAMPLE_SCALE_CONSTANT, style->dist_trans, 0.f);
    if (!rc && ids) rc = ptref_pyramid_create(c, &sid, PT_ELEM_I32, SW, SH, 0, style->object_ids, 0.f);
    ptref_pyramid_destroy(st);
    frame_pyr *fp = (frame_pyr *)calloc(F, sizeof(frame_pyr));
    for (int f = 0; f < F && !rc; f++) rc = build_frame_pyramids(c, W, H, &frames[f], &fp[f], ids);
    /* FBMO saves its two directed passes, their merge and the fixups; TCTF a
     * consecutive pass and, per level, one refine pass (+ fixups + 1 from
     * three frames), numbered on across levels */
    int passes = scheme == PT_S_FBMO ? fixup_passes + 3
                 : scheme == PT_S_TCTF ? 1 + levels * (1 + (F > 2 ? fixup_passes + 1 : 0))
                                       : fixup_passes + 1;
    snap_t *snaps = (snap_t *)calloc((size_t)levels * F * passes, sizeof(snap_t));
#define SNAP(l, f, p) snaps[((size_t)(l) * F + (f)) * passes + (p)]
    pt_params fparams = *params;
    /* renderAnimationIND (synthesisProcessor.cpp:806-839): every frame on its
     * own, levels inside the frame loop, synthesizeConsecutive with
     * is_first_frame and forward time for every pass, no layer flips */
    for (int f = 0; f < F && !rc && scheme == PT_S_IND; f++) {
        for (int level = levels - 1; level >= 0 && !rc; level--) {
            rc = ptref_initialize_exemplars(c, SW >> level, SH >> level, 1);
            if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_EXEMPLAR_BASE, sb, 0);
            if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_EXEMPLAR_OUTPUT, so, 0);
            if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_EXEMPLAR_ORIENTATION, sor, 0);
            if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_EXEMPLAR_DIST_TRANS, sdt, 0);
            if (!rc && sid) rc = ptref_upload_image(c, level, PT_LAYER_EXEMPLAR_OBJECT_IDS, sid, 0);
            if (!rc) rc = ptref_reset_advected_buffers(c);
            for (int pass = 0; pass < passes && !rc; pass++) {
                rc = ptref_upload_image(c, level, PT_LAYER_INPUT_COLOR, fp[f].color, 0);
                if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_ORIENTATION, fp[f].orient, 0);
                if (!rc && fp[f].id) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_ID, fp[f].id, 0);
                if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_VEL_B, fp[f].velb, 0);
                if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_VEL_F, fp[f].velf, 0);
                if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_DIST_TRANS, fp[f].dt, 0);
                if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_SCALE, fp[f].scale, 0);
                if (rc) break;
                fparams.direction = PT_FORWARD;
                fparams.advection_jitter = (f % 3 == 0) ? params->advection_jitter : 0.f;
                fparams.frame_number = f;
                rc = ptref_upload_parameters(c, &fparams);
                if (rc) break;
                if (pass == 0) {
                    /* firstPass: random pyramid, then random offsets (coarsest)
                     * or this frame's pass-0 result of the coarser level */
                    rc = ptref_create_random_pyramid(c, level, f, fp[f].color, fp[f].orient, so, sor);
                    if (!rc && level == levels - 1) {
                        rc = ptref_randomize_offsets(c, level);
                    } else if (!rc) {
                        snap_t *s = &SNAP(level + 1, f, 0);
                        rc = ptref_load_offsets(c, level + 1, s->off);
                        if (!rc) rc = ptref_upsample(c, level);
                        if (!rc) rc = ptref_sync_buffers(c, level, 0);
                    }
                } else {
                    /* later passes: advect (from the never-loaded other layer),
                     * previous pass, merge by residual */
                    snap_t *s = &SNAP(level, f, pass - 1);
                    rc = ptref_advect_offsets(c, level, 1, 1, PT_ADVECT_UPDATE_RESIDUAL | PT_ADVECT_IGNORE_RANDOM);
                    if (!rc) rc = ptref_load_offsets(c, level, s->off);
                    if (!rc) rc = ptref_load_residual(c, level, s->res);
                    if (!rc) rc = ptref_sync_buffers(c, level, PT_DONT_COPY_OFFSETS | PT_DONT_UPDATE_RESIDUAL);
                    if (!rc) rc = ptref_merge_using_residual(c, level, 1);
                }
                if (!rc) rc = ptref_sync_buffers(c, level, 0);
                if (!rc) rc = propagate_and_scatter(c, level, iterations);
                if (!rc) snap_save(c, level, &SNAP(level, f, pass));
            }
        }
    }
    /* renderAnimationFBMO (synthesisProcessor.cpp:902-970) and its fixup
     * passes (renderInterpolatedFixupPass :771-805, synthesizeInterpolated
     * and advectNeighbors, synthesisFrame.cpp:325-407) */
    int *last = (int *)calloc(F, sizeof(int)); /* SynthesisFrame::_last_saved_pass at this level */
    for (int level = levels - 1; level >= 0 && !rc && scheme == PT_S_FBMO; level--) {
        int merged = fixup_passes + 2;
        rc = ptref_initialize_exemplars(c, SW >> level, SH >> level, 1);
        if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_EXEMPLAR_BASE, sb, 0);
        if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_EXEMPLAR_OUTPUT, so, 0);
        if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_EXEMPLAR_ORIENTATION, sor, 0);
        if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_EXEMPLAR_DIST_TRANS, sdt, 0);
        if (!rc && sid) rc = ptref_upload_image(c, level, PT_LAYER_EXEMPLAR_OBJECT_IDS, sid, 0);
        /* passes 0 and 1: forward, then backward, every frame through
         * synthesizeConsecutive's first-pass branch */
        for (int pass = 0; pass < 2 && !rc; pass++) {
            rc = ptref_reset_advected_buffers(c);
            int fwd = pass == params->first_pass_direction; /* synthesisProcessor.cpp:938 */
            for (int k = 0; k < F && !rc; k++) {
                int f = fwd ? k : F - 1 - k;
                int is_first = k == 0;
                if (!is_first || pass > 0) ptref_flip_anim_buffers(c);
                rc = ptref_upload_image(c, level, PT_LAYER_INPUT_COLOR, fp[f].color, 0);
                if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_ORIENTATION, fp[f].orient, 0);
                if (!rc && fp[f].id) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_ID, fp[f].id, 0);
                if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_VEL_B, fp[f].velb, 0);
                if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_VEL_F, fp[f].velf, 0);
                if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_DIST_TRANS, fp[f].dt, 0);
                if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_SCALE, fp[f].scale, 0);
                if (rc) break;
                fparams.direction = fwd ? PT_FORWARD : PT_BACKWARD;
                fparams.advection_jitter = (f % 3 == 0) ? params->advection_jitter : 0.f;
                fparams.frame_number = f;
                rc = ptref_upload_parameters(c, &fparams);
                if (!rc) rc = ptref_create_random_pyramid(c, level, f, fp[f].color, fp[f].orient, so, sor);
                if (rc) break;
                if (is_first) {
                    if (level == levels - 1) {
                        rc = ptref_randomize_offsets(c, level);
                    } else {
                        rc = ptref_load_offsets(c, level + 1, SNAP(level + 1, f, merged).off);
                        if (!rc) rc = ptref_upsample(c, level);
                        if (!rc) rc = ptref_sync_buffers(c, level, 0);
                    }
                } else if (level == levels - 1) {
                    rc = ptref_advect_offsets(c, level, 1, fwd, PT_ADVECT_UPDATE_RESIDUAL);
                    if (!rc) memcpy(c->off_work, c->off_base, sizeof(nnf) * c->W * c->H);
                } else {
                    snap_t *s = &SNAP(level + 1, f, merged);
                    rc = ptref_advect_offsets(c, level, 1, fwd, PT_ADVECT_UPDATE_RESIDUAL | PT_ADVECT_IGNORE_RANDOM);
                    if (!rc) rc = ptref_load_offsets(c, level + 1, s->off);
                    if (!rc) rc = ptref_load_residual(c, level + 1, s->res);
                    if (!rc) rc = ptref_upsample(c, level);
                    if (!rc) rc = ptref_sync_buffers(c, level, 0);
                    if (!rc) rc = ptref_merge_using_residual(c, level, fwd);
                }
                if (!rc) rc = ptref_sync_buffers(c, level, 0);
                if (!rc) rc = propagate_and_scatter(c, level, iterations);
                if (!rc) snap_save(c, level, &SNAP(level, f, pass));
                last[f] = pass;
            }
        }
        /* pass 2: the forward result becomes the advected-forward offsets and
         * residual, the backward result the current ones, force-valid merge */
        for (int f = 0; f < F && !rc; f++) {
            size_t n = (size_t)c->W * c->H;
            rc = ptref_load_offsets(c, level, SNAP(level, f, 0).off);
            if (!rc) rc = ptref_load_residual(c, level, SNAP(level, f, 0).res);
            if (rc) break;
            memcpy(c->adv[0].off, c->off_base, sizeof(nnf) * n);
            memcpy(c->adv[0].res, c->res_base, sizeof(float) * n);
            rc = ptref_sync_buffers(c, level, PT_DONT_COPY_OFFSETS | PT_DONT_UPDATE_RESIDUAL);
            if (!rc) rc = ptref_load_offsets(c, level, SNAP(level, f, 1).off);
            if (!rc) rc = ptref_load_residual(c, level, SNAP(level, f, 1).res);
            if (!rc) rc = ptref_sync_buffers(c, level, PT_DONT_COPY_OFFSETS | PT_DONT_UPDATE_RESIDUAL);
            if (!rc) rc = ptref_merge_randomly(c, level, 1, 0.5f, 1);
            if (!rc) snap_save(c, level, &SNAP(level, f, 2));
        }
        fparams.direction = PT_BIDIRECTIONAL;
        if (!rc) rc = ptref_upload_parameters(c, &fparams);
        int temporal = F > 1 && (params->time_derivative_input_weight > 0.f ||
                                 params->time_derivative_output_weight > 0.f ||
                                 params->temporal_coherence_weight > 0.f);
        for (int pass = 3; pass < fixup_passes + 3 && !rc; pass++) {
            for (int f = 0; f < F && !rc; f++) {
                rc = ptref_create_random_pyramid(c, level, f, fp[f].color, fp[f].orient, so, sor);
                fparams.direction = PT_BIDIRECTIONAL;
                fparams.advection_jitter = (f % 3 == 0) ? params->advection_jitter : 0.f;
                fparams.frame_number = f;
                if (!rc) rc = ptref_upload_parameters(c, &fparams);
                if (!rc) rc = ptref_reset_advected_buffers(c);
                for (int side = 0; side < 2 && !rc && temporal; side++) {
                    /* prior frame advected forward, next frame backward */
                    int g = side == 0 ? f - 1 : f + 1;
                    if (g < 0 || g >= F) continue;
                    rc = ptref_upload_image(c, level, PT_LAYER_INPUT_COLOR, fp[g].color, 0);
                    if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_ORIENTATION, fp[g].orient, 0);
                    if (!rc && fp[g].id) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_ID, fp[g].id, 0);
                    if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_VEL_B, fp[g].velb, 0);
                    if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_VEL_F, fp[g].velf, 0);
                    if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_DIST_TRANS, fp[g].dt, 0);
                    if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_SCALE, fp[g].scale, 0);
                    if (!rc) ptref_flip_anim_buffers(c);
                    if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_VEL_B, fp[f].velb, 0);
                    if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_VEL_F, fp[f].velf, 0);
                    if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_DIST_TRANS, fp[f].dt, 0);
                    if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_SCALE, fp[f].scale, 0);
                    if (!rc) rc = ptref_load_offsets(c, level, SNAP(level, g, last[g]).off);
                    if (!rc) rc = ptref_advect_offsets(c, level, 1, side == 0, 0);
                }
                if (!rc && !temporal) {
                    rc = ptref_upload_image(c, level, PT_LAYER_INPUT_VEL_B, fp[f].velb, 0);
                    if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_VEL_F, fp[f].velf, 0);
                    if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_DIST_TRANS, fp[f].dt, 0);
                    if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_SCALE, fp[f].scale, 0);
                }
                if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_COLOR, fp[f].color, 0);
                if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_ORIENTATION, fp[f].orient, 0);
                if (!rc && fp[f].id) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_ID, fp[f].id, 0);
                if (!rc) rc = ptref_load_offsets(c, level, SNAP(level, f, pass - 1).off);
                if (!rc) rc = ptref_sync_buffers(c, level, PT_DONT_COPY_OFFSETS | PT_DONT_UPDATE_RESIDUAL);
                if (!rc) rc = propagate_and_scatter(c, level, iterations);
                if (!rc) snap_save(c, level, &SNAP(level, f, pass));
                last[f] = pass;
            }
        }
    }
    free(last);
    /* renderAnimationTCTF (synthesisProcessor.cpp:1151-1312), direct
     * execution, no keyframes or ribbons (largest time step 1): a forward
     * consecutive pass on the coarsest level, then per level coarse to fine
     * refine passes over all frames (refine, synthesisFrame.cpp:409-457) */
    if (scheme == PT_S_TCTF && !rc) {
        int L = levels - 1;
        int *lastp = (int *)calloc((size_t)F * levels, sizeof(int)); /* [f * levels + level] */
        int temporal = F > 1 && (params->time_derivative_input_weight > 0.f ||
                                 params->time_derivative_output_weight > 0.f ||
                                 params->temporal_coherence_weight > 0.f);
        int passes_this = 0, passes_last = 0, last_level = -1;
        rc = ptref_reset_advected_buffers(c);
#define ENTER_LEVEL(l)                                                                               \
        do {                                                                                         \
            if ((l) != last_level && !rc) {                                                          \
                rc = ptref_initialize_exemplars(c, SW >> (l), SH >> (l), 1);                          \
                if (!rc) rc = ptref_upload_image(c, (l), PT_LAYER_EXEMPLAR_BASE, sb, 0);              \
                if (!rc) rc = ptref_upload_image(c, (l), PT_LAYER_EXEMPLAR_OUTPUT, so, 0);            \
                if (!rc) rc = ptref_upload_image(c, (l), PT_LAYER_EXEMPLAR_ORIENTATION, sor, 0);      \
                if (!rc) rc = ptref_upload_image(c, (l), PT_LAYER_EXEMPLAR_DIST_TRANS, sdt, 0);       \
                if (!rc && sid) rc = ptref_upload_image(c, (l), PT_LAYER_EXEMPLAR_OBJECT_IDS, sid, 0); \
                passes_last = passes_this;                                                           \
                last_level = (l);                                                                    \
            }                                                                                        \
        } while (0)
#define LOAD_IMAGES(l, g)                                                                            \
        do {                                                                                         \
            if (!rc) rc = ptref_upload_image(c, (l), PT_LAYER_INPUT_COLOR, fp[g].color, 0);           \
            if (!rc) rc = ptref_upload_image(c, (l), PT_LAYER_INPUT_ORIENTATION, fp[g].orient, 0);    \
            if (!rc && fp[g].id) rc = ptref_upload_image(c, (l), PT_LAYER_INPUT_ID, fp[g].id, 0);     \
        } while (0)
#define LOAD_ANIM(l, g)                                                                              \
        do {                                                                                         \
            if (!rc) rc = ptref_upload_image(c, (l), PT_LAYER_INPUT_VEL_B, fp[g].velb, 0);            \
            if (!rc) rc = ptref_upload_image(c, (l), PT_LAYER_INPUT_VEL_F, fp[g].velf, 0);            \
            if (!rc) rc = ptref_upload_image(c, (l), PT_LAYER_INPUT_DIST_TRANS, fp[g].dt, 0);         \
            if (!rc) rc = ptref_upload_image(c, (l), PT_LAYER_INPUT_SCALE, fp[g].scale, 0);           \
        } while (0)
        /* a pass that was never saved loads nothing (missing temporary file) */
#define LOAD_IF_SAVED(l, g, p)                                                                       \
        do {                                                                                         \
            if (!rc && SNAP((l), (g), (p)).off) rc = ptref_load_offsets(c, (l), SNAP((l), (g), (p)).off); \
        } while (0)
        ENTER_LEVEL(L);
        for (int f = 0; f < F && !rc; f++) {
            int is_first = f == 0;
            if (!is_first) ptref_flip_anim_buffers(c);
            LOAD_IMAGES(L, f);
            LOAD_ANIM(L, f);
            fparams.direction = PT_FORWARD;
            fparams.advection_jitter = (f % 3 == 0) ? params->advection_jitter : 0.f;
            fparams.frame_number = f;
            if (!rc) rc = ptref_upload_parameters(c, &fparams);
            if (!rc) rc = ptref_create_random_pyramid(c, L, f, fp[f].color, fp[f].orient, so, sor);
            if (!rc && is_first) {
                rc = ptref_randomize_offsets(c, L);
            } else if (!rc) {
                rc = ptref_advect_offsets(c, L, 1, 1, PT_ADVECT_UPDATE_RESIDUAL);
                if (!rc) memcpy(c->off_work, c->off_base, sizeof(nnf) * c->W * c->H);
            }
            if (!rc) rc = ptref_sync_buffers(c, L, 0);
            if (!rc) rc = propagate_and_scatter(c, L, iterations);
            if (!rc) snap_save(c, L, &SNAP(L, f, 0));
            lastp[f * levels + L] = 0;
        }
        passes_this++;
        for (int level = L; level >= 0 && !rc; level--) {
            int count = 1 + (F > 2 ? fixup_passes + 1 : 0);
            for (int r = 0; r < count && !rc; r++) {
                ENTER_LEVEL(level);
                int this_pass = passes_this, last_coarse = passes_last - 1;
                int spatial = last_coarse == this_pass - 1;
                for (int f = 0; f < F && !rc; f++) {
                    fparams.direction = PT_BIDIRECTIONAL;
                    fparams.advection_jitter = (f % 3 == 0) ? params->advection_jitter : 0.f;
                    fparams.frame_number = f;
                    rc = ptref_upload_parameters(c, &fparams);
                    if (!rc) rc = ptref_reset_advected_buffers(c);
                    if (temporal && (f > 0 || f < F - 1)) {
                        for (int side = 0; side < 2 && !rc; side++) {
                            int g = side == 0 ? f - 1 : f + 1;
                            if (g < 0 || g >= F || this_pass - 1 < 0) continue;
                            LOAD_IMAGES(level, g);
                            LOAD_ANIM(level, g);
                            if (!rc) ptref_flip_anim_buffers(c);
                            LOAD_ANIM(level, f);
                            LOAD_IF_SAVED(level, g, lastp[g * levels + level]);
                            if (!rc) rc = ptref_advect_offsets(c, level, 1, side == 0, 0);
                        }
                    } else {
                        LOAD_ANIM(level, f);
                    }
                    if (spatial) {
                        LOAD_IF_SAVED(level + 1, f, lastp[f * levels + level + 1]);
                        if (!rc) rc = ptref_upsample(c, level);
                    } else {
                        LOAD_IF_SAVED(level, f, lastp[f * levels + level]);
                    }
                    if (!rc) rc = ptref_sync_buffers(c, level, PT_DONT_COPY_OFFSETS | PT_DONT_UPDATE_RESIDUAL);
                    LOAD_IMAGES(level, f);
                    if (!rc) rc = propagate_and_scatter(c, level, iterations);
                    if (!rc) snap_save(c, level, &SNAP(level, f, this_pass));
                    lastp[f * levels + level] = this_pass;
                }
                passes_this++;
            }
        }
#undef ENTER_LEVEL
#undef LOAD_IMAGES
#undef LOAD_ANIM
#undef LOAD_IF_SAVED
        free(lastp);
    }
    for (int level = levels - 1; level >= params->max_level && !rc && scheme != PT_S_IND && scheme != PT_S_FBMO &&
                                 scheme != PT_S_TCTF; level--) {
        rc = ptref_initialize_exemplars(c, SW >> level, SH >> level, 1);
        if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_EXEMPLAR_BASE, sb, 0);
        if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_EXEMPLAR_OUTPUT, so, 0);
        if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_EXEMPLAR_ORIENTATION, sor, 0);
        if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_EXEMPLAR_DIST_TRANS, sdt, 0);
        if (!rc && sid) rc = ptref_upload_image(c, level, PT_LAYER_EXEMPLAR_OBJECT_IDS, sid, 0);
        if (!rc) rc = ptref_reset_advected_buffers(c);
        for (int pass = 0; pass < passes && !rc; pass++) {
            /* renderConsecutivePass (synthesisProcessor.cpp:703-753) */
            int fwd = (pass % 2 == params->first_pass_direction); /* synthesisProcessor.cpp:707 */
            int order[4096];
            int nf = 0;
            if (F > 4096) { rc = fail(c, PT_ERR_INVALID_ARG, "too many frames"); break; }
            for (int i = 0; i < F; i++) order[nf++] = fwd ? i : F - 1 - i;
            int start = 0;
            if (pass > 0) {
                snap_save(c, level, &SNAP(level, order[0], pass));
                start = 1;
            }
            for (int k = start; k < nf && !rc; k++) {
                int f = order[k];
                int is_first = (k == 0) && (pass == 0);
                if (!is_first) ptref_flip_anim_buffers(c);
                /* loadImages / loadAnimImages (synthesisFrame.cpp:562-656) */
                rc = ptref_upload_image(c, level, PT_LAYER_INPUT_COLOR, fp[f].color, 0);
                if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_ORIENTATION, fp[f].orient, 0);
                if (!rc && fp[f].id) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_ID, fp[f].id, 0);
                if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_VEL_B, fp[f].velb, 0);
                if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_VEL_F, fp[f].velf, 0);
                if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_DIST_TRANS, fp[f].dt, 0);
                if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_SCALE, fp[f].scale, 0);
                if (rc) break;
                /* synthesizeConsecutive (synthesisFrame.cpp:214-323) */
                fparams.direction = fwd ? PT_FORWARD : PT_BACKWARD;
                fparams.advection_jitter = (f % 3 == 0) ? params->advection_jitter : 0.f;
                fparams.frame_number = f;
                rc = ptref_upload_parameters(c, &fparams);
                if (rc) break;
                int ops = (scheme != PT_S_FMBM_RM) ? PT_ADVECT_UPDATE_RESIDUAL : 0;
                if (pass == 0) {
                    rc = ptref_create_random_pyramid(c, level, f, fp[f].color, fp[f].orient, so, sor);
                    if (rc) break;
                    if (is_first) {
                        if (level == levels - 1) {
                            rc = ptref_randomize_offsets(c, level);
                        } else {
                            snap_t *s = &SNAP(level + 1, f, passes - 1);
                            rc = ptref_load_offsets(c, level + 1, s->off);
                            if (!rc) rc = ptref_upsample(c, level);
                            if (!rc) rc = ptref_sync_buffers(c, level, 0);
                        }
                    }
                    if (!rc && !is_first && level == levels - 1) {
                        rc = ptref_advect_offsets(c, level, 1, fwd, ops);
                        if (!rc) memcpy(c->off_work, c->off_base, sizeof(nnf) * c->W * c->H);
                    } else if (!rc && !is_first && level < levels - 1) {
                        snap_t *s = &SNAP(level + 1, f, passes - 1);
                        rc = ptref_advect_offsets(c, level, 1, fwd, ops | PT_ADVECT_IGNORE_RANDOM);
                        if (!rc) rc = ptref_load_offsets(c, level + 1, s->off);
                        if (!rc) rc = ptref_load_residual(c, level + 1, s->res);
                        if (!rc) rc = ptref_upsample(c, level);
                        if (!rc) rc = ptref_sync_buffers(c, level, 0);
                        if (!rc) rc = (scheme == PT_S_FMBM_RM) ? ptref_merge_randomly(c, level, fwd, 0.5f, 0)
                                                               : ptref_merge_using_residual(c, level, fwd);
                    }
                } else {
                    snap_t *s = &SNAP(level, f, pass - 1);
                    rc = ptref_advect_offsets(c, level, 1, fwd, ops | PT_ADVECT_IGNORE_RANDOM);
                    if (!rc) rc = ptref_load_offsets(c, level, s->off);
                    if (!rc) rc = ptref_load_residual(c, level, s->res);
                    if (!rc) rc = ptref_sync_buffers(c, level, PT_DONT_COPY_OFFSETS | PT_DONT_UPDATE_RESIDUAL);
                    if (!rc) rc = (scheme == PT_S_FMBM_RM) ? ptref_merge_randomly(c, level, fwd, 0.5f, 0)
                                                           : ptref_merge_using_residual(c, level, fwd);
                }
                if (!rc) rc = ptref_sync_buffers(c, level, 0);
                if (!rc) rc = propagate_and_scatter(c, level, iterations);
                if (!rc) snap_save(c, level, &SNAP(level, f, pass));
            }
        }
    }
    /* maxLevel > 0 (FMBM, FMBM_RM): renderAnimationFMBM's tail
     * (synthesisProcessor.cpp:863-870), SynthesisFrame::spatialUpsample
     * (synthesisFrame.cpp:465-481) for every frame, level by level */
    for (int level = params->max_level; level > 0 && !rc; level--) {
        rc = ptref_initialize_exemplars(c, SW >> (level - 1), SH >> (level - 1), 1);
        if (!rc) rc = ptref_upload_image(c, level - 1, PT_LAYER_EXEMPLAR_BASE, sb, 0);
        if (!rc) rc = ptref_upload_image(c, level - 1, PT_LAYER_EXEMPLAR_OUTPUT, so, 0);
        if (!rc) rc = ptref_upload_image(c, level - 1, PT_LAYER_EXEMPLAR_ORIENTATION, sor, 0);
        if (!rc) rc = ptref_upload_image(c, level - 1, PT_LAYER_EXEMPLAR_DIST_TRANS, sdt, 0);
        if (!rc && sid) rc = ptref_upload_image(c, level - 1, PT_LAYER_EXEMPLAR_OBJECT_IDS, sid, 0);
        for (int f = 0; f < F && !rc; f++) {
            rc = ptref_load_offsets(c, level, SNAP(level, f, level == params->max_level ? passes - 1 : 0).off);
            if (!rc) rc = ptref_upsample(c, level - 1);
            if (!rc) rc = ptref_sync_buffers(c, level - 1, PT_DONT_COPY_OFFSETS | PT_DONT_UPDATE_RESIDUAL);
            if (!rc) snap_save(c, level - 1, &SNAP(level - 1, f, 0));
        }
    }
    if (!rc) {
        size_t n = (size_t)W * H;
        for (int f = 0; f < F; f++) {
            snap_t *s = &SNAP(0, f, params->max_level > 0 ? 0 : passes - 1); /* copyOutputToFinal (:872) */
            if (out_off) memcpy(out_off + f * n, s->off, sizeof(pt_patchxf) * n);
            if (out_canvas) memcpy(out_canvas + f * n * 4, s->canvas, sizeof(float) * 4 * n);
            if (out_res) memcpy(out_res + f * n, s->res, sizeof(float) * n);
        }
    }
#undef SNAP
    for (size_t i = 0; i < (size_t)levels * F * passes; i++) {
        free(snaps[i].off); free(snaps[i].res); free(snaps[i].canvas);
    }
    free(snaps);
    for (int f = 0; f < F; f++) free_frame_pyramids(&fp[f]);
    free(fp);
    ptref_pyramid_destroy(sb); ptref_pyramid_destroy(so); ptref_pyramid_destroy(sor); ptref_pyramid_destroy(sdt);
    ptref_pyramid_destroy(sid);
    return rc;
}

/* ------------------------------------------------------------------ */
/* per-frame sequence schedule                                         */
/* ------------------------------------------------------------------ */
static void free_sequence(ptref_ctx *c)
{
    for (int k = 0; k < 2; k++) {
        if (c->seq_snap[k]) {
            for (int l = 0; l < c->seq_levels; l++) {
                free(c->seq_snap[k][l].off); free(c->seq_snap[k][l].res); free(c->seq_snap[k][l].canvas);
            }
            free(c->seq_snap[k]);
            c->seq_snap[k] = NULL;
        }
        free_frame_pyramids(&c->seq_pyr[k]);
        memset(&c->seq_pyr[k], 0, sizeof(frame_pyr));
    }
    for (int k = 0; k < 5; k++) {
        ptref_pyramid_destroy(c->seq_style[k]);
        c->seq_style[k] = NULL;
    }
    c->seq_levels = 0;
}

int ptref_sequence_reset(ptref_ctx *c)
{
    if (!c) return PT_ERR_INVALID_ARG;
    c->seq_frame = 0;
    return PT_OK;
}

/*
 * The frame-major schedule of pt_synthesize_next_frame: every level of one
 * frame before the next frame.  Per level, the forward first pass of
 * SynthesisFrame::synthesizeConsecutive (synthesisFrame.cpp:214-323) with the
 * FMBM_RM merge, after restoring the previous frame's state of that level
 * (its offsets and residual, and its images in the other texture layer).
 */
int ptref_synthesize_next_frame(ptref_ctx *c, const pt_params *params, int iterations, int W, int H,
                                const pt_frame_guides *frame, int SW, int SH, const pt_style_guides *style,
                                pt_patchxf *out_off, float *out_canvas, float *out_res)
{
    if (!c || !params || !frame || !style || iterations < 1) return fail(c, PT_ERR_INVALID_ARG, "bad frame arguments");
    if (params->first_pass_direction != 0 || params->max_level != 0)
        return fail(c, PT_ERR_UNSUPPORTED, "forwards and down to level 0 only");
    int levels = params->levels;
    int img_lv = (int)fmin(floor(log2f((float)W)), floor(log2f((float)H)));
    int sty_lv = (int)fmin(floor(log2f((float)SW)), floor(log2f((float)SH)));
    if (levels < 1 || levels > img_lv || levels > sty_lv) return fail(c, PT_ERR_INVALID_ARG, "levels");
    int f = c->seq_frame, rc;
    if (f == 0) {
        rc = ptref_setup(c, params);
        if (rc) return rc;
        free_sequence(c);
        free_buffers(c);
        rc = ptref_initialize(c, W, H, SW, SH);
        if (rc) return rc;
        c->scatter_calls = 0;
        for (int k = 0; k < 2; k++) c->seq_snap[k] = (snap_t *)calloc(levels, sizeof(snap_t));
        c->seq_levels = levels;
        ptref_pyramid *ten = NULL;
        rc = ptref_pyramid_create(c, &c->seq_style[0], PT_ELEM_F32X4, SW, SH, PT_DOWNSAMPLE_SCALE_CONSTANT,
                                  style->base_rgba, 0.f);
        if (!rc) rc = ptref_pyramid_create(c, &c->seq_style[1], PT_ELEM_F32X4, SW, SH, PT_DOWNSAMPLE_SCALE_CONSTANT,
                                           style->output_rgba, 0.f);
        if (!rc) rc = ptref_pyramid_create(c, &ten, PT_ELEM_F32X3, SW, SH, PT_DOWNSAMPLE_SCALE_CONSTANT,
                                           style->orientation_xx_xy_yy, 0.f);
        if (!rc) rc = ptref_pyramid_orientation(c, &c->seq_style[2], ten);
        ptref_pyramid_destroy(ten);
        if (!rc) rc = ptref_pyramid_create(c, &c->seq_style[3], PT_ELEM_F32, SW, SH, PT_DOWNSAMPLE_SCALE_CONSTANT,
                                           style->dist_trans, 0.f);
        if (!rc && params->style_obj_id_weight > 0.f)
            rc = ptref_pyramid_create(c, &c->seq_style[4], PT_ELEM_I32, SW, SH, 0, style->object_ids, 0.f);
        if (rc) return rc;
    } else if (c->W != W || c->H != H || c->SW != SW || c->SH != SH || c->seq_levels != levels) {
        return fail(c, PT_ERR_INVALID_ARG, "frame shape differs from the sequence");
    }
    frame_pyr *cur = &c->seq_pyr[f & 1], *prev = &c->seq_pyr[(f + 1) & 1];
    free_frame_pyramids(cur);
    memset(cur, 0, sizeof(frame_pyr));
    rc = build_frame_pyramids(c, W, H, frame, cur, params->style_obj_id_weight > 0.f);
    if (!rc) rc = ptref_setup(c, params);
    if (rc) return rc;
    pt_params fp = *params;
    fp.direction = PT_FORWARD;
    fp.advection_jitter = (f % 3 == 0) ? params->advection_jitter : 0.f;
    fp.frame_number = f;
    rc = ptref_upload_parameters(c, &fp);
    snap_t *now = c->seq_snap[f & 1], *before = c->seq_snap[(f + 1) & 1];
    ptref_pyramid **sp = c->seq_style;
    for (int level = levels - 1; level >= 0 && !rc; level--) {
        rc = ptref_initialize_exemplars(c, SW >> level, SH >> level, 1);
        if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_EXEMPLAR_BASE, sp[0], 0);
        if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_EXEMPLAR_OUTPUT, sp[1], 0);
        if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_EXEMPLAR_ORIENTATION, sp[2], 0);
        if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_EXEMPLAR_DIST_TRANS, sp[3], 0);
        if (!rc && sp[4]) rc = ptref_upload_image(c, level, PT_LAYER_EXEMPLAR_OBJECT_IDS, sp[4], 0);
        if (!rc) rc = ptref_reset_advected_buffers(c);
        if (!rc && f > 0) {
            c->cur = (f + 1) % 2;
            rc = ptref_upload_image(c, level, PT_LAYER_INPUT_COLOR, prev->color, 0);
            if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_VEL_B, prev->velb, 0);
            if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_VEL_F, prev->velf, 0);
            if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_DIST_TRANS, prev->dt, 0);
            if (!rc) rc = ptref_load_offsets(c, level, before[level].off);
            if (!rc) rc = ptref_load_residual(c, level, before[level].res);
        }
        c->cur = f % 2;
        if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_COLOR, cur->color, 0);
        if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_ORIENTATION, cur->orient, 0);
        if (!rc && cur->id) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_ID, cur->id, 0);
        if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_VEL_B, cur->velb, 0);
        if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_VEL_F, cur->velf, 0);
        if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_DIST_TRANS, cur->dt, 0);
        if (!rc) rc = ptref_upload_image(c, level, PT_LAYER_INPUT_SCALE, cur->scale, 0);
        if (!rc) rc = ptref_create_random_pyramid(c, level, f, cur->color, cur->orient, sp[1], sp[2]);
        if (rc) break;
        if (f == 0) {
            if (level == levels - 1) {
                rc = ptref_randomize_offsets(c, level);
            } else {
                rc = ptref_load_offsets(c, level + 1, now[level + 1].off);
                if (!rc) rc = ptref_upsample(c, level);
                if (!rc) rc = ptref_sync_buffers(c, level, 0);
            }
        } else if (level == levels - 1) {
            rc = ptref_advect_offsets(c, level, 1, 1, 0);
            if (!rc) memcpy(c->off_work, c->off_base, sizeof(nnf) * c->W * c->H);
        } else {
            rc = ptref_advect_offsets(c, level, 1, 1, PT_ADVECT_IGNORE_RANDOM);
            if (!rc) rc = ptref_load_offsets(c, level + 1, now[level + 1].off);
            if (!rc) rc = ptref_load_residual(c, level + 1, now[level + 1].res);
            if (!rc) rc = ptref_upsample(c, level);
            if (!rc) rc = ptref_sync_buffers(c, level, 0);
            if (!rc) rc = ptref_merge_randomly(c, level, 1, 0.5f, 0);
        }
        if (!rc) rc = ptref_sync_buffers(c, level, 0);
        if (!rc) rc = propagate_and_scatter(c, level, iterations);
        if (!rc) snap_save(c, level, &now[level]);
    }
    if (rc) return rc;
    size_t n = (size_t)W * H;
    if (out_off) memcpy(out_off, now[0].off, sizeof(pt_patchxf) * n);
    if (out_canvas) memcpy(out_canvas, now[0].canvas, sizeof(float) * 4 * n);
    if (out_res) memcpy(out_res, now[0].res, sizeof(float) * n);
    c->seq_frame = f + 1;
    return PT_OK;
}
