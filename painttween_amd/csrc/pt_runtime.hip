/*
 * pt_runtime.hip -- host runtime behind include/pt_api.h: contexts, device
 * buffers, device pyramids, the TexSynth operators and the FMBM / FMBM_RM
 * frame schedule.
 *
 * Buffer model (cudaImageBuffer.h:59-369, workingBuffers.cpp:48-123): every
 * working buffer is level-0 sized and zeroed only when allocated or reset;
 * kernels touch the active level rectangle only, so content outside it is
 * stale from earlier levels and frames, and the reference reads it (pixelWrap
 * modulo level-0 dimensions, the one-past-the-end upsample read).  The
 * reference copies whole buffers working -> base after every step; here the
 * runtime tracks the rectangle in which base and working can differ and
 * copies only that, which leaves both buffers byte-identical to the
 * whole-buffer copy.
 */
#include "pt_internal.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

namespace {

thread_local char g_last_error[512] = "";

struct Rect {
    int w = 0, h = 0;
    void grow(int ww, int hh)
    {
        w = std::max(w, ww);
        h = std::max(h, hh);
    }
    bool empty() const { return w <= 0 || h <= 0; }
};

struct DevAdv {
    float4 *xyth = nullptr;
    float2 *sc = nullptr;
    float4 *base = nullptr;
    float4 *out = nullptr;
    float *res = nullptr;
    int32_t *ts = nullptr;
};

struct FramePyr {
    pt_pyramid *color = nullptr, *tensor = nullptr, *orient = nullptr, *dt = nullptr, *scale = nullptr,
               *velb = nullptr, *velf = nullptr;
};

struct StylePyr {
    pt_pyramid *base = nullptr, *out = nullptr, *tensor = nullptr, *orient = nullptr, *dt = nullptr;
};

/* Kernel classes timed by pt_profile_*; names in kProfNames. */
enum ProfKind { PK_PROPAGATE = 0, PK_SCATTER, PK_RESIDUAL, PK_CANVAS, PK_UPSAMPLE, PK_ADVECT, PK_MERGE, PK_RANDOM,
                PK_COPY, PK_PYRAMID, PK_COUNT };
static const char *kProfNames[PK_COUNT] = {"propagate", "scatter", "update_residual", "canvas", "upsample",
                                           "advect", "merge", "random", "copy", "pyramid"};

struct ProfRec {
    int kind;
    int ev0, ev1;
    uint64_t pixels;
};

struct Snapshot {
    float4 *xyth = nullptr;
    float2 *sc = nullptr;
    float *res = nullptr;
    float4 *canvas = nullptr;
    int w = 0, h = 0;
};

} // namespace

struct pt_pyramid {
    pt_ctx *ctx;
    int elem, w, h, levels;
    int lw[32], lh[32];
    size_t off[32];
    float *data; /* one allocation holding every level */
};

struct pt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    pt_params p;
    bool have_params = false;
    float gauss[2 * PT_MAX_R + 1];
    int gauss_r = 0;
    /* working buffers */
    int W = 0, H = 0, SW = 0, SH = 0;
    PtNnf base{}, work{};
    float *res_base = nullptr, *res_work = nullptr;
    float4 *canvas = nullptr;
    DevAdv adv[2];
    bool adv_written[2] = {false, false}; /* advected set k may hold valid offsets */
    bool no_early_exit = false;            /* force the plain candidate kernels (tests) */
    /* Early rejection lengthens each lane's dependent chain; it pays only on
     * levels with enough waves to hide the latency. */
    int64_t early_exit_min_pixels = 512 * 512;
    bool unit_parallel = true; /* candidate units over workgroup lanes */
    Rect off_diff, res_diff;
    /* image textures */
    float4 *color[2] = {nullptr, nullptr};
    float *dt[2] = {nullptr, nullptr};
    float2 *velb[2] = {nullptr, nullptr}, *velf[2] = {nullptr, nullptr};
    int cur = 0;
    float *orient = nullptr;
    int ow = 0, oh = 0;
    float2 *scale = nullptr;
    int scw = 0, sch = 0;
    float4 *random = nullptr;
    int rw = 0, rh = 0;
    /* exemplar textures (capacity ex_cap texels) */
    float4 *ex = nullptr;
    float *ex_orient = nullptr, *ex_dt = nullptr;
    size_t ex_cap = 0;
    int ew = 0, eh = 0;
    uint32_t scatter_calls = 0;
    uint64_t pde = 0;
    /* schedule results */
    std::vector<Snapshot> snaps;
    int snap_levels = 0, snap_frames = 0, snap_passes = 0;
    /* device pyramids and staging reused across pt_synthesize calls */
    std::vector<FramePyr> frame_pyrs;
    StylePyr style_pyr;
    pt_patchxf *pack_tmp = nullptr;
    size_t pack_cap = 0;
    /* optional per-launch timing with HIP events on the context stream */
    bool prof_on = false;
    std::vector<hipEvent_t> prof_events;
    std::vector<ProfRec> prof_recs;
};

static void free_frame(FramePyr &f);
static void free_style(StylePyr &sp);
static void prof_clear(pt_ctx *c);

static int set_err(pt_ctx *c, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    snprintf(g_last_error, sizeof(g_last_error), "%s", buf);
    return code;
}

#define HIP_CHECK(c, expr)                                                                          \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return set_err((c), PT_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));         \
    } while (0)

#define RC_CHECK(expr)                                                                              \
    do {                                                                                            \
        int rc_ = (expr);                                                                           \
        if (rc_ != PT_OK) return rc_;                                                               \
    } while (0)

template <class T> static int dalloc(pt_ctx *c, T **p, size_t n)
{
    *p = nullptr;
    if (n == 0) return PT_OK;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return set_err(c, PT_ERR_OUT_OF_MEMORY, "hipMalloc of %zu bytes failed: %s", n * sizeof(T),
                       hipGetErrorString(e));
    }
    e = hipMemsetAsync(*p, 0, n * sizeof(T), c->stream);
    if (e != hipSuccess) return set_err(c, PT_ERR_HIP, "hipMemsetAsync failed: %s", hipGetErrorString(e));
    return PT_OK;
}

template <class T> static void dfree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

static int elem_channels(int elem) { return (elem >= 1 && elem <= 4) ? elem : 0; }

static int check_supported(pt_ctx *c, const pt_params *p)
{
    if (p->residual_window_size < 1 || p->residual_window_size > PT_MAX_R)
        return set_err(c, PT_ERR_INVALID_ARG, "residual_window_size must be in [1, %d]", PT_MAX_R);
    if (p->num_styles != 1) return set_err(c, PT_ERR_UNSUPPORTED, "only one style layer is supported");
    if (p->coherence_weight != 0.f) return set_err(c, PT_ERR_UNSUPPORTED, "coherence term is not supported");
    if (p->colorspace != 0) return set_err(c, PT_ERR_UNSUPPORTED, "only the RGB colourspace is supported");
    if (p->lum_shift_enable) return set_err(c, PT_ERR_UNSUPPORTED, "luminance shift is not supported");
    if (p->style_obj_id_weight != 0.f) return set_err(c, PT_ERR_UNSUPPORTED, "object-id term is not supported");
    if (p->interpolate_key_frame)
        return set_err(c, PT_ERR_UNSUPPORTED, "keyframe interpolation is not supported");
    for (int i = 0; i < 5; i++)
        if (p->offsets_histogram_slope[i] != 0.f)
            return set_err(c, PT_ERR_UNSUPPORTED, "offset histogram term is not supported");
    if (p->rng_mode != PT_RNG_REFERENCE && p->rng_mode != PT_RNG_COUNTER)
        return set_err(c, PT_ERR_INVALID_ARG, "unknown rng_mode");
    return PT_OK;
}

/* ------------------------------------------------------------------ */
/* parameters and context                                              */
/* ------------------------------------------------------------------ */
extern "C" void pt_params_default(pt_params *p)
{
    if (!p) return;
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
    p->dist_trans_cull = 50;
    p->residual_window_scale_u = 1;
    p->residual_window_scale_v = 1;
    p->first_key_index = -1;
    p->last_key_index = -1;
    p->direction = PT_FORWARD;
    p->rng_mode = PT_RNG_REFERENCE;
}

extern "C" const char *pt_global_last_error(void) { return g_last_error; }

extern "C" size_t pt_struct_size(int which)
{
    switch (which) {
    case 0: return sizeof(pt_params);
    case 1: return sizeof(pt_patchxf);
    case 2: return sizeof(pt_frame_guides);
    case 3: return sizeof(pt_style_guides);
    default: return 0;
    }
}

extern "C" const char *pt_last_error(pt_ctx *c) { return c ? c->err.c_str() : g_last_error; }

extern "C" int pt_create(pt_ctx **out, int device)
{
    if (!out) return set_err(nullptr, PT_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        return set_err(nullptr, PT_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= n) return set_err(nullptr, PT_ERR_INVALID_ARG, "device %d out of range", device);
    e = hipSetDevice(device);
    if (e != hipSuccess) return set_err(nullptr, PT_ERR_HIP, "hipSetDevice failed: %s", hipGetErrorString(e));
    pt_ctx *c = new pt_ctx();
    c->device = device;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return set_err(nullptr, PT_ERR_HIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
    pt_params_default(&c->p);
    *out = c;
    return PT_OK;
}

static void free_snaps(pt_ctx *c)
{
    for (auto &s : c->snaps) {
        dfree(s.xyth);
        dfree(s.sc);
        dfree(s.res);
        dfree(s.canvas);
    }
    c->snaps.clear();
    c->snap_levels = c->snap_frames = c->snap_passes = 0;
}

static void free_working(pt_ctx *c)
{
    dfree(c->base.xyth);
    dfree(c->base.sc);
    dfree(c->work.xyth);
    dfree(c->work.sc);
    dfree(c->res_base);
    dfree(c->res_work);
    dfree(c->canvas);
    for (int k = 0; k < 2; k++) {
        dfree(c->adv[k].xyth);
        dfree(c->adv[k].sc);
        dfree(c->adv[k].base);
        dfree(c->adv[k].out);
        dfree(c->adv[k].res);
        dfree(c->adv[k].ts);
        dfree(c->color[k]);
        dfree(c->dt[k]);
        dfree(c->velb[k]);
        dfree(c->velf[k]);
    }
    dfree(c->orient);
    dfree(c->scale);
    dfree(c->random);
}

extern "C" int pt_destroy(pt_ctx *c)
{
    if (!c) return PT_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_snaps(c);
    free_working(c);
    for (auto &f : c->frame_pyrs) free_frame(f);
    free_style(c->style_pyr);
    dfree(c->pack_tmp);
    prof_clear(c);
    dfree(c->ex);
    dfree(c->ex_orient);
    dfree(c->ex_dt);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return PT_OK;
}

static int sync_stream(pt_ctx *c)
{
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return PT_OK;
}

static int prof_event(pt_ctx *c)
{
    hipEvent_t e;
    HIP_CHECK(c, hipEventCreate(&e));
    c->prof_events.push_back(e);
    return (int)c->prof_events.size() - 1;
}

/* Records the start of a timed launch; returns the record index or -1. */
static int prof_begin(pt_ctx *c, int kind, uint64_t pixels)
{
    if (!c->prof_on) return -1;
    ProfRec r;
    r.kind = kind;
    r.pixels = pixels;
    r.ev0 = prof_event(c);
    r.ev1 = prof_event(c);
    if (r.ev0 < 0 || r.ev1 < 0) return -1;
    if (hipEventRecord(c->prof_events[r.ev0], c->stream) != hipSuccess) return -1;
    c->prof_recs.push_back(r);
    return (int)c->prof_recs.size() - 1;
}

static void prof_end(pt_ctx *c, int rec)
{
    if (rec < 0) return;
    (void)hipEventRecord(c->prof_events[c->prof_recs[rec].ev1], c->stream);
}

static void prof_clear(pt_ctx *c)
{
    (void)hipStreamSynchronize(c->stream);
    for (auto e : c->prof_events) (void)hipEventDestroy(e);
    c->prof_events.clear();
    c->prof_recs.clear();
}

extern "C" int pt_profile_enable(pt_ctx *c, int enable)
{
    if (!c) return set_err(nullptr, PT_ERR_INVALID_ARG, "NULL context");
    HIP_CHECK(c, hipSetDevice(c->device));
    prof_clear(c);
    c->prof_on = enable != 0;
    return PT_OK;
}

extern "C" int pt_profile_get(pt_ctx *c, const char *kernel, uint64_t *launches, double *total_ms, uint64_t *pixels)
{
    if (!c || !kernel || !launches || !total_ms || !pixels) return set_err(c, PT_ERR_INVALID_ARG, "NULL argument");
    int kind = -1;
    for (int k = 0; k < PK_COUNT; k++)
        if (strcmp(kernel, kProfNames[k]) == 0) kind = k;
    if (kind < 0) return set_err(c, PT_ERR_INVALID_ARG, "unknown kernel class '%s'", kernel);
    HIP_CHECK(c, hipSetDevice(c->device));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    *launches = 0;
    *total_ms = 0.0;
    *pixels = 0;
    for (const auto &r : c->prof_recs) {
        if (r.kind != kind) continue;
        float ms = 0.f;
        HIP_CHECK(c, hipEventElapsedTime(&ms, c->prof_events[r.ev0], c->prof_events[r.ev1]));
        *launches += 1;
        *total_ms += ms;
        *pixels += r.pixels;
    }
    return PT_OK;
}

extern "C" int pt_set_tuning(pt_ctx *c, int key, int value)
{
    if (!c) return set_err(nullptr, PT_ERR_INVALID_ARG, "NULL context");
    switch (key) {
    case PT_TUNE_EARLY_EXIT:
        c->no_early_exit = (value == 0);
        return PT_OK;
    case PT_TUNE_UNIT_PARALLEL:
        c->unit_parallel = value != 0;
        return PT_OK;
    case PT_TUNE_EARLY_EXIT_MIN_PIXELS:
        if (value < 0) return set_err(c, PT_ERR_INVALID_ARG, "negative pixel count");
        c->early_exit_min_pixels = value;
        return PT_OK;
    default:
        return set_err(c, PT_ERR_INVALID_ARG, "unknown tuning key %d", key);
    }
}

extern "C" int pt_upload_parameters(pt_ctx *c, const pt_params *p)
{
    if (!c || !p) return set_err(c, PT_ERR_INVALID_ARG, "NULL argument");
    RC_CHECK(check_supported(c, p));
    c->p = *p;
    c->have_params = true;
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

extern "C" int pt_setup(pt_ctx *c, const pt_params *p)
{
    RC_CHECK(pt_upload_parameters(c, p));
    if (c->gauss_r != p->residual_window_size) {
        gaussian(c->gauss, p->residual_window_size);
        c->gauss_r = p->residual_window_size;
    }
    return PT_OK;
}

extern "C" int pt_initialize(pt_ctx *c, int iw, int ih, int sw, int sh)
{
    if (!c || iw < 2 || ih < 2 || sw < 2 || sh < 2) return set_err(c, PT_ERR_INVALID_ARG, "bad dimensions");
    HIP_CHECK(c, hipSetDevice(c->device));
    if (c->base.xyth && c->W == iw && c->H == ih && c->SW == sw && c->SH == sh) return PT_OK;
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    free_working(c);
    free_snaps(c);
    size_t n = (size_t)iw * ih;
    c->W = iw;
    c->H = ih;
    c->SW = sw;
    c->SH = sh;
    RC_CHECK(dalloc(c, &c->base.xyth, n));
    RC_CHECK(dalloc(c, &c->base.sc, n));
    RC_CHECK(dalloc(c, &c->work.xyth, n));
    RC_CHECK(dalloc(c, &c->work.sc, n));
    RC_CHECK(dalloc(c, &c->res_base, n));
    RC_CHECK(dalloc(c, &c->res_work, n));
    RC_CHECK(dalloc(c, &c->canvas, n));
    for (int k = 0; k < 2; k++) {
        RC_CHECK(dalloc(c, &c->adv[k].xyth, n));
        RC_CHECK(dalloc(c, &c->adv[k].sc, n));
        RC_CHECK(dalloc(c, &c->adv[k].base, n));
        RC_CHECK(dalloc(c, &c->adv[k].out, n));
        RC_CHECK(dalloc(c, &c->adv[k].res, n));
        RC_CHECK(dalloc(c, &c->adv[k].ts, n));
        RC_CHECK(dalloc(c, &c->color[k], n));
        RC_CHECK(dalloc(c, &c->dt[k], n));
        RC_CHECK(dalloc(c, &c->velb[k], n));
        RC_CHECK(dalloc(c, &c->velf[k], n));
    }
    /* level-sized textures never exceed level 0 */
    RC_CHECK(dalloc(c, &c->orient, n));
    RC_CHECK(dalloc(c, &c->scale, n));
    RC_CHECK(dalloc(c, &c->random, n));
    c->ow = c->oh = c->scw = c->sch = c->rw = c->rh = 0;
    c->off_diff = Rect();
    c->res_diff = Rect();
    c->adv_written[0] = c->adv_written[1] = false;
    c->cur = 0;
    return sync_stream(c);
}

extern "C" int pt_initialize_exemplars(pt_ctx *c, int w, int h, int num_layers)
{
    if (!c || w < 1 || h < 1) return set_err(c, PT_ERR_INVALID_ARG, "bad exemplar dimensions");
    if (num_layers != 1) return set_err(c, PT_ERR_UNSUPPORTED, "one style layer only");
    HIP_CHECK(c, hipSetDevice(c->device));
    if (c->ex && c->ew == w && c->eh == h) return PT_OK;
    size_t n = (size_t)w * h;
    if (n > c->ex_cap) {
        HIP_CHECK(c, hipStreamSynchronize(c->stream));
        dfree(c->ex);
        dfree(c->ex_orient);
        dfree(c->ex_dt);
        RC_CHECK(dalloc(c, &c->ex, 2 * n));
        RC_CHECK(dalloc(c, &c->ex_orient, n));
        RC_CHECK(dalloc(c, &c->ex_dt, n));
        c->ex_cap = n;
    } else {
        /* a fresh texture allocation of the new size: zero it */
        HIP_CHECK(c, hipMemsetAsync(c->ex, 0, 2 * n * sizeof(float4), c->stream));
        HIP_CHECK(c, hipMemsetAsync(c->ex_orient, 0, n * sizeof(float), c->stream));
        HIP_CHECK(c, hipMemsetAsync(c->ex_dt, 0, n * sizeof(float), c->stream));
    }
    c->ew = w;
    c->eh = h;
    return sync_stream(c);
}

static int reset_advected(pt_ctx *c)
{
    size_t n = (size_t)c->W * c->H;
    for (int k = 0; k < 2; k++) {
        HIP_CHECK(c, hipMemsetAsync(c->adv[k].xyth, 0, n * sizeof(float4), c->stream));
        HIP_CHECK(c, hipMemsetAsync(c->adv[k].sc, 0, n * sizeof(float2), c->stream));
        HIP_CHECK(c, hipMemsetAsync(c->adv[k].base, 0, n * sizeof(float4), c->stream));
        HIP_CHECK(c, hipMemsetAsync(c->adv[k].out, 0, n * sizeof(float4), c->stream));
        HIP_CHECK(c, hipMemsetAsync(c->adv[k].res, 0, n * sizeof(float), c->stream));
        HIP_CHECK(c, hipMemsetAsync(c->adv[k].ts, 0, n * sizeof(int32_t), c->stream));
        c->adv_written[k] = false;
    }
    return PT_OK;
}

extern "C" int pt_reset_advected_buffers(pt_ctx *c)
{
    if (!c || !c->base.xyth) return set_err(c, PT_ERR_STATE, "not initialized");
    RC_CHECK(reset_advected(c));
    return sync_stream(c);
}

extern "C" int pt_flip_anim_buffers(pt_ctx *c)
{
    if (!c) return set_err(nullptr, PT_ERR_INVALID_ARG, "NULL context");
    c->cur = (c->cur + 1) % 2;
    return PT_OK;
}

/* ------------------------------------------------------------------ */
/* pyramids                                                            */
/* ------------------------------------------------------------------ */
static int pyramid_rebuild(pt_ctx *c, pt_pyramid *p, int scale_mode)
{
    int ch = elem_channels(p->elem);
    for (int l = 1; l < p->levels; l++) {
        HIP_CHECK(c, pt_launch_downsample(p->data + p->off[l - 1], p->lw[l - 1], p->data + p->off[l], p->lw[l],
                                          p->lh[l], ch, scale_mode, c->stream));
    }
    return PT_OK;
}

static int pyramid_alloc(pt_ctx *c, pt_pyramid **out, int elem, int w, int h)
{
    int ch = elem_channels(elem);
    if (!out || !ch || w < 2 || h < 2) return set_err(c, PT_ERR_INVALID_ARG, "bad pyramid arguments");
    pt_pyramid *p = new pt_pyramid();
    p->ctx = c;
    p->elem = elem;
    p->w = w;
    p->h = h;
    p->levels = (int)fmin(floor(log2f((float)w)), floor(log2f((float)h)));
    size_t total = 0;
    for (int l = 0; l < p->levels; l++) {
        p->lw[l] = w >> l;
        p->lh[l] = h >> l;
        p->off[l] = total;
        total += (size_t)p->lw[l] * p->lh[l] * ch;
        total = (total + 63) & ~(size_t)63; /* 256-byte aligned levels */
    }
    int rc = dalloc(c, &p->data, total);
    if (rc) {
        delete p;
        return rc;
    }
    *out = p;
    return PT_OK;
}

extern "C" int pt_pyramid_create(pt_ctx *c, pt_pyramid **out, int elem, int w, int h, int scale_mode,
                                 const void *host0, float defv)
{
    if (!c) return set_err(nullptr, PT_ERR_INVALID_ARG, "NULL context");
    HIP_CHECK(c, hipSetDevice(c->device));
    pt_pyramid *p = nullptr;
    RC_CHECK(pyramid_alloc(c, &p, elem, w, h));
    size_t n0 = (size_t)w * h * elem_channels(elem);
    hipError_t e = host0 ? hipMemcpyAsync(p->data, host0, n0 * sizeof(float), hipMemcpyHostToDevice, c->stream)
                         : pt_launch_fill(p->data, n0, defv, c->stream);
    if (e != hipSuccess) {
        pt_pyramid_destroy(p);
        return set_err(c, PT_ERR_HIP, "level-0 upload failed: %s", hipGetErrorString(e));
    }
    int rc = pyramid_rebuild(c, p, scale_mode);
    if (!rc) rc = sync_stream(c);
    if (rc) {
        pt_pyramid_destroy(p);
        return rc;
    }
    *out = p;
    return PT_OK;
}

extern "C" int pt_pyramid_orientation(pt_ctx *c, pt_pyramid **out, const pt_pyramid *t)
{
    if (!c || !out || !t || t->elem != PT_ELEM_F32X3) return set_err(c, PT_ERR_INVALID_ARG, "need a tensor pyramid");
    HIP_CHECK(c, hipSetDevice(c->device));
    pt_pyramid *p = nullptr;
    RC_CHECK(pyramid_alloc(c, &p, PT_ELEM_F32, t->w, t->h));
    for (int l = 0; l < t->levels; l++) {
        hipError_t e = pt_launch_tensor_angle(t->data + t->off[l], p->data + p->off[l], (size_t)t->lw[l] * t->lh[l],
                                              c->stream);
        if (e != hipSuccess) {
            pt_pyramid_destroy(p);
            return set_err(c, PT_ERR_HIP, "tensor angle failed: %s", hipGetErrorString(e));
        }
    }
    int rc = sync_stream(c);
    if (rc) {
        pt_pyramid_destroy(p);
        return rc;
    }
    *out = p;
    return PT_OK;
}

extern "C" int pt_pyramid_destroy(pt_pyramid *p)
{
    if (!p) return PT_OK;
    if (p->ctx) {
        (void)hipSetDevice(p->ctx->device);
        (void)hipStreamSynchronize(p->ctx->stream);
    }
    dfree(p->data);
    delete p;
    return PT_OK;
}

extern "C" int pt_pyramid_levels(const pt_pyramid *p) { return p ? p->levels : 0; }

extern "C" int pt_pyramid_dims(const pt_pyramid *p, int level, int *w, int *h)
{
    if (!p || !w || !h || level < 0 || level >= p->levels) return PT_ERR_INVALID_ARG;
    *w = p->lw[level];
    *h = p->lh[level];
    return PT_OK;
}

extern "C" int pt_pyramid_download(const pt_pyramid *p, int level, void *host)
{
    if (!p || !host || level < 0 || level >= p->levels) return PT_ERR_INVALID_ARG;
    pt_ctx *c = p->ctx;
    HIP_CHECK(c, hipSetDevice(c->device));
    size_t n = (size_t)p->lw[level] * p->lh[level] * elem_channels(p->elem);
    HIP_CHECK(c, hipMemcpyAsync(host, p->data + p->off[level], n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    return sync_stream(c);
}

/* TexSynth::uploadImage_* (texSynth_kernel.cu:2128-2253), device to device */
static int upload_image(pt_ctx *c, int level, int layer, const pt_pyramid *p, int style_layer)
{
    if (!c || !p || level < 0 || level >= p->levels) return set_err(c, PT_ERR_INVALID_ARG, "bad upload");
    if (!c->base.xyth) return set_err(c, PT_ERR_STATE, "not initialized");
    if (style_layer != 0) return set_err(c, PT_ERR_UNSUPPORTED, "one style layer only");
    int w = p->lw[level], h = p->lh[level];
    const float *src = p->data + p->off[level];
    int prev = (c->cur + 1) % 2;
    hipStream_t s = c->stream;
    auto rect = [&](void *dst, int ch) -> int {
        if (w > c->W || h > c->H) return set_err(c, PT_ERR_INVALID_ARG, "level larger than image");
        HIP_CHECK(c, hipMemcpy2DAsync(dst, (size_t)c->W * ch * sizeof(float), src, (size_t)w * ch * sizeof(float),
                                      (size_t)w * ch * sizeof(float), h, hipMemcpyDeviceToDevice, s));
        return PT_OK;
    };
    auto tight = [&](void *dst, int ch) -> int {
        HIP_CHECK(c, hipMemcpyAsync(dst, src, (size_t)w * h * ch * sizeof(float), hipMemcpyDeviceToDevice, s));
        return PT_OK;
    };
    switch (layer) {
    case PT_LAYER_INPUT_COLOR:
        if (p->elem != 4) break;
        return rect(c->color[c->cur], 4);
    case PT_LAYER_INPUT_DIST_TRANS:
        if (p->elem != 1) break;
        return rect(c->dt[c->cur], 1);
    case PT_LAYER_INPUT_VEL_B:
        if (p->elem != 2) break;
        return rect(c->velb[c->cur], 2);
    case PT_LAYER_INPUT_VEL_F:
        if (p->elem != 2) break;
        return rect(c->velf[c->cur], 2);
    case PT_LAYER_INPUT_VEL_B_PREVIOUS:
        if (p->elem != 2) break;
        return rect(c->velb[prev], 2);
    case PT_LAYER_INPUT_VEL_F_PREVIOUS:
        if (p->elem != 2) break;
        return rect(c->velf[prev], 2);
    case PT_LAYER_INPUT_ORIENTATION:
        if (p->elem != 1 || w > c->W || h > c->H) break;
        c->ow = w;
        c->oh = h;
        return tight(c->orient, 1);
    case PT_LAYER_INPUT_SCALE:
        if (p->elem != 2 || w > c->W || h > c->H) break;
        c->scw = w;
        c->sch = h;
        return tight(c->scale, 2);
    case PT_LAYER_RANDOM:
        if (p->elem != 4 || w > c->W || h > c->H) break;
        c->rw = w;
        c->rh = h;
        return tight(c->random, 4);
    case PT_LAYER_EXEMPLAR_BASE:
    case PT_LAYER_EXEMPLAR_OUTPUT:
        if (p->elem != 4 || !c->ex || w > c->ew || h > c->eh) break;
        HIP_CHECK(c, pt_launch_interleave((const float4 *)src, w, h, c->ex, c->ew,
                                          layer == PT_LAYER_EXEMPLAR_BASE ? 0 : 1, s));
        return PT_OK;
    case PT_LAYER_EXEMPLAR_ORIENTATION:
    case PT_LAYER_EXEMPLAR_DIST_TRANS: {
        if (p->elem != 1 || !c->ex || w > c->ew || h > c->eh) break;
        float *dst = layer == PT_LAYER_EXEMPLAR_ORIENTATION ? c->ex_orient : c->ex_dt;
        HIP_CHECK(c, hipMemcpy2DAsync(dst, (size_t)c->ew * sizeof(float), src, (size_t)w * sizeof(float),
                                      (size_t)w * sizeof(float), h, hipMemcpyDeviceToDevice, s));
        return PT_OK;
    }
    default:
        return set_err(c, PT_ERR_UNSUPPORTED, "layer %d not supported", layer);
    }
    return set_err(c, PT_ERR_INVALID_ARG, "pyramid kind or size does not fit layer %d", layer);
}

extern "C" int pt_upload_image(pt_ctx *c, int level, int layer, const pt_pyramid *p, int style_layer)
{
    if (!c) return set_err(nullptr, PT_ERR_INVALID_ARG, "NULL context");
    HIP_CHECK(c, hipSetDevice(c->device));
    RC_CHECK(upload_image(c, level, layer, p, style_layer));
    return sync_stream(c);
}

static int random_pyramid(pt_ctx *c, int level, int frame, const pt_pyramid *icol, const pt_pyramid *iori,
                          const pt_pyramid *eout, const pt_pyramid *eori)
{
    if (!c || !icol || !iori || !eout || !eori) return set_err(c, PT_ERR_INVALID_ARG, "NULL pyramid");
    if (!c->base.xyth || !c->have_params) return set_err(c, PT_ERR_STATE, "not initialized");
    if (icol->elem != 4 || iori->elem != 1 || eout->elem != 4 || eori->elem != 1)
        return set_err(c, PT_ERR_INVALID_ARG, "pyramid kinds");
    if (level < 0 || level >= icol->levels || level >= iori->levels || level >= eout->levels ||
        level >= eori->levels)
        return set_err(c, PT_ERR_INVALID_ARG, "level out of range");
    int w = icol->lw[level], h = icol->lh[level], ew = eout->lw[level], eh = eout->lh[level];
    if (iori->lw[level] != w || iori->lh[level] != h || eori->lw[level] != ew || eori->lh[level] != eh)
        return set_err(c, PT_ERR_INVALID_ARG, "pyramid sizes differ");
    if (ew < 2 || eh < 2) return set_err(c, PT_ERR_INVALID_ARG, "exemplar level too small");
    if (w > c->W || h > c->H) return set_err(c, PT_ERR_INVALID_ARG, "level larger than image");
    HIP_CHECK(c, pt_launch_random_pyramid((const float4 *)(icol->data + icol->off[level]), iori->data + iori->off[level],
                                          w, h, (const float4 *)(eout->data + eout->off[level]),
                                          eori->data + eori->off[level], ew, eh, c->p.transparency_ok,
                                          c->p.orientation_enabled, c->p.seed, (uint32_t)(frame * level), c->random,
                                          c->stream));
    c->rw = w;
    c->rh = h;
    return PT_OK;
}

extern "C" int pt_create_random_pyramid(pt_ctx *c, int level, int frame, const pt_pyramid *icol,
                                        const pt_pyramid *iori, const pt_pyramid *eout, const pt_pyramid *eori)
{
    if (!c) return set_err(nullptr, PT_ERR_INVALID_ARG, "NULL context");
    HIP_CHECK(c, hipSetDevice(c->device));
    RC_CHECK(random_pyramid(c, level, frame, icol, iori, eout, eori));
    return sync_stream(c);
}

/* ------------------------------------------------------------------ */
/* operators                                                           */
/* ------------------------------------------------------------------ */
static int need_level(pt_ctx *c, int level)
{
    if (!c) return set_err(nullptr, PT_ERR_INVALID_ARG, "NULL context");
    if (!c->base.xyth) return set_err(c, PT_ERR_STATE, "not initialized");
    if (!c->have_params || c->gauss_r != c->p.residual_window_size)
        return set_err(c, PT_ERR_STATE, "setup not called for these parameters");
    if (!c->ex) return set_err(c, PT_ERR_STATE, "no exemplar");
    if (level < 0 || level > 30 || (c->W >> level) < 1 || (c->H >> level) < 1)
        return set_err(c, PT_ERR_INVALID_ARG, "bad level %d", level);
    if (c->ow == 0 || c->scw == 0) return set_err(c, PT_ERR_STATE, "orientation/scale not uploaded");
    return PT_OK;
}

static PtArgs make_args(const pt_ctx *c, int level)
{
    PtArgs a;
    memset(&a, 0, sizeof(a));
    const pt_params &p = c->p;
    a.W = c->W;
    a.H = c->H;
    a.level = level;
    a.lw = c->W >> level;
    a.lh = c->H >> level;
    a.hw = c->SW >> level;
    a.hh = (c->SH * p.num_styles) >> level;
    a.ew = c->ew;
    a.eh = c->eh;
    a.ow = c->ow;
    a.oh = c->oh;
    a.scw = c->scw;
    a.sch = c->sch;
    a.rw = c->rw;
    a.rh = c->rh;
    a.cur = c->cur;
    a.r = p.residual_window_size;
    for (int i = 0; i < 2 * a.r + 1; i++) a.gauss[i] = c->gauss[i];
    a.alpha_w = p.alpha_weight;
    a.in_w = p.input_analogy_weight;
    a.canvas_w = p.canvas_output_weight;
    a.dt_w = p.dist_trans_weight;
    a.su = p.residual_window_scale_u;
    a.sv = p.residual_window_scale_v;
    a.cull = p.dist_trans_cull;
    a.transp = p.transparency_ok;
    a.orient_en = p.orientation_enabled;
    a.direction = p.direction;
    a.tdi = p.time_derivative_input_weight;
    a.tdo = p.time_derivative_output_weight;
    a.tcw = p.temporal_coherence_weight;
    a.hyst_w = p.hysteresis_weight;
    a.max_dist = p.max_distortion;
    a.jitter = p.advection_jitter;
    a.samples = p.scatter_samples_count;
    a.rng_mode = p.rng_mode;
    a.key0 = p.seed ^ 0x5CA77E12u;
    a.key1 = (uint32_t)(p.frame_number * level);
    a.time_active = (c->adv_written[0] || c->adv_written[1]) ? 1 : 0;
    a.early_exit = (p.hysteresis_weight == 0.f && p.input_analogy_weight >= 0.f && p.canvas_output_weight >= 0.f &&
                    p.alpha_weight >= 0.f && p.dist_trans_weight >= 0.f && p.time_derivative_input_weight >= 0.f &&
                    p.time_derivative_output_weight >= 0.f && p.temporal_coherence_weight >= 0.f && !c->no_early_exit &&
                    (int64_t)a.lw * a.lh >= c->early_exit_min_pixels)
                       ? 1
                       : 0;
    a.unit_parallel = c->unit_parallel ? 1 : 0;
    a.base = c->base;
    a.work = c->work;
    a.res_base = c->res_base;
    a.res_work = c->res_work;
    a.canvas = c->canvas;
    for (int k = 0; k < 2; k++) {
        a.adv[k].xyth = c->adv[k].xyth;
        a.adv[k].sc = c->adv[k].sc;
        a.adv[k].base = c->adv[k].base;
        a.adv[k].out = c->adv[k].out;
        a.adv[k].res = c->adv[k].res;
        a.adv[k].ts = c->adv[k].ts;
        a.color[k] = c->color[k];
        a.dt[k] = c->dt[k];
        a.velb[k] = c->velb[k];
        a.velf[k] = c->velf[k];
    }
    a.orient = c->orient;
    a.scale = c->scale;
    a.random = c->random;
    a.ex = c->ex;
    a.ex_orient = c->ex_orient;
    a.ex_dt = c->ex_dt;
    return a;
}

/* copyWorkingToBase over the rectangle where the pair can differ */
static int copy_nnf_rect(pt_ctx *c, const PtNnf &dst, const PtNnf &src, Rect r)
{
    if (r.empty()) return PT_OK;
    HIP_CHECK(c, hipMemcpy2DAsync(dst.xyth, (size_t)c->W * sizeof(float4), src.xyth, (size_t)c->W * sizeof(float4),
                                  (size_t)r.w * sizeof(float4), r.h, hipMemcpyDeviceToDevice, c->stream));
    HIP_CHECK(c, hipMemcpy2DAsync(dst.sc, (size_t)c->W * sizeof(float2), src.sc, (size_t)c->W * sizeof(float2),
                                  (size_t)r.w * sizeof(float2), r.h, hipMemcpyDeviceToDevice, c->stream));
    return PT_OK;
}

static int copy_working_to_base(pt_ctx *c)
{
    RC_CHECK(copy_nnf_rect(c, c->base, c->work, c->off_diff));
    c->off_diff = Rect();
    return PT_OK;
}

static int copy_base_to_working(pt_ctx *c)
{
    RC_CHECK(copy_nnf_rect(c, c->work, c->base, c->off_diff));
    c->off_diff = Rect();
    return PT_OK;
}

static int update_residual(pt_ctx *c, int level)
{
    PtArgs a = make_args(c, level);
    int pr = prof_begin(c, PK_RESIDUAL, (uint64_t)a.lw * a.lh);
    HIP_CHECK(c, pt_launch_update_residual(a, c->stream));
    prof_end(c, pr);
    c->res_diff.grow(a.lw, a.lh);
    c->pde += (uint64_t)a.lw * a.lh;
    return PT_OK;
}

/* TexSynth::syncBuffers (texSynth_kernel.cu:1240-1303) */
static int sync_buffers(pt_ctx *c, int level, int flags)
{
    PtArgs a = make_args(c, level);
    int pr = prof_begin(c, PK_CANVAS, (uint64_t)a.lw * a.lh);
    if (!(flags & PT_DONT_COPY_OFFSETS) && !c->off_diff.empty()) {
        /* copy of the differing rectangle fused with the canvas pass */
        HIP_CHECK(c, pt_launch_sync_canvas(a, c->off_diff.w, c->off_diff.h, c->stream));
        c->off_diff = Rect();
    } else {
        HIP_CHECK(c, pt_launch_canvas(a, c->stream));
    }
    prof_end(c, pr);
    if (!(flags & PT_DONT_UPDATE_RESIDUAL)) RC_CHECK(update_residual(c, level));
    return PT_OK;
}

static int randomize_offsets(pt_ctx *c, int level)
{
    if (c->rw == 0) return set_err(c, PT_ERR_STATE, "no random texture");
    PtArgs a = make_args(c, level);
    int pr = prof_begin(c, PK_RANDOM, (uint64_t)a.lw * a.lh);
    HIP_CHECK(c, pt_launch_copy_random(a, c->stream));
    prof_end(c, pr);
    c->off_diff.grow(a.lw, a.lh);
    return sync_buffers(c, level, 0);
}

static int propagate(pt_ctx *c, int level, int jump)
{
    PtArgs a = make_args(c, level);
    int pr = prof_begin(c, PK_PROPAGATE, (uint64_t)a.lw * a.lh);
    HIP_CHECK(c, pt_launch_propagate(a, jump, c->stream));
    prof_end(c, pr);
    c->off_diff.grow(a.lw, a.lh);
    c->pde += (uint64_t)a.lw * a.lh * 4;
    return sync_buffers(c, level, 0);
}

static int scatter_samples(int ew, int eh, int samples)
{
    float dist = (float)(ew > eh ? ew : eh);
    int n = 0;
    for (int i = 0; (i < samples) && (dist * 0.9f >= 10.f); i++) {
        n++;
        dist *= 0.9f;
    }
    return n;
}

static int scatter(pt_ctx *c, int level)
{
    PtArgs a = make_args(c, level);
    a.call = c->scatter_calls++;
    int pr = prof_begin(c, PK_SCATTER, (uint64_t)a.lw * a.lh);
    HIP_CHECK(c, pt_launch_scatter(a, c->stream));
    prof_end(c, pr);
    c->off_diff.grow(a.lw, a.lh);
    c->pde += (uint64_t)a.lw * a.lh * (uint64_t)(scatter_samples(c->ew, c->eh, a.samples) + 1);
    return sync_buffers(c, level, 0);
}

static int upsample(pt_ctx *c, int fine_level)
{
    PtArgs a = make_args(c, fine_level);
    int pr = prof_begin(c, PK_UPSAMPLE, (uint64_t)a.lw * a.lh);
    HIP_CHECK(c, pt_launch_upsample(a, c->stream));
    prof_end(c, pr);
    c->off_diff.grow(a.lw, a.lh);
    c->res_diff.grow(a.lw, a.lh);
    RC_CHECK(copy_working_to_base(c));
    if (!c->res_diff.empty()) {
        HIP_CHECK(c, hipMemcpy2DAsync(c->res_base, (size_t)c->W * sizeof(float), c->res_work,
                                      (size_t)c->W * sizeof(float), (size_t)c->res_diff.w * sizeof(float),
                                      c->res_diff.h, hipMemcpyDeviceToDevice, c->stream));
    }
    c->res_diff = Rect();
    return PT_OK;
}

static int advect(pt_ctx *c, int level, int time_step, int forwards, int flags)
{
    if (c->rw == 0) return set_err(c, PT_ERR_STATE, "no random texture");
    PtArgs a = make_args(c, level);
    int pr = prof_begin(c, PK_ADVECT, (uint64_t)a.lw * a.lh);
    HIP_CHECK(c, pt_launch_advect(a, forwards, time_step, flags, c->stream));
    prof_end(c, pr);
    c->adv_written[forwards ? 0 : 1] = true;
    DevAdv &d = c->adv[forwards ? 0 : 1];
    size_t n = (size_t)c->W * c->H;
    /* dest.offsets -> base, whole buffer (texSynth_kernel.cu:1716) */
    HIP_CHECK(c, hipMemcpyAsync(c->base.xyth, d.xyth, n * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
    HIP_CHECK(c, hipMemcpyAsync(c->base.sc, d.sc, n * sizeof(float2), hipMemcpyDeviceToDevice, c->stream));
    c->off_diff.grow(c->W, c->H);
    RC_CHECK(sync_buffers(c, level, PT_DONT_COPY_OFFSETS | PT_DONT_UPDATE_RESIDUAL));
    if (flags & PT_ADVECT_UPDATE_RESIDUAL) RC_CHECK(update_residual(c, level));
    /* residual -> dest.residual, whole buffer (:1724) */
    HIP_CHECK(c, hipMemcpyAsync(d.res, c->res_base, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    return PT_OK;
}

static int merge_randomly(pt_ctx *c, int level, int forwards, float w, int force)
{
    if (c->rw == 0) return set_err(c, PT_ERR_STATE, "no random texture");
    PtArgs a = make_args(c, level);
    HIP_CHECK(c, pt_launch_merge_randomly(a, forwards, w, force, c->stream));
    c->off_diff.grow(a.lw, a.lh);
    return sync_buffers(c, level, 0);
}

static int merge_residual(pt_ctx *c, int level, int forwards)
{
    PtArgs a = make_args(c, level);
    HIP_CHECK(c, pt_launch_merge_residual(a, forwards, c->stream));
    c->off_diff.grow(a.lw, a.lh);
    return sync_buffers(c, level, 0);
}

#define PUBLIC_OP(call)                                                                             \
    do {                                                                                            \
        HIP_CHECK(c, hipSetDevice(c->device));                                                      \
        RC_CHECK(call);                                                                             \
        return sync_stream(c);                                                                      \
    } while (0)

extern "C" int pt_randomize_offsets(pt_ctx *c, int level)
{
    RC_CHECK(need_level(c, level));
    PUBLIC_OP(randomize_offsets(c, level));
}

extern "C" int pt_propagate(pt_ctx *c, int level, int jump, int key_index)
{
    RC_CHECK(need_level(c, level));
    if (key_index >= 0) return set_err(c, PT_ERR_UNSUPPORTED, "keyframes are not supported");
    PUBLIC_OP(propagate(c, level, jump));
}

extern "C" int pt_scatter(pt_ctx *c, int level, int key_index)
{
    RC_CHECK(need_level(c, level));
    if (key_index >= 0) return set_err(c, PT_ERR_UNSUPPORTED, "keyframes are not supported");
    PUBLIC_OP(scatter(c, level));
}

extern "C" int pt_upsample(pt_ctx *c, int fine_level)
{
    RC_CHECK(need_level(c, fine_level));
    PUBLIC_OP(upsample(c, fine_level));
}

extern "C" int pt_sync_buffers(pt_ctx *c, int level, int flags)
{
    RC_CHECK(need_level(c, level));
    PUBLIC_OP(sync_buffers(c, level, flags));
}

extern "C" int pt_update_residual_cache(pt_ctx *c, int level)
{
    RC_CHECK(need_level(c, level));
    PUBLIC_OP(update_residual(c, level));
}

extern "C" int pt_advect_offsets(pt_ctx *c, int level, int time_step, int forwards, int flags)
{
    RC_CHECK(need_level(c, level));
    PUBLIC_OP(advect(c, level, time_step, forwards, flags));
}

extern "C" int pt_merge_randomly(pt_ctx *c, int level, int forwards, float w, int force)
{
    RC_CHECK(need_level(c, level));
    PUBLIC_OP(merge_randomly(c, level, forwards, w, force));
}

extern "C" int pt_merge_using_residual(pt_ctx *c, int level, int forwards)
{
    RC_CHECK(need_level(c, level));
    PUBLIC_OP(merge_residual(c, level, forwards));
}

/* ------------------------------------------------------------------ */
/* readback and load                                                   */
/* ------------------------------------------------------------------ */
static int get_nnf(pt_ctx *c, const PtNnf &src, int lw, int lh, pt_patchxf *host)
{
    pt_patchxf *tmp = nullptr;
    RC_CHECK(dalloc(c, &tmp, (size_t)lw * lh));
    hipError_t e = pt_launch_pack_nnf(src, c->W, lw, lh, tmp, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(host, tmp, (size_t)lw * lh * sizeof(pt_patchxf), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(tmp);
    if (e != hipSuccess) return set_err(c, PT_ERR_HIP, "offset readback failed: %s", hipGetErrorString(e));
    return PT_OK;
}

template <class T> static int get_plane(pt_ctx *c, const T *src, int lw, int lh, T *host)
{
    HIP_CHECK(c, hipMemcpy2DAsync(host, (size_t)lw * sizeof(T), src, (size_t)c->W * sizeof(T), (size_t)lw * sizeof(T),
                                  lh, hipMemcpyDeviceToHost, c->stream));
    return PT_OK;
}

extern "C" int pt_get_output(pt_ctx *c, int level, pt_patchxf *offsets, float *canvas, float *residual)
{
    if (!c || !c->base.xyth || level < 0 || (c->W >> level) < 1 || (c->H >> level) < 1)
        return set_err(c, PT_ERR_INVALID_ARG, "bad get_output");
    HIP_CHECK(c, hipSetDevice(c->device));
    int lw = c->W >> level, lh = c->H >> level;
    if (offsets) RC_CHECK(get_nnf(c, c->base, lw, lh, offsets));
    if (canvas) RC_CHECK(get_plane(c, c->canvas, lw, lh, (float4 *)canvas));
    if (residual) RC_CHECK(get_plane(c, c->res_base, lw, lh, residual));
    return sync_stream(c);
}

extern "C" int pt_get_advected(pt_ctx *c, int level, int which, pt_patchxf *offsets, float *base, float *output,
                               float *residual, int32_t *ts)
{
    if (!c || !c->base.xyth || level < 0 || which < 0 || which > 1 || (c->W >> level) < 1 || (c->H >> level) < 1)
        return set_err(c, PT_ERR_INVALID_ARG, "bad get_advected");
    HIP_CHECK(c, hipSetDevice(c->device));
    int lw = c->W >> level, lh = c->H >> level;
    DevAdv &d = c->adv[which];
    PtNnf n{d.xyth, d.sc};
    if (offsets) RC_CHECK(get_nnf(c, n, lw, lh, offsets));
    if (base) RC_CHECK(get_plane(c, d.base, lw, lh, (float4 *)base));
    if (output) RC_CHECK(get_plane(c, d.out, lw, lh, (float4 *)output));
    if (residual) RC_CHECK(get_plane(c, d.res, lw, lh, residual));
    if (ts) RC_CHECK(get_plane(c, d.ts, lw, lh, ts));
    return sync_stream(c);
}

extern "C" int pt_load_offsets(pt_ctx *c, int level, const pt_patchxf *in)
{
    if (!c || !c->base.xyth || !in || level < 0 || (c->W >> level) < 1 || (c->H >> level) < 1)
        return set_err(c, PT_ERR_INVALID_ARG, "bad load_offsets");
    HIP_CHECK(c, hipSetDevice(c->device));
    int lw = c->W >> level, lh = c->H >> level;
    for (size_t i = 0; i < (size_t)lw * lh; i++)
        if (in[i].layer != 0 || in[i].lum_y != 0.f || in[i].lum_a != 0.f)
            return set_err(c, PT_ERR_UNSUPPORTED, "offsets with layer or luminance shift");
    pt_patchxf *tmp = nullptr;
    RC_CHECK(dalloc(c, &tmp, (size_t)lw * lh));
    hipError_t e = hipMemcpyAsync(tmp, in, (size_t)lw * lh * sizeof(pt_patchxf), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = pt_launch_unpack_nnf(tmp, lw, lh, c->base, c->W, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(tmp);
    if (e != hipSuccess) return set_err(c, PT_ERR_HIP, "offset load failed: %s", hipGetErrorString(e));
    c->off_diff.grow(lw, lh);
    return PT_OK;
}

extern "C" int pt_load_residual(pt_ctx *c, int level, const float *in)
{
    if (!c || !c->base.xyth || !in || level < 0 || (c->W >> level) < 1 || (c->H >> level) < 1)
        return set_err(c, PT_ERR_INVALID_ARG, "bad load_residual");
    HIP_CHECK(c, hipSetDevice(c->device));
    int lw = c->W >> level, lh = c->H >> level;
    HIP_CHECK(c, hipMemcpy2DAsync(c->res_base, (size_t)c->W * sizeof(float), in, (size_t)lw * sizeof(float),
                                  (size_t)lw * sizeof(float), lh, hipMemcpyHostToDevice, c->stream));
    c->res_diff.grow(lw, lh);
    return sync_stream(c);
}

extern "C" int pt_pde_count(pt_ctx *c, uint64_t *out)
{
    if (!c || !out) return set_err(c, PT_ERR_INVALID_ARG, "NULL argument");
    *out = c->pde;
    return PT_OK;
}

/* ------------------------------------------------------------------ */
/* schedule                                                            */
/* ------------------------------------------------------------------ */
static int snap_alloc(pt_ctx *c, int levels, int frames, int passes)
{
    if (c->snap_levels == levels && c->snap_frames == frames && c->snap_passes == passes && !c->snaps.empty() &&
        c->snaps[0].w == (c->W >> 0) && c->snaps[0].h == (c->H >> 0))
        return PT_OK;
    free_snaps(c);
    c->snaps.resize((size_t)levels * frames * passes);
    c->snap_levels = levels;
    c->snap_frames = frames;
    c->snap_passes = passes;
    for (int l = 0; l < levels; l++)
        for (int f = 0; f < frames; f++)
            for (int p = 0; p < passes; p++) {
                Snapshot &s = c->snaps[((size_t)l * frames + f) * passes + p];
                s.w = c->W >> l;
                s.h = c->H >> l;
                size_t n = (size_t)s.w * s.h;
                RC_CHECK(dalloc(c, &s.xyth, n));
                RC_CHECK(dalloc(c, &s.sc, n));
                RC_CHECK(dalloc(c, &s.res, n));
                RC_CHECK(dalloc(c, &s.canvas, n));
            }
    return PT_OK;
}

static Snapshot &snap(pt_ctx *c, int l, int f, int p)
{
    return c->snaps[((size_t)l * c->snap_frames + f) * c->snap_passes + p];
}

/* WorkingBuffers::saveTempImages (workingBuffers.cpp:242-296), kept on the device */
static int snap_save(pt_ctx *c, int level, Snapshot &s)
{
    hipStream_t st = c->stream;
    HIP_CHECK(c, hipMemcpy2DAsync(s.xyth, (size_t)s.w * sizeof(float4), c->base.xyth, (size_t)c->W * sizeof(float4),
                                  (size_t)s.w * sizeof(float4), s.h, hipMemcpyDeviceToDevice, st));
    HIP_CHECK(c, hipMemcpy2DAsync(s.sc, (size_t)s.w * sizeof(float2), c->base.sc, (size_t)c->W * sizeof(float2),
                                  (size_t)s.w * sizeof(float2), s.h, hipMemcpyDeviceToDevice, st));
    HIP_CHECK(c, hipMemcpy2DAsync(s.res, (size_t)s.w * sizeof(float), c->res_base, (size_t)c->W * sizeof(float),
                                  (size_t)s.w * sizeof(float), s.h, hipMemcpyDeviceToDevice, st));
    HIP_CHECK(c, hipMemcpy2DAsync(s.canvas, (size_t)s.w * sizeof(float4), c->canvas, (size_t)c->W * sizeof(float4),
                                  (size_t)s.w * sizeof(float4), s.h, hipMemcpyDeviceToDevice, st));
    (void)level;
    return PT_OK;
}

/* WorkingBuffers::loadOffsets (workingBuffers.cpp:190-207) from a snapshot */
static int snap_load_offsets(pt_ctx *c, const Snapshot &s)
{
    hipStream_t st = c->stream;
    HIP_CHECK(c, hipMemcpy2DAsync(c->base.xyth, (size_t)c->W * sizeof(float4), s.xyth, (size_t)s.w * sizeof(float4),
                                  (size_t)s.w * sizeof(float4), s.h, hipMemcpyDeviceToDevice, st));
    HIP_CHECK(c, hipMemcpy2DAsync(c->base.sc, (size_t)c->W * sizeof(float2), s.sc, (size_t)s.w * sizeof(float2),
                                  (size_t)s.w * sizeof(float2), s.h, hipMemcpyDeviceToDevice, st));
    c->off_diff.grow(s.w, s.h);
    return PT_OK;
}

/* WorkingBuffers::loadResidual (workingBuffers.cpp:166-188) from a snapshot */
static int snap_load_residual(pt_ctx *c, const Snapshot &s)
{
    HIP_CHECK(c, hipMemcpy2DAsync(c->res_base, (size_t)c->W * sizeof(float), s.res, (size_t)s.w * sizeof(float),
                                  (size_t)s.w * sizeof(float), s.h, hipMemcpyDeviceToDevice, c->stream));
    c->res_diff.grow(s.w, s.h);
    return PT_OK;
}

/* propagateAndScatter (synthesisFrame.cpp:483-508) */
static int propagate_and_scatter(pt_ctx *c, int level, int iterations)
{
    for (int itr = 0; itr < iterations; itr++) {
        for (int i = 128; i > 1; i /= 2) RC_CHECK(propagate(c, level, i));
        RC_CHECK(propagate(c, level, 1));
        RC_CHECK(propagate(c, level, 1));
        if (itr < iterations - 1) RC_CHECK(scatter(c, level));
    }
    return PT_OK;
}

static int run_schedule(pt_ctx *c, const pt_params *params, int scheme, int iterations, int fixup_passes, int F,
                        const FramePyr *fp, const StylePyr &sp)
{
    int levels = params->levels;
    int passes = fixup_passes + 1;
    RC_CHECK(pt_setup(c, params));
    RC_CHECK(snap_alloc(c, levels, F, passes));
    pt_params fparams = *params;
    for (int level = levels - 1; level >= 0; level--) {
        /* Style::load (style.cpp:62-140) */
        RC_CHECK(pt_initialize_exemplars(c, c->SW >> level, c->SH >> level, 1));
        RC_CHECK(upload_image(c, level, PT_LAYER_EXEMPLAR_BASE, sp.base, 0));
        RC_CHECK(upload_image(c, level, PT_LAYER_EXEMPLAR_OUTPUT, sp.out, 0));
        RC_CHECK(upload_image(c, level, PT_LAYER_EXEMPLAR_ORIENTATION, sp.orient, 0));
        RC_CHECK(upload_image(c, level, PT_LAYER_EXEMPLAR_DIST_TRANS, sp.dt, 0));
        RC_CHECK(reset_advected(c));
        for (int pass = 0; pass < passes; pass++) {
            /* renderConsecutivePass (synthesisProcessor.cpp:703-753), firstPassDirection 0 */
            int fwd = (pass % 2 == 0);
            std::vector<int> order;
            for (int i = 0; i < F; i++) order.push_back(fwd ? i : F - 1 - i);
            size_t start = 0;
            if (pass > 0) {
                RC_CHECK(snap_save(c, level, snap(c, level, order[0], pass)));
                start = 1;
            }
            for (size_t k = start; k < order.size(); k++) {
                int f = order[k];
                bool is_first = (k == 0) && (pass == 0);
                if (!is_first) c->cur = (c->cur + 1) % 2;
                /* loadImages / loadAnimImages (synthesisFrame.cpp:562-656) */
                RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_COLOR, fp[f].color, 0));
                RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_ORIENTATION, fp[f].orient, 0));
                RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_VEL_B, fp[f].velb, 0));
                RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_VEL_F, fp[f].velf, 0));
                RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_DIST_TRANS, fp[f].dt, 0));
                RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_SCALE, fp[f].scale, 0));
                /* synthesizeConsecutive (synthesisFrame.cpp:214-323) */
                fparams.direction = fwd ? PT_FORWARD : PT_BACKWARD;
                fparams.advection_jitter = (f % 3 == 0) ? params->advection_jitter : 0.f;
                fparams.frame_number = f;
                RC_CHECK(pt_upload_parameters(c, &fparams));
                int ops = (scheme != PT_S_FMBM_RM) ? PT_ADVECT_UPDATE_RESIDUAL : 0;
                if (pass == 0) {
                    RC_CHECK(random_pyramid(c, level, f, fp[f].color, fp[f].orient, sp.out, sp.orient));
                    if (is_first) {
                        if (level == levels - 1) {
                            RC_CHECK(randomize_offsets(c, level));
                        } else {
                            RC_CHECK(snap_load_offsets(c, snap(c, level + 1, f, passes - 1)));
                            RC_CHECK(upsample(c, level));
                            RC_CHECK(sync_buffers(c, level, 0));
                        }
                    }
                    if (!is_first && level == levels - 1) {
                        RC_CHECK(advect(c, level, 1, fwd, ops));
                        RC_CHECK(copy_base_to_working(c));
                    } else if (!is_first && level < levels - 1) {
                        Snapshot &s = snap(c, level + 1, f, passes - 1);
                        RC_CHECK(advect(c, level, 1, fwd, ops | PT_ADVECT_IGNORE_RANDOM));
                        RC_CHECK(snap_load_offsets(c, s));
                        RC_CHECK(snap_load_residual(c, s));
                        RC_CHECK(upsample(c, level));
                        RC_CHECK(sync_buffers(c, level, 0));
                        if (scheme == PT_S_FMBM_RM) RC_CHECK(merge_randomly(c, level, fwd, 0.5f, 0));
                        else RC_CHECK(merge_residual(c, level, fwd));
                    }
                } else {
                    Snapshot &s = snap(c, level, f, pass - 1);
                    RC_CHECK(advect(c, level, 1, fwd, ops | PT_ADVECT_IGNORE_RANDOM));
                    RC_CHECK(snap_load_offsets(c, s));
                    RC_CHECK(snap_load_residual(c, s));
                    RC_CHECK(sync_buffers(c, level, PT_DONT_COPY_OFFSETS | PT_DONT_UPDATE_RESIDUAL));
                    if (scheme == PT_S_FMBM_RM) RC_CHECK(merge_randomly(c, level, fwd, 0.5f, 0));
                    else RC_CHECK(merge_residual(c, level, fwd));
                }
                RC_CHECK(sync_buffers(c, level, 0));
                RC_CHECK(propagate_and_scatter(c, level, iterations));
                RC_CHECK(snap_save(c, level, snap(c, level, f, pass)));
            }
        }
    }
    return PT_OK;
}

static void free_frame(FramePyr &f)
{
    pt_pyramid_destroy(f.color);
    pt_pyramid_destroy(f.orient);
    pt_pyramid_destroy(f.tensor);
    pt_pyramid_destroy(f.dt);
    pt_pyramid_destroy(f.scale);
    pt_pyramid_destroy(f.velb);
    pt_pyramid_destroy(f.velf);
    f = FramePyr();
}

static void free_style(StylePyr &sp)
{
    pt_pyramid_destroy(sp.base);
    pt_pyramid_destroy(sp.out);
    pt_pyramid_destroy(sp.orient);
    pt_pyramid_destroy(sp.tensor);
    pt_pyramid_destroy(sp.dt);
    sp = StylePyr();
}

/* Fills level 0 of `p` from host or device memory, or with a constant, and
 * rebuilds the other levels on the device (downsampleAverage). */
static int pyramid_fill(pt_ctx *c, pt_pyramid *p, const float *src, int on_device, float defv, int scale_mode)
{
    size_t n0 = (size_t)p->w * p->h * elem_channels(p->elem);
    if (src)
        HIP_CHECK(c, hipMemcpyAsync(p->data, src, n0 * sizeof(float),
                                    on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->stream));
    else
        HIP_CHECK(c, pt_launch_fill(p->data, n0, defv, c->stream));
    return pyramid_rebuild(c, p, scale_mode);
}

static int orientation_fill(pt_ctx *c, pt_pyramid *tensor, pt_pyramid *angle, const float *src, int on_device)
{
    RC_CHECK(pyramid_fill(c, tensor, src, on_device, 0.f, PT_DOWNSAMPLE_SCALE_CONSTANT));
    for (int l = 0; l < tensor->levels; l++)
        HIP_CHECK(c, pt_launch_tensor_angle(tensor->data + tensor->off[l], angle->data + angle->off[l],
                                            (size_t)tensor->lw[l] * tensor->lh[l], c->stream));
    return PT_OK;
}

static int ensure_pyr(pt_ctx *c, pt_pyramid **p, int elem, int w, int h)
{
    if (*p && (*p)->w == w && (*p)->h == h && (*p)->elem == elem) return PT_OK;
    pt_pyramid_destroy(*p);
    *p = nullptr;
    return pyramid_alloc(c, p, elem, w, h);
}

static int prepare_frame(pt_ctx *c, int w, int h, const pt_frame_guides &g, int on_device, FramePyr &fp)
{
    if (!g.color_rgba) return set_err(c, PT_ERR_INVALID_ARG, "frame colour is required");
    RC_CHECK(ensure_pyr(c, &fp.color, PT_ELEM_F32X4, w, h));
    RC_CHECK(ensure_pyr(c, &fp.tensor, PT_ELEM_F32X3, w, h));
    RC_CHECK(ensure_pyr(c, &fp.orient, PT_ELEM_F32, w, h));
    RC_CHECK(ensure_pyr(c, &fp.dt, PT_ELEM_F32, w, h));
    RC_CHECK(ensure_pyr(c, &fp.scale, PT_ELEM_F32X2, w, h));
    RC_CHECK(ensure_pyr(c, &fp.velb, PT_ELEM_F32X2, w, h));
    RC_CHECK(ensure_pyr(c, &fp.velf, PT_ELEM_F32X2, w, h));
    RC_CHECK(pyramid_fill(c, fp.color, g.color_rgba, on_device, 0.f, PT_DOWNSAMPLE_SCALE_CONSTANT));
    RC_CHECK(orientation_fill(c, fp.tensor, fp.orient, g.orientation_xx_xy_yy, on_device));
    RC_CHECK(pyramid_fill(c, fp.dt, g.dist_trans, on_device, 1.0f, PT_DOWNSAMPLE_SCALE_CONSTANT));
    RC_CHECK(pyramid_fill(c, fp.scale, g.scale_uv, on_device, 1.0f, PT_DOWNSAMPLE_SCALE_CONSTANT));
    RC_CHECK(pyramid_fill(c, fp.velb, g.vel_b, on_device, 0.f, PT_DOWNSAMPLE_SCALE_WITH_LEVELS));
    RC_CHECK(pyramid_fill(c, fp.velf, g.vel_f, on_device, 0.f, PT_DOWNSAMPLE_SCALE_WITH_LEVELS));
    return PT_OK;
}

static int prepare_style(pt_ctx *c, int w, int h, const pt_style_guides &g, int on_device, StylePyr &sp)
{
    if (!g.base_rgba || !g.output_rgba) return set_err(c, PT_ERR_INVALID_ARG, "exemplar pair required");
    RC_CHECK(ensure_pyr(c, &sp.base, PT_ELEM_F32X4, w, h));
    RC_CHECK(ensure_pyr(c, &sp.out, PT_ELEM_F32X4, w, h));
    RC_CHECK(ensure_pyr(c, &sp.tensor, PT_ELEM_F32X3, w, h));
    RC_CHECK(ensure_pyr(c, &sp.orient, PT_ELEM_F32, w, h));
    RC_CHECK(ensure_pyr(c, &sp.dt, PT_ELEM_F32, w, h));
    RC_CHECK(pyramid_fill(c, sp.base, g.base_rgba, on_device, 0.f, PT_DOWNSAMPLE_SCALE_CONSTANT));
    RC_CHECK(pyramid_fill(c, sp.out, g.output_rgba, on_device, 0.f, PT_DOWNSAMPLE_SCALE_CONSTANT));
    RC_CHECK(orientation_fill(c, sp.tensor, sp.orient, g.orientation_xx_xy_yy, on_device));
    RC_CHECK(pyramid_fill(c, sp.dt, g.dist_trans, on_device, 0.f, PT_DOWNSAMPLE_SCALE_CONSTANT));
    return PT_OK;
}

static int read_results(pt_ctx *c, int F, int last_pass, pt_patchxf *out_off, float *out_canvas, float *out_res,
                        int on_device)
{
    hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    for (int f = 0; f < F; f++) {
        Snapshot &s = snap(c, 0, f, last_pass);
        size_t n = (size_t)s.w * s.h;
        if (out_off) {
            if (!c->pack_tmp || c->pack_cap < n) {
                dfree(c->pack_tmp);
                RC_CHECK(dalloc(c, &c->pack_tmp, n));
                c->pack_cap = n;
            }
            PtNnf nn{s.xyth, s.sc};
            HIP_CHECK(c, pt_launch_pack_nnf(nn, s.w, s.w, s.h, c->pack_tmp, c->stream));
            HIP_CHECK(c, hipMemcpyAsync(out_off + f * n, c->pack_tmp, n * sizeof(pt_patchxf), kind, c->stream));
        }
        if (out_canvas)
            HIP_CHECK(c, hipMemcpyAsync(out_canvas + f * n * 4, s.canvas, n * sizeof(float4), kind, c->stream));
        if (out_res) HIP_CHECK(c, hipMemcpyAsync(out_res + f * n, s.res, n * sizeof(float), kind, c->stream));
    }
    return PT_OK;
}

/* A one-shot synthesis starts from freshly allocated (zeroed) working
 * buffers, as a new WorkingBuffers does (workingBuffers.cpp:48-95), so
 * repeated calls on one context give identical results. */
static int reset_working(pt_ctx *c)
{
    size_t n = (size_t)c->W * c->H;
    hipStream_t s = c->stream;
    HIP_CHECK(c, hipMemsetAsync(c->base.xyth, 0, n * sizeof(float4), s));
    HIP_CHECK(c, hipMemsetAsync(c->base.sc, 0, n * sizeof(float2), s));
    HIP_CHECK(c, hipMemsetAsync(c->work.xyth, 0, n * sizeof(float4), s));
    HIP_CHECK(c, hipMemsetAsync(c->work.sc, 0, n * sizeof(float2), s));
    HIP_CHECK(c, hipMemsetAsync(c->res_base, 0, n * sizeof(float), s));
    HIP_CHECK(c, hipMemsetAsync(c->res_work, 0, n * sizeof(float), s));
    HIP_CHECK(c, hipMemsetAsync(c->canvas, 0, n * sizeof(float4), s));
    for (int k = 0; k < 2; k++) {
        HIP_CHECK(c, hipMemsetAsync(c->color[k], 0, n * sizeof(float4), s));
        HIP_CHECK(c, hipMemsetAsync(c->dt[k], 0, n * sizeof(float), s));
        HIP_CHECK(c, hipMemsetAsync(c->velb[k], 0, n * sizeof(float2), s));
        HIP_CHECK(c, hipMemsetAsync(c->velf[k], 0, n * sizeof(float2), s));
    }
    RC_CHECK(reset_advected(c));
    c->off_diff = Rect();
    c->res_diff = Rect();
    c->cur = 0;
    c->scatter_calls = 0;
    return PT_OK;
}

static int synthesize(pt_ctx *c, const pt_params *params, int scheme, int iterations, int fixup_passes, int W, int H,
                      int F, const pt_frame_guides *frames, int SW, int SH, const pt_style_guides *style,
                      int guides_on_device, pt_patchxf *out_off, float *out_canvas, float *out_res,
                      int outputs_on_device)
{
    if (!params || !frames || !style || F < 1 || iterations < 1 || fixup_passes < 0)
        return set_err(c, PT_ERR_INVALID_ARG, "bad synthesize arguments");
    if (scheme != PT_S_FMBM && scheme != PT_S_FMBM_RM) return set_err(c, PT_ERR_UNSUPPORTED, "scheme");
    int img_lv = (int)fmin(floor(log2f((float)W)), floor(log2f((float)H)));
    int sty_lv = (int)fmin(floor(log2f((float)SW)), floor(log2f((float)SH)));
    if (params->levels < 1 || params->levels > img_lv || params->levels > sty_lv)
        return set_err(c, PT_ERR_INVALID_ARG, "levels exceeds the pyramid depth");
    RC_CHECK(check_supported(c, params));
    HIP_CHECK(c, hipSetDevice(c->device));
    RC_CHECK(pt_initialize(c, W, H, SW, SH));
    RC_CHECK(reset_working(c));
    RC_CHECK(prepare_style(c, SW, SH, *style, guides_on_device, c->style_pyr));
    if ((int)c->frame_pyrs.size() > F) {
        for (size_t i = F; i < c->frame_pyrs.size(); i++) free_frame(c->frame_pyrs[i]);
    }
    c->frame_pyrs.resize(F);
    for (int f = 0; f < F; f++) RC_CHECK(prepare_frame(c, W, H, frames[f], guides_on_device, c->frame_pyrs[f]));
    RC_CHECK(run_schedule(c, params, scheme, iterations, fixup_passes, F, c->frame_pyrs.data(), c->style_pyr));
    RC_CHECK(read_results(c, F, fixup_passes, out_off, out_canvas, out_res, outputs_on_device));
    return sync_stream(c);
}

extern "C" int pt_synthesize(pt_ctx *c, const pt_params *params, int scheme, int iterations, int fixup_passes,
                             int W, int H, int F, const pt_frame_guides *frames, int SW, int SH,
                             const pt_style_guides *style, pt_patchxf *out_off, float *out_canvas, float *out_res)
{
    if (!c) return set_err(nullptr, PT_ERR_INVALID_ARG, "NULL context");
    return synthesize(c, params, scheme, iterations, fixup_passes, W, H, F, frames, SW, SH, style, 0, out_off,
                      out_canvas, out_res, 0);
}

extern "C" int pt_synthesize_device(pt_ctx *c, const pt_params *params, int scheme, int iterations,
                                    int fixup_passes, int W, int H, int F, const pt_frame_guides *frames, int SW,
                                    int SH, const pt_style_guides *style, pt_patchxf *out_off, float *out_canvas,
                                    float *out_res)
{
    if (!c) return set_err(nullptr, PT_ERR_INVALID_ARG, "NULL context");
    return synthesize(c, params, scheme, iterations, fixup_passes, W, H, F, frames, SW, SH, style, 1, out_off,
                      out_canvas, out_res, 1);
}
