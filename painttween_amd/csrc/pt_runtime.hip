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
#include "pt_math.h"
#include "pt_plan.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <mutex>
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
    float *lum = nullptr;
    int32_t *la = nullptr;
};

/* `id` (object IDs) exists only while the object-ID term is on */
struct FramePyr {
    pt_pyramid *color = nullptr, *tensor = nullptr, *orient = nullptr, *dt = nullptr, *scale = nullptr,
               *velb = nullptr, *velf = nullptr, *id = nullptr;
};

struct StylePyr {
    pt_pyramid *base = nullptr, *out = nullptr, *tensor = nullptr, *orient = nullptr, *dt = nullptr, *id = nullptr;
};

/* Kernel classes timed by pt_profile_*; names in kProfNames. */
enum ProfKind { PK_PROPAGATE = 0, PK_SCATTER, PK_RESIDUAL, PK_CANVAS, PK_UPSAMPLE, PK_ADVECT, PK_MERGE, PK_RANDOM,
                PK_COPY, PK_PYRAMID, PK_COUNT };
/* Algorithmic HBM bytes per pixel of each profiled launch class (every
 * operand read or written once; the exemplar counted once per pixel): the
 * NNF searches 96 (offsets 16+16+8, residual 4, canvas 16, target 16,
 * orientation 4, distance transform 4, exemplar amortised), the residual
 * refresh 60, random init 40 (random texel 16, offsets 24), upsample 40,
 * advection 152, merges 64; the fused sync/canvas pass and the pyramid are
 * counted per launch (sync_buffers, pyramid_rebuild). */
enum : uint64_t { PB_NNF = 96, PB_RESIDUAL = 60, PB_RANDOM = 40, PB_UPSAMPLE = 40, PB_ADVECT = 152, PB_MERGE = 64 };

static const char *kProfNames[PK_COUNT] = {"propagate", "scatter", "update_residual", "canvas", "upsample",
                                           "advect", "merge", "random", "copy", "pyramid"};

struct ProfRec {
    int kind;
    int ev0, ev1;
    uint64_t pixels;
    uint64_t bytes; /* algorithmic HBM bytes of the launch */
};

struct Snapshot {
    float4 *xyth = nullptr;
    float2 *sc = nullptr;
    float *lum = nullptr; /* luminance shift, saved and loaded while lum_live */
    float *res = nullptr;
    float4 *canvas = nullptr;
    int w = 0, h = 0;
    bool saved = false; /* written in the current run (the reference's temporary file exists) */
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
    /* time tables (k_time_table): for advected set k, the exemplar base and
     * output at every tap of every pixel's advected offset, gathered once and
     * read by the temporal terms of every candidate afterwards.  Valid while
     * the set, the exemplar textures, the image orientation, the parameters
     * and the level stay as they were when it was built. */
    float4 *tt[2] = {nullptr, nullptr};
    size_t tt_cap[2] = {0, 0}; /* float4 elements */
    bool tt_valid[2] = {false, false};
    int tt_level[2] = {-1, -1}, tt_r[2] = {-1, -1};
    int tt_lab[2] = {0, 0}; /* entries stored in LAB (the colour space of the launch that built them) */
    int64_t tt_max_taps = 81; /* tables for patches of at most this many taps (0: never) */
    /* luminance-shift planes are maintained from the first parameters with
     * lum_shift_enable (or loaded nonzero shifts) until the working buffers
     * are reset; before that every shift is 0 */
    bool lum_live = false;
    /* style layers: once live, every operator maintains the layer planes
     * (PtNnf::la) and the searches run their layered variant.  Set by the
     * first operator that runs with several exemplar layers, keyframe
     * interpolation or a keyframe index, by keyframe fields and ribbons, and
     * by offsets loaded with a layer; cleared with the working buffers (the
     * planes are zero until then). */
    bool layered_live = false;
    bool no_early_exit = false;            /* force the plain candidate kernels (tests) */
    /* Early rejection lengthens each lane's dependent chain; it pays only on
     * levels with enough waves to hide the latency. */
    int64_t early_exit_min_pixels = 256 * 256;
    bool unit_parallel = true; /* candidate units over workgroup lanes */
    /* residual refresh: res_base is exact for the current offsets, canvas
     * and inputs of level res_level (-1: unknown); `changed` marks the
     * offsets the last sync changed */
    uint8_t *changed = nullptr;
    int res_level = -1;
    bool incremental_residual = true;
    int64_t tap_max_pixels = 128 * 128; /* tap-parallel shapes up to this level size */
    int tap_evals = 1024;               /* tap-parallel: tap evaluations per workgroup */
    int64_t ex_window_min_pixels = 200000; /* XL refresh on levels of at least this size (0: never) */
    bool xcd_tiles = true; /* contiguous tile ranges per XCD in single-job launches */
    int prop_window_jump = 8; /* XL propagate for jumps up to this (0: never) ... */
    int64_t prop_window_min_pixels = 200000; /* ... on levels of at least this size (0: never) */
    int time_unit_order = 0; /* PT_TUNE_TEMPORAL_UNIT_ORDER */
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
    int32_t *img_id = nullptr; /* object IDs, idw x idh: one 0 texel until uploaded */
    int idw = 0, idh = 0;
    /* exemplar textures (capacity ex_cap texels) */
    float4 *ex = nullptr;
    float *ex_orient = nullptr, *ex_dt = nullptr;
    int32_t *ex_id = nullptr;
    size_t ex_cap = 0;
    int ew = 0, eh = 0;
    int nl = 1; /* exemplar style layers */
    /* keyframe interpolation: cumulative key-to-key fields (cl layers of
     * cw x ch; one zero texel until initialize_cumulative_fields), the
     * frame-to-key ribbons (fw x fh; one zero texel until uploaded) and the
     * ribbon field pair (W x H, allocated invalid on first use) */
    float4 *cumul[2] = {nullptr, nullptr};
    int cw = 1, ch = 1, cl = 1;
    float4 *ftk[2] = {nullptr, nullptr};
    size_t ftk_cap[2] = {0, 0};
    int fw[2] = {1, 1}, fh[2] = {1, 1};
    pt_ribbonp *rib_base = nullptr, *rib_work = nullptr;
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
    /* when set, stream work is appended here instead of issued (plans for
     * the multi-job executor); host-side state advances as if it ran */
    std::vector<RecOp> *rec = nullptr;
    /* per-frame sequence state (pt_synthesize_next_frame) */
    int seq_frame = 0;
    int seq_levels = 0;
    std::vector<Snapshot> seq_snap[2]; /* [frame parity][level] */
    FramePyr seq_pyr[2];                /* [frame parity] */
    /* optional per-launch timing with HIP events on the context stream */
    bool prof_on = false;
    std::vector<hipEvent_t> prof_events;
    std::vector<ProfRec> prof_recs;
};

static void free_frame(FramePyr &f);
static void free_style(StylePyr &sp);
static void free_seq(pt_ctx *c);
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

static int elem_channels(int elem) { return elem == PT_ELEM_I32 ? 1 : (elem >= 1 && elem <= 4) ? elem : 0; }

static int check_supported(pt_ctx *c, const pt_params *p)
{
    if (p->residual_window_size < 1 || p->residual_window_size > PT_MAX_R)
        return set_err(c, PT_ERR_INVALID_ARG, "residual_window_size must be in [1, %d]", PT_MAX_R);
    /* NUM_MAX_STYLES 5, numStyles < 5 (synthesisProcessor.cpp:192) */
    if (p->num_styles < 1 || p->num_styles > 4) return set_err(c, PT_ERR_INVALID_ARG, "num_styles must be in [1, 4]");
    if (p->coherence_weight > 0.f && !(p->coherence_ratio_range > 1.f && p->coherence_angle_range > 0.f))
        return set_err(c, PT_ERR_INVALID_ARG, "coherence needs ratio range > 1 and angle range > 0");
    if (p->colorspace != 0 && p->colorspace != 1)
        return set_err(c, PT_ERR_UNSUPPORTED, "colourspace must be RGB (0) or LAB (1)");
    if (p->interpolate_key_frame &&
        (p->first_key_index < 0 || p->last_key_index < p->first_key_index || p->last_key_index >= PT_MAX_STYLE_LAYERS))
        return set_err(c, PT_ERR_INVALID_ARG, "keyframe interpolation needs 0 <= first_key_index <= last_key_index < %d",
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
            return set_err(c, PT_ERR_UNSUPPORTED, "offsets histogram with a negative threshold is not supported");
    }
    if (p->rng_mode != PT_RNG_REFERENCE && p->rng_mode != PT_RNG_COUNTER)
        return set_err(c, PT_ERR_INVALID_ARG, "unknown rng_mode");
    if (p->first_pass_direction != 0 && p->first_pass_direction != 1)
        return set_err(c, PT_ERR_INVALID_ARG, "first_pass_direction must be 0 or 1");
    if (p->max_level < 0 || p->max_level >= p->levels)
        return set_err(c, PT_ERR_INVALID_ARG, "max_level must be in [0, levels)");
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
    case 4: return sizeof(pt_job);
    case 5: return sizeof(pt_batch_stats);
    case 6: return sizeof(pt_sequence_desc);
    case 7: return sizeof(pt_server_stats);
    default: return 0;
    }
}

extern "C" const char *pt_last_error(pt_ctx *c) { return c ? c->err.c_str() : g_last_error; }

/* ------------------------------------------------------------------ */
/* allocation-free readback (pt_get_output_noalloc)                    */
/* ------------------------------------------------------------------ */
/* Staging for the packed offsets of pt_get_output_noalloc: one array per
 * GPU, held until the process ends.  Host code never passes the array's
 * name to a kernel or a copy (on the host it is only a stand-in); it uses
 * the device address hipGetSymbolAddress returns for each device, resolved
 * once per device by pt_create with that device current. */
static __device__ __align__(256) unsigned char scratch[PT_NOALLOC_SCRATCH_BYTES];

#define PT_SCRATCH_MAX_DEVICES 64
static std::mutex g_scratch_mu;                        /* guards the table and the arrays' users */
static void *g_scratch_ptr[PT_SCRATCH_MAX_DEVICES];    /* device address of `scratch` per device */

/* With `device` current: records that device's address of `scratch`. */
static hipError_t scratch_resolve(int device)
{
    if (device < 0 || device >= PT_SCRATCH_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}

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
    e = scratch_resolve(device);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return set_err(nullptr, PT_ERR_HIP, "resolving the readback staging array failed: %s", hipGetErrorString(e));
    }
    *out = c;
    return PT_OK;
}

static void free_snaps(pt_ctx *c)
{
    for (auto &s : c->snaps) {
        dfree(s.xyth);
        dfree(s.sc);
        dfree(s.lum);
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
    dfree(c->base.lum);
    dfree(c->work.lum);
    dfree(c->base.la);
    dfree(c->work.la);
    dfree(c->rib_base);
    dfree(c->rib_work);
    dfree(c->res_base);
    dfree(c->res_work);
    dfree(c->canvas);
    dfree(c->changed);
    for (int k = 0; k < 2; k++) {
        dfree(c->adv[k].xyth);
        dfree(c->adv[k].sc);
        dfree(c->adv[k].base);
        dfree(c->adv[k].out);
        dfree(c->adv[k].res);
        dfree(c->adv[k].ts);
        dfree(c->adv[k].lum);
        dfree(c->adv[k].la);
        dfree(c->color[k]);
        dfree(c->dt[k]);
        dfree(c->velb[k]);
        dfree(c->velf[k]);
    }
    dfree(c->orient);
    dfree(c->scale);
    dfree(c->random);
    dfree(c->img_id);
}

extern "C" int pt_destroy(pt_ctx *c)
{
    if (!c) return PT_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_snaps(c);
    free_seq(c);
    free_working(c);
    for (auto &f : c->frame_pyrs) free_frame(f);
    free_style(c->style_pyr);
    dfree(c->pack_tmp);
    dfree(c->tt[0]);
    dfree(c->tt[1]);
    prof_clear(c);
    dfree(c->ex);
    dfree(c->ex_orient);
    dfree(c->ex_dt);
    dfree(c->ex_id);
    for (int k = 0; k < 2; k++) {
        dfree(c->cumul[k]);
        dfree(c->ftk[k]);
    }
    (void)hipStreamDestroy(c->stream);
    delete c;
    return PT_OK;
}

/* Marks the residual cache as not known to be exact (any change of offsets,
 * canvas inputs, textures or parameters other than a sync's copy). */
static void res_invalidate(pt_ctx *c) { c->res_level = -1; }

/* Marks the time tables stale (an advected set, an exemplar texture, the
 * image orientation or the parameters changed). */
static void tt_invalidate(pt_ctx *c) { c->tt_valid[0] = c->tt_valid[1] = false; }

static int sync_stream(pt_ctx *c)
{
    if (c->rec) return PT_OK;
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return PT_OK;
}

/* Issues stream work now, or appends it to the plan being recorded. */
static int issue(pt_ctx *c, const char *what, std::function<hipError_t(hipStream_t)> fn)
{
    if (c->rec) {
        RecOp o;
        o.op = -1;
        o.fn = std::move(fn);
        o.what = what;
        c->rec->push_back(std::move(o));
        return PT_OK;
    }
    hipError_t e = fn(c->stream);
    if (e != hipSuccess) return set_err(c, PT_ERR_HIP, "%s failed: %s", what, hipGetErrorString(e));
    return PT_OK;
}

/* An operator launch with a multi-job form (PT_OP_*). */
static int issue_op(pt_ctx *c, int op, const PtArgs &a)
{
    if (c->rec) {
        RecOp o;
        o.op = op;
        o.a = a;
        c->rec->push_back(std::move(o));
        return PT_OK;
    }
    hipError_t e = pt_run_op(op, a, c->stream);
    if (e != hipSuccess) return set_err(c, PT_ERR_HIP, "operator %d launch failed: %s", op, hipGetErrorString(e));
    return PT_OK;
}

static int zero_fill(pt_ctx *c, void *p, size_t bytes)
{
    float *f = (float *)p;
    size_t n = bytes / sizeof(float);
    return issue(c, "zero fill", [f, n](hipStream_t s) { return pt_launch_fill(f, n, 0.f, s); });
}

static int prof_event(pt_ctx *c)
{
    hipEvent_t e;
    HIP_CHECK(c, hipEventCreate(&e));
    c->prof_events.push_back(e);
    return (int)c->prof_events.size() - 1;
}

/* Records the start of a timed launch; returns the record index or -1. */
static int prof_begin(pt_ctx *c, int kind, uint64_t pixels, uint64_t bytes)
{
    if (!c->prof_on || c->rec) return -1;
    ProfRec r;
    r.kind = kind;
    r.pixels = pixels;
    r.bytes = bytes;
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

extern "C" int pt_profile_bytes(pt_ctx *c, const char *kernel, uint64_t *bytes)
{
    if (!c || !kernel || !bytes) return set_err(c, PT_ERR_INVALID_ARG, "NULL argument");
    int kind = -1;
    for (int k = 0; k < PK_COUNT; k++)
        if (strcmp(kernel, kProfNames[k]) == 0) kind = k;
    if (kind < 0) return set_err(c, PT_ERR_INVALID_ARG, "unknown kernel class '%s'", kernel);
    *bytes = 0;
    for (const auto &r : c->prof_recs)
        if (r.kind == kind) *bytes += r.bytes;
    return PT_OK;
}

extern "C" int pt_profile_records(pt_ctx *c, const char *kernel, int max_records, uint64_t *pixels, uint64_t *bytes,
                                  float *ms, int *count)
{
    if (!c || !kernel || !count || max_records < 0 || (max_records > 0 && (!pixels || !bytes || !ms)))
        return set_err(c, PT_ERR_INVALID_ARG, "bad profile record arguments");
    int kind = -1;
    for (int k = 0; k < PK_COUNT; k++)
        if (strcmp(kernel, kProfNames[k]) == 0) kind = k;
    if (kind < 0) return set_err(c, PT_ERR_INVALID_ARG, "unknown kernel class '%s'", kernel);
    HIP_CHECK(c, hipSetDevice(c->device));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    int n = 0;
    for (const auto &r : c->prof_recs) {
        if (r.kind != kind) continue;
        if (n < max_records) {
            float t = 0.f;
            HIP_CHECK(c, hipEventElapsedTime(&t, c->prof_events[r.ev0], c->prof_events[r.ev1]));
            pixels[n] = r.pixels;
            bytes[n] = r.bytes;
            ms[n] = t;
        }
        n++;
    }
    *count = n;
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
    case PT_TUNE_INCREMENTAL_RESIDUAL:
        c->incremental_residual = value != 0;
        return PT_OK;
    case PT_TUNE_TAP_PARALLEL_MAX_PIXELS:
        if (value < 0) return set_err(c, PT_ERR_INVALID_ARG, "negative pixel count");
        c->tap_max_pixels = value;
        return PT_OK;
    case PT_TUNE_TAP_PARALLEL_EVALS:
        if (value < 1 || value > PT_TAP_EVALS)
            return set_err(c, PT_ERR_INVALID_ARG, "tap evaluations per workgroup must be 1..%d", PT_TAP_EVALS);
        c->tap_evals = (int)value;
        return PT_OK;
    case PT_TUNE_EXEMPLAR_WINDOW:
        if (value < 0) return set_err(c, PT_ERR_INVALID_ARG, "negative pixel count");
        c->ex_window_min_pixels = value;
        return PT_OK;
    case PT_TUNE_TIME_TABLE_MAX_TAPS:
        if (value < 0) return set_err(c, PT_ERR_INVALID_ARG, "negative tap count");
        c->tt_max_taps = value;
        tt_invalidate(c);
        return PT_OK;
    case PT_TUNE_XCD_TILES:
        c->xcd_tiles = value != 0;
        return PT_OK;
    case PT_TUNE_PROPAGATE_WINDOW_JUMP:
        if (value < 0) return set_err(c, PT_ERR_INVALID_ARG, "negative jump");
        c->prop_window_jump = (int)std::min<int64_t>(value, 1 << 30);
        return PT_OK;
    case PT_TUNE_PROPAGATE_WINDOW_MIN_PIXELS:
        if (value < 0) return set_err(c, PT_ERR_INVALID_ARG, "negative pixel count");
        c->prop_window_min_pixels = value;
        return PT_OK;
    case PT_TUNE_TEMPORAL_UNIT_ORDER:
        if (value < 0 || value > 2) return set_err(c, PT_ERR_INVALID_ARG, "unit order must be 0, 1 or 2");
        c->time_unit_order = value;
        return PT_OK;
    default:
        return set_err(c, PT_ERR_INVALID_ARG, "unknown tuning key %d", key);
    }
}

extern "C" int pt_upload_parameters(pt_ctx *c, const pt_params *p)
{
    if (!c || !p) return set_err(c, PT_ERR_INVALID_ARG, "NULL argument");
    RC_CHECK(check_supported(c, p));
    /* what the time tables depend on: the patch basis and the angle fold */
    if (!c->have_params || p->residual_window_size != c->p.residual_window_size ||
        p->residual_window_scale_u != c->p.residual_window_scale_u ||
        p->residual_window_scale_v != c->p.residual_window_scale_v ||
        p->orientation_enabled != c->p.orientation_enabled)
        tt_invalidate(c);
    c->p = *p;
    c->have_params = true;
    if (p->lum_shift_enable) c->lum_live = true;
    res_invalidate(c);
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

/* Frame-to-key ribbon k back to one zero texel (an unbound texture). */
static int ftk_reset(pt_ctx *c, int k)
{
    if (c->ftk_cap[k] < 1) {
        dfree(c->ftk[k]);
        RC_CHECK(dalloc(c, &c->ftk[k], 1));
        c->ftk_cap[k] = 1;
    }
    RC_CHECK(zero_fill(c, c->ftk[k], sizeof(float4)));
    c->fw[k] = c->fh[k] = 1;
    return PT_OK;
}

/* Cumulative fields default to one zero texel (unbound). */
static int cumul_defaults(pt_ctx *c)
{
    for (int k = 0; k < 2; k++) {
        if (c->cumul[k]) continue;
        RC_CHECK(dalloc(c, &c->cumul[k], 1));
        RC_CHECK(zero_fill(c, c->cumul[k], sizeof(float4)));
        c->cw = c->ch = c->cl = 1;
    }
    return PT_OK;
}

/* The ribbon field pair, allocated on first use with every ribbon invalid
 * (RibbonP()). */
static int ribbons_ready(pt_ctx *c)
{
    if (c->rib_base) return PT_OK;
    if (c->rec) return set_err(c, PT_ERR_STATE, "ribbon fields must exist before recording");
    size_t n = (size_t)c->W * c->H;
    RC_CHECK(dalloc(c, &c->rib_base, n));
    RC_CHECK(dalloc(c, &c->rib_work, n));
    pt_ribbonp *a = c->rib_base, *b = c->rib_work;
    return issue(c, "ribbon fill", [=](hipStream_t s) {
        hipError_t e = pt_launch_fill_ribbon(a, n, s);
        return e != hipSuccess ? e : pt_launch_fill_ribbon(b, n, s);
    });
}

extern "C" int pt_initialize(pt_ctx *c, int iw, int ih, int sw, int sh)
{
    if ((int64_t)iw * ih > PT_MAX_IMAGE_PIXELS)
        return set_err(c, PT_ERR_UNSUPPORTED, "image larger than PT_MAX_IMAGE_PIXELS");
    if ((int64_t)sw * sh > PT_MAX_EXEMPLAR_TEXELS)
        return set_err(c, PT_ERR_UNSUPPORTED, "exemplar larger than PT_MAX_EXEMPLAR_TEXELS");
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
    RC_CHECK(dalloc(c, &c->base.lum, n));
    RC_CHECK(dalloc(c, &c->work.lum, n));
    RC_CHECK(dalloc(c, &c->base.la, n));
    RC_CHECK(dalloc(c, &c->work.la, n));
    RC_CHECK(zero_fill(c, c->base.la, n * sizeof(int32_t)));
    RC_CHECK(zero_fill(c, c->work.la, n * sizeof(int32_t)));
    RC_CHECK(dalloc(c, &c->res_base, n));
    RC_CHECK(dalloc(c, &c->res_work, n));
    RC_CHECK(dalloc(c, &c->canvas, n));
    RC_CHECK(dalloc(c, &c->changed, n));
    for (int k = 0; k < 2; k++) {
        RC_CHECK(dalloc(c, &c->adv[k].xyth, n));
        RC_CHECK(dalloc(c, &c->adv[k].sc, n));
        RC_CHECK(dalloc(c, &c->adv[k].base, n));
        RC_CHECK(dalloc(c, &c->adv[k].out, n));
        RC_CHECK(dalloc(c, &c->adv[k].res, n));
        RC_CHECK(dalloc(c, &c->adv[k].ts, n));
        RC_CHECK(dalloc(c, &c->adv[k].lum, n));
        RC_CHECK(dalloc(c, &c->adv[k].la, n));
        RC_CHECK(zero_fill(c, c->adv[k].la, n * sizeof(int32_t)));
        RC_CHECK(dalloc(c, &c->color[k], n));
        RC_CHECK(dalloc(c, &c->dt[k], n));
        RC_CHECK(dalloc(c, &c->velb[k], n));
        RC_CHECK(dalloc(c, &c->velf[k], n));
    }
    /* level-sized textures never exceed level 0 */
    RC_CHECK(dalloc(c, &c->orient, n));
    RC_CHECK(dalloc(c, &c->scale, n));
    RC_CHECK(dalloc(c, &c->random, n));
    RC_CHECK(dalloc(c, &c->img_id, n));
    RC_CHECK(zero_fill(c, c->img_id, sizeof(int32_t)));
    c->idw = c->idh = 1;
    /* unbound frame-to-key ribbons read as one zero texel */
    for (int k = 0; k < 2; k++) RC_CHECK(ftk_reset(c, k));
    RC_CHECK(cumul_defaults(c));
    c->layered_live = false;
    c->ow = c->oh = c->scw = c->sch = c->rw = c->rh = 0;
    c->off_diff = Rect();
    c->res_diff = Rect();
    c->adv_written[0] = c->adv_written[1] = false;
    tt_invalidate(c);
    c->cur = 0;
    c->res_level = -1;
    c->lum_live = c->have_params && c->p.lum_shift_enable;
    return sync_stream(c);
}

extern "C" int pt_initialize_exemplars(pt_ctx *c, int w, int h, int num_layers)
{
    /* the gathers address the interleaved 32-byte texels with 32-bit byte
     * offsets (ex_rsrc, ex_pair); checked before anything else */
    if ((int64_t)w * h * (num_layers > 1 ? num_layers : 1) > PT_MAX_EXEMPLAR_TEXELS)
        return set_err(c, PT_ERR_UNSUPPORTED, "exemplar layers larger than PT_MAX_EXEMPLAR_TEXELS");
    if (!c || w < 1 || h < 1) return set_err(c, PT_ERR_INVALID_ARG, "bad exemplar dimensions");
    if (num_layers < 1 || num_layers > PT_MAX_STYLE_LAYERS)
        return set_err(c, PT_ERR_INVALID_ARG, "num_layers must be in [1, %d]", PT_MAX_STYLE_LAYERS);
    HIP_CHECK(c, hipSetDevice(c->device));
    res_invalidate(c);
    tt_invalidate(c);
    if (c->ex && c->ew == w && c->eh == h && c->nl == num_layers) return PT_OK;
    /* style layer l occupies texels [l * w * h, (l + 1) * w * h) */
    size_t n = (size_t)w * h * num_layers;
    if (n > c->ex_cap) {
        if (c->rec) return set_err(c, PT_ERR_STATE, "exemplar buffers must be sized before recording");
        HIP_CHECK(c, hipStreamSynchronize(c->stream));
        dfree(c->ex);
        dfree(c->ex_orient);
        dfree(c->ex_dt);
        dfree(c->ex_id);
        RC_CHECK(dalloc(c, &c->ex, 2 * n));
        RC_CHECK(dalloc(c, &c->ex_orient, n));
        RC_CHECK(dalloc(c, &c->ex_dt, n));
        RC_CHECK(dalloc(c, &c->ex_id, n));
        c->ex_cap = n;
    } else {
        /* a fresh texture allocation of the new size: zero it */
        RC_CHECK(zero_fill(c, c->ex, 2 * n * sizeof(float4)));
        RC_CHECK(zero_fill(c, c->ex_orient, n * sizeof(float)));
        RC_CHECK(zero_fill(c, c->ex_dt, n * sizeof(float)));
        RC_CHECK(zero_fill(c, c->ex_id, n * sizeof(int32_t)));
    }
    c->ew = w;
    c->eh = h;
    c->nl = num_layers;
    return sync_stream(c);
}

/* TexSynth::initialize_cumulativeFields (texSynth_kernel.cu:2136-2140):
 * zero-filled layered fields of w x h texels. */
extern "C" int pt_initialize_cumulative_fields(pt_ctx *c, int w, int h, int num_layers)
{
    if (!c || w < 1 || h < 1 || num_layers < 1 || num_layers > PT_MAX_STYLE_LAYERS)
        return set_err(c, PT_ERR_INVALID_ARG, "bad cumulative field dimensions");
    HIP_CHECK(c, hipSetDevice(c->device));
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    size_t n = (size_t)w * h * num_layers;
    for (int k = 0; k < 2; k++) {
        dfree(c->cumul[k]);
        RC_CHECK(dalloc(c, &c->cumul[k], n));
        RC_CHECK(zero_fill(c, c->cumul[k], n * sizeof(float4)));
    }
    c->cw = w;
    c->ch = h;
    c->cl = num_layers;
    c->layered_live = true;
    res_invalidate(c);
    return sync_stream(c);
}

static int reset_advected(pt_ctx *c)
{
    res_invalidate(c);
    size_t n = (size_t)c->W * c->H;
    for (int k = 0; k < 2; k++) {
        RC_CHECK(zero_fill(c, c->adv[k].xyth, n * sizeof(float4)));
        RC_CHECK(zero_fill(c, c->adv[k].sc, n * sizeof(float2)));
        RC_CHECK(zero_fill(c, c->adv[k].base, n * sizeof(float4)));
        RC_CHECK(zero_fill(c, c->adv[k].out, n * sizeof(float4)));
        RC_CHECK(zero_fill(c, c->adv[k].res, n * sizeof(float)));
        RC_CHECK(zero_fill(c, c->adv[k].ts, n * sizeof(int32_t)));
        RC_CHECK(zero_fill(c, c->adv[k].lum, n * sizeof(float)));
        RC_CHECK(zero_fill(c, c->adv[k].la, n * sizeof(int32_t)));
        c->adv_written[k] = false;
    }
    tt_invalidate(c);
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
    res_invalidate(c);
    return PT_OK;
}

/* ------------------------------------------------------------------ */
/* pyramids                                                            */
/* ------------------------------------------------------------------ */
static int pyramid_rebuild(pt_ctx *c, pt_pyramid *p, int scale_mode)
{
    int ch = elem_channels(p->elem);
    const bool vote = p->elem == PT_ELEM_I32;
    for (int l = 1; l < p->levels; l++) {
        const float *src = p->data + p->off[l - 1];
        float *dst = p->data + p->off[l];
        int pw = p->lw[l - 1], lw = p->lw[l], lh = p->lh[l];
        /* a 2x2 block read and one element written per output pixel */
        int pr = prof_begin(c, PK_PYRAMID, (uint64_t)lw * lh, (uint64_t)lw * lh * ch * 4 * 5);
        RC_CHECK(issue(c, "downsample", [=](hipStream_t s) {
            if (vote) return pt_launch_downsample_vote((const int *)src, pw, (int *)dst, lw, lh, s);
            if (scale_mode == PT_DOWNSAMPLE_RIBBON) return pt_launch_downsample_ribbon(src, pw, dst, lw, lh, s);
            return pt_launch_downsample(src, pw, dst, lw, lh, ch, scale_mode, s);
        }));
        prof_end(c, pr);
    }
    return PT_OK;
}

/* The fill pattern of a default level 0: the value itself, or for object-ID
 * pyramids the bits of (int)v. */
static float fill_bits(int elem, float v)
{
    if (elem != PT_ELEM_I32) return v;
    int32_t i = (int32_t)v;
    float f;
    memcpy(&f, &i, sizeof(f));
    return f;
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
    if (scale_mode == PT_DOWNSAMPLE_RIBBON && elem != PT_ELEM_F32X4)
        return set_err(c, PT_ERR_INVALID_ARG, "ribbon pyramids are F32X4");
    HIP_CHECK(c, hipSetDevice(c->device));
    pt_pyramid *p = nullptr;
    RC_CHECK(pyramid_alloc(c, &p, elem, w, h));
    size_t n0 = (size_t)w * h * elem_channels(elem);
    hipError_t e = host0 ? hipMemcpyAsync(p->data, host0, n0 * sizeof(float), hipMemcpyHostToDevice, c->stream)
                         : pt_launch_fill(p->data, n0, fill_bits(elem, defv), c->stream);
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
    if (style_layer < 0 || style_layer >= PT_MAX_STYLE_LAYERS)
        return set_err(c, PT_ERR_INVALID_ARG, "style layer must be in [0, %d]", PT_MAX_STYLE_LAYERS - 1);
    int w = p->lw[level], h = p->lh[level];
    /* first texel of the exemplar style layer */
    const size_t exl = (size_t)style_layer * c->ew * c->eh;
    const float *src = p->data + p->off[level];
    int prev = (c->cur + 1) % 2;
    res_invalidate(c);
    if (layer == PT_LAYER_EXEMPLAR_BASE || layer == PT_LAYER_EXEMPLAR_OUTPUT || layer == PT_LAYER_EXEMPLAR_ORIENTATION ||
        layer == PT_LAYER_INPUT_ORIENTATION)
        tt_invalidate(c);
    const int W = c->W;
    auto rect = [&](void *dst, int ch) -> int {
        if (w > c->W || h > c->H) return set_err(c, PT_ERR_INVALID_ARG, "level larger than image");
        return issue(c, "image upload", [=](hipStream_t s) {
            return pt_copy2d(dst, (size_t)W * ch * sizeof(float), src, (size_t)w * ch * sizeof(float),
                                    (size_t)w * ch * sizeof(float), h, hipMemcpyDeviceToDevice, s);
        });
    };
    auto tight = [&](void *dst, int ch) -> int {
        return issue(c, "texture upload", [=](hipStream_t s) {
            return pt_copy(dst, src, (size_t)w * h * ch * sizeof(float), hipMemcpyDeviceToDevice, s);
        });
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
    case PT_LAYER_INPUT_ID:
        if (p->elem != PT_ELEM_I32 || w > c->W || h > c->H) break;
        c->idw = w;
        c->idh = h;
        return tight(c->img_id, 1);
    case PT_LAYER_EXEMPLAR_OBJECT_IDS: {
        if (p->elem != PT_ELEM_I32 || !c->ex || w > c->ew || h > c->eh || style_layer >= c->nl) break;
        int32_t *dst = c->ex_id + exl;
        int ew = c->ew;
        return issue(c, "exemplar ID upload", [=](hipStream_t s) {
            return pt_copy2d(dst, (size_t)ew * sizeof(int32_t), src, (size_t)w * sizeof(int32_t),
                                    (size_t)w * sizeof(int32_t), h, hipMemcpyDeviceToDevice, s);
        });
    }
    case PT_LAYER_EXEMPLAR_BASE:
    case PT_LAYER_EXEMPLAR_OUTPUT:
        if (p->elem != 4 || !c->ex || w > c->ew || h > c->eh || style_layer >= c->nl) break;
    {
        float4 *ex = c->ex + 2 * exl;
        int ew = c->ew, slot = layer == PT_LAYER_EXEMPLAR_BASE ? 0 : 1;
        return issue(c, "exemplar upload", [=](hipStream_t s) {
            return pt_launch_interleave((const float4 *)src, w, h, ex, ew, slot, s);
        });
    }
    case PT_LAYER_EXEMPLAR_ORIENTATION:
    case PT_LAYER_EXEMPLAR_DIST_TRANS: {
        if (p->elem != 1 || !c->ex || w > c->ew || h > c->eh || style_layer >= c->nl) break;
        float *dst = (layer == PT_LAYER_EXEMPLAR_ORIENTATION ? c->ex_orient : c->ex_dt) + exl;
        int ew = c->ew;
        return issue(c, "exemplar guide upload", [=](hipStream_t s) {
            return pt_copy2d(dst, (size_t)ew * sizeof(float), src, (size_t)w * sizeof(float),
                                    (size_t)w * sizeof(float), h, hipMemcpyDeviceToDevice, s);
        });
    }
    case PT_LAYER_CUMUL_FIELD_B:
    case PT_LAYER_CUMUL_FIELD_F: {
        /* into style layer `style_layer` of the cumulative field */
        int k = layer == PT_LAYER_CUMUL_FIELD_B ? 0 : 1;
        if (p->elem != 4 || w > c->cw || h > c->ch || style_layer >= c->cl) break;
        float4 *dst = c->cumul[k] + (size_t)style_layer * c->cw * c->ch;
        int cw = c->cw;
        c->layered_live = true;
        return issue(c, "cumulative field upload", [=](hipStream_t s) {
            return pt_copy2d(dst, (size_t)cw * sizeof(float4), src, (size_t)w * sizeof(float4),
                                    (size_t)w * sizeof(float4), h, hipMemcpyDeviceToDevice, s);
        });
    }
    case PT_LAYER_FRAME_TO_KEY_RIBBON_B:
    case PT_LAYER_FRAME_TO_KEY_RIBBON_F: {
        /* a texture of the level's size */
        int k = layer == PT_LAYER_FRAME_TO_KEY_RIBBON_B ? 0 : 1;
        if (p->elem != 4) break;
        size_t need = (size_t)w * h;
        if (need > c->ftk_cap[k]) {
            if (c->rec) return set_err(c, PT_ERR_STATE, "ribbon textures must be sized before recording");
            HIP_CHECK(c, hipStreamSynchronize(c->stream));
            dfree(c->ftk[k]);
            c->ftk_cap[k] = 0;
            RC_CHECK(dalloc(c, &c->ftk[k], need));
            c->ftk_cap[k] = need;
        }
        c->fw[k] = w;
        c->fh[k] = h;
        c->layered_live = true;
        return tight(c->ftk[k], 4);
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
    {
        const float4 *ic = (const float4 *)(icol->data + icol->off[level]);
        const float *io = iori->data + iori->off[level];
        const float4 *eo = (const float4 *)(eout->data + eout->off[level]);
        const float *er = eori->data + eori->off[level];
        int transp = c->p.transparency_ok, oe = c->p.orientation_enabled;
        uint32_t seed = c->p.seed, k1 = (uint32_t)(frame * level);
        float4 *out = c->random;
        RC_CHECK(issue(c, "random pyramid", [=](hipStream_t s) {
            return pt_launch_random_pyramid(ic, io, w, h, eo, er, ew, eh, transp, oe, seed, k1, out, s);
        }));
    }
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

/* createRandomPyramid with keyframes (synthesisFrame.cpp:102-143): each
 * pixel draws its style among the keys first_key_index..last_key_index and
 * samples that style's exemplar; the texel's w holds the style. */
extern "C" int pt_create_random_pyramid_keys(pt_ctx *c, int level, int frame, const pt_pyramid *icol,
                                             const pt_pyramid *iori, int num_styles,
                                             const pt_pyramid *const *eouts, const pt_pyramid *const *eoris)
{
    if (!c) return set_err(nullptr, PT_ERR_INVALID_ARG, "NULL context");
    if (!icol || !iori || !eouts || !eoris) return set_err(c, PT_ERR_INVALID_ARG, "NULL pyramid");
    if (!c->base.xyth || !c->have_params) return set_err(c, PT_ERR_STATE, "not initialized");
    if (num_styles < 1 || num_styles > PT_MAX_STYLE_LAYERS)
        return set_err(c, PT_ERR_INVALID_ARG, "num_styles must be in [1, %d]", PT_MAX_STYLE_LAYERS);
    if (icol->elem != 4 || iori->elem != 1) return set_err(c, PT_ERR_INVALID_ARG, "pyramid kinds");
    if (level < 0 || level >= icol->levels || level >= iori->levels) return set_err(c, PT_ERR_INVALID_ARG, "level out of range");
    int w = icol->lw[level], h = icol->lh[level];
    if (iori->lw[level] != w || iori->lh[level] != h) return set_err(c, PT_ERR_INVALID_ARG, "pyramid sizes differ");
    if (w > c->W || h > c->H) return set_err(c, PT_ERR_INVALID_ARG, "level larger than image");
    int k0 = c->p.first_key_index, nk = c->p.last_key_index - c->p.first_key_index + 1;
    if (nk < 1 || k0 < 0 || k0 + nk > num_styles) return set_err(c, PT_ERR_INVALID_ARG, "keys outside the styles");
    const float4 *eo[PT_MAX_STYLE_LAYERS];
    const float *er[PT_MAX_STYLE_LAYERS];
    int ews[PT_MAX_STYLE_LAYERS], ehs[PT_MAX_STYLE_LAYERS];
    for (int k = 0; k < num_styles; k++) {
        const pt_pyramid *a = eouts[k], *b = eoris[k];
        if (!a || !b || a->elem != 4 || b->elem != 1 || level >= a->levels || level >= b->levels)
            return set_err(c, PT_ERR_INVALID_ARG, "style %d: pyramid kind or level", k);
        if (a->lw[level] != b->lw[level] || a->lh[level] != b->lh[level])
            return set_err(c, PT_ERR_INVALID_ARG, "style %d: pyramid sizes differ", k);
        if (a->lw[level] < 2 || a->lh[level] < 2) return set_err(c, PT_ERR_INVALID_ARG, "exemplar level too small");
        eo[k] = (const float4 *)(a->data + a->off[level]);
        er[k] = b->data + b->off[level];
        ews[k] = a->lw[level];
        ehs[k] = a->lh[level];
    }
    HIP_CHECK(c, hipSetDevice(c->device));
    const float4 *ic = (const float4 *)(icol->data + icol->off[level]);
    const float *io = iori->data + iori->off[level];
    int transp = c->p.transparency_ok, oe = c->p.orientation_enabled;
    uint32_t seed = c->p.seed, k1 = (uint32_t)(frame * level);
    float4 *out = c->random;
    HIP_CHECK(c, pt_launch_random_pyramid_keys(ic, io, w, h, num_styles, eo, er, ews, ehs, k0, nk, transp, oe, seed,
                                               k1, out, c->stream));
    c->rw = w;
    c->rh = h;
    c->layered_live = true;
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
    if (c->nl > 1 || c->p.interpolate_key_frame) c->layered_live = true;
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
    a.coh_w = p.coherence_weight > 0.f ? p.coherence_weight : 0.f;
    a.coh_ratio = p.coherence_ratio_range;
    a.coh_angle = p.coherence_angle_range;
    {
        float lz = pt_logf(a.coh_ratio);
        a.coh_inv_ratio = 1.0f / a.coh_ratio;
        a.coh_z = lz * lz;
        a.coh_lim = (double)a.coh_angle * 3.14159265358979323846;
    }
    a.lab = p.colorspace == 1 ? 1 : 0;
    a.obj_w = p.style_obj_id_weight > 0.f ? p.style_obj_id_weight : 0.f;
    a.lum_live = c->lum_live ? 1 : 0;
    a.lum_enable = p.lum_shift_enable ? 1 : 0;
    a.idw = c->idw;
    a.idh = c->idh;
    a.hyst_w = p.hysteresis_weight;
    a.max_dist = p.max_distortion;
    a.jitter = p.advection_jitter;
    a.samples = p.scatter_samples_count;
    a.rng_mode = p.rng_mode;
    a.key0 = p.seed ^ 0x5CA77E12u;
    a.key1 = (uint32_t)(p.frame_number * level);
    a.time_active = (c->adv_written[0] || c->adv_written[1]) ? 1 : 0;
    a.ee_ok = (p.hysteresis_weight == 0.f && p.input_analogy_weight >= 0.f && p.canvas_output_weight >= 0.f &&
               p.alpha_weight >= 0.f && p.dist_trans_weight >= 0.f && p.time_derivative_input_weight >= 0.f &&
               p.time_derivative_output_weight >= 0.f && p.temporal_coherence_weight >= 0.f && !c->no_early_exit)
                  ? 1
                  : 0;
    a.early_exit = (a.ee_ok && (int64_t)a.lw * a.lh >= c->early_exit_min_pixels) ? 1 : 0;
    a.unit_parallel = c->unit_parallel ? 1 : 0;
    a.res_incremental = 0;
    a.tap_max_px = (int)std::min<int64_t>(c->tap_max_pixels, 1 << 30);
    pt_tap_configure(a, c->tap_evals);
    a.ex_window = (c->ex_window_min_pixels > 0 && (int64_t)a.lw * a.lh >= c->ex_window_min_pixels) ? 1 : 0;
    a.xcd_tiles = c->xcd_tiles ? 1 : 0;
    a.prop_window_jump = c->prop_window_jump;
    a.prop_window = (c->prop_window_min_pixels > 0 && (int64_t)a.lw * a.lh >= c->prop_window_min_pixels) ? 1 : 0;
    a.time_unit_order = c->time_unit_order;
    a.changed = c->changed;
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
        a.adv[k].lum = c->adv[k].lum;
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
    a.img_id = c->img_id;
    a.ex_id = c->ex_id;
    a.layered = c->layered_live ? 1 : 0;
    a.nl = c->nl;
    a.keys = p.interpolate_key_frame ? 1 : 0;
    a.key_first = p.first_key_index;
    a.key_last = p.last_key_index;
    a.key_index = -1;
    for (int k = 0; k < 2; k++) {
        a.adv[k].la = c->adv[k].la;
        a.cumul[k] = c->cumul[k];
        a.ftk[k] = c->ftk[k];
        a.fw[k] = c->fw[k];
        a.fh[k] = c->fh[k];
    }
    a.cw = c->cw;
    a.ch = c->ch;
    a.cl = c->cl;
    return a;
}

/* copyWorkingToBase over the rectangle where the pair can differ */
static int copy_nnf_rect(pt_ctx *c, const PtNnf &dst, const PtNnf &src, Rect r)
{
    if (r.empty()) return PT_OK;
    const size_t W = c->W;
    const bool lum = c->lum_live, lay = c->layered_live;
    return issue(c, "offset copy", [=](hipStream_t s) {
        hipError_t e = pt_copy2d(dst.xyth, W * sizeof(float4), src.xyth, W * sizeof(float4),
                                        (size_t)r.w * sizeof(float4), r.h, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess && lum)
            e = pt_copy2d(dst.lum, W * sizeof(float), src.lum, W * sizeof(float), (size_t)r.w * sizeof(float),
                                 r.h, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess && lay)
            e = pt_copy2d(dst.la, W * sizeof(int32_t), src.la, W * sizeof(int32_t),
                                 (size_t)r.w * sizeof(int32_t), r.h, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return e;
        return pt_copy2d(dst.sc, W * sizeof(float2), src.sc, W * sizeof(float2), (size_t)r.w * sizeof(float2),
                                r.h, hipMemcpyDeviceToDevice, s);
    });
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

/* Sets a.tt[k] to the time table of advected set k for this launch,
 * building it first when it is stale: only for sets the temporal terms read
 * (valid offsets written, the pass direction uses them, a temporal weight
 * positive), patches of at most tt_max_taps taps, and tables that fit a
 * 32-bit element index.  While a plan is recorded an existing table is
 * never reallocated (recorded work may hold its address); a launch without a
 * table gathers the same samples itself, with the same result. */
static int time_tables(pt_ctx *c, int level, PtArgs &a)
{
    a.tt[0] = a.tt[1] = nullptr;
    if (!a.time_active || a.layered) return PT_OK; /* tables hold layer-0 samples */
    if (!(a.tdi > 0.f || a.tdo > 0.f || a.tcw > 0.f)) return PT_OK;
    const int64_t taps = (int64_t)(2 * a.r + 1) * (2 * a.r + 1);
    if (taps > c->tt_max_taps) return PT_OK;
    const bool dir[2] = {a.direction == PT_FORWARD || a.direction == PT_BIDIRECTIONAL,
                         a.direction == PT_BACKWARD || a.direction == PT_BIDIRECTIONAL};
    for (int k = 0; k < 2; k++) {
        if (!dir[k] || !c->adv_written[k]) continue;
        const size_t need = (size_t)taps * a.lw * a.lh * 2;
        if (need > (size_t)INT32_MAX) continue;
        if (c->tt_cap[k] < need) {
            if (c->rec && c->tt[k]) continue; /* recorded work may hold the old table */
            /* capacity for level 0, so finer levels of the same radius fit */
            const size_t cap = std::max(need, (size_t)taps * c->W * c->H * 2);
            if (cap > (size_t)INT32_MAX) continue;
            HIP_CHECK(c, hipStreamSynchronize(c->stream));
            dfree(c->tt[k]);
            c->tt[k] = nullptr;
            c->tt_cap[k] = 0;
            c->tt_valid[k] = false;
            RC_CHECK(dalloc(c, &c->tt[k], cap));
            c->tt_cap[k] = cap;
        }
        if (!c->tt_valid[k] || c->tt_level[k] != level || c->tt_r[k] != a.r || c->tt_lab[k] != a.lab) {
            PtArgs b = a;
            float4 *t = c->tt[k];
            const int dk = k;
            RC_CHECK(issue(c, "time table", [=](hipStream_t st) { return pt_launch_time_table(b, dk, t, st); }));
            c->tt_valid[k] = true;
            c->tt_level[k] = level;
            c->tt_r[k] = a.r;
            c->tt_lab[k] = a.lab;
        }
        a.tt[k] = c->tt[k];
    }
    return PT_OK;
}

static int update_residual(pt_ctx *c, int level, bool allow_incremental = false)
{
    PtArgs a = make_args(c, level);
    RC_CHECK(time_tables(c, level, a));
    a.res_incremental = (allow_incremental && c->incremental_residual && c->res_level == level) ? 1 : 0;
    c->res_level = level;
    int pr = prof_begin(c, PK_RESIDUAL, (uint64_t)a.lw * a.lh, PB_RESIDUAL * a.lw * a.lh);
    RC_CHECK(issue_op(c, PT_OP_RESIDUAL, a));
    prof_end(c, pr);
    c->res_diff.grow(a.lw, a.lh);
    c->pde += (uint64_t)a.lw * a.lh;
    return PT_OK;
}

/* TexSynth::syncBuffers (texSynth_kernel.cu:1240-1303) */
static int sync_buffers(pt_ctx *c, int level, int flags)
{
    PtArgs a = make_args(c, level);
    if (!(flags & PT_DONT_COPY_OFFSETS) && !c->off_diff.empty()) {
        /* copy of the differing rectangle fused with the canvas pass */
        a.dw = c->off_diff.w;
        a.dh = c->off_diff.h;
        c->off_diff = Rect();
    } else {
        a.dw = a.dh = 0; /* canvas only */
    }
    /* copied pixels: working offsets read, base offsets read (change
     * flags) and written, 24 B each; level pixels: canvas written, one
     * exemplar texel, the change flag, and the base offset read where it
     * was not copied */
    uint64_t lvl = (uint64_t)a.lw * a.lh, cp = (uint64_t)a.dw * a.dh;
    uint64_t cpl = (uint64_t)std::min(a.dw, a.lw) * std::min(a.dh, a.lh);
    int pr = prof_begin(c, PK_CANVAS, lvl, cp * 72 + lvl * 33 + (lvl - cpl) * 16);
    RC_CHECK(issue_op(c, PT_OP_SYNC_CANVAS, a));
    prof_end(c, pr);
    if (!(flags & PT_DONT_UPDATE_RESIDUAL)) RC_CHECK(update_residual(c, level, true));
    else res_invalidate(c);
    return PT_OK;
}

static int randomize_offsets(pt_ctx *c, int level)
{
    if (c->rw == 0) return set_err(c, PT_ERR_STATE, "no random texture");
    res_invalidate(c);
    PtArgs a = make_args(c, level);
    int pr = prof_begin(c, PK_RANDOM, (uint64_t)a.lw * a.lh, PB_RANDOM * a.lw * a.lh);
    RC_CHECK(issue(c, "random copy", [=](hipStream_t s) { return pt_launch_copy_random(a, s); }));
    prof_end(c, pr);
    c->off_diff.grow(a.lw, a.lh);
    return sync_buffers(c, level, 0);
}

static int propagate(pt_ctx *c, int level, int jump, int key_index = -1)
{
    if (key_index >= 0) c->layered_live = true;
    PtArgs a = make_args(c, level);
    RC_CHECK(time_tables(c, level, a));
    a.jump = jump;
    a.key_index = key_index;
    /* the cache is known to equal a recomputation for the current offsets
     * (the kernels then skip candidates equal to a pixel's offset) */
    a.res_exact = c->res_level == level ? 1 : 0;
    int pr = prof_begin(c, PK_PROPAGATE, (uint64_t)a.lw * a.lh, PB_NNF * a.lw * a.lh);
    RC_CHECK(issue_op(c, PT_OP_PROPAGATE, a));
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

static int scatter(pt_ctx *c, int level, int key_index = -1)
{
    if (key_index >= 0) c->layered_live = true;
    PtArgs a = make_args(c, level);
    RC_CHECK(time_tables(c, level, a));
    a.call = c->scatter_calls++;
    a.key_index = key_index;
    /* with keyframes the samples are shared by the keys
     * (texSynth_kernel.cu:765) and there is no averaged candidate */
    if (a.keys) a.samples /= (a.key_last - a.key_first + 1);
    a.nscat = scatter_samples(c->ew, c->eh, a.samples);
    int pr = prof_begin(c, PK_SCATTER, (uint64_t)a.lw * a.lh, PB_NNF * a.lw * a.lh);
    RC_CHECK(issue_op(c, PT_OP_SCATTER, a));
    prof_end(c, pr);
    c->off_diff.grow(a.lw, a.lh);
    c->pde += (uint64_t)a.lw * a.lh * (uint64_t)(scatter_samples(c->ew, c->eh, a.samples) + (a.keys ? 0 : 1));
    return sync_buffers(c, level, 0);
}

static int upsample(pt_ctx *c, int fine_level)
{
    res_invalidate(c);
    PtArgs a = make_args(c, fine_level);
    int pr = prof_begin(c, PK_UPSAMPLE, (uint64_t)a.lw * a.lh, PB_UPSAMPLE * a.lw * a.lh);
    RC_CHECK(issue(c, "upsample", [=](hipStream_t s) { return pt_launch_upsample(a, s); }));
    prof_end(c, pr);
    c->off_diff.grow(a.lw, a.lh);
    c->res_diff.grow(a.lw, a.lh);
    RC_CHECK(copy_working_to_base(c));
    if (!c->res_diff.empty()) {
        float *rb = c->res_base, *rw = c->res_work;
        size_t W = c->W;
        Rect r = c->res_diff;
        RC_CHECK(issue(c, "residual copy", [=](hipStream_t s) {
            return pt_copy2d(rb, W * sizeof(float), rw, W * sizeof(float), (size_t)r.w * sizeof(float), r.h,
                                    hipMemcpyDeviceToDevice, s);
        }));
    }
    c->res_diff = Rect();
    return PT_OK;
}

static int advect(pt_ctx *c, int level, int time_step, int forwards, int flags)
{
    if (c->rw == 0) return set_err(c, PT_ERR_STATE, "no random texture");
    res_invalidate(c);
    if (flags & PT_ADVECT_CHECK_FRAME_TO_KEY) c->layered_live = true;
    PtArgs a = make_args(c, level);
    int pr = prof_begin(c, PK_ADVECT, (uint64_t)a.lw * a.lh, PB_ADVECT * a.lw * a.lh);
    RC_CHECK(issue(c, "advect", [=](hipStream_t s) { return pt_launch_advect(a, forwards, time_step, flags, s); }));
    prof_end(c, pr);
    c->adv_written[forwards ? 0 : 1] = true;
    tt_invalidate(c);
    DevAdv d = c->adv[forwards ? 0 : 1];
    size_t n = (size_t)c->W * c->H;
    /* dest.offsets -> base, whole buffer (texSynth_kernel.cu:1716) */
    PtNnf base = c->base;
    const bool lum = c->lum_live, lay = c->layered_live;
    RC_CHECK(issue(c, "advected offset copy", [=](hipStream_t s) {
        hipError_t e = pt_copy(base.xyth, d.xyth, n * sizeof(float4), hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess && lum) e = pt_copy(base.lum, d.lum, n * sizeof(float), hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess && lay)
            e = pt_copy(base.la, d.la, n * sizeof(int32_t), hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return e;
        return pt_copy(base.sc, d.sc, n * sizeof(float2), hipMemcpyDeviceToDevice, s);
    }));
    c->off_diff.grow(c->W, c->H);
    RC_CHECK(sync_buffers(c, level, PT_DONT_COPY_OFFSETS | PT_DONT_UPDATE_RESIDUAL));
    if (flags & PT_ADVECT_UPDATE_RESIDUAL) RC_CHECK(update_residual(c, level));
    /* residual -> dest.residual, whole buffer (:1724) */
    float *rb = c->res_base;
    RC_CHECK(issue(c, "advected residual copy", [=](hipStream_t s) {
        return pt_copy(d.res, rb, n * sizeof(float), hipMemcpyDeviceToDevice, s);
    }));
    return PT_OK;
}

static int merge_randomly(pt_ctx *c, int level, int forwards, float w, int force)
{
    if (c->rw == 0) return set_err(c, PT_ERR_STATE, "no random texture");
    res_invalidate(c);
    PtArgs a = make_args(c, level);
    int pr = prof_begin(c, PK_MERGE, (uint64_t)a.lw * a.lh, PB_MERGE * a.lw * a.lh);
    RC_CHECK(issue(c, "merge", [=](hipStream_t s) { return pt_launch_merge_randomly(a, forwards, w, force, s); }));
    prof_end(c, pr);
    c->off_diff.grow(a.lw, a.lh);
    return sync_buffers(c, level, 0);
}

static int merge_residual(pt_ctx *c, int level, int forwards)
{
    res_invalidate(c);
    PtArgs a = make_args(c, level);
    int pr = prof_begin(c, PK_MERGE, (uint64_t)a.lw * a.lh, PB_MERGE * a.lw * a.lh);
    RC_CHECK(issue(c, "merge", [=](hipStream_t s) { return pt_launch_merge_residual(a, forwards, s); }));
    prof_end(c, pr);
    c->off_diff.grow(a.lw, a.lh);
    return sync_buffers(c, level, 0);
}

static int merge_advected(pt_ctx *c, int level, int fallback)
{
    if (c->rw == 0) return set_err(c, PT_ERR_STATE, "no random texture");
    res_invalidate(c);
    c->layered_live = true;
    PtArgs a = make_args(c, level);
    int pr = prof_begin(c, PK_MERGE, (uint64_t)a.lw * a.lh, PB_MERGE * a.lw * a.lh);
    RC_CHECK(issue(c, "merge", [=](hipStream_t s) { return pt_launch_merge_advected(a, fallback, s); }));
    prof_end(c, pr);
    c->off_diff.grow(a.lw, a.lh);
    return sync_buffers(c, level, 0);
}

static int offsets_from_ribbon(pt_ctx *c, int level, int look_forwards)
{
    if (c->rw == 0) return set_err(c, PT_ERR_STATE, "no random texture");
    res_invalidate(c);
    c->layered_live = true;
    PtArgs a = make_args(c, level);
    int pr = prof_begin(c, PK_RANDOM, (uint64_t)a.lw * a.lh, PB_RANDOM * a.lw * a.lh);
    RC_CHECK(issue(c, "ribbon offsets", [=](hipStream_t s) { return pt_launch_offsets_from_ribbon(a, look_forwards, s); }));
    prof_end(c, pr);
    c->off_diff.grow(a.lw, a.lh);
    return sync_buffers(c, level, 0);
}

/* TexSynth::accumulateRibbonField (texSynth_kernel.cu:1828-1854): one step
 * into the working ribbon buffer, copied whole to the base, then the canvas
 * texels the ribbons reach into the forward advected output colours. */
static int accumulate_ribbon(pt_ctx *c, int level, int forwards, int time_step, int overwrite)
{
    RC_CHECK(ribbons_ready(c));
    PtArgs a = make_args(c, level);
    pt_ribbonp *rb = c->rib_base, *rw = c->rib_work;
    size_t n = (size_t)c->W * c->H;
    int pr = prof_begin(c, PK_ADVECT, (uint64_t)a.lw * a.lh, 80 * (uint64_t)a.lw * a.lh);
    RC_CHECK(issue(c, "ribbon accumulation", [=](hipStream_t s) {
        hipError_t e = pt_launch_accumulate_ribbon(a, rb, rw, forwards, time_step, overwrite, s);
        if (e == hipSuccess) e = pt_copy(rb, rw, n * sizeof(pt_ribbonp), hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = pt_launch_ribbon_output(a, rb, s);
        return e;
    }));
    prof_end(c, pr);
    c->adv_written[0] = true;
    tt_invalidate(c);
    res_invalidate(c);
    return PT_OK;
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

static int check_key_index(pt_ctx *c, int key_index)
{
    if (key_index >= PT_MAX_STYLE_LAYERS)
        return set_err(c, PT_ERR_INVALID_ARG, "key_index must be below %d", PT_MAX_STYLE_LAYERS);
    return PT_OK;
}

extern "C" int pt_propagate(pt_ctx *c, int level, int jump, int key_index)
{
    RC_CHECK(need_level(c, level));
    RC_CHECK(check_key_index(c, key_index));
    PUBLIC_OP(propagate(c, level, jump, key_index));
}

extern "C" int pt_scatter(pt_ctx *c, int level, int key_index)
{
    RC_CHECK(need_level(c, level));
    RC_CHECK(check_key_index(c, key_index));
    PUBLIC_OP(scatter(c, level, key_index));
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

extern "C" int pt_merge_advected_buffers(pt_ctx *c, int level, int merge_fallback)
{
    RC_CHECK(need_level(c, level));
    PUBLIC_OP(merge_advected(c, level, merge_fallback));
}

extern "C" int pt_offsets_from_frame_to_key_ribbon(pt_ctx *c, int level, int look_forwards)
{
    RC_CHECK(need_level(c, level));
    PUBLIC_OP(offsets_from_ribbon(c, level, look_forwards));
}

extern "C" int pt_accumulate_ribbon_field(pt_ctx *c, int level, int forwards, int time_step, int overwrite)
{
    if (!c) return set_err(nullptr, PT_ERR_INVALID_ARG, "NULL context");
    if (!c->base.xyth) return set_err(c, PT_ERR_STATE, "not initialized");
    if (level < 0 || level > 30 || (c->W >> level) < 1 || (c->H >> level) < 1)
        return set_err(c, PT_ERR_INVALID_ARG, "bad level %d", level);
    PUBLIC_OP(accumulate_ribbon(c, level, forwards, time_step, overwrite));
}

/* The ribbon field (base) of a level, lw x lh ribbons, to or from the host. */
extern "C" int pt_load_ribbon_field(pt_ctx *c, int level, const pt_ribbonp *ribbons)
{
    if (!c || !c->base.xyth || !ribbons || level < 0 || level > 30 || (c->W >> level) < 1 || (c->H >> level) < 1)
        return set_err(c, PT_ERR_INVALID_ARG, "bad load_ribbon_field");
    HIP_CHECK(c, hipSetDevice(c->device));
    RC_CHECK(ribbons_ready(c));
    int lw = c->W >> level, lh = c->H >> level;
    HIP_CHECK(c, hipMemcpy2DAsync(c->rib_base, (size_t)c->W * sizeof(pt_ribbonp), ribbons, (size_t)lw * sizeof(pt_ribbonp),
                                  (size_t)lw * sizeof(pt_ribbonp), lh, hipMemcpyHostToDevice, c->stream));
    return sync_stream(c);
}

extern "C" int pt_get_ribbon_field(pt_ctx *c, int level, pt_ribbonp *ribbons)
{
    if (!c || !c->base.xyth || !ribbons || level < 0 || level > 30 || (c->W >> level) < 1 || (c->H >> level) < 1)
        return set_err(c, PT_ERR_INVALID_ARG, "bad get_ribbon_field");
    HIP_CHECK(c, hipSetDevice(c->device));
    RC_CHECK(ribbons_ready(c));
    int lw = c->W >> level, lh = c->H >> level;
    HIP_CHECK(c, hipMemcpy2DAsync(ribbons, (size_t)lw * sizeof(pt_ribbonp), c->rib_base, (size_t)c->W * sizeof(pt_ribbonp),
                                  (size_t)lw * sizeof(pt_ribbonp), lh, hipMemcpyDeviceToHost, c->stream));
    return sync_stream(c);
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
    PtNnf b = c->base;
    if (!c->lum_live) b.lum = nullptr;
    if (!c->layered_live) b.la = nullptr;
    if (offsets) RC_CHECK(get_nnf(c, b, lw, lh, offsets));
    if (canvas) RC_CHECK(get_plane(c, c->canvas, lw, lh, (float4 *)canvas));
    if (residual) RC_CHECK(get_plane(c, c->res_base, lw, lh, residual));
    return sync_stream(c);
}

extern "C" int pt_get_output_noalloc(pt_ctx *c, int level, pt_patchxf *offsets, float *canvas, float *residual)
{
    if (!c || !c->base.xyth || level < 0 || (c->W >> level) < 1 || (c->H >> level) < 1)
        return set_err(c, PT_ERR_INVALID_ARG, "bad get_output_noalloc");
    int lw = c->W >> level, lh = c->H >> level;
    size_t bytes = (size_t)lw * lh * sizeof(pt_patchxf);
    if (offsets && bytes > PT_NOALLOC_SCRATCH_BYTES)
        return set_err(c, PT_ERR_UNSUPPORTED, "%d x %d offsets need %zu staging bytes, above "
                       "PT_NOALLOC_SCRATCH_BYTES (%u)", lw, lh, bytes, (unsigned)PT_NOALLOC_SCRATCH_BYTES);
    HIP_CHECK(c, hipSetDevice(c->device));
    PtNnf b = c->base;
    if (!c->lum_live) b.lum = nullptr;
    if (!c->layered_live) b.la = nullptr;
    if (offsets) {
        /* every call on this GPU shares its array: one user at a time */
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        pt_patchxf *stage = (pt_patchxf *)g_scratch_ptr[c->device];
        if (!stage) return set_err(c, PT_ERR_STATE, "no staging array for device %d", c->device);
        /* the pack writes every staged element: nothing needs clearing */
        hipError_t e = pt_launch_pack_nnf(b, c->W, lw, lh, stage, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(offsets, stage, bytes, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return set_err(c, PT_ERR_HIP, "offset readback failed: %s", hipGetErrorString(e));
    }
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
    PtNnf n{d.xyth, d.sc, c->lum_live ? d.lum : nullptr, c->layered_live ? d.la : nullptr};
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
    res_invalidate(c);
    int lw = c->W >> level, lh = c->H >> level;
    bool shifted = false, layered = false;
    for (size_t i = 0; i < (size_t)lw * lh; i++) {
        if (in[i].lum_a != 0.f)
            return set_err(c, PT_ERR_UNSUPPORTED, "offsets with an alpha luminance shift");
        shifted = shifted || in[i].lum_y != 0.f;
        layered = layered || in[i].layer != 0;
    }
    if (shifted) c->lum_live = true;
    if (layered) c->layered_live = true;
    PtNnf b = c->base;
    if (!c->lum_live) b.lum = nullptr;
    if (!c->layered_live) b.la = nullptr;
    pt_patchxf *tmp = nullptr;
    RC_CHECK(dalloc(c, &tmp, (size_t)lw * lh));
    hipError_t e = hipMemcpyAsync(tmp, in, (size_t)lw * lh * sizeof(pt_patchxf), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = pt_launch_unpack_nnf(tmp, lw, lh, b, c->W, c->stream);
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
    res_invalidate(c);
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
                RC_CHECK(dalloc(c, &s.lum, n));
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
static int snap_save(pt_ctx *c, int level, Snapshot &sn)
{
    sn.saved = true;
    Snapshot d = sn;
    PtNnf base = c->base;
    float *rb = c->res_base;
    float4 *cv = c->canvas;
    size_t W = c->W;
    const bool lum = c->lum_live;
    (void)level;
    return issue(c, "snapshot save", [=](hipStream_t st) {
        hipError_t e = pt_copy2d(d.xyth, (size_t)d.w * sizeof(float4), base.xyth, W * sizeof(float4),
                                        (size_t)d.w * sizeof(float4), d.h, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess && lum)
            e = pt_copy2d(d.lum, (size_t)d.w * sizeof(float), base.lum, W * sizeof(float),
                                 (size_t)d.w * sizeof(float), d.h, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess)
            e = pt_copy2d(d.sc, (size_t)d.w * sizeof(float2), base.sc, W * sizeof(float2),
                                 (size_t)d.w * sizeof(float2), d.h, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess)
            e = pt_copy2d(d.res, (size_t)d.w * sizeof(float), rb, W * sizeof(float), (size_t)d.w * sizeof(float),
                                 d.h, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess)
            e = pt_copy2d(d.canvas, (size_t)d.w * sizeof(float4), cv, W * sizeof(float4),
                                 (size_t)d.w * sizeof(float4), d.h, hipMemcpyDeviceToDevice, st);
        return e;
    });
}

/* WorkingBuffers::loadOffsets (workingBuffers.cpp:190-207) from a snapshot */
static int snap_load_offsets(pt_ctx *c, const Snapshot &sn)
{
    res_invalidate(c);
    Snapshot d = sn;
    PtNnf base = c->base;
    size_t W = c->W;
    const bool lum = c->lum_live;
    RC_CHECK(issue(c, "snapshot offset load", [=](hipStream_t st) {
        hipError_t e = pt_copy2d(base.xyth, W * sizeof(float4), d.xyth, (size_t)d.w * sizeof(float4),
                                        (size_t)d.w * sizeof(float4), d.h, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess && lum)
            e = pt_copy2d(base.lum, W * sizeof(float), d.lum, (size_t)d.w * sizeof(float),
                                 (size_t)d.w * sizeof(float), d.h, hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return e;
        return pt_copy2d(base.sc, W * sizeof(float2), d.sc, (size_t)d.w * sizeof(float2),
                                (size_t)d.w * sizeof(float2), d.h, hipMemcpyDeviceToDevice, st);
    }));
    c->off_diff.grow(sn.w, sn.h);
    return PT_OK;
}

/* WorkingBuffers::loadResidual (workingBuffers.cpp:166-188) from a snapshot */
static int snap_load_residual(pt_ctx *c, const Snapshot &sn)
{
    res_invalidate(c);
    Snapshot d = sn;
    float *rb = c->res_base;
    size_t W = c->W;
    RC_CHECK(issue(c, "snapshot residual load", [=](hipStream_t st) {
        return pt_copy2d(rb, W * sizeof(float), d.res, (size_t)d.w * sizeof(float), (size_t)d.w * sizeof(float),
                                d.h, hipMemcpyDeviceToDevice, st);
    }));
    c->res_diff.grow(sn.w, sn.h);
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

/* Style::load (style.cpp:62-140): the exemplar layers of one level */
static int load_style_level(pt_ctx *c, int level, const StylePyr &sp)
{
    RC_CHECK(pt_initialize_exemplars(c, c->SW >> level, c->SH >> level, 1));
    RC_CHECK(upload_image(c, level, PT_LAYER_EXEMPLAR_BASE, sp.base, 0));
    RC_CHECK(upload_image(c, level, PT_LAYER_EXEMPLAR_OUTPUT, sp.out, 0));
    RC_CHECK(upload_image(c, level, PT_LAYER_EXEMPLAR_ORIENTATION, sp.orient, 0));
    RC_CHECK(upload_image(c, level, PT_LAYER_EXEMPLAR_DIST_TRANS, sp.dt, 0));
    return sp.id ? upload_image(c, level, PT_LAYER_EXEMPLAR_OBJECT_IDS, sp.id, 0) : PT_OK;
}

/* SynthesisFrame::loadImages / loadAnimImages (synthesisFrame.cpp:562-656):
 * frame f's guides of one level into the current layer */
static int load_frame_level(pt_ctx *c, int level, const FramePyr &f)
{
    RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_COLOR, f.color, 0));
    RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_ORIENTATION, f.orient, 0));
    if (f.id) RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_ID, f.id, 0));
    RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_VEL_B, f.velb, 0));
    RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_VEL_F, f.velf, 0));
    RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_DIST_TRANS, f.dt, 0));
    return upload_image(c, level, PT_LAYER_INPUT_SCALE, f.scale, 0);
}

/* Saved passes per (level, frame) of a scheme; the last one is the output
 * (copyOutputToFinal): FBMO saves its two directed passes, their merge and
 * the fixup passes. */
static int scheme_passes(int scheme, int fixup_passes, int levels, int F)
{
    if (scheme == PT_S_TCTF) /* a consecutive pass, then per level one refine (+ fixups + 1 from 3 frames) */
        return 1 + levels * (1 + (F > 2 ? fixup_passes + 1 : 0));
    return scheme == PT_S_FBMO ? fixup_passes + 3 : fixup_passes + 1;
}

/* SynthesisFrame::loadImages (synthesisFrame.cpp:562-578): colour,
 * orientation and object IDs of frame f into the current layer */
static int load_images(pt_ctx *c, int level, const FramePyr &f)
{
    RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_COLOR, f.color, 0));
    RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_ORIENTATION, f.orient, 0));
    return f.id ? upload_image(c, level, PT_LAYER_INPUT_ID, f.id, 0) : PT_OK;
}

/* SynthesisFrame::loadAnimImages (synthesisFrame.cpp:581-656), time step 1:
 * velocities, distance transform and scale of frame f */
static int load_anim_images(pt_ctx *c, int level, const FramePyr &f)
{
    RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_VEL_B, f.velb, 0));
    RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_VEL_F, f.velf, 0));
    RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_DIST_TRANS, f.dt, 0));
    return upload_image(c, level, PT_LAYER_INPUT_SCALE, f.scale, 0);
}

/* offsets.base().copyTo(advectedF.offsets()) and
 * residualCache.base().copyTo(advectedF.residual())
 * (synthesisProcessor.cpp:939-940): whole level-0-sized buffers */
static int copy_base_to_advected_f(pt_ctx *c)
{
    PtNnf base = c->base;
    DevAdv d = c->adv[0];
    float *rb = c->res_base;
    size_t n = (size_t)c->W * c->H;
    const bool lum = c->lum_live, lay = c->layered_live;
    RC_CHECK(issue(c, "copy to advected", [=](hipStream_t st) {
        hipError_t e = pt_copy(d.xyth, base.xyth, n * sizeof(float4), hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess && lum) e = pt_copy(d.lum, base.lum, n * sizeof(float), hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess && lay) e = pt_copy(d.la, base.la, n * sizeof(int32_t), hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess) e = pt_copy(d.sc, base.sc, n * sizeof(float2), hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess) e = pt_copy(d.res, rb, n * sizeof(float), hipMemcpyDeviceToDevice, st);
        return e;
    }));
    c->adv_written[0] = true;
    res_invalidate(c);
    tt_invalidate(c);
    return PT_OK;
}

/* renderAnimationFBMO (synthesisProcessor.cpp:902-970) with its fixup passes
 * (renderInterpolatedFixupPass :771-805, synthesizeInterpolated and
 * advectNeighbors, synthesisFrame.cpp:325-407).  Passes 0 and 1 are
 * independent forward and backward consecutive passes (every frame through
 * synthesizeConsecutive's first-pass branch, coarser levels read at pass
 * fixup+2); pass 2 merges them per frame (forward result as the advected
 * offsets, force-valid random merge); passes 3.. re-optimise every frame
 * with both neighbours advected into it (their last saved pass at this
 * level: the neighbour before has already run the current pass, the one
 * after has not). */
static int run_schedule_fbmo(pt_ctx *c, const pt_params *params, int iterations, int fixup, int F,
                             const FramePyr *fp, const StylePyr &sp)
{
    int levels = params->levels;
    int merged = fixup + 2;
    pt_params fparams = *params;
    std::vector<int> last(F, 0);
    for (int level = levels - 1; level >= 0; level--) {
        RC_CHECK(load_style_level(c, level, sp));
        for (int pass = 0; pass < 2; pass++) {
            RC_CHECK(reset_advected(c));
            bool fwd = pass == params->first_pass_direction; /* synthesisProcessor.cpp:938 */
            for (int k = 0; k < F; k++) {
                int f = fwd ? k : F - 1 - k;
                bool is_first = k == 0;
                if (!is_first || pass > 0) RC_CHECK(pt_flip_anim_buffers(c));
                RC_CHECK(load_images(c, level, fp[f]));
                RC_CHECK(load_anim_images(c, level, fp[f]));
                /* synthesizeConsecutive(level, pass, 1, is_first, fwd, fixup+2, firstPass=true) */
                fparams.direction = fwd ? PT_FORWARD : PT_BACKWARD;
                fparams.advection_jitter = (f % 3 == 0) ? params->advection_jitter : 0.f;
                fparams.frame_number = f;
                RC_CHECK(pt_upload_parameters(c, &fparams));
                RC_CHECK(random_pyramid(c, level, f, fp[f].color, fp[f].orient, sp.out, sp.orient));
                if (is_first) {
                    if (level == levels - 1) {
                        RC_CHECK(randomize_offsets(c, level));
                    } else {
                        RC_CHECK(snap_load_offsets(c, snap(c, level + 1, f, merged)));
                        RC_CHECK(upsample(c, level));
                        RC_CHECK(sync_buffers(c, level, 0));
                    }
                } else if (level == levels - 1) {
                    RC_CHECK(advect(c, level, 1, fwd, PT_ADVECT_UPDATE_RESIDUAL));
                    RC_CHECK(copy_base_to_working(c));
                } else {
                    Snapshot &s = snap(c, level + 1, f, merged);
                    RC_CHECK(advect(c, level, 1, fwd, PT_ADVECT_UPDATE_RESIDUAL | PT_ADVECT_IGNORE_RANDOM));
                    RC_CHECK(snap_load_offsets(c, s));
                    RC_CHECK(snap_load_residual(c, s));
                    RC_CHECK(upsample(c, level));
                    RC_CHECK(sync_buffers(c, level, 0));
                    RC_CHECK(merge_residual(c, level, fwd));
                }
                RC_CHECK(sync_buffers(c, level, 0));
                RC_CHECK(propagate_and_scatter(c, level, iterations));
                RC_CHECK(snap_save(c, level, snap(c, level, f, pass)));
                last[f] = pass;
            }
        }
        /* pass 2: merge the forward and backward passes */
        for (int f = 0; f < F; f++) {
            RC_CHECK(snap_load_offsets(c, snap(c, level, f, 0)));
            RC_CHECK(snap_load_residual(c, snap(c, level, f, 0)));
            RC_CHECK(copy_base_to_advected_f(c));
            RC_CHECK(sync_buffers(c, level, PT_DONT_COPY_OFFSETS | PT_DONT_UPDATE_RESIDUAL));
            RC_CHECK(snap_load_offsets(c, snap(c, level, f, 1)));
            RC_CHECK(snap_load_residual(c, snap(c, level, f, 1)));
            RC_CHECK(sync_buffers(c, level, PT_DONT_COPY_OFFSETS | PT_DONT_UPDATE_RESIDUAL));
            RC_CHECK(merge_randomly(c, level, 1, 0.5f, 1));
            RC_CHECK(snap_save(c, level, snap(c, level, f, 2)));
        }
        fparams.direction = PT_BIDIRECTIONAL;
        RC_CHECK(pt_upload_parameters(c, &fparams));
        const bool temporal = F > 1 && (params->time_derivative_input_weight > 0.f ||
                                        params->time_derivative_output_weight > 0.f ||
                                        params->temporal_coherence_weight > 0.f);
        for (int pass = 3; pass < fixup + 3; pass++) {
            for (int f = 0; f < F; f++) {
                RC_CHECK(random_pyramid(c, level, f, fp[f].color, fp[f].orient, sp.out, sp.orient));
                /* synthesizeInterpolated(level, pass, pass-1, 1, prior, next) */
                fparams.direction = PT_BIDIRECTIONAL;
                fparams.advection_jitter = (f % 3 == 0) ? params->advection_jitter : 0.f;
                fparams.frame_number = f;
                RC_CHECK(pt_upload_parameters(c, &fparams));
                RC_CHECK(reset_advected(c));
                if (temporal) {
                    if (f > 0) {
                        RC_CHECK(load_images(c, level, fp[f - 1]));
                        RC_CHECK(load_anim_images(c, level, fp[f - 1]));
                        RC_CHECK(pt_flip_anim_buffers(c));
                        RC_CHECK(load_anim_images(c, level, fp[f]));
                        RC_CHECK(snap_load_offsets(c, snap(c, level, f - 1, last[f - 1])));
                        RC_CHECK(advect(c, level, 1, 1, 0));
                    }
                    if (f < F - 1) {
                        RC_CHECK(load_images(c, level, fp[f + 1]));
                        RC_CHECK(load_anim_images(c, level, fp[f + 1]));
                        RC_CHECK(pt_flip_anim_buffers(c));
                        RC_CHECK(load_anim_images(c, level, fp[f]));
                        RC_CHECK(snap_load_offsets(c, snap(c, level, f + 1, last[f + 1])));
                        RC_CHECK(advect(c, level, 1, 0, 0));
                    }
                } else {
                    RC_CHECK(load_anim_images(c, level, fp[f]));
                }
                RC_CHECK(load_images(c, level, fp[f]));
                RC_CHECK(snap_load_offsets(c, snap(c, level, f, pass - 1)));
                RC_CHECK(sync_buffers(c, level, PT_DONT_COPY_OFFSETS | PT_DONT_UPDATE_RESIDUAL));
                RC_CHECK(propagate_and_scatter(c, level, iterations));
                RC_CHECK(snap_save(c, level, snap(c, level, f, pass)));
                last[f] = pass;
            }
        }
    }
    return PT_OK;
}

/* WorkingBuffers::loadOffsets of a pass that may never have been saved: the
 * reference reads a temporary file and, when it is missing, leaves the
 * buffers as they are (synthesisFrame.cpp:769-772 ignores the failure). */
static int snap_load_offsets_if_saved(pt_ctx *c, const Snapshot &sn)
{
    return sn.saved ? snap_load_offsets(c, sn) : PT_OK;
}

/* renderAnimationTCTF (synthesisProcessor.cpp:1151-1312), direct execution,
 * without keyframes or ribbons (so the largest time step is 1): a
 * consecutive forward pass on the coarsest level, then on every level from
 * the coarsest one refine pass over all frames (from the coarser level's
 * last pass, upsampled, or from this level's previous pass), followed by
 * fixup+1 more when there are more than two frames.  Pass indices run on
 * across levels.  refine (synthesisFrame.cpp:409-457) advects both
 * neighbours into the frame from their last saved pass at this level; a
 * neighbour that has not saved one here yet (its record still 0) loads
 * nothing, as the reference's missing temporary file does. */
static int run_schedule_tctf(pt_ctx *c, const pt_params *params, int iterations, int fixup, int F,
                             const FramePyr *fp, const StylePyr &sp)
{
    int levels = params->levels, L = levels - 1;
    pt_params fparams = *params;
    std::vector<int> last((size_t)F * levels, 0); /* [f * levels + level] */
    auto LAST = [&](int f, int l) -> int & { return last[(size_t)f * levels + l]; };
    const bool temporal = F > 1 && (params->time_derivative_input_weight > 0.f ||
                                    params->time_derivative_output_weight > 0.f ||
                                    params->temporal_coherence_weight > 0.f);
    RC_CHECK(reset_advected(c));
    int passes_this = 0, passes_last = 0, last_level = -1;
    auto enter = [&](int level) -> int {
        if (level != last_level) {
            RC_CHECK(load_style_level(c, level, sp));
            passes_last = passes_this;
            last_level = level;
        }
        return PT_OK;
    };
    /* renderConsecutivePass(L, 0, -1, 1): forward, firstPass */
    RC_CHECK(enter(L));
    for (int f = 0; f < F; f++) {
        bool is_first = f == 0;
        if (!is_first) RC_CHECK(pt_flip_anim_buffers(c));
        RC_CHECK(load_images(c, L, fp[f]));
        RC_CHECK(load_anim_images(c, L, fp[f]));
        fparams.direction = PT_FORWARD;
        fparams.advection_jitter = (f % 3 == 0) ? params->advection_jitter : 0.f;
        fparams.frame_number = f;
        RC_CHECK(pt_upload_parameters(c, &fparams));
        RC_CHECK(random_pyramid(c, L, f, fp[f].color, fp[f].orient, sp.out, sp.orient));
        if (is_first) {
            RC_CHECK(randomize_offsets(c, L));
        } else {
            RC_CHECK(advect(c, L, 1, 1, PT_ADVECT_UPDATE_RESIDUAL));
            RC_CHECK(copy_base_to_working(c));
        }
        RC_CHECK(sync_buffers(c, L, 0));
        RC_CHECK(propagate_and_scatter(c, L, iterations));
        RC_CHECK(snap_save(c, L, snap(c, L, f, 0)));
        LAST(f, L) = 0;
    }
    passes_this++;
    /* refine passes, coarse to fine */
    for (int level = L; level >= 0; level--) {
        int count = 1 + (F > 2 ? fixup + 1 : 0);
        for (int r = 0; r < count; r++) {
            RC_CHECK(enter(level));
            int this_pass = passes_this, last_coarse = passes_last - 1;
            bool spatial = last_coarse == this_pass - 1;
            for (int f = 0; f < F; f++) {
                fparams.direction = PT_BIDIRECTIONAL;
                fparams.advection_jitter = (f % 3 == 0) ? params->advection_jitter : 0.f;
                fparams.frame_number = f;
                RC_CHECK(pt_upload_parameters(c, &fparams));
                /* advectNeighbors(level, this_pass - 1, 1, prior, next) */
                RC_CHECK(reset_advected(c));
                if (temporal && (f > 0 || f < F - 1)) {
                    for (int side = 0; side < 2; side++) {
                        int g = side == 0 ? f - 1 : f + 1;
                        if (g < 0 || g >= F || this_pass - 1 < 0) continue;
                        RC_CHECK(load_images(c, level, fp[g]));
                        RC_CHECK(load_anim_images(c, level, fp[g]));
                        RC_CHECK(pt_flip_anim_buffers(c));
                        RC_CHECK(load_anim_images(c, level, fp[f]));
                        RC_CHECK(snap_load_offsets_if_saved(c, snap(c, level, g, LAST(g, level))));
                        RC_CHECK(advect(c, level, 1, side == 0, 0));
                    }
                } else {
                    RC_CHECK(load_anim_images(c, level, fp[f]));
                }
                if (spatial) { /* spatialUpsample(level+1, level, ...) */
                    RC_CHECK(snap_load_offsets_if_saved(c, snap(c, level + 1, f, LAST(f, level + 1))));
                    RC_CHECK(upsample(c, level));
                } else {
                    RC_CHECK(snap_load_offsets_if_saved(c, snap(c, level, f, LAST(f, level))));
                }
                RC_CHECK(sync_buffers(c, level, PT_DONT_COPY_OFFSETS | PT_DONT_UPDATE_RESIDUAL));
                RC_CHECK(load_images(c, level, fp[f]));
                RC_CHECK(propagate_and_scatter(c, level, iterations));
                RC_CHECK(snap_save(c, level, snap(c, level, f, this_pass)));
                LAST(f, level) = this_pass;
            }
            passes_this++;
        }
    }
    return PT_OK;
}

/* renderAnimationIND (synthesisProcessor.cpp:806-839): every frame on its
 * own, levels coarse to fine inside the frame loop, each pass through
 * synthesizeConsecutive with is_first_frame and forward time
 * (synthesisFrame.cpp:214-323).  The image layers are never flipped, so a
 * later pass advects from the other (never loaded) layer, as in the
 * reference; the merge is mergeUsingResidual (the scheme is not FMBM_RM). */
static int run_schedule_ind(pt_ctx *c, const pt_params *params, int iterations, int passes, int F,
                            const FramePyr *fp, const StylePyr &sp)
{
    int levels = params->levels;
    pt_params fparams = *params;
    for (int f = 0; f < F; f++) {
        for (int level = levels - 1; level >= 0; level--) {
            RC_CHECK(load_style_level(c, level, sp));
            RC_CHECK(reset_advected(c));
            for (int pass = 0; pass < passes; pass++) {
                RC_CHECK(load_frame_level(c, level, fp[f]));
                fparams.direction = PT_FORWARD;
                fparams.advection_jitter = (f % 3 == 0) ? params->advection_jitter : 0.f;
                fparams.frame_number = f;
                RC_CHECK(pt_upload_parameters(c, &fparams));
                if (pass == 0) {
                    RC_CHECK(random_pyramid(c, level, f, fp[f].color, fp[f].orient, sp.out, sp.orient));
                    if (level == levels - 1) {
                        RC_CHECK(randomize_offsets(c, level));
                    } else {
                        RC_CHECK(snap_load_offsets(c, snap(c, level + 1, f, 0)));
                        RC_CHECK(upsample(c, level));
                        RC_CHECK(sync_buffers(c, level, 0));
                    }
                } else {
                    Snapshot &s = snap(c, level, f, pass - 1);
                    RC_CHECK(advect(c, level, 1, 1, PT_ADVECT_UPDATE_RESIDUAL | PT_ADVECT_IGNORE_RANDOM));
                    RC_CHECK(snap_load_offsets(c, s));
                    RC_CHECK(snap_load_residual(c, s));
                    RC_CHECK(sync_buffers(c, level, PT_DONT_COPY_OFFSETS | PT_DONT_UPDATE_RESIDUAL));
                    RC_CHECK(merge_residual(c, level, 1));
                }
                RC_CHECK(sync_buffers(c, level, 0));
                RC_CHECK(propagate_and_scatter(c, level, iterations));
                RC_CHECK(snap_save(c, level, snap(c, level, f, pass)));
            }
        }
    }
    return PT_OK;
}

/* The scheme and the schedule settings it implements: a backward first pass
 * with FMBM, FMBM_RM and FBMO, maxLevel above 0 with FMBM and FMBM_RM. */
static int check_scheme(pt_ctx *c, const pt_params *params, int scheme)
{
    if (scheme < PT_S_FMBM || scheme > PT_S_IND) return set_err(c, PT_ERR_UNSUPPORTED, "scheme %d is not supported", scheme);
    if (scheme == PT_S_TCTF && params->first_pass_direction != 0)
        return set_err(c, PT_ERR_UNSUPPORTED, "TCTF with a backward first pass is not supported");
    if (params->max_level != 0 && scheme != PT_S_FMBM && scheme != PT_S_FMBM_RM)
        return set_err(c, PT_ERR_UNSUPPORTED, "max_level above 0 is implemented for FMBM and FMBM_RM only");
    if (params->interpolate_key_frame)
        return set_err(c, PT_ERR_UNSUPPORTED, "keyframe schedules run through the operator entry points only");
    return PT_OK;
}

static int run_schedule(pt_ctx *c, const pt_params *params, int scheme, int iterations, int fixup_passes, int F,
                        const FramePyr *fp, const StylePyr &sp)
{
    int levels = params->levels;
    int passes = scheme_passes(scheme, fixup_passes, levels, F);
    RC_CHECK(pt_setup(c, params));
    RC_CHECK(snap_alloc(c, levels, F, passes));
    for (auto &sn : c->snaps) sn.saved = false;
    if (scheme == PT_S_TCTF) return run_schedule_tctf(c, params, iterations, fixup_passes, F, fp, sp);
    if (scheme == PT_S_IND) return run_schedule_ind(c, params, iterations, passes, F, fp, sp);
    if (scheme == PT_S_FBMO) return run_schedule_fbmo(c, params, iterations, fixup_passes, F, fp, sp);
    pt_params fparams = *params;
    const int final_level = params->max_level;
    for (int level = levels - 1; level >= final_level; level--) {
        /* Style::load (style.cpp:62-140) */
        RC_CHECK(pt_initialize_exemplars(c, c->SW >> level, c->SH >> level, 1));
        RC_CHECK(upload_image(c, level, PT_LAYER_EXEMPLAR_BASE, sp.base, 0));
        RC_CHECK(upload_image(c, level, PT_LAYER_EXEMPLAR_OUTPUT, sp.out, 0));
        RC_CHECK(upload_image(c, level, PT_LAYER_EXEMPLAR_ORIENTATION, sp.orient, 0));
        RC_CHECK(upload_image(c, level, PT_LAYER_EXEMPLAR_DIST_TRANS, sp.dt, 0));
        if (sp.id) RC_CHECK(upload_image(c, level, PT_LAYER_EXEMPLAR_OBJECT_IDS, sp.id, 0));
        RC_CHECK(reset_advected(c));
        for (int pass = 0; pass < passes; pass++) {
            /* renderConsecutivePass (synthesisProcessor.cpp:703-753) */
            int fwd = (pass % 2 == params->first_pass_direction);
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
                if (fp[f].id) RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_ID, fp[f].id, 0));
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
    /* renderAnimationFMBM's tail (synthesisProcessor.cpp:863-870) with
     * renderSpatialUpsample / SynthesisFrame::spatialUpsample
     * (synthesisFrame.cpp:465-481): from maxLevel down, every frame's last
     * saved offsets are upsampled one level, their canvas synced and the
     * result saved as pass 0 of the finer level.  No frame images are loaded,
     * so the upsample reads the image orientation and scale textures as the
     * last synthesized frame left them. */
    for (int level = final_level; level > 0; level--) {
        RC_CHECK(pt_initialize_exemplars(c, c->SW >> (level - 1), c->SH >> (level - 1), 1));
        RC_CHECK(upload_image(c, level - 1, PT_LAYER_EXEMPLAR_BASE, sp.base, 0));
        RC_CHECK(upload_image(c, level - 1, PT_LAYER_EXEMPLAR_OUTPUT, sp.out, 0));
        RC_CHECK(upload_image(c, level - 1, PT_LAYER_EXEMPLAR_ORIENTATION, sp.orient, 0));
        RC_CHECK(upload_image(c, level - 1, PT_LAYER_EXEMPLAR_DIST_TRANS, sp.dt, 0));
        if (sp.id) RC_CHECK(upload_image(c, level - 1, PT_LAYER_EXEMPLAR_OBJECT_IDS, sp.id, 0));
        for (int f = 0; f < F; f++) {
            RC_CHECK(snap_load_offsets(c, snap(c, level, f, level == final_level ? passes - 1 : 0)));
            RC_CHECK(upsample(c, level - 1));
            RC_CHECK(sync_buffers(c, level - 1, PT_DONT_COPY_OFFSETS | PT_DONT_UPDATE_RESIDUAL));
            RC_CHECK(snap_save(c, level - 1, snap(c, level - 1, f, 0)));
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
    pt_pyramid_destroy(f.id);
    f = FramePyr();
}

static void free_style(StylePyr &sp)
{
    pt_pyramid_destroy(sp.base);
    pt_pyramid_destroy(sp.out);
    pt_pyramid_destroy(sp.orient);
    pt_pyramid_destroy(sp.tensor);
    pt_pyramid_destroy(sp.dt);
    pt_pyramid_destroy(sp.id);
    sp = StylePyr();
}

/* Fills level 0 of `p` from host or device memory, or with a constant, and
 * rebuilds the other levels on the device (downsampleAverage). */
static int pyramid_fill(pt_ctx *c, pt_pyramid *p, const float *src, int on_device, float defv, int scale_mode)
{
    size_t n0 = (size_t)p->w * p->h * elem_channels(p->elem);
    float *dst = p->data;
    if (src) {
        hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        if (c->rec && !on_device) return set_err(c, PT_ERR_STATE, "recorded plans take device inputs only");
        RC_CHECK(issue(c, "pyramid level-0 copy",
                       [=](hipStream_t s) { return pt_copy(dst, src, n0 * sizeof(float), kind, s); }));
    } else {
        float v = fill_bits(p->elem, defv);
        RC_CHECK(issue(c, "pyramid fill", [=](hipStream_t s) { return pt_launch_fill(dst, n0, v, s); }));
    }
    return pyramid_rebuild(c, p, scale_mode);
}

static int orientation_fill(pt_ctx *c, pt_pyramid *tensor, pt_pyramid *angle, const float *src, int on_device)
{
    RC_CHECK(pyramid_fill(c, tensor, src, on_device, 0.f, PT_DOWNSAMPLE_SCALE_CONSTANT));
    for (int l = 0; l < tensor->levels; l++) {
        const float *t = tensor->data + tensor->off[l];
        float *a = angle->data + angle->off[l];
        size_t n = (size_t)tensor->lw[l] * tensor->lh[l];
        RC_CHECK(issue(c, "tensor angle", [=](hipStream_t s) { return pt_launch_tensor_angle(t, a, n, s); }));
    }
    return PT_OK;
}

static bool wants_ids(const pt_params *p) { return p->style_obj_id_weight > 0.f; }

static int ensure_pyr(pt_ctx *c, pt_pyramid **p, int elem, int w, int h)
{
    if (*p && (*p)->w == w && (*p)->h == h && (*p)->elem == elem) return PT_OK;
    pt_pyramid_destroy(*p);
    *p = nullptr;
    return pyramid_alloc(c, p, elem, w, h);
}

/* Object-ID pyramids are kept only while the term is on (`ids`); an absent
 * ID image is all 0 (NLDataAccess::getInputElement). */
static int prepare_ids(pt_ctx *c, int w, int h, const int32_t *src, int on_device, bool ids, pt_pyramid **p)
{
    if (!ids) {
        pt_pyramid_destroy(*p);
        *p = nullptr;
        return PT_OK;
    }
    RC_CHECK(ensure_pyr(c, p, PT_ELEM_I32, w, h));
    return pyramid_fill(c, *p, (const float *)src, on_device, 0.f, PT_DOWNSAMPLE_SCALE_CONSTANT);
}

static int prepare_frame(pt_ctx *c, int w, int h, const pt_frame_guides &g, int on_device, FramePyr &fp, bool ids)
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
    return prepare_ids(c, w, h, g.object_id, on_device, ids, &fp.id);
}

static int prepare_style(pt_ctx *c, int w, int h, const pt_style_guides &g, int on_device, StylePyr &sp, bool ids)
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
    return prepare_ids(c, w, h, g.object_ids, on_device, ids, &sp.id);
}

/* One level-0 snapshot to the caller's buffers (each may be NULL). */
static int read_snapshot(pt_ctx *c, const Snapshot &s, pt_patchxf *out_off, float *out_canvas, float *out_res,
                         int on_device)
{
    hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    size_t n = (size_t)s.w * s.h;
    if (out_off) {
        if (!on_device && (!c->pack_tmp || c->pack_cap < n)) {
            if (c->rec) return set_err(c, PT_ERR_STATE, "recorded plans write device outputs only");
            HIP_CHECK(c, hipStreamSynchronize(c->stream));
            dfree(c->pack_tmp);
            RC_CHECK(dalloc(c, &c->pack_tmp, n));
            c->pack_cap = n;
        }
        PtNnf nn{s.xyth, s.sc, c->lum_live ? s.lum : nullptr, nullptr};
        int w = s.w, h = s.h;
        pt_patchxf *tmp = c->pack_tmp, *dst = out_off;
        if (on_device) {
            /* pack straight into the caller's device buffer */
            RC_CHECK(issue(c, "offset pack", [=](hipStream_t st) { return pt_launch_pack_nnf(nn, w, w, h, dst, st); }));
        } else {
            RC_CHECK(issue(c, "offset pack", [=](hipStream_t st) {
                hipError_t e = pt_launch_pack_nnf(nn, w, w, h, tmp, st);
                if (e != hipSuccess) return e;
                return pt_copy(dst, tmp, n * sizeof(pt_patchxf), kind, st);
            }));
        }
    }
    if (out_canvas) {
        float *dst = out_canvas;
        const float4 *src = s.canvas;
        RC_CHECK(issue(c, "canvas readback",
                       [=](hipStream_t st) { return pt_copy(dst, src, n * sizeof(float4), kind, st); }));
    }
    if (out_res) {
        float *dst = out_res;
        const float *src = s.res;
        RC_CHECK(issue(c, "residual readback",
                       [=](hipStream_t st) { return pt_copy(dst, src, n * sizeof(float), kind, st); }));
    }
    return PT_OK;
}

static int read_results(pt_ctx *c, int F, int last_pass, pt_patchxf *out_off, float *out_canvas, float *out_res,
                        int on_device)
{
    for (int f = 0; f < F; f++) {
        Snapshot &s = snap(c, 0, f, last_pass);
        size_t n = (size_t)s.w * s.h;
        RC_CHECK(read_snapshot(c, s, out_off ? out_off + f * n : nullptr, out_canvas ? out_canvas + f * n * 4 : nullptr,
                               out_res ? out_res + f * n : nullptr, on_device));
    }
    return PT_OK;
}

/* A one-shot synthesis starts from freshly allocated (zeroed) working
 * buffers, as a new WorkingBuffers does (workingBuffers.cpp:48-95), so
 * repeated calls on one context give identical results. */
static int reset_working(pt_ctx *c)
{
    size_t n = (size_t)c->W * c->H;
    RC_CHECK(zero_fill(c, c->base.xyth, n * sizeof(float4)));
    RC_CHECK(zero_fill(c, c->base.sc, n * sizeof(float2)));
    RC_CHECK(zero_fill(c, c->work.xyth, n * sizeof(float4)));
    RC_CHECK(zero_fill(c, c->work.sc, n * sizeof(float2)));
    RC_CHECK(zero_fill(c, c->base.lum, n * sizeof(float)));
    RC_CHECK(zero_fill(c, c->work.lum, n * sizeof(float)));
    c->lum_live = false;
    RC_CHECK(zero_fill(c, c->base.la, n * sizeof(int32_t)));
    RC_CHECK(zero_fill(c, c->work.la, n * sizeof(int32_t)));
    c->layered_live = false;
    RC_CHECK(zero_fill(c, c->res_base, n * sizeof(float)));
    RC_CHECK(zero_fill(c, c->res_work, n * sizeof(float)));
    RC_CHECK(zero_fill(c, c->canvas, n * sizeof(float4)));
    for (int k = 0; k < 2; k++) {
        RC_CHECK(zero_fill(c, c->color[k], n * sizeof(float4)));
        RC_CHECK(zero_fill(c, c->dt[k], n * sizeof(float)));
        RC_CHECK(zero_fill(c, c->velb[k], n * sizeof(float2)));
        RC_CHECK(zero_fill(c, c->velf[k], n * sizeof(float2)));
    }
    RC_CHECK(reset_advected(c));
    c->off_diff = Rect();
    c->res_diff = Rect();
    c->cur = 0;
    c->scatter_calls = 0;
    res_invalidate(c);
    return PT_OK;
}

static int check_synthesis(pt_ctx *c, const pt_params *params, int W, int H, int SW, int SH)
{
    if (!params) return set_err(c, PT_ERR_INVALID_ARG, "NULL parameters");
    if (W < 2 || H < 2 || SW < 2 || SH < 2) return set_err(c, PT_ERR_INVALID_ARG, "bad dimensions");
    if ((int64_t)W * H > PT_MAX_IMAGE_PIXELS)
        return set_err(c, PT_ERR_UNSUPPORTED, "image larger than PT_MAX_IMAGE_PIXELS");
    if ((int64_t)SW * SH > PT_MAX_EXEMPLAR_TEXELS)
        return set_err(c, PT_ERR_UNSUPPORTED, "exemplar larger than PT_MAX_EXEMPLAR_TEXELS");
    int img_lv = (int)fmin(floor(log2f((float)W)), floor(log2f((float)H)));
    int sty_lv = (int)fmin(floor(log2f((float)SW)), floor(log2f((float)SH)));
    if (params->levels < 1 || params->levels > img_lv || params->levels > sty_lv)
        return set_err(c, PT_ERR_INVALID_ARG, "levels exceeds the pyramid depth");
    if (params->interpolate_key_frame)
        return set_err(c, PT_ERR_UNSUPPORTED, "keyframe schedules run through the operator entry points only");
    return check_supported(c, params);
}

/* Buffers of a one-shot synthesis, allocated before any work is issued or
 * recorded (a recorded plan must not see its buffers move). */
static int synthesis_alloc(pt_ctx *c, const pt_params *params, int W, int H, int F, int passes, int SW, int SH)
{
    RC_CHECK(pt_initialize(c, W, H, SW, SH));
    RC_CHECK(pt_initialize_exemplars(c, SW, SH, 1));
    RC_CHECK(snap_alloc(c, params->levels, F, passes));
    return PT_OK;
}

static int synthesize_issue(pt_ctx *c, const pt_params *params, int scheme, int iterations, int fixup_passes,
                            int W, int H, int F, const pt_frame_guides *frames, int SW, int SH,
                            const pt_style_guides *style, int guides_on_device, pt_patchxf *out_off,
                            float *out_canvas, float *out_res, int outputs_on_device)
{
    RC_CHECK(reset_working(c));
    RC_CHECK(prepare_style(c, SW, SH, *style, guides_on_device, c->style_pyr, wants_ids(params)));
    if ((int)c->frame_pyrs.size() > F) {
        for (size_t i = F; i < c->frame_pyrs.size(); i++) free_frame(c->frame_pyrs[i]);
    }
    c->frame_pyrs.resize(F);
    for (int f = 0; f < F; f++) RC_CHECK(prepare_frame(c, W, H, frames[f], guides_on_device, c->frame_pyrs[f], wants_ids(params)));
    RC_CHECK(run_schedule(c, params, scheme, iterations, fixup_passes, F, c->frame_pyrs.data(), c->style_pyr));
    /* copyOutputToFinal(final_pass): pass 0 when maxLevel > 0 (synthesisProcessor.cpp:872) */
    int final_pass = params->max_level > 0 ? 0 : scheme_passes(scheme, fixup_passes, params->levels, F) - 1;
    return read_results(c, F, final_pass, out_off, out_canvas, out_res,
                        outputs_on_device);
}

static int synthesize(pt_ctx *c, const pt_params *params, int scheme, int iterations, int fixup_passes, int W, int H,
                      int F, const pt_frame_guides *frames, int SW, int SH, const pt_style_guides *style,
                      int guides_on_device, pt_patchxf *out_off, float *out_canvas, float *out_res,
                      int outputs_on_device)
{
    if (!params || !frames || !style || F < 1 || iterations < 1 || fixup_passes < 0)
        return set_err(c, PT_ERR_INVALID_ARG, "bad synthesize arguments");
    RC_CHECK(check_scheme(c, params, scheme));
    RC_CHECK(check_synthesis(c, params, W, H, SW, SH));
    HIP_CHECK(c, hipSetDevice(c->device));
    RC_CHECK(synthesis_alloc(c, params, W, H, F, scheme_passes(scheme, fixup_passes, params->levels, F), SW, SH));
    RC_CHECK(synthesize_issue(c, params, scheme, iterations, fixup_passes, W, H, F, frames, SW, SH, style,
                              guides_on_device, out_off, out_canvas, out_res, outputs_on_device));
    return sync_stream(c);
}

/* Runs `body` with the context recording into *ops; the context's own
 * stream is drained first so that allocation-time clears are complete
 * before any executor stream uses the buffers. */
template <class F> static int record_into(pt_ctx *c, std::vector<RecOp> *ops, F body)
{
    if (!c || !ops) return set_err(c, PT_ERR_INVALID_ARG, "NULL argument");
    if (c->rec) return set_err(c, PT_ERR_STATE, "already recording");
    c->rec = ops;
    int rc = body();
    c->rec = nullptr;
    if (rc != PT_OK) return rc;
    HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return PT_OK;
}

int pt_record_synthesis(pt_ctx *c, std::vector<RecOp> *ops, const pt_params *params, int scheme, int iterations,
                        int fixup_passes, int W, int H, int F, const pt_frame_guides *frames, int SW, int SH,
                        const pt_style_guides *style, pt_patchxf *out_off, float *out_canvas, float *out_res)
{
    if (!c) return set_err(nullptr, PT_ERR_INVALID_ARG, "NULL context");
    if (!params || !frames || !style || F < 1 || iterations < 1 || fixup_passes < 0)
        return set_err(c, PT_ERR_INVALID_ARG, "bad synthesize arguments");
    RC_CHECK(check_scheme(c, params, scheme));
    RC_CHECK(check_synthesis(c, params, W, H, SW, SH));
    HIP_CHECK(c, hipSetDevice(c->device));
    RC_CHECK(synthesis_alloc(c, params, W, H, F, scheme_passes(scheme, fixup_passes, params->levels, F), SW, SH));
    /* the frame and style pyramids are allocated here too, by a dry pass of
     * their preparation with recording into a scratch list */
    {
        std::vector<RecOp> scratch;
        RC_CHECK(record_into(c, &scratch, [&] {
            RC_CHECK(prepare_style(c, SW, SH, *style, 1, c->style_pyr, wants_ids(params)));
            if ((int)c->frame_pyrs.size() > F)
                for (size_t i = F; i < c->frame_pyrs.size(); i++) free_frame(c->frame_pyrs[i]);
            c->frame_pyrs.resize(F);
            for (int f = 0; f < F; f++) RC_CHECK(prepare_frame(c, W, H, frames[f], 1, c->frame_pyrs[f], wants_ids(params)));
            return PT_OK;
        }));
    }
    return record_into(c, ops, [&] {
        return synthesize_issue(c, params, scheme, iterations, fixup_passes, W, H, F, frames, SW, SH, style, 1,
                                out_off, out_canvas, out_res, 1);
    });
}

hipError_t pt_run_op(int op, const PtArgs &a, hipStream_t s)
{
    switch (op) {
    case PT_OP_PROPAGATE: return pt_launch_propagate(a, a.jump, s);
    case PT_OP_SCATTER: return pt_launch_scatter(a, s);
    case PT_OP_RESIDUAL: return pt_launch_update_residual(a, s);
    case PT_OP_SYNC_CANVAS:
        if (a.dw > 0 && a.dh > 0) return pt_launch_sync_canvas(a, a.dw, a.dh, s);
        return pt_launch_canvas(a, s);
    default: return hipErrorInvalidValue;
    }
}

int pt_ctx_device(const pt_ctx *c) { return c ? c->device : -1; }

int pt_ctx_fail(pt_ctx *c, int code, const char *msg) { return set_err(c, code, "%s", msg); }

/* ------------------------------------------------------------------ */
/* per-frame sequence schedule (streaming)                             */
/* ------------------------------------------------------------------ */
static void free_seq(pt_ctx *c)
{
    for (int k = 0; k < 2; k++) {
        for (auto &s : c->seq_snap[k]) {
            dfree(s.xyth);
            dfree(s.sc);
            dfree(s.lum);
            dfree(s.res);
            dfree(s.canvas);
        }
        c->seq_snap[k].clear();
        free_frame(c->seq_pyr[k]);
    }
    c->seq_levels = 0;
}

static int seq_alloc(pt_ctx *c, int levels)
{
    if (c->seq_levels == levels && !c->seq_snap[0].empty() && c->seq_snap[0][0].w == c->W &&
        c->seq_snap[0][0].h == c->H)
        return PT_OK;
    for (int k = 0; k < 2; k++) {
        for (auto &s : c->seq_snap[k]) {
            dfree(s.xyth);
            dfree(s.sc);
            dfree(s.lum);
            dfree(s.res);
            dfree(s.canvas);
        }
        c->seq_snap[k].assign(levels, Snapshot());
        for (int l = 0; l < levels; l++) {
            Snapshot &s = c->seq_snap[k][l];
            s.w = c->W >> l;
            s.h = c->H >> l;
            size_t n = (size_t)s.w * s.h;
            RC_CHECK(dalloc(c, &s.xyth, n));
            RC_CHECK(dalloc(c, &s.sc, n));
            RC_CHECK(dalloc(c, &s.lum, n));
            RC_CHECK(dalloc(c, &s.res, n));
            RC_CHECK(dalloc(c, &s.canvas, n));
        }
    }
    c->seq_levels = levels;
    return PT_OK;
}

/*
 * One new frame of a sequence, all levels, before the next frame arrives
 * (frame-major order).  Per level it runs the forward first pass of
 * SynthesisFrame::synthesizeConsecutive (synthesisFrame.cpp:214-323) with
 * FMBM_RM: the coarsest level advects the previous frame's result and keeps
 * it; finer levels advect, upsample this frame's coarser result and merge
 * the two at random (mergeRandomly, weight 0.5), then propagateAndScatter.
 * The state the level-major reference schedule would hold at (frame, level)
 * is restored from resident per-level snapshots: the previous frame's
 * offsets and residual at this level, and its images in the other texture
 * layer (flipAnimBuffers).  Level-0-sized buffers keep what earlier levels
 * and frames of this sequence left outside the active rectangle, so a frame
 * depends only on its sequence's own history.
 */
static int next_frame_schedule(pt_ctx *c, const pt_params *params, int iterations, const FramePyr &cur,
                               const FramePyr *prev, const StylePyr &sp)
{
    const int levels = params->levels;
    const int f = c->seq_frame;
    std::vector<Snapshot> &now = c->seq_snap[f & 1];
    std::vector<Snapshot> &before = c->seq_snap[(f + 1) & 1];
    RC_CHECK(pt_setup(c, params));
    pt_params fp = *params;
    fp.direction = PT_FORWARD;
    fp.advection_jitter = (f % 3 == 0) ? params->advection_jitter : 0.f;
    fp.frame_number = f;
    RC_CHECK(pt_upload_parameters(c, &fp));
    for (int level = levels - 1; level >= 0; level--) {
        RC_CHECK(pt_initialize_exemplars(c, c->SW >> level, c->SH >> level, 1));
        RC_CHECK(upload_image(c, level, PT_LAYER_EXEMPLAR_BASE, sp.base, 0));
        RC_CHECK(upload_image(c, level, PT_LAYER_EXEMPLAR_OUTPUT, sp.out, 0));
        RC_CHECK(upload_image(c, level, PT_LAYER_EXEMPLAR_ORIENTATION, sp.orient, 0));
        RC_CHECK(upload_image(c, level, PT_LAYER_EXEMPLAR_DIST_TRANS, sp.dt, 0));
        if (sp.id) RC_CHECK(upload_image(c, level, PT_LAYER_EXEMPLAR_OBJECT_IDS, sp.id, 0));
        RC_CHECK(reset_advected(c));
        if (f > 0) {
            c->cur = (f + 1) % 2;
            RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_COLOR, prev->color, 0));
            RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_VEL_B, prev->velb, 0));
            RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_VEL_F, prev->velf, 0));
            RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_DIST_TRANS, prev->dt, 0));
            RC_CHECK(snap_load_offsets(c, before[level]));
            RC_CHECK(snap_load_residual(c, before[level]));
        }
        c->cur = f % 2;
        RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_COLOR, cur.color, 0));
        RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_ORIENTATION, cur.orient, 0));
        if (cur.id) RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_ID, cur.id, 0));
        RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_VEL_B, cur.velb, 0));
        RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_VEL_F, cur.velf, 0));
        RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_DIST_TRANS, cur.dt, 0));
        RC_CHECK(upload_image(c, level, PT_LAYER_INPUT_SCALE, cur.scale, 0));
        RC_CHECK(random_pyramid(c, level, f, cur.color, cur.orient, sp.out, sp.orient));
        if (f == 0) {
            if (level == levels - 1) {
                RC_CHECK(randomize_offsets(c, level));
            } else {
                RC_CHECK(snap_load_offsets(c, now[level + 1]));
                RC_CHECK(upsample(c, level));
                RC_CHECK(sync_buffers(c, level, 0));
            }
        } else if (level == levels - 1) {
            RC_CHECK(advect(c, level, 1, 1, 0));
            RC_CHECK(copy_base_to_working(c));
        } else {
            RC_CHECK(advect(c, level, 1, 1, PT_ADVECT_IGNORE_RANDOM));
            RC_CHECK(snap_load_offsets(c, now[level + 1]));
            RC_CHECK(snap_load_residual(c, now[level + 1]));
            RC_CHECK(upsample(c, level));
            RC_CHECK(sync_buffers(c, level, 0));
            RC_CHECK(merge_randomly(c, level, 1, 0.5f, 0));
        }
        RC_CHECK(sync_buffers(c, level, 0));
        RC_CHECK(propagate_and_scatter(c, level, iterations));
        RC_CHECK(snap_save(c, level, now[level]));
    }
    return PT_OK;
}

static int next_frame_issue(pt_ctx *c, const pt_params *params, int iterations, int W, int H,
                            const pt_frame_guides *frame, int SW, int SH, const pt_style_guides *style,
                            pt_patchxf *out_off, float *out_canvas, float *out_res)
{
    const int f = c->seq_frame;
    if (f == 0) {
        RC_CHECK(reset_working(c));
        RC_CHECK(prepare_style(c, SW, SH, *style, 1, c->style_pyr, wants_ids(params)));
    }
    RC_CHECK(prepare_frame(c, W, H, *frame, 1, c->seq_pyr[f & 1], wants_ids(params)));
    RC_CHECK(next_frame_schedule(c, params, iterations, c->seq_pyr[f & 1], f > 0 ? &c->seq_pyr[(f + 1) & 1] : nullptr,
                                 c->style_pyr));
    /* results: this frame's level-0 snapshot */
    RC_CHECK(read_snapshot(c, c->seq_snap[f & 1][0], out_off, out_canvas, out_res, 1));
    c->seq_frame = f + 1;
    return PT_OK;
}

static int next_frame_check(pt_ctx *c, const pt_params *params, int iterations, int W, int H,
                            const pt_frame_guides *frame, int SW, int SH, const pt_style_guides *style)
{
    if (!params || !frame || !style || iterations < 1) return set_err(c, PT_ERR_INVALID_ARG, "bad frame arguments");
    if (params->first_pass_direction != 0 || params->max_level != 0)
        return set_err(c, PT_ERR_UNSUPPORTED, "the per-frame schedule runs forwards and down to level 0 only");
    RC_CHECK(check_synthesis(c, params, W, H, SW, SH));
    if (c->seq_frame > 0 && (c->W != W || c->H != H || c->SW != SW || c->SH != SH || c->seq_levels != params->levels))
        return set_err(c, PT_ERR_INVALID_ARG, "frame shape differs from the sequence; call pt_sequence_reset");
    HIP_CHECK(c, hipSetDevice(c->device));
    if (c->seq_frame == 0) {
        RC_CHECK(pt_initialize(c, W, H, SW, SH));
        RC_CHECK(pt_initialize_exemplars(c, SW, SH, 1));
        RC_CHECK(seq_alloc(c, params->levels));
    }
    return PT_OK;
}

/* Allocates this frame's pyramids (and on frame 0 the style's) without
 * issuing anything. */
static int next_frame_pyramids(pt_ctx *c, const pt_params *params, int W, int H, const pt_frame_guides *frame,
                               int SW, int SH, const pt_style_guides *style)
{
    std::vector<RecOp> scratch;
    const int f = c->seq_frame;
    return record_into(c, &scratch, [&] {
        if (f == 0) RC_CHECK(prepare_style(c, SW, SH, *style, 1, c->style_pyr, wants_ids(params)));
        return prepare_frame(c, W, H, *frame, 1, c->seq_pyr[f & 1], wants_ids(params));
    });
}

int pt_record_next_frame(pt_ctx *c, std::vector<RecOp> *ops, const pt_params *params, int iterations, int W, int H,
                         const pt_frame_guides *frame, int SW, int SH, const pt_style_guides *style,
                         pt_patchxf *out_off, float *out_canvas, float *out_res)
{
    if (!c) return set_err(nullptr, PT_ERR_INVALID_ARG, "NULL context");
    RC_CHECK(next_frame_check(c, params, iterations, W, H, frame, SW, SH, style));
    RC_CHECK(next_frame_pyramids(c, params, W, H, frame, SW, SH, style));
    return record_into(c, ops, [&] {
        return next_frame_issue(c, params, iterations, W, H, frame, SW, SH, style, out_off, out_canvas, out_res);
    });
}

extern "C" int pt_synthesize_next_frame(pt_ctx *c, const pt_params *params, int iterations, int W, int H,
                                        const pt_frame_guides *frame, int SW, int SH, const pt_style_guides *style,
                                        pt_patchxf *out_off, float *out_canvas, float *out_res)
{
    if (!c) return set_err(nullptr, PT_ERR_INVALID_ARG, "NULL context");
    RC_CHECK(next_frame_check(c, params, iterations, W, H, frame, SW, SH, style));
    RC_CHECK(next_frame_issue(c, params, iterations, W, H, frame, SW, SH, style, out_off, out_canvas, out_res));
    return sync_stream(c);
}

extern "C" int pt_sequence_reset(pt_ctx *c)
{
    if (!c) return set_err(nullptr, PT_ERR_INVALID_ARG, "NULL context");
    c->seq_frame = 0;
    c->scatter_calls = 0;
    return PT_OK;
}

extern "C" int pt_sequence_frame(pt_ctx *c, int *frame)
{
    if (!c || !frame) return set_err(c, PT_ERR_INVALID_ARG, "NULL argument");
    *frame = c->seq_frame;
    return PT_OK;
}

size_t pt_sequence_bytes(int W, int H, int SW, int SH, int levels, int r)
{
    (void)levels;
    const double px = (double)W * H, ex = (double)SW * SH;
    /* the forward time table: 32 B per tap and pixel (patches up to 81 taps) */
    const double taps = (double)(2 * r + 1) * (2 * r + 1);
    const double tt = taps <= 81 ? 32.0 * taps : 0.0;
    /* working buffers 344 B/px (luminance-shift and style-layer planes and
     * the image-ID texture included); two frame pyramids 2 x 60 B/px x 4/3;
     * two snapshot sets 2 x 48 B/px x 4/3; exemplar textures 40 B/texel;
     * style pyramids 52 B/texel x 4/3; plus 10% slack */
    double b = px * (344.0 + tt + 2 * 60.0 * 4.0 / 3.0 + 2 * 48.0 * 4.0 / 3.0) + ex * (40.0 + 52.0 * 4.0 / 3.0);
    return (size_t)(b * 1.1) + ((size_t)8 << 20);
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
