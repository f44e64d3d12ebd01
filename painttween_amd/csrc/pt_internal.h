/*
 * pt_internal.h -- types shared by the HIP kernels (pt_kernels.hip) and the
 * host runtime (pt_runtime.hip).  Not part of the public C ABI.
 */
#ifndef PT_INTERNAL_H
#define PT_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/pt_api.h"

/* Nearest-neighbour field as structure-of-arrays planes: {x, y, theta,
 * hysteresis}, {scaleU, scaleV}, the luminance shift and the style layer
 * (PatchXF, types.h:98-114). */
struct PtNnf {
    float4 *xyth;
    float2 *sc;
    float *lum;  /* luminance shift y (PatchXF::luminanceShift); read only while lum_live */
    int32_t *la; /* style layer (PatchXF::layer); read and written only while layered */
};

/* AdvectedBuffersDevice (advectedBuffers.h:37-52). */
struct PtAdv {
    float4 *xyth;
    float2 *sc;
    float4 *base;
    float4 *out;
    float *res;
    int32_t *ts;
    float *lum;
    int32_t *la;
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
    float coh_w, coh_ratio, coh_angle; /* coherence term: weight (0 = off), ratio and angle ranges */
    float coh_inv_ratio, coh_z;         /* its constants: 1 / ratio range, log(ratio range)^2 */
    double coh_lim;                     /* and angle range * pi */
    int lab;           /* colour distances in the LAB space (colorspace 1) */
    float obj_w;       /* object-ID term weight (0 = off)              */
    int lum_live;      /* luminance-shift planes maintained (some shift may be nonzero) */
    int lum_enable;    /* lumShiftEnable: advection accumulates the shift */
    int idw, idh;      /* image object-ID texture size                 */
    int samples, rng_mode;
    uint32_t key0, key1, call;
    int time_active;   /* some advected offset may be valid */
    int ee_ok;         /* all energy terms non-negative: exact early rejection is exact */
    int res_exact;     /* propagate: res_base holds the exact residual of the base offsets */
    int early_exit;    /* default policy: use early rejection on this launch */
    int unit_parallel; /* default policy: candidate units over the lanes of a workgroup */
    /* per-launch operands, so that one multi-job launch can serve jobs whose
     * operands differ */
    int jump;          /* propagate: jump distance                     */
    int nscat;         /* scatter: random samples per pixel            */
    int dw, dh;        /* sync: rectangle where working and base differ */
    int res_incremental; /* residual refresh: recompute only pixels whose
                            patch window holds a changed offset */
    int tap_max_px;    /* default policy: tap-parallel shapes on levels up to this size */
    int tap_np_res;    /* tap-parallel refresh: pixels per workgroup (pt_tap_configure)   */
    int tap_np_prop;   /* tap-parallel propagate: pixels per workgroup (pt_tap_configure) */
    int ex_window;     /* XL kernel variants: exemplar windows in LDS */
    int xcd_tiles;     /* single-job launches: contiguous tile ranges per XCD */
    int prop_window;      /* unit-parallel propagate: exemplar windows on this level ... */
    int prop_window_jump; /* ... for jumps up to this */
    int time_unit_order;  /* PT_TUNE_TEMPORAL_UNIT_ORDER */

    /* style layers (keyframe interpolation): with `layered` set every
     * offset carries its layer (PtNnf::la) and the per-pixel kernels' RT < 0
     * variant samples the exemplar layer of each candidate.  The exemplar
     * holds nl layers of ew x eh texels, layer l from texel l * ew * eh. */
    int layered, nl;
    int keys;          /* interpolateKeyFrame: the *_keys search kernels     */
    int key_first, key_last;
    int key_index;     /* propagate/scatter: keyframe layer whose painted texels are fixed (-1: none) */
    /* cumulative key-to-key fields (layered, cl layers of cw x ch, bilinear,
     * [0] = backward, [1] = forward) and the frame-to-key ribbons (fw x fh,
     * point fetches); a context without them binds one zero texel */
    const float4 *cumul[2];
    int cw, ch, cl;
    const float4 *ftk[2];
    int fw[2], fh[2];

    uint8_t *changed;  /* per pixel: base offset changed by the last sync */
    /* time tables of the advected sets (k_time_table) or null: entry
     * time_table_index(x, y, i, j) holds the exemplar base and output at tap
     * (i, j) of pixel (x, y)'s advected offset */
    const float4 *tt[2];

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
    const int *img_id; /* image object IDs, idw x idh                  */
    const int *ex_id;  /* exemplar object IDs, pitch ew                */
};

/* Operators with several kernel shapes and a multi-job form. */
enum PtOp { PT_OP_PROPAGATE = 0, PT_OP_SCATTER = 1, PT_OP_RESIDUAL = 2, PT_OP_SYNC_CANVAS = 3, PT_OP_COUNT = 4 };

/* Kernel shapes of an operator.  Every valid shape gives bit-identical
 * results; they differ only in how (pixel, candidate, tap) work maps onto
 * lanes.  Propagate/scatter: PIXEL = one lane per pixel, candidates in
 * sequence; PIXEL_EE = the same with exact early rejection; UP_* =
 * candidate units spread over a 256-lane workgroup owning a pixel tile
 * (EE_A: propagate 16x8 / scatter 16x8 with early rejection; A: 8x8 / 4x4;
 * EE_B: 8x8 / 8x8 with early rejection; B: 16x8 / 8x8).  TAP (propagate
 * and residual refresh, for small levels): every tap of every candidate is a
 * lane, the workgroup then sums each candidate's taps in the reference
 * order; it trades work per lane for latency.  Sync has the PIXEL shape
 * only. */
enum PtShape {
    PT_SHAPE_PIXEL = 0,
    PT_SHAPE_PIXEL_EE = 1,
    PT_SHAPE_UP_EE_A = 2,
    PT_SHAPE_UP_A = 3,
    PT_SHAPE_UP_EE_B = 4,
    PT_SHAPE_UP_B = 5,
    PT_SHAPE_TAP = 6, /* propagate and residual: one lane per (candidate, tap) */
    PT_SHAPE_COUNT = 7
};

/* A multi-job launch: workgroup g serves job j with prefix[j] <= g <
 * prefix[j+1], tile (g - prefix[j]) of that job in row-major order over
 * tiles_x[j] tiles per row.  All arrays live in device memory. */
struct PtJobTable {
    const PtArgs *args;
    const int *prefix;  /* n + 1 entries */
    const int *tiles_x; /* n entries     */
    int n;
    int xcd_remap;      /* spread contiguous tile ranges over the 8 XCDs */
};

/* Tap evaluations per tap-parallel workgroup: the LDS capacity of those
 * kernels.  pt_tap_configure sets a.tap_np_* for a target of `evals`
 * evaluations per workgroup (clamped to 1 .. PT_TAP_EVALS). */
#define PT_TAP_EVALS 1024
void pt_tap_configure(PtArgs &a, int evals);
int pt_default_shape(int op, const PtArgs &a);
int pt_shape_valid(int op, int shape, const PtArgs &a);
/* Tiles of job `a` for (op, shape); *tiles_x receives tiles per row. */
int pt_shape_tiles(int op, int shape, const PtArgs &a, int *tiles_x);
/* Launch key: jobs may share one multi-job launch iff their keys match. */
int pt_shape_key(int op, int shape, const PtArgs &a);
hipError_t pt_launch_multi(int op, int shape, const PtArgs &first, const PtJobTable &t, int ntiles, hipStream_t s);

/* Host-side launchers (pt_kernels.hip).  All run on `stream`. */
hipError_t pt_launch_update_residual(const PtArgs &a, hipStream_t s);
hipError_t pt_launch_canvas(const PtArgs &a, hipStream_t s);
hipError_t pt_launch_sync_canvas(const PtArgs &a, int dw, int dh, hipStream_t s);
hipError_t pt_launch_copy_random(const PtArgs &a, hipStream_t s);
hipError_t pt_launch_propagate(const PtArgs &a, int jump, hipStream_t s);
hipError_t pt_launch_scatter(const PtArgs &a, hipStream_t s);
hipError_t pt_launch_upsample(const PtArgs &a, hipStream_t s);
hipError_t pt_launch_advect(const PtArgs &a, int forwards, int time_step, int flags, hipStream_t s);
/* Time table of advected set k at a's level: (2r+1)^2 * lw * lh * 2 float4. */
hipError_t pt_launch_time_table(const PtArgs &a, int k, float4 *tt, hipStream_t s);
hipError_t pt_launch_merge_randomly(const PtArgs &a, int forwards, float weight, int force, hipStream_t s);
hipError_t pt_launch_merge_residual(const PtArgs &a, int forwards, hipStream_t s);
hipError_t pt_launch_downsample(const float *src, int pw, float *dst, int lw, int lh, int ch,
                                int scale_mode, hipStream_t s);
hipError_t pt_launch_downsample_vote(const int *src, int pw, int *dst, int lw, int lh, hipStream_t s);
hipError_t pt_launch_fill(float *dst, size_t n, float v, hipStream_t s);
hipError_t pt_launch_tensor_angle(const float *tensor, float *angle, size_t n, hipStream_t s);
hipError_t pt_launch_interleave(const float4 *src, int w, int h, float4 *ex, int ew, int slot, hipStream_t s);
hipError_t pt_launch_random_pyramid(const float4 *color, const float *iorient, int w, int h,
                                    const float4 *eout, const float *eorient, int ew, int eh,
                                    int transp, int orient_en, uint32_t key0, uint32_t key1,
                                    float4 *out, hipStream_t s);
hipError_t pt_launch_random_pyramid_keys(const float4 *color, const float *iorient, int w, int h, int nstyles,
                                         const float4 *const *eouts, const float *const *eorients, const int *ews,
                                         const int *ehs, int key_first, int num_keys, int transp, int orient_en,
                                         uint32_t key0, uint32_t key1, float4 *out, hipStream_t s);
/* keyframe interpolation: ribbons are pt_ribbonp, level-0 pitch W */
hipError_t pt_launch_merge_advected(const PtArgs &a, int fallback, hipStream_t s);
hipError_t pt_launch_offsets_from_ribbon(const PtArgs &a, int look_forwards, hipStream_t s);
hipError_t pt_launch_accumulate_ribbon(const PtArgs &a, const pt_ribbonp *src, pt_ribbonp *dst, int forwards,
                                       int time_step, int overwrite, hipStream_t s);
hipError_t pt_launch_ribbon_output(const PtArgs &a, const pt_ribbonp *rib, hipStream_t s);
hipError_t pt_launch_downsample_ribbon(const float *src, int pw, float *dst, int lw, int lh, hipStream_t s);
hipError_t pt_launch_fill_ribbon(pt_ribbonp *dst, size_t n, hipStream_t s);
hipError_t pt_launch_pack_nnf(const PtNnf &src, int W, int lw, int lh, void *dst_patchxf, hipStream_t s);
hipError_t pt_launch_unpack_nnf(const void *src_patchxf, int lw, int lh, PtNnf dst, int W, hipStream_t s);

/* Device-to-device copies of recorded stream work go through these (same
 * arguments as hipMemcpyAsync / hipMemcpy2DAsync), so that a task sink can
 * capture them. */
hipError_t pt_copy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t s);
hipError_t pt_copy2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height,
                     hipMemcpyKind kind, hipStream_t s);

/* ------------------------------------------------------------------ */
/* persistent batch launch                                             */
/* ------------------------------------------------------------------ */
/* One step of a job's device program: every workgroup tile of one kernel
 * the plain path would launch.  The first four kinds are the PtOp values. */
enum PtTaskKind {
    PT_TASK_PROPAGATE = 0,
    PT_TASK_SCATTER = 1,
    PT_TASK_RESIDUAL = 2,
    PT_TASK_SYNC = 3,
    PT_TASK_COPY_RANDOM = 4,
    PT_TASK_UPSAMPLE = 5,
    PT_TASK_COPY = 6,
    PT_TASK_FILL = 7,
    PT_TASK_DOWNSAMPLE = 8,
    PT_TASK_TENSOR_ANGLE = 9,
    PT_TASK_INTERLEAVE = 10,
    PT_TASK_RANDOM_PYRAMID = 11,
    PT_TASK_PACK_NNF = 12,
    PT_TASK_COUNT = 13
};

struct PtTask {
    int kind;    /* PtTaskKind */
    int variant; /* NNF kinds: shape << 4 | exemplar window << 3 | radius; copy: bytes per element;
                    downsample: channels */
    int ntiles;  /* workgroup tiles (> 0) */
    int tiles_x; /* tiles per row of a 2-D tiling */
    int arg;     /* index of the step's PtArgs in the argument table, or -1 */
    int i[8];    /* integer operands */
    uint32_t u[2];
    float f;
    int chunk;        /* tiles per claim in the persistent launch */
    const void *p[4]; /* source operands */
    void *q;          /* destination operand */
    int64_t n[3];     /* counts and pitches */
    /* persistent launch: the tiles are split into one contiguous band per
     * XCD the job runs on (band b: tiles band[b] .. band[b + 1]); per band
     * the job's tiles through this step, and the job's nonempty bands
     * through this step (cumulative counts) */
    int band[9];
    int band_end[8];
    int parts_end;
};

/* Task capture: while a sink is active on the calling thread, the
 * launchers of pt_kernels.hip and pt_copy / pt_copy2d append tasks (and the
 * PtArgs they take) instead of issuing work; a launcher without a task form
 * records its name in `unsupported` and fails. */
struct PtTaskSink {
    std::vector<PtTask> tasks;
    std::vector<PtArgs> args; /* args[k]: the PtArgs of a task with arg == k */
    const char *unsupported = nullptr;
};
void pt_sink_begin(PtTaskSink *s);
void pt_sink_end();

/* The task of one multi-job operator step (op, shape) of job `a`, its
 * arguments at `arg`; PT_ERR_UNSUPPORTED when the persistent kernel has no
 * body for it (the message names what). */
int pt_task_op(int op, int shape, const PtArgs &a, int arg, PtTask *t, const char **why);

/* Claim word of a job: step << 44 | claims (runs of PtTask::chunk tiles)
 * in the step << 22 | claims taken. */
__host__ __device__ inline unsigned long long pt_claim_word(int step, int nclaims)
{
    return ((unsigned long long)step << 44) | ((unsigned long long)nclaims << 22);
}
__host__ __device__ inline int pt_cw_step(unsigned long long w) { return (int)(w >> 44); }
__host__ __device__ inline int pt_cw_claims(unsigned long long w) { return (int)((w >> 22) & 0x3FFFFFull); }
__host__ __device__ inline int pt_cw_taken(unsigned long long w) { return (int)(w & 0x3FFFFFull); }
#define PT_PERSIST_MAX_CLAIMS (1 << 21) /* claims of one step (the word's field holds 22 bits) */
#define PT_PERSIST_MAX_STEPS (1 << 19)  /* steps of one job */

/* Device state of one persistent launch. */
struct PtPersist {
    const PtTask *tasks;       /* every job's program, concatenated        */
    const int *first;          /* njobs + 1: job j's steps are tasks[first[j] .. first[j+1]);
                                  then njobs more: job j's XCDs (bit x of first[njobs + 1 + j]) */
    const PtArgs *args;
    unsigned long long *claim; /* 8 per job: claim word of each band (pt_claim_word) */
    int *done;                 /* 8 per job: tiles of each band completed, all steps */
    int *parts;                /* per job: bands completed, all steps      */
    unsigned long long *cost;  /* per job: shader clocks its claims took (placement feedback) */
    int *ctl;                  /* [0] jobs finished, [1] steps completed, [2] gave up */
    int njobs;
    long long timeout_ticks;   /* wall-clock ticks without a completed step before giving up */
};
/* Workgroups of the persistent kernel that fit on the device at once. */
int pt_persistent_grid(int device);
hipError_t pt_launch_persistent(const PtPersist &p, int grid, hipStream_t s);
/* ORs 1 << (XCD id) into *mask (device memory) for every XCD that runs one
 * of `grid` workgroups. */
hipError_t pt_launch_xcd_probe(int grid, unsigned *mask, hipStream_t s);

#endif
