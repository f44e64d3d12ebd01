/*
 * pt_plan.h -- recorded stream work ("plans") shared by the host runtime
 * (pt_runtime.hip) and the multi-job executor (pt_engine.hip).  Not part of
 * the public C ABI.
 *
 * A plan is the exact sequence of stream operations one synthesis call would
 * issue on its context, captured instead of issued.  Operators that have a
 * multi-job kernel form (propagate, scatter, residual refresh, sync/canvas)
 * are kept as data (operator id plus the launch's PtArgs) so that the
 * executor can merge the same step of many jobs into one launch; everything
 * else (copies, pyramid kernels, uploads, advection, merges) is a closure
 * that issues the work on a given stream.
 */
#ifndef PT_PLAN_H
#define PT_PLAN_H

#include <functional>
#include <vector>

#include "pt_internal.h"

struct RecOp {
    int op = -1; /* PT_OP_* with a multi-job form, or -1 for fn */
    PtArgs a{};
    std::function<hipError_t(hipStream_t)> fn;
    const char *what = nullptr; /* what fn issues (error messages) */
};

/* Launches one recorded operator as a single-job kernel with the default
 * shape (the plain path). */
hipError_t pt_run_op(int op, const PtArgs &a, hipStream_t s);

/* Records the stream work of pt_synthesize_device(ctx, ...) into *ops
 * (allocations happen now; nothing is issued).  Guides and outputs are
 * device pointers. */
int pt_record_synthesis(pt_ctx *ctx, std::vector<RecOp> *ops, const pt_params *params, int scheme,
                        int iterations, int fixup_passes, int image_w, int image_h, int num_frames,
                        const pt_frame_guides *frames, int style_w, int style_h, const pt_style_guides *style,
                        pt_patchxf *out_offsets, float *out_canvas_rgba, float *out_residual);

/* Records the stream work of pt_synthesize_next_frame(ctx, ...) into *ops
 * and advances the context's sequence state as if it had run. */
int pt_record_next_frame(pt_ctx *ctx, std::vector<RecOp> *ops, const pt_params *params, int iterations,
                         int image_w, int image_h, const pt_frame_guides *frame, int style_w, int style_h,
                         const pt_style_guides *style, pt_patchxf *out_offsets, float *out_canvas_rgba,
                         float *out_residual);

/* Device bytes a context holds for a sequence of this shape and patch
 * radius (working buffers, forward time table, pyramids, per-level
 * snapshots): the serving admission estimate. */
size_t pt_sequence_bytes(int image_w, int image_h, int style_w, int style_h, int levels, int r);

/* The context's HIP device and last error (for the executor). */
int pt_ctx_device(const pt_ctx *ctx);
int pt_ctx_fail(pt_ctx *ctx, int code, const char *msg);

#endif
