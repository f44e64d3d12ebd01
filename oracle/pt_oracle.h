/*
 * pt_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the PaintTween per-frame synthesis hot path, used as the
 * correctness oracle for the HIP implementation.  Only the test suite, the
 * smoke check and the benchmark's CPU-baseline leg load it; the product
 * library never links or calls it.
 *
 * The API mirrors include/pt_api.h one to one with a `ptref_` prefix and a
 * host-memory context, so a test can drive both through the same sequence of
 * calls and compare results bit for bit.
 */
#ifndef PT_ORACLE_H
#define PT_ORACLE_H

#include "../include/pt_api.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ptref_ctx ptref_ctx;
typedef struct ptref_pyramid ptref_pyramid;

void ptref_params_default(pt_params *out);
int ptref_create(ptref_ctx **out, int num_threads);
int ptref_destroy(ptref_ctx *ctx);
const char *ptref_last_error(ptref_ctx *ctx);
int ptref_setup(ptref_ctx *ctx, const pt_params *params);
int ptref_upload_parameters(ptref_ctx *ctx, const pt_params *params);
int ptref_initialize(ptref_ctx *ctx, int image_w, int image_h, int style_w, int style_h);
int ptref_initialize_exemplars(ptref_ctx *ctx, int w, int h, int num_layers);
int ptref_reset_advected_buffers(ptref_ctx *ctx);
int ptref_flip_anim_buffers(ptref_ctx *ctx);

int ptref_pyramid_create(ptref_ctx *ctx, ptref_pyramid **out, int elem, int w, int h,
                         int scale_mode, const void *host_level0, float default_value);
int ptref_pyramid_orientation(ptref_ctx *ctx, ptref_pyramid **out, const ptref_pyramid *tensor);
int ptref_pyramid_destroy(ptref_pyramid *pyr);
int ptref_pyramid_levels(const ptref_pyramid *pyr);
int ptref_pyramid_dims(const ptref_pyramid *pyr, int level, int *w, int *h);
int ptref_pyramid_download(const ptref_pyramid *pyr, int level, void *host);
int ptref_upload_image(ptref_ctx *ctx, int level, int layer, const ptref_pyramid *pyr, int style_layer);
int ptref_create_random_pyramid(ptref_ctx *ctx, int level, int frame_number,
                                const ptref_pyramid *image_color, const ptref_pyramid *image_orientation,
                                const ptref_pyramid *exemplar_output,
                                const ptref_pyramid *exemplar_orientation);

int ptref_create_random_pyramid_keys(ptref_ctx *ctx, int level, int frame_number, const ptref_pyramid *image_color,
                                     const ptref_pyramid *image_orientation, int num_styles,
                                     const ptref_pyramid *const *exemplar_outputs,
                                     const ptref_pyramid *const *exemplar_orientations);
int ptref_initialize_cumulative_fields(ptref_ctx *ctx, int w, int h, int num_layers);
int ptref_offsets_from_frame_to_key_ribbon(ptref_ctx *ctx, int level, int look_forwards);
int ptref_merge_advected_buffers(ptref_ctx *ctx, int level, int merge_fallback);
int ptref_accumulate_ribbon_field(ptref_ctx *ctx, int level, int forwards, int time_step, int overwrite);
int ptref_load_ribbon_field(ptref_ctx *ctx, int level, const pt_ribbonp *ribbons);
int ptref_get_ribbon_field(ptref_ctx *ctx, int level, pt_ribbonp *ribbons);
int ptref_randomize_offsets(ptref_ctx *ctx, int level);
int ptref_propagate(ptref_ctx *ctx, int level, int jump, int key_index);
int ptref_scatter(ptref_ctx *ctx, int level, int key_index);
int ptref_upsample(ptref_ctx *ctx, int fine_level);
int ptref_sync_buffers(ptref_ctx *ctx, int level, int flags);
int ptref_update_residual_cache(ptref_ctx *ctx, int level);
int ptref_advect_offsets(ptref_ctx *ctx, int level, int time_step, int forwards, int flags);
int ptref_merge_randomly(ptref_ctx *ctx, int level, int forwards, float weight_advected, int force_valid);
int ptref_merge_using_residual(ptref_ctx *ctx, int level, int forwards);
int ptref_get_output(ptref_ctx *ctx, int level, pt_patchxf *offsets, float *canvas_rgba, float *residual);
int ptref_load_offsets(ptref_ctx *ctx, int level, const pt_patchxf *offsets);
int ptref_load_residual(ptref_ctx *ctx, int level, const float *residual);
int ptref_get_advected(ptref_ctx *ctx, int level, int which, pt_patchxf *offsets,
                       float *base_rgba, float *output_rgba, float *residual, int32_t *time_step);
int ptref_pde_count(ptref_ctx *ctx, uint64_t *out);

int ptref_synthesize(ptref_ctx *ctx, const pt_params *params, int scheme, int iterations,
                     int fixup_passes, int image_w, int image_h, int num_frames,
                     const pt_frame_guides *frames, int style_w, int style_h,
                     const pt_style_guides *style, pt_patchxf *out_offsets,
                     float *out_canvas_rgba, float *out_residual);

/* The frame-major per-frame schedule of pt_synthesize_next_frame, with the
 * sequence state kept in ctx; ptref_sequence_reset starts a new sequence. */
int ptref_synthesize_next_frame(ptref_ctx *ctx, const pt_params *params, int iterations, int image_w,
                                int image_h, const pt_frame_guides *frame, int style_w, int style_h,
                                const pt_style_guides *style, pt_patchxf *out_offsets, float *out_canvas_rgba,
                                float *out_residual);
int ptref_sequence_reset(ptref_ctx *ctx);

/* Single patch distance at (x, y) for a candidate, for unit tests. */
float ptref_patch_distance(ptref_ctx *ctx, int x, int y, int level, const pt_patchxf *cand);
/* screenToExemplarAngle (texSynth_kernel.cu:425-436) at image (x, y) and
 * exemplar (ex, ey) with the textures bound now, for unit tests. */
float ptref_screen_to_exemplar_angle(ptref_ctx *ctx, float x, float y, float ex, float ey);
/* Gaussian tap weights of the current setup (2r+1 floats). */
int ptref_gaussian_weights(ptref_ctx *ctx, float *out, int n);
/* Deterministic math entry points for unit tests. */
void ptref_sincosf(float x, float *s, float *c);
float ptref_atan2f(float y, float x);
float ptref_logf(float x);
float ptref_acosf(float x);
double ptref_log_d(double x);
double ptref_exp_d(double x);
void ptref_rgb_to_lab(float r, float g, float b, float *out);
float ptref_luminance_delta(float cr, float cg, float cb, float pr, float pg, float pb);
float ptref_coherence_term(float cx, float cy, float tcx, float tcy, float tx, float ty, int i, int j, int level,
                           float ratio_range, float angle_range);
float ptref_tensor_to_angle(float xx, float xy, float yy);
void ptref_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

#ifdef __cplusplus
}
#endif

#endif
