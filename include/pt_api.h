/*
 * pt_api.h -- C ABI of the MI355X (gfx950) PaintTween synthesis hot path.
 *
 * The entry points mirror the reference's TexSynth namespace
 * (src/PaintTweenCUDA/texSynth_kernel.h:38-93) one to one, with three
 * differences forced by a plain C boundary:
 *   - an explicit context handle replaces the process-global texture
 *     references and __constant__ parameters (texSynth_kernel.cu:64-96);
 *   - host-side ImagePyramid<T> objects (imagePyramid.h) become device-resident
 *     pt_pyramid handles that are built on the GPU from level 0;
 *   - errors are returned as negative status codes with pt_last_error()
 *     instead of checkCUDAError() -> exit(1) (cudaHostUtil.cpp:46-66).
 * All calls on one context are synchronous with respect to the host, as every
 * reference launcher ends in cudaThreadSynchronize().  A context is used by
 * one host thread at a time; distinct contexts are independent.
 *
 * Structs are fixed-width PODs (no bool or enum members) so the layout is the
 * same from C, C++, ctypes and cgo.
 */
#ifndef PT_API_H
#define PT_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define PT_OK 0
#define PT_ERR_INVALID_ARG (-1)
#define PT_ERR_UNSUPPORTED (-2)
#define PT_ERR_HIP (-3)
#define PT_ERR_OUT_OF_MEMORY (-4)
#define PT_ERR_STATE (-5)
#define PT_ERR_BUDGET (-6)

/* ---- size limits (no reference counterpart) ---------------------------
 * The exemplar of one level is gathered through a buffer resource with
 * 32-bit byte offsets over its interleaved 32-byte texels, and image buffers
 * are indexed with 32-bit pixel offsets: pt_initialize_exemplars and the
 * one-shot schedules refuse larger sizes with PT_ERR_UNSUPPORTED. */
#define PT_MAX_EXEMPLAR_TEXELS 67108863 /* (2^31 - 1) / 32: below 8192 x 8192 */
#define PT_MAX_IMAGE_PIXELS 2147483647
/* Style layers of the exemplar textures (pt_initialize_exemplars,
 * pt_initialize_cumulative_fields); layers * texels of one exemplar level
 * stays below PT_MAX_EXEMPLAR_TEXELS. */
#define PT_MAX_STYLE_LAYERS 4

/* ---- PassDirection (types.h:128-132) ---------------------------------- */
#define PT_FORWARD 0
#define PT_BACKWARD 1
#define PT_BIDIRECTIONAL 2

/* ---- SynthesisScheme (types.h:134-140) -------------------------------- */
#define PT_S_FMBM 0
#define PT_S_FMBM_RM 1
#define PT_S_FBMO 2 /* forward and backward passes, merge, fixups (renderAnimationFBMO) */
#define PT_S_TCTF 3 /* consecutive pass, then refine passes coarse to fine (renderAnimationTCTF) */
#define PT_S_IND 4  /* every frame on its own (renderAnimationIND) */

/* ---- TsImageLayer (texSynth_interface.h:52-75), same values ----------- */
#define PT_LAYER_EXEMPLAR_BASE 0
#define PT_LAYER_EXEMPLAR_OUTPUT 1
#define PT_LAYER_EXEMPLAR_ORIENTATION 2
#define PT_LAYER_EXEMPLAR_DIST_TRANS 3
#define PT_LAYER_EXEMPLAR_OBJECT_IDS 4
#define PT_LAYER_INPUT_COLOR 5
#define PT_LAYER_INPUT_ORIENTATION 6
#define PT_LAYER_INPUT_DIST_TRANS 7
#define PT_LAYER_INPUT_SCALE 8
#define PT_LAYER_INPUT_ID 9
#define PT_LAYER_INPUT_SURF_ID 10
#define PT_LAYER_INPUT_VEL_B 11
#define PT_LAYER_INPUT_VEL_F 12
#define PT_LAYER_INPUT_VEL_B_PREVIOUS 13
#define PT_LAYER_INPUT_VEL_F_PREVIOUS 14
#define PT_LAYER_CUMUL_FIELD_B 17         /* key-to-key ribbons, layer = style_layer */
#define PT_LAYER_CUMUL_FIELD_F 18
#define PT_LAYER_FRAME_TO_KEY_RIBBON_B 19 /* frame-to-key ribbons of the frame */
#define PT_LAYER_FRAME_TO_KEY_RIBBON_F 20
#define PT_LAYER_RANDOM 21

/* ---- operation flags (texSynth_kernel.h:79-91), same bits ------------- */
#define PT_DONT_COPY_OFFSETS 0x01
#define PT_DONT_UPDATE_RESIDUAL 0x10
#define PT_MERGE_FALLBACK_RANDOM 0x20
#define PT_MERGE_FALLBACK_LAST_PASS 0x40
#define PT_ADVECT_IGNORE_RANDOM 0x1
#define PT_ADVECT_DO_NOT_CHECK_PREVIOUS 0x10
#define PT_ADVECT_CHECK_FRAME_TO_KEY 0x20
#define PT_ADVECT_UPDATE_RESIDUAL 0x80

/* ---- scatter random-number modes (no counterpart in the reference) ----
 * PT_RNG_REFERENCE reproduces the reference scatter exactly: its randFloat()
 * returns an unsigned int that truncates to 0 (texSynth_util.h:155-163), so
 * every sample is the shift (+0.9*dist, 0).  PT_RNG_COUNTER draws r and theta
 * from Philox4x32-10 keyed by (seed, frame*level), counter (x, y, call, i). */
#define PT_RNG_REFERENCE 0
#define PT_RNG_COUNTER 1

/* ---- pyramid element kinds and downsample modes ----------------------- */
#define PT_ELEM_F32 1   /* float          */
#define PT_ELEM_F32X2 2 /* float2         */
#define PT_ELEM_F32X3 3 /* float3 tensor  */
#define PT_ELEM_F32X4 4 /* Color4 / float4 */
#define PT_ELEM_I32 5   /* int32 object IDs (ImagePyramid<int>) */
/* DownsampleScaleMode (downsample.h:35-38), same values. */
#define PT_DOWNSAMPLE_SCALE_WITH_LEVELS 0
#define PT_DOWNSAMPLE_SCALE_CONSTANT 1
/* Ribbon pyramids (initializeRibbonPyramid, imagePyramid.cpp:211-235):
 * PT_ELEM_F32X4 texels packed as RibbonP's Color4 (time_step, y, x,
 * layer); level L+1 takes the top-left texel of each 2x2 block with x and
 * y halved. */
#define PT_DOWNSAMPLE_RIBBON 2

/*
 * pt_params -- TsParameters (texSynth_interface.h:79-152) as a fixed-width
 * POD.  Field order and meaning follow the reference; bool fields are int32
 * (0/1).  The last three fields are extensions with no reference
 * counterpart.  Defaults (G/nlParameters.cpp:63-101) come from
 * pt_params_default().
 */
typedef struct pt_params {
    int32_t residual_window_size;      /* r: patch is (2r+1)^2 taps          */
    float coherence_weight;
    float dist_trans_weight;
    float lum_shift_weight;
    int32_t lum_shift_enable;
    float input_analogy_weight;
    float canvas_output_weight;
    int32_t num_styles;
    float offsets_histogram_slope[5];     /* offsetsHistogramSlope: any value;
                                             the reference's histogram stays
                                             zero, so the term is 0 unless the
                                             layer-0 threshold is negative
                                             (PT_ERR_UNSUPPORTED then)      */
    float offsets_histogram_threshold[5];
    float max_distortion;
    float alpha_weight;
    int32_t transparency_ok;
    float style_obj_id_weight;
    int32_t orientation_enabled;
    int32_t colorspace;
    float coherence_ratio_range;
    float coherence_angle_range;
    int32_t levels;
    int32_t scatter_samples_count;
    float advection_jitter;
    float time_derivative_input_weight;
    float time_derivative_output_weight;
    float temporal_coherence_weight;
    float hysteresis_weight;
    int32_t interpolate_key_frame;
    int32_t use_guide;
    int32_t current_interpolation_index;
    int32_t num_frames;
    int32_t num_key_frames;
    int32_t first_key_index;
    int32_t last_key_index;
    int32_t dist_trans_cull;
    float residual_window_scale_u;
    float residual_window_scale_v;
    int32_t direction;                 /* PT_FORWARD / PT_BACKWARD / ...     */
    /* extensions */
    int32_t rng_mode;                  /* PT_RNG_REFERENCE / PT_RNG_COUNTER  */
    uint32_t seed;                     /* Philox key word 0                  */
    int32_t frame_number;              /* frame used to key the RNG          */
    int32_t first_pass_direction;      /* firstPassDirection of the schedules
                                          (SynthesisProcessor, 0 = PT_FORWARD
                                          first, 1 = backward first): pass p
                                          runs forwards iff p % 2 equals it
                                          (FMBM, FMBM_RM, FBMO; TCTF and the
                                          per-frame schedule take 0 only)   */
    int32_t max_level;                 /* maxLevel (SynthesisProcessor::
                                          _finalLevel): FMBM / FMBM_RM
                                          synthesize levels down to it, then
                                          upsample the result to level 0
                                          without synthesis; 0 = all levels,
                                          other schedules take 0 only       */
} pt_params;

/*
 * pt_patchxf -- PatchXF (types.h:98-114), 36 bytes, same field order:
 * the nearest-neighbour-field element.  Used only at the host boundary;
 * on the device the field is stored as structure-of-arrays planes.
 */
typedef struct pt_patchxf {
    float x, y, theta, hysteresis;
    int32_t layer;
    float scale_u, scale_v;
    float lum_y, lum_a;
} pt_patchxf;

/* pt_ribbonp -- RibbonP (types.h:116-126), 16 bytes: the offset from a pixel
 * to where a keyframe's texel travelled, its time step, and the keyframe's
 * style layer (-1: invalid). */
typedef struct pt_ribbonp {
    float x, y;
    int32_t time_step, layer;
} pt_ribbonp;

typedef struct pt_ctx pt_ctx;
typedef struct pt_pyramid pt_pyramid;

/* ---- parameters --------------------------------------------------------- */

/* NLParameters::getTsDefaultParams (G/nlParameters.cpp:63-101) plus
 * distTransCull=50 (synthesisProcessor.cpp:220), direction FORWARD,
 * rng_mode PT_RNG_REFERENCE, seed 0, frame 0. */
void pt_params_default(pt_params *out);

/* sizeof of the ABI structs as compiled into the library, for bindings to
 * check their layouts: 0 pt_params, 1 pt_patchxf, 2 pt_frame_guides,
 * 3 pt_style_guides, 4 pt_job, 5 pt_batch_stats, 6 pt_sequence_desc,
 * 7 pt_server_stats; 0 for an unknown id.  No reference counterpart. */
size_t pt_struct_size(int which);

/* ---- context -------------------------------------------------------------- */

/* Create a context on HIP device `device`.  No reference counterpart (the
 * reference uses process globals). */
int pt_create(pt_ctx **out, int device);
int pt_destroy(pt_ctx *ctx);
/* Message of the last failing call on ctx ("" if none). */
const char *pt_last_error(pt_ctx *ctx);
/* Message of the last failure that had no context to record it on. */
const char *pt_global_last_error(void);

/* TexSynth::setup (texSynth_kernel.cu:109-117): Gaussian tap weights
 * (convolution.cu:37-60) and parameters. */
int pt_setup(pt_ctx *ctx, const pt_params *params);
/* TexSynth::uploadParameters (texSynth_kernel.cu:119-123). */
int pt_upload_parameters(pt_ctx *ctx, const pt_params *params);
/* WorkingBuffers::initialize (workingBuffers.cpp:48-95) together with
 * TexSynth::initialize_images (texSynth_kernel.cu:2111-2117): level-0-sized
 * working buffers, zero-filled, and the two-layer image textures.
 * style_w/style_h are the level-0 exemplar dimensions (they size the
 * offsets-histogram buffer whose dimensions drive the patch wrap). */
int pt_initialize(pt_ctx *ctx, int image_w, int image_h, int style_w, int style_h);
/* TexSynth::initialize_exemplars (texSynth_kernel.cu:2119-2126): layered
 * exemplar textures of one level (w, h are that level's dimensions),
 * num_layers in 1..PT_MAX_STYLE_LAYERS (the styles, which are the keyframes
 * with keyframe interpolation).  An offset samples the layer it carries
 * (PatchXF::layer), the index clamped into the range as a layered texture
 * does.  Without keyframe interpolation every offset stays in layer 0
 * (random initialisation, scatter and the average candidate use style 0,
 * synthesisFrame.cpp:107-110, texSynth_kernel.cu:925,939); more styles then
 * only enlarge the patch wrap height to style_h * num_styles, as the
 * reference's offsets histogram does. */
int pt_initialize_exemplars(pt_ctx *ctx, int w, int h, int num_layers);
/* TexSynth::initialize_cumulativeFields (texSynth_kernel.cu:2136-2140): the
 * layered key-to-key fields B and F (w, h: their level dimensions), zero
 * filled; PT_LAYER_CUMUL_FIELD_B/F uploads fill layer style_layer. */
int pt_initialize_cumulative_fields(pt_ctx *ctx, int w, int h, int num_layers);
/* WorkingBuffers::resetAdvectedBuffers (workingBuffers.cpp:113-117). */
int pt_reset_advected_buffers(pt_ctx *ctx);
/* TexSynth::flipAnimBuffers (texSynth_kernel.cu:1534-1538). */
int pt_flip_anim_buffers(pt_ctx *ctx);

/* ---- device image pyramids ------------------------------------------------ */

/* ImagePyramid<T>::initialize (imagePyramid.cpp:81-106) with
 * downsampleAverage (downsample.cpp:41-160) run on the GPU: level L+1 is the
 * 2x2 box mean of level L, times 0.5 more for PT_DOWNSAMPLE_SCALE_WITH_LEVELS.
 * PT_ELEM_I32 pyramids use downsampleVote (downsample.cpp:162-210) instead:
 * the most frequent of the 2x2 IDs, a tie going to the first of them in the
 * order top-left, top-right, bottom-left, bottom-right (the reference breaks
 * ties by QHash iteration order, which depends on a per-process hash seed).
 * `host_level0` is w*h elements of kind `elem`, tightly packed; NULL fills
 * level 0 with `default_value` in every channel (the absent-input defaults of
 * G/nlDataAccess.cpp:96-155; converted to int for PT_ELEM_I32).  The pyramid
 * has min(floor(log2 w), floor(log2 h)) levels. */
int pt_pyramid_create(pt_ctx *ctx, pt_pyramid **out, int elem, int w, int h,
                      int scale_mode, const void *host_level0, float default_value);
/* NLDataAccess::toOrientationPyramid (G/nlDataAccess.cpp:82-93): converts a
 * PT_ELEM_F32X3 structure-tensor pyramid into a PT_ELEM_F32 angle pyramid
 * (strucTensorToAngle, imageIO.h:136-139) on the GPU. */
int pt_pyramid_orientation(pt_ctx *ctx, pt_pyramid **out, const pt_pyramid *tensor);
/* A pyramid belongs to the context that built it: destroy it before that
 * context. */
int pt_pyramid_destroy(pt_pyramid *pyr);
int pt_pyramid_levels(const pt_pyramid *pyr);
int pt_pyramid_dims(const pt_pyramid *pyr, int level, int *w, int *h);
/* Copies one level to the host (w*h elements, tightly packed). */
int pt_pyramid_download(const pt_pyramid *pyr, int level, void *host);

/* TexSynth::uploadImage_{float,float2,Color4,int} (texSynth_kernel.cu:2128-2253):
 * binds `level` of `pyr` to image layer `layer`, device to device.  The
 * element kind must match the layer (color/exemplar/random: F32X4, velocity
 * and scale: F32X2, orientation and distance transform: F32, INPUT_ID and
 * EXEMPLAR_OBJECT_IDS: I32, cumulative fields and frame-to-key ribbons:
 * F32X4 RibbonP texels).  Until INPUT_ID is uploaded the image ID texture
 * is a single 0 texel, and so are the frame-to-key ribbons (a valid layer-0
 * ribbon, as the reference's unbound textures read).  style_layer (0..3)
 * selects the layer of an exemplar texture or a cumulative field; image
 * layers ignore it. */
int pt_upload_image(pt_ctx *ctx, int level, int layer, const pt_pyramid *pyr, int style_layer);

/* SynthesisFrame::createRandomPyramid (synthesisFrame.cpp:66-150) on the GPU
 * for one level, then uploaded as TS_LAYER_RANDOM.  glibc rand() seeded with
 * srand(frame*level) is replaced by Philox4x32-10 keyed (seed, frame*level)
 * with counter (x, y, draw, 0).  Needs the level's image color and
 * orientation, and the exemplar output and orientation pyramids. */
int pt_create_random_pyramid(pt_ctx *ctx, int level, int frame_number,
                             const pt_pyramid *image_color, const pt_pyramid *image_orientation,
                             const pt_pyramid *exemplar_output, const pt_pyramid *exemplar_orientation);
/* The same with keyframe interpolation (setupWithKeys, synthesisFrame.cpp:
 * 94-110): each pixel first draws its style thisStyle = rand() % (number of
 * keys) + first_key_index (params of the context), then its texel in that
 * style; styles[k] are the output and orientation pyramids of style k
 * (num_styles of them).  The style draw is Philox counter (x, y, 51, 0). */
int pt_create_random_pyramid_keys(pt_ctx *ctx, int level, int frame_number, const pt_pyramid *image_color,
                                  const pt_pyramid *image_orientation, int num_styles,
                                  const pt_pyramid *const *exemplar_outputs,
                                  const pt_pyramid *const *exemplar_orientations);

/* ---- NNF operators (TexSynth::*) --------------------------------------- */

/* TexSynth::randomizeOffsets (texSynth_kernel.cu:2100-2109). */
int pt_randomize_offsets(pt_ctx *ctx, int level);
/* TexSynth::offsetsFromFrameToKeyRibbon (texSynth_kernel.cu:2061-2118):
 * offsets from the frame-to-key ribbon F (look_forwards) or B, the random
 * texel where the ribbon is invalid, then syncBuffers. */
int pt_offsets_from_frame_to_key_ribbon(pt_ctx *ctx, int level, int look_forwards);
/* TexSynth::propagate (texSynth_kernel.cu:1306-1357): propagate_kernel, or
 * propagate_kernel_keys when params.interpolate_key_frame is set (every
 * neighbour candidate also followed through the cumulative fields to the
 * other keys, :984-1110).  key_index >= 0: pixels where that style layer's
 * output is painted (alpha > 0) take their own position in it. */
int pt_propagate(pt_ctx *ctx, int level, int jump, int key_index);
/* TexSynth::scatter (texSynth_kernel.cu:1360-1401): scatter_kernel, or
 * scatter_kernel_keys with params.interpolate_key_frame (:723-839); key_index
 * as for pt_propagate (painted pixels keep their offset). */
int pt_scatter(pt_ctx *ctx, int level, int key_index);
/* TexSynth::upsample (texSynth_kernel.cu:1513-1532). */
int pt_upsample(pt_ctx *ctx, int fine_level);
/* TexSynth::syncBuffers (texSynth_kernel.cu:1240-1303). */
int pt_sync_buffers(pt_ctx *ctx, int level, int flags);
/* TexSynth::updateResidualCache (texSynth_kernel.cu:1221-1238). */
int pt_update_residual_cache(pt_ctx *ctx, int level);
/* TexSynth::advectOffsets (texSynth_kernel.cu:1695-1726). */
int pt_advect_offsets(pt_ctx *ctx, int level, int time_step, int forwards, int flags);
/* TexSynth::mergeRandomly (texSynth_kernel.cu:2003-2024). */
int pt_merge_randomly(pt_ctx *ctx, int level, int forwards, float weight_advected, int force_valid);
/* TexSynth::mergeUsingResidual (texSynth_kernel.cu:1978-2001). */
int pt_merge_using_residual(pt_ctx *ctx, int level, int forwards);
/* TexSynth::mergeAdvectedBuffers (texSynth_kernel.cu:1938-2042): the forward
 * and backward advected offsets merged, merge_fallback
 * PT_MERGE_FALLBACK_RANDOM or PT_MERGE_FALLBACK_LAST_PASS where neither is
 * valid. */
int pt_merge_advected_buffers(pt_ctx *ctx, int level, int merge_fallback);
/* TexSynth::accumulateRibbonField (texSynth_kernel.cu:1728-1854): the
 * ribbon field advected one step of time_step frames along the uploaded
 * velocities (overwrite: replace every texel; else keep the shorter ribbon),
 * and the canvas texels the ribbons reach written to the forward advected
 * output colours (clamped into the canvas). */
int pt_accumulate_ribbon_field(pt_ctx *ctx, int level, int forwards, int time_step, int overwrite);
/* The ribbon field (WorkingBuffers::ribbonField, level-0 sized, invalid
 * after pt_initialize): level-rectangle copies in and out. */
int pt_load_ribbon_field(pt_ctx *ctx, int level, const pt_ribbonp *ribbons);
int pt_get_ribbon_field(pt_ctx *ctx, int level, pt_ribbonp *ribbons);

/* WorkingBuffers::getOutput (workingBuffers.cpp:125-142): level-sized copies
 * of the base offsets, the canvas (RGBA) and the residual; any pointer may be
 * NULL. */
int pt_get_output(pt_ctx *ctx, int level, pt_patchxf *offsets, float *canvas_rgba, float *residual);
/* pt_get_output without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  pt_get_output stages the packed offsets in a buffer it
 * allocates and frees per call; this variant stages them in a file-scope
 * device array of PT_NOALLOC_SCRATCH_BYTES per GPU (allocated when the
 * library's code loads on that GPU, held until the process ends), so a
 * level is limited to PT_NOALLOC_SCRATCH_BYTES / 36 offsets (3 728 270:
 * 2048 x 1820, larger levels return PT_ERR_UNSUPPORTED).  All calls on one
 * GPU share that array; the library serializes them with a process-wide
 * lock, and callers in other processes have arrays of their own.  Results
 * equal pt_get_output's. */
#define PT_NOALLOC_SCRATCH_BYTES (128u << 20)
int pt_get_output_noalloc(pt_ctx *ctx, int level, pt_patchxf *offsets, float *canvas_rgba, float *residual);
/* WorkingBuffers::loadOffsets / loadResidual (workingBuffers.cpp:166-207):
 * writes the level rectangle of the base offsets / base residual.  Offsets
 * must have lum_a 0 (PT_ERR_UNSUPPORTED otherwise); a nonzero lum_y
 * (luminance shift) is kept and added to the canvas; the style layer is
 * kept. */
int pt_load_offsets(pt_ctx *ctx, int level, const pt_patchxf *offsets);
int pt_load_residual(pt_ctx *ctx, int level, const float *residual);
/* Reads an advected buffer set (advectedBuffers.h) for tests: which = 0 for
 * F, 1 for B.  Any pointer may be NULL. */
int pt_get_advected(pt_ctx *ctx, int level, int which, pt_patchxf *offsets,
                    float *base_rgba, float *output_rgba, float *residual, int32_t *time_step);

/* Patch-distance evaluations launched by this context so far (nominal count:
 * every candidate tested, rejected ones included, plus every residual
 * refresh). */
int pt_pde_count(pt_ctx *ctx, uint64_t *out);

/* ---- execution options (no reference counterpart) ----------------------
 * PT_TUNE_EARLY_EXIT (default 1): propagate and scatter reject a candidate
 * as soon as a partial patch energy provably exceeds the best so far (only
 * when every energy weight is non-negative and the hysteresis weight is 0);
 * results are bit-identical either way.  0 forces the plain kernels. */
#define PT_TUNE_EARLY_EXIT 1
/* Early rejection is used on levels with at least this many pixels
 * (default 256*256); smaller levels have too few waves to hide the longer
 * per-lane dependency chain. */
#define PT_TUNE_EARLY_EXIT_MIN_PIXELS 2
/* PT_TUNE_UNIT_PARALLEL (default 1): propagate and scatter spread every
 * (pixel, candidate) unit of a workgroup's tile over its lanes through an LDS
 * queue and resolve each pixel in candidate order afterwards; 0 evaluates
 * each pixel's candidates sequentially on its own lane.  Same results. */
#define PT_TUNE_UNIT_PARALLEL 3
/* PT_TUNE_INCREMENTAL_RESIDUAL (default 1): the residual refresh after a
 * sync recomputes only pixels whose patch window holds an offset that sync
 * changed; the others would reproduce their cached value bit for bit. */
#define PT_TUNE_INCREMENTAL_RESIDUAL 4
/* PT_TUNE_TAP_PARALLEL_MAX_PIXELS (default 128*128): on levels up to this
 * many pixels, propagate and the residual refresh give every tap of every
 * candidate its own lane and sum each candidate's taps in the reference
 * order afterwards (shorter dependency chains where the GPU would idle);
 * 0 disables it.  Same results. */
#define PT_TUNE_TAP_PARALLEL_MAX_PIXELS 5
/* PT_TUNE_TAP_PARALLEL_EVALS (default 1024, range 1..1024): tap evaluations
 * per tap-parallel workgroup; a workgroup covers this many / (2r+1)^2 pixels
 * in the refresh and / (4 (2r+1)^2) in propagate (at least one).  Same
 * results. */
#define PT_TUNE_TAP_PARALLEL_EVALS 6
/* PT_TUNE_EXEMPLAR_WINDOW (default 200000): on levels of at least this
 * many pixels the per-pixel residual refresh stages an exemplar window around
 * each tile's offsets in LDS and reads the taps of patches inside it from
 * there (smaller levels measured slower with it); 0 reads every tap from
 * global memory.  Same results. */
#define PT_TUNE_EXEMPLAR_WINDOW 7
/* PT_TUNE_TIME_TABLE_MAX_TAPS (default 81): with temporal terms active and
 * patches of at most this many taps, the exemplar samples at every tap of
 * every pixel's advected offset (the candidate-independent half of
 * timeDerivativeTerms, texSynth_kernel.cu:448-452) are gathered once per
 * level into a table of 32 B per tap and pixel, which every candidate then
 * reads instead of gathering them again; 0 disables it.  Same results. */
#define PT_TUNE_TIME_TABLE_MAX_TAPS 8
/* PT_TUNE_XCD_TILES (default 1): the single-job NNF kernels (propagate,
 * scatter, residual refresh) deal the image's tiles to the 8 XCDs in
 * contiguous row-major ranges instead of round-robin, so neighbouring tiles
 * share one XCD's L2.  Same results. */
#define PT_TUNE_XCD_TILES 9
/* PT_TUNE_PROPAGATE_WINDOW_JUMP (default 8) and
 * PT_TUNE_PROPAGATE_WINDOW_MIN_PIXELS (default 200000): on levels of at
 * least that many pixels, the unit-parallel propagate steps with a jump up
 * to this stage a 24x24-texel exemplar window in LDS (20x20 with temporal
 * terms), anchored on each tile's current offsets, and candidates whose
 * whole patch lies in it read it there (short jumps propose offsets near the
 * current ones); 0 disables it.  Same results. */
#define PT_TUNE_PROPAGATE_WINDOW_JUMP 10
#define PT_TUNE_PROPAGATE_WINDOW_MIN_PIXELS 11
/* PT_TUNE_TEMPORAL_UNIT_ORDER: with temporal terms and time tables, the
 * order in which a unit-parallel workgroup enumerates its (pixel,
 * candidate) units: 0 candidate-major (neighbouring lanes evaluate the same
 * candidate of neighbouring pixels: coherent exemplar gathers), 1
 * pixel-major in the propagate steps with an exemplar window (the lanes of
 * one pixel read the same time-table entries in the same instruction), 2
 * pixel-major in every propagate and scatter step.  Same results. */
#define PT_TUNE_TEMPORAL_UNIT_ORDER 12
int pt_set_tuning(pt_ctx *ctx, int key, int value);

/* ---- timing (no reference counterpart; the reference has no GPU timers) --
 * pt_profile_enable(ctx, 1) clears the log and starts recording a pair of HIP
 * events on the context's stream around every synthesis-path launch;
 * pt_profile_get() sums the launches of one kernel class ("propagate",
 * "scatter", "update_residual", "canvas" (the fused sync/canvas pass),
 * "upsample", "advect", "merge", "random", "pyramid") since then: launch
 * count, total GPU milliseconds, total pixels covered.  pt_profile_bytes()
 * gives the same launches' algorithmic HBM bytes (each operand read or
 * written once, the exemplar counted once per pixel), from which a caller
 * forms achieved GB/s against the HBM roofline. */
int pt_profile_enable(pt_ctx *ctx, int enable);
int pt_profile_get(pt_ctx *ctx, const char *kernel, uint64_t *launches, double *total_ms,
                   uint64_t *pixels);
int pt_profile_bytes(pt_ctx *ctx, const char *kernel, uint64_t *bytes);
/* The recorded launches of one class one by one, in issue order: pixels,
 * algorithmic bytes and GPU milliseconds of the first max_records of them;
 * *count receives how many there are. */
int pt_profile_records(pt_ctx *ctx, const char *kernel, int max_records, uint64_t *pixels, uint64_t *bytes,
                       float *ms, int *count);

/* ---- one-shot schedules ------------------------------------------------- */

/*
 * pt_frame_guides -- one frame's level-0 guides, host memory, tightly packed
 * row-major.  NULL entries take the reference's absent-input defaults
 * (G/nlDataAccess.cpp:96-155): orientation tensor 0 (angle -pi/2),
 * distance transform 1.0, scale (1,1), velocities (0,0).
 */
typedef struct pt_frame_guides {
    const float *color_rgba;        /* required, w*h*4                     */
    const float *orientation_xx_xy_yy; /* w*h*3 structure tensor or NULL   */
    const float *dist_trans;        /* w*h or NULL                          */
    const float *scale_uv;          /* w*h*2 or NULL                        */
    const float *vel_b;             /* w*h*2 pixels/frame or NULL           */
    const float *vel_f;             /* w*h*2 pixels/frame or NULL           */
    const int32_t *object_id;       /* w*h object IDs or NULL (all 0); read only
                                       when style_obj_id_weight > 0           */
} pt_frame_guides;

/* The exemplar ("style") pair with its guides.  NULL orientation or distance
 * transform take the absent-style-element default, a zero image
 * (G/nlDataAccess.cpp:335-339). */
typedef struct pt_style_guides {
    const float *base_rgba;         /* required, w*h*4 exemplar input      */
    const float *output_rgba;       /* required, w*h*4 exemplar output     */
    const float *orientation_xx_xy_yy; /* w*h*3 or NULL                     */
    const float *dist_trans;        /* w*h or NULL                          */
    const int32_t *object_ids;      /* w*h object IDs or NULL (all 0); read only
                                       when style_obj_id_weight > 0           */
} pt_style_guides;

/*
 * Synthesis of a sequence of `num_frames` frames (1 = a still) with one of
 * the reference schedules (`scheme`, PT_S_*): FMBM / FMBM_RM,
 * SynthesisProcessor::renderAnimationFMBM (synthesisProcessor.cpp:846-878)
 * over renderConsecutivePass (:703-753) and SynthesisFrame::synthesizeConsecutive
 * (synthesisFrame.cpp:214-323) with propagateAndScatter (:483-508); IND
 * (:806-839), FBMO (:902-970) or TCTF (:1151-1312).  params->max_level > 0
 * is taken by FMBM / FMBM_RM only (their tail, :863-870: every frame's
 * result upsampled level by level to level 0 without synthesis).
 * Frames are processed in the reference order of the scheme.
 * out_offsets / out_canvas / out_residual (each may be NULL) receive every
 * frame's level-0 result of the final pass: num_frames*w*h elements.
 * Each call starts from zeroed working buffers (a fresh WorkingBuffers), so
 * repeated calls with the same inputs give identical results.
 */
int pt_synthesize(pt_ctx *ctx, const pt_params *params, int scheme, int iterations,
                  int fixup_passes, int image_w, int image_h, int num_frames,
                  const pt_frame_guides *frames, int style_w, int style_h,
                  const pt_style_guides *style, pt_patchxf *out_offsets,
                  float *out_canvas_rgba, float *out_residual);

/*
 * pt_synthesize with every guide pointer and output pointer in device memory
 * of the context's GPU (no counterpart in the reference, whose inputs always
 * arrive from host pyramids).  Used when the frames already live in HBM.
 */
int pt_synthesize_device(pt_ctx *ctx, const pt_params *params, int scheme, int iterations,
                         int fixup_passes, int image_w, int image_h, int num_frames,
                         const pt_frame_guides *frames, int style_w, int style_h,
                         const pt_style_guides *style, pt_patchxf *out_offsets,
                         float *out_canvas_rgba, float *out_residual);

/* ---- streaming: one frame of a sequence at a time ------------------------ */

/*
 * Synthesizes the next frame of the sequence held by ctx (frame 0 after
 * pt_create or pt_sequence_reset) and returns its level-0 result.  Frames are
 * processed frame-major: every level of this frame, coarse to fine, before
 * the next frame exists.  Per level this is the forward first pass of
 * SynthesisFrame::synthesizeConsecutive (synthesisFrame.cpp:214-323) with the
 * FMBM_RM merge (mergeRandomly, weight 0.5): frame 0 starts from random
 * offsets / the upsampled coarser level; later frames advect the previous
 * frame's offsets of the same level (advectOffsets) and merge them with the
 * upsampled coarser result.  The previous frame's per-level offsets,
 * residual and images stay resident on the device between calls.  The
 * reference has no frame-major schedule (all of renderAnimationFMBM, FBMO,
 * TCTF and IND loop levels outside frames, synthesisProcessor.cpp:807-1312);
 * this one exists for callers that receive frames one at a time.  All
 * pointers are device pointers; frame size, exemplar size and levels are
 * fixed for a sequence.
 */
int pt_synthesize_next_frame(pt_ctx *ctx, const pt_params *params, int iterations, int image_w, int image_h,
                             const pt_frame_guides *frame, int style_w, int style_h,
                             const pt_style_guides *style, pt_patchxf *out_offsets, float *out_canvas_rgba,
                             float *out_residual);
/* The next pt_synthesize_next_frame call starts a new sequence (frame 0). */
int pt_sequence_reset(pt_ctx *ctx);
/* Index of the frame the next call will synthesize. */
int pt_sequence_frame(pt_ctx *ctx, int *frame);

/* ---- batched form: many independent jobs ------------------------------ */

/* One job of a batch: the arguments of pt_synthesize_device (all pointers
 * are device pointers of the batch's GPU; outputs may be NULL). */
typedef struct pt_job {
    pt_params params;
    int32_t scheme, iterations, fixup_passes;
    int32_t image_w, image_h, num_frames;
    const pt_frame_guides *frames; /* num_frames entries */
    int32_t style_w, style_h;
    const pt_style_guides *style;
    pt_patchxf *out_offsets;
    float *out_canvas_rgba;
    float *out_residual;
} pt_job;

typedef struct pt_batch pt_batch;

/*
 * Plans a batch: every job gets its own device state, its stream work is
 * recorded once, and the same step of all jobs (propagate at one jump,
 * scatter, residual refresh, sync/canvas) runs as ONE multi-job kernel
 * launch over the concatenated tiles of every job, whatever their sizes,
 * levels, patch radii and exemplars.  The launch sequence is captured in a
 * HIP graph on the first run and replayed afterwards.  Each job's outputs
 * are bit-identical to pt_synthesize_device on its own.  No reference
 * counterpart (the reference synthesizes one working set per process).
 */
int pt_batch_create(pt_batch **out, int device, int num_jobs, const pt_job *jobs);
int pt_batch_destroy(pt_batch *b);
const char *pt_batch_last_error(pt_batch *b);
/* Runs every job to completion (synchronous).  Repeated runs recompute the
 * same outputs from the same inputs. */
int pt_batch_run(pt_batch *b);
/*
 * Autotuner: for each operator with several kernel shapes (propagate,
 * scatter) and each level-size class (< 128^2, < 512^2, >= 512^2 pixels),
 * times every valid shape inside whole-batch runs with HIP events and keeps
 * the fastest; drops the captured graph so the next run captures the tuned
 * plan.  `rounds` whole-batch runs per shape (>= 1).
 */
int pt_batch_tune(pt_batch *b, int rounds);
#define PT_BATCH_USE_GRAPH 1 /* default 1: capture/replay the plan as a HIP graph */
#define PT_BATCH_XCD_REMAP 2 /* default 1: deal contiguous tile ranges to each XCD */
#define PT_BATCH_SHAPE 3     /* value = op*256 + class*16 + shape (shape 15 = default policy) */
/* default 2048: a job step of at least this many workgroup tiles fills the
 * GPU alone and gets a launch of its own instead of sharing one */
#define PT_BATCH_ISOLATE_TILES 4
/*
 * Persistent form (default 0): pt_batch_run makes ONE kernel launch that
 * runs every job to completion.  Each job's recorded plan becomes a device
 * program (one step per kernel its plain path launches, each step's
 * workgroup tiles claimed from per-job counters); a job's steps run in
 * order on the XCDs it is placed on, different jobs overlap freely, and the
 * grid holds only as many workgroups as fit on the GPU at once.  Outputs are bit-identical to the
 * plain path.  Jobs whose plans need a kernel without a device-program form
 * (temporal terms, extended patch-distance terms, style layers, keyframes,
 * patch radii other than 2 and 3) make pt_batch_run return
 * PT_ERR_UNSUPPORTED naming the step.  pt_batch_tune then picks the tiles
 * per claim.  A launch that sees no tile complete for the timeout gives up
 * and pt_batch_run returns PT_ERR_HIP (its outputs are then undefined).
 */
#define PT_BATCH_PERSISTENT 5
#define PT_BATCH_PERSIST_CHUNK 6      /* default 8: most tiles one claim takes */
#define PT_BATCH_PERSIST_TIMEOUT_MS 7 /* default 2000 */
/* value 1: the next persistent launch starts with its give-up flag set, so
 * every workgroup stops at its first look for work and pt_batch_run returns
 * PT_ERR_HIP (exercises the give-up path; later runs are unaffected) */
#define PT_BATCH_PERSIST_ABORT 8
int pt_batch_set(pt_batch *b, int key, int value);

typedef struct pt_batch_stats {
    int64_t jobs, ops, launches, multi_launches, ticks;
    double last_run_ms;  /* host wall time of the last pt_batch_run */
    double tune_ms;      /* total time spent in pt_batch_tune */
    int32_t shape[2][3]; /* chosen shape per (propagate, scatter) x size class; -1 = default */
    int32_t xcd_remap, isolate_tiles;
    int32_t persistent;           /* PT_BATCH_PERSISTENT */
    int32_t persistent_grid;      /* workgroups of the persistent launch */
    int64_t persistent_tasks;     /* steps of all jobs' device programs */
    int64_t persistent_tiles;     /* workgroup tiles of all steps */
    int32_t persistent_max_chunk; /* tiles per claim */
    int32_t persistent_spanned;   /* jobs placed on more than one XCD */
} pt_batch_stats;
int pt_batch_get_stats(pt_batch *b, pt_batch_stats *out);

/* One-shot: plan, run and free a batch. */
int pt_submit_batch(int device, int num_jobs, const pt_job *jobs);

/* ---- serving form: streamed sequences with continuous batching ---------- */

typedef struct pt_server pt_server;

/* A sequence to serve: parameters, iterations per level and the frame and
 * exemplar shapes; the exemplar guides (device pointers) must stay valid
 * until the sequence is closed. */
typedef struct pt_sequence_desc {
    pt_params params;
    int32_t iterations;
    int32_t image_w, image_h;
    int32_t style_w, style_h;
    const pt_style_guides *style;
} pt_sequence_desc;

/*
 * A request loop on one GPU.  Frames submitted for open sequences are
 * admitted as in-flight slots free up (max_in_flight frames at once, at most
 * one per sequence, as frames of a sequence depend on each other) and while
 * the device memory of the resident sequences stays within mem_budget_bytes
 * (a sequence becomes resident, with its offset field, previous frame and
 * per-level state on the device, when its first frame is admitted, and stays
 * so until closed).  Every engine tick advances every in-flight frame by one
 * step, merging the same step of different frames into one multi-job launch,
 * so frames that arrive while others run join the running batch.  Each
 * frame's outputs are bit-identical to pt_synthesize_next_frame on a context
 * that served the same sequence alone.
 */
int pt_server_create(pt_server **out, int device, uint64_t mem_budget_bytes, int max_in_flight);
int pt_server_destroy(pt_server *s);
const char *pt_server_last_error(pt_server *s);
int pt_server_open(pt_server *s, const pt_sequence_desc *desc, int32_t *seq);
/* Queues the next frame of `seq` (device guides and outputs); *ticket
 * identifies it in pt_server_poll. */
int pt_server_submit(pt_server *s, int32_t seq, const pt_frame_guides *frame, pt_patchxf *out_offsets,
                     float *out_canvas_rgba, float *out_residual, uint64_t *ticket);
/* Runs up to max_ticks engine ticks (0: only collect completions) and
 * returns up to max_done completed tickets with their host-observed latency
 * from submit to completion in ms (latency_ms may be NULL). */
int pt_server_poll(pt_server *s, int max_ticks, uint64_t *done, double *latency_ms, int max_done, int *n_done);
/* Runs ticks until every submitted frame has completed. */
int pt_server_drain(pt_server *s);
/* Releases a sequence's device state once its queued frames are done. */
int pt_server_close(pt_server *s, int32_t seq);
/* Latency policy (default 2048): each engine tick advances the in-flight
 * frames oldest first, the oldest always, younger ones while the workgroup
 * tiles taken so far are below this value, so a step that fills the GPU runs
 * alone and small (coarse-level) steps are joined by younger frames' steps;
 * 0 advances every in-flight frame each tick (processor sharing). */
#define PT_SERVER_FILL_TILES 5
/* Launch settings of the engine, keys as pt_batch_set: PT_BATCH_XCD_REMAP
 * (default 0 here), PT_BATCH_SHAPE, PT_BATCH_ISOLATE_TILES, and
 * PT_SERVER_FILL_TILES. */
int pt_server_set(pt_server *s, int key, int value);

typedef struct pt_server_stats {
    int64_t submitted, completed, ticks, launches, multi_launches;
    int64_t resident_sequences, waiting_frames, in_flight;
    uint64_t resident_bytes, budget_bytes;
} pt_server_stats;
int pt_server_get_stats(pt_server *s, pt_server_stats *out);

#ifdef __cplusplus
}
#endif

#endif /* PT_API_H */
