"""ctypes mirror of include/pt_api.h.

The structs here must match the C definitions field for field; a CPU test
checks their sizes against the values the library reports for itself.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpainttween.so")

PT_OK = 0
PT_ERR_INVALID_ARG, PT_ERR_UNSUPPORTED, PT_ERR_HIP, PT_ERR_OUT_OF_MEMORY, PT_ERR_STATE, PT_ERR_BUDGET = \
    -1, -2, -3, -4, -5, -6
PT_MAX_EXEMPLAR_TEXELS = 67108863
PT_MAX_IMAGE_PIXELS = 2147483647
PT_NOALLOC_SCRATCH_BYTES = 128 << 20
PT_FORWARD, PT_BACKWARD, PT_BIDIRECTIONAL = 0, 1, 2
PT_S_FMBM, PT_S_FMBM_RM, PT_S_FBMO, PT_S_TCTF, PT_S_IND = 0, 1, 2, 3, 4
PT_RNG_REFERENCE, PT_RNG_COUNTER = 0, 1
PT_DONT_COPY_OFFSETS = 0x01
PT_DONT_UPDATE_RESIDUAL = 0x10
PT_ADVECT_IGNORE_RANDOM = 0x1
PT_ADVECT_DO_NOT_CHECK_PREVIOUS = 0x10
PT_ADVECT_UPDATE_RESIDUAL = 0x80
PT_ADVECT_CHECK_FRAME_TO_KEY = 0x20
PT_MERGE_FALLBACK_RANDOM = 0x20
PT_MERGE_FALLBACK_LAST_PASS = 0x40
PT_MAX_STYLE_LAYERS = 4
PT_ELEM_F32, PT_ELEM_F32X2, PT_ELEM_F32X3, PT_ELEM_F32X4, PT_ELEM_I32 = 1, 2, 3, 4, 5
PT_TUNE_EARLY_EXIT = 1
PT_TUNE_EARLY_EXIT_MIN_PIXELS = 2
PT_TUNE_UNIT_PARALLEL = 3
PT_TUNE_INCREMENTAL_RESIDUAL = 4
PT_TUNE_TAP_PARALLEL_MAX_PIXELS = 5
PT_TUNE_TAP_PARALLEL_EVALS = 6
PT_TUNE_EXEMPLAR_WINDOW = 7
PT_TUNE_TIME_TABLE_MAX_TAPS = 8
PT_TUNE_XCD_TILES = 9
PT_TUNE_PROPAGATE_WINDOW_JUMP = 10
PT_TUNE_PROPAGATE_WINDOW_MIN_PIXELS = 11
PT_TUNE_TEMPORAL_UNIT_ORDER = 12
PT_SERVER_FILL_TILES = 5
PT_DOWNSAMPLE_SCALE_WITH_LEVELS, PT_DOWNSAMPLE_SCALE_CONSTANT, PT_DOWNSAMPLE_RIBBON = 0, 1, 2

LAYER = dict(
    EXEMPLAR_BASE=0, EXEMPLAR_OUTPUT=1, EXEMPLAR_ORIENTATION=2, EXEMPLAR_DIST_TRANS=3, EXEMPLAR_OBJECT_IDS=4,
    INPUT_COLOR=5, INPUT_ORIENTATION=6, INPUT_DIST_TRANS=7, INPUT_SCALE=8, INPUT_ID=9,
    INPUT_VEL_B=11, INPUT_VEL_F=12, INPUT_VEL_B_PREVIOUS=13, INPUT_VEL_F_PREVIOUS=14,
    CUMUL_FIELD_B=17, CUMUL_FIELD_F=18, FRAME_TO_KEY_RIBBON_B=19, FRAME_TO_KEY_RIBBON_F=20,
    RANDOM=21,
)


class PtParams(C.Structure):
    _fields_ = [
        ("residual_window_size", C.c_int32),
        ("coherence_weight", C.c_float),
        ("dist_trans_weight", C.c_float),
        ("lum_shift_weight", C.c_float),
        ("lum_shift_enable", C.c_int32),
        ("input_analogy_weight", C.c_float),
        ("canvas_output_weight", C.c_float),
        ("num_styles", C.c_int32),
        ("offsets_histogram_slope", C.c_float * 5),
        ("offsets_histogram_threshold", C.c_float * 5),
        ("max_distortion", C.c_float),
        ("alpha_weight", C.c_float),
        ("transparency_ok", C.c_int32),
        ("style_obj_id_weight", C.c_float),
        ("orientation_enabled", C.c_int32),
        ("colorspace", C.c_int32),
        ("coherence_ratio_range", C.c_float),
        ("coherence_angle_range", C.c_float),
        ("levels", C.c_int32),
        ("scatter_samples_count", C.c_int32),
        ("advection_jitter", C.c_float),
        ("time_derivative_input_weight", C.c_float),
        ("time_derivative_output_weight", C.c_float),
        ("temporal_coherence_weight", C.c_float),
        ("hysteresis_weight", C.c_float),
        ("interpolate_key_frame", C.c_int32),
        ("use_guide", C.c_int32),
        ("current_interpolation_index", C.c_int32),
        ("num_frames", C.c_int32),
        ("num_key_frames", C.c_int32),
        ("first_key_index", C.c_int32),
        ("last_key_index", C.c_int32),
        ("dist_trans_cull", C.c_int32),
        ("residual_window_scale_u", C.c_float),
        ("residual_window_scale_v", C.c_float),
        ("direction", C.c_int32),
        ("rng_mode", C.c_int32),
        ("seed", C.c_uint32),
        ("frame_number", C.c_int32),
        ("first_pass_direction", C.c_int32),
        ("max_level", C.c_int32),
    ]


class PatchXF(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("theta", C.c_float), ("hysteresis", C.c_float),
                ("layer", C.c_int32), ("scale_u", C.c_float), ("scale_v", C.c_float),
                ("lum_y", C.c_float), ("lum_a", C.c_float)]


PATCHXF_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("theta", "<f4"), ("hysteresis", "<f4"),
                          ("layer", "<i4"), ("scale_u", "<f4"), ("scale_v", "<f4"),
                          ("lum_y", "<f4"), ("lum_a", "<f4")])
assert PATCHXF_DTYPE.itemsize == C.sizeof(PatchXF) == 36

# pt_ribbonp: a pixel's ribbon to a keyframe (RibbonP, types.h:116-126)
RIBBON_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("time_step", "<i4"), ("layer", "<i4")])

FP = C.POINTER(C.c_float)
IP = C.POINTER(C.c_int32)


class FrameGuides(C.Structure):
    _fields_ = [("color_rgba", FP), ("orientation_xx_xy_yy", FP), ("dist_trans", FP),
                ("scale_uv", FP), ("vel_b", FP), ("vel_f", FP), ("object_id", IP)]


class StyleGuides(C.Structure):
    _fields_ = [("base_rgba", FP), ("output_rgba", FP), ("orientation_xx_xy_yy", FP), ("dist_trans", FP),
                ("object_ids", IP)]


class PtJob(C.Structure):
    _fields_ = [("params", PtParams), ("scheme", C.c_int32), ("iterations", C.c_int32),
                ("fixup_passes", C.c_int32), ("image_w", C.c_int32), ("image_h", C.c_int32),
                ("num_frames", C.c_int32), ("frames", C.POINTER(FrameGuides)), ("style_w", C.c_int32),
                ("style_h", C.c_int32), ("style", C.POINTER(StyleGuides)), ("out_offsets", C.c_void_p),
                ("out_canvas_rgba", C.c_void_p), ("out_residual", C.c_void_p)]


class BatchStats(C.Structure):
    _fields_ = [("jobs", C.c_int64), ("ops", C.c_int64), ("launches", C.c_int64),
                ("multi_launches", C.c_int64), ("ticks", C.c_int64), ("last_run_ms", C.c_double),
                ("tune_ms", C.c_double), ("shape", (C.c_int32 * 3) * 2), ("xcd_remap", C.c_int32),
                ("isolate_tiles", C.c_int32), ("persistent", C.c_int32), ("persistent_grid", C.c_int32),
                ("persistent_tasks", C.c_int64), ("persistent_tiles", C.c_int64),
                ("persistent_max_chunk", C.c_int32), ("persistent_spanned", C.c_int32)]


class SequenceDesc(C.Structure):
    _fields_ = [("params", PtParams), ("iterations", C.c_int32), ("image_w", C.c_int32),
                ("image_h", C.c_int32), ("style_w", C.c_int32), ("style_h", C.c_int32),
                ("style", C.POINTER(StyleGuides))]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("ticks", C.c_int64),
                ("launches", C.c_int64), ("multi_launches", C.c_int64), ("resident_sequences", C.c_int64),
                ("waiting_frames", C.c_int64), ("in_flight", C.c_int64), ("resident_bytes", C.c_uint64),
                ("budget_bytes", C.c_uint64)]


# struct ids of pt_struct_size()
STRUCTS = {0: PtParams, 1: PatchXF, 2: FrameGuides, 3: StyleGuides, 4: PtJob, 5: BatchStats,
           6: SequenceDesc, 7: ServerStats}

PT_BATCH_USE_GRAPH, PT_BATCH_XCD_REMAP, PT_BATCH_SHAPE, PT_BATCH_ISOLATE_TILES = 1, 2, 3, 4
PT_BATCH_PERSISTENT, PT_BATCH_PERSIST_CHUNK, PT_BATCH_PERSIST_TIMEOUT_MS, PT_BATCH_PERSIST_ABORT = 5, 6, 7, 8
PT_OP_PROPAGATE, PT_OP_SCATTER = 0, 1
SHAPE_NAMES = ["pixel", "pixel_ee", "unit_ee_a", "unit_a", "unit_ee_b", "unit_b", "tap"]


def fptr(a):
    """float* into a C-contiguous float32 array, or NULL for None."""
    if a is None:
        return FP()
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(FP)


def iptr(a):
    """int32* into a C-contiguous int32 array, or NULL for None."""
    if a is None:
        return IP()
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(IP)


def vptr(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


VP = C.c_void_p
I = C.c_int
F = C.c_float

# (name, restype, argtypes) with the prefix left off; shared by the HIP
# library ("pt_") and the CPU oracle ("ptref_").
SIGNATURES = [
    ("setup", I, [VP, C.POINTER(PtParams)]),
    ("upload_parameters", I, [VP, C.POINTER(PtParams)]),
    ("initialize", I, [VP, I, I, I, I]),
    ("initialize_exemplars", I, [VP, I, I, I]),
    ("reset_advected_buffers", I, [VP]),
    ("flip_anim_buffers", I, [VP]),
    ("pyramid_create", I, [VP, C.POINTER(VP), I, I, I, I, VP, F]),
    ("pyramid_orientation", I, [VP, C.POINTER(VP), VP]),
    ("pyramid_destroy", I, [VP]),
    ("pyramid_levels", I, [VP]),
    ("pyramid_dims", I, [VP, I, C.POINTER(I), C.POINTER(I)]),
    ("pyramid_download", I, [VP, I, VP]),
    ("upload_image", I, [VP, I, I, VP, I]),
    ("create_random_pyramid", I, [VP, I, I, VP, VP, VP, VP]),
    ("randomize_offsets", I, [VP, I]),
    ("propagate", I, [VP, I, I, I]),
    ("scatter", I, [VP, I, I]),
    ("upsample", I, [VP, I]),
    ("sync_buffers", I, [VP, I, I]),
    ("update_residual_cache", I, [VP, I]),
    ("advect_offsets", I, [VP, I, I, I, I]),
    ("merge_randomly", I, [VP, I, I, F, I]),
    ("merge_using_residual", I, [VP, I, I]),
    ("get_output", I, [VP, I, VP, VP, VP]),
    ("load_offsets", I, [VP, I, VP]),
    ("load_residual", I, [VP, I, VP]),
    ("get_advected", I, [VP, I, I, VP, VP, VP, VP, VP]),
    ("pde_count", I, [VP, C.POINTER(C.c_uint64)]),
    ("initialize_cumulative_fields", I, [VP, I, I, I]),
    ("create_random_pyramid_keys", I, [VP, I, I, VP, VP, I, C.POINTER(VP), C.POINTER(VP)]),
    ("offsets_from_frame_to_key_ribbon", I, [VP, I, I]),
    ("merge_advected_buffers", I, [VP, I, I]),
    ("accumulate_ribbon_field", I, [VP, I, I, I, I]),
    ("load_ribbon_field", I, [VP, I, VP]),
    ("get_ribbon_field", I, [VP, I, VP]),
    ("synthesize", I, [VP, C.POINTER(PtParams), I, I, I, I, I, I, C.POINTER(FrameGuides), I, I,
                       C.POINTER(StyleGuides), VP, VP, VP]),
    ("last_error", C.c_char_p, [VP]),
    ("destroy", I, [VP]),
]


def bind(lib, prefix):
    for name, res, args in SIGNATURES:
        fn = getattr(lib, prefix + name)
        fn.restype = res
        fn.argtypes = args
    getattr(lib, prefix + "params_default").restype = None
    getattr(lib, prefix + "params_default").argtypes = [C.POINTER(PtParams)]
    return lib


_LIB = None


def load_library():
    """Load the in-tree HIP library; fail loudly if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "painttween_amd/libpainttween.so is missing: build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'`")
        # One HIP runtime per process: when PyTorch is installed, let it load
        # its HIP runtime first so that the library and torch's device
        # allocations (used for device-resident inputs) share it, whichever
        # of the two the caller touches first.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = C.CDLL(LIB_PATH)
        bind(lib, "pt_")
        lib.pt_create.restype = I
        lib.pt_create.argtypes = [C.POINTER(VP), I]
        lib.pt_global_last_error.restype = C.c_char_p
        lib.pt_global_last_error.argtypes = []
        lib.pt_synthesize_device.restype = I
        lib.pt_synthesize_device.argtypes = [VP, C.POINTER(PtParams), I, I, I, I, I, I, C.POINTER(FrameGuides),
                                             I, I, C.POINTER(StyleGuides), VP, VP, VP]
        U64 = C.c_uint64
        extra = [
            ("pt_synthesize_next_frame", I, [VP, C.POINTER(PtParams), I, I, I, C.POINTER(FrameGuides), I, I,
                                             C.POINTER(StyleGuides), VP, VP, VP]),
            ("pt_sequence_reset", I, [VP]),
            ("pt_sequence_frame", I, [VP, C.POINTER(I)]),
            ("pt_struct_size", C.c_size_t, [I]),
            ("pt_batch_create", I, [C.POINTER(VP), I, I, C.POINTER(PtJob)]),
            ("pt_batch_destroy", I, [VP]),
            ("pt_batch_last_error", C.c_char_p, [VP]),
            ("pt_batch_run", I, [VP]),
            ("pt_batch_tune", I, [VP, I]),
            ("pt_batch_set", I, [VP, I, I]),
            ("pt_batch_get_stats", I, [VP, C.POINTER(BatchStats)]),
            ("pt_submit_batch", I, [I, I, C.POINTER(PtJob)]),
            ("pt_server_create", I, [C.POINTER(VP), I, U64, I]),
            ("pt_server_destroy", I, [VP]),
            ("pt_server_last_error", C.c_char_p, [VP]),
            ("pt_server_open", I, [VP, C.POINTER(SequenceDesc), C.POINTER(C.c_int32)]),
            ("pt_server_submit", I, [VP, C.c_int32, C.POINTER(FrameGuides), VP, VP, VP, C.POINTER(U64)]),
            ("pt_server_poll", I, [VP, I, C.POINTER(U64), C.POINTER(C.c_double), I, C.POINTER(I)]),
            ("pt_server_drain", I, [VP]),
            ("pt_server_close", I, [VP, C.c_int32]),
            ("pt_server_get_stats", I, [VP, C.POINTER(ServerStats)]),
            ("pt_server_set", I, [VP, I, I]),
        ]
        for name, res, args in extra:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        lib.pt_set_tuning.restype = I
        lib.pt_set_tuning.argtypes = [VP, I, I]
        lib.pt_profile_enable.restype = I
        lib.pt_profile_enable.argtypes = [VP, I]
        lib.pt_profile_bytes.restype = I
        lib.pt_profile_bytes.argtypes = [VP, C.c_char_p, C.POINTER(C.c_uint64)]
        lib.pt_profile_records.restype = I
        lib.pt_profile_records.argtypes = [VP, C.c_char_p, I, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                           C.POINTER(C.c_float), C.POINTER(I)]
        lib.pt_profile_get.restype = I
        lib.pt_profile_get.argtypes = [VP, C.c_char_p, C.POINTER(C.c_uint64), C.POINTER(C.c_double),
                                       C.POINTER(C.c_uint64)]
        _LIB = lib
    return _LIB
