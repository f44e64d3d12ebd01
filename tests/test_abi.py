"""The C ABI: the built HIP library loads without a GPU, exports every entry
point include/pt_api.h declares, and agrees with the ctypes mirror and the
oracle on struct layouts and parameter defaults."""
import ctypes as C
import os
import re

import pytest

from painttween_amd import _ffi
from tests.oracle_ctx import load_oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pt_api.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pt_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_texsynth_surface():
    names = declared_functions()
    for n in ("pt_setup", "pt_upload_parameters", "pt_initialize", "pt_initialize_exemplars", "pt_upload_image",
              "pt_flip_anim_buffers", "pt_randomize_offsets", "pt_propagate", "pt_scatter", "pt_upsample",
              "pt_sync_buffers", "pt_update_residual_cache", "pt_advect_offsets", "pt_merge_randomly",
              "pt_merge_using_residual", "pt_get_output", "pt_load_offsets", "pt_synthesize"):
        assert n in names, n


def test_library_exports_every_declared_symbol():
    lib = _ffi.load_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts():
    lib = _ffi.load_library()
    lib.pt_struct_size.restype = C.c_size_t
    lib.pt_struct_size.argtypes = [C.c_int]
    assert C.sizeof(_ffi.PtParams) == 49 * 4 == lib.pt_struct_size(0)  # 39 scalars + two float[5]
    assert C.sizeof(_ffi.PatchXF) == 36 == lib.pt_struct_size(1)      # PatchXF, types.h:98-114
    assert C.sizeof(_ffi.FrameGuides) == lib.pt_struct_size(2)
    assert C.sizeof(_ffi.StyleGuides) == lib.pt_struct_size(3)
    for which, cls in _ffi.STRUCTS.items():
        assert C.sizeof(cls) == lib.pt_struct_size(which), cls.__name__
    assert lib.pt_struct_size(99) == 0


def test_defaults_match_reference_and_oracle():
    lib = _ffi.load_library()
    o = load_oracle()
    a, b = _ffi.PtParams(), _ffi.PtParams()
    lib.pt_params_default(C.byref(a))
    o.ptref_params_default(C.byref(b))
    assert bytes(a) == bytes(b)
    # G/nlParameters.cpp:63-101 and synthesisProcessor.cpp:220
    assert a.residual_window_size == 4 and a.levels == 4 and a.scatter_samples_count == 25
    assert a.dist_trans_cull == 50 and a.orientation_enabled == 1 and a.transparency_ok == 0
    assert a.temporal_coherence_weight == pytest.approx(0.2) and a.max_distortion == 4.0


def test_errors_are_returned_not_fatal():
    lib = _ffi.load_library()
    h = C.c_void_p()
    rc = lib.pt_create(C.byref(h), 99)   # no such device here or on a one-GPU box
    assert rc < 0
    assert lib.pt_global_last_error()
    assert lib.pt_propagate(None, 0, 1, -1) < 0
    # batch and server entry points reject bad arguments without a GPU call
    b = C.c_void_p()
    assert lib.pt_batch_create(C.byref(b), 0, 0, None) < 0
    assert lib.pt_batch_run(None) < 0
    assert lib.pt_batch_last_error(None)
    s = C.c_void_p()
    assert lib.pt_server_create(C.byref(s), 0, 0, 1) < 0
    assert lib.pt_server_submit(None, 0, None, None, None, None, None) < 0


def test_batch_and_server_surface_declared():
    names = declared_functions()
    for n in ("pt_batch_create", "pt_batch_run", "pt_batch_tune", "pt_submit_batch", "pt_server_create",
              "pt_server_open", "pt_server_submit", "pt_server_poll", "pt_server_drain", "pt_server_close",
              "pt_synthesize_next_frame", "pt_sequence_reset"):
        assert n in names, n


def test_exemplar_and_image_size_limits_are_refused_before_any_device_work():
    """The exemplar gathers use 32-bit byte offsets over 32-byte texels, so an
    exemplar of PT_MAX_EXEMPLAR_TEXELS + 1 texels (8192 x 8192 is above it) is
    refused with PT_ERR_UNSUPPORTED; the check runs before the context is
    touched, so it needs no GPU."""
    lib = _ffi.load_library()
    lib.pt_global_last_error.restype = C.c_char_p
    assert _ffi.PT_MAX_EXEMPLAR_TEXELS == (2 ** 31 - 1) // 32
    assert lib.pt_initialize_exemplars(None, 8192, 8192, 1) == _ffi.PT_ERR_UNSUPPORTED
    assert b"PT_MAX_EXEMPLAR_TEXELS" in lib.pt_global_last_error()
    assert lib.pt_initialize_exemplars(None, 8192, 8193, 1) == _ffi.PT_ERR_UNSUPPORTED
    assert lib.pt_initialize(None, 64, 64, 8192, 8192) == _ffi.PT_ERR_UNSUPPORTED
    assert lib.pt_initialize(None, 65536, 32768, 64, 64) == _ffi.PT_ERR_UNSUPPORTED
    assert b"PT_MAX_IMAGE_PIXELS" in lib.pt_global_last_error()
    # sizes inside the limits reach the context check; 8191 x 8193 is the limit exactly
    assert 8191 * 8193 == _ffi.PT_MAX_EXEMPLAR_TEXELS
    assert lib.pt_initialize_exemplars(None, 8191, 8193, 1) == _ffi.PT_ERR_INVALID_ARG
    assert lib.pt_initialize(None, 64, 64, 64, 64) == _ffi.PT_ERR_INVALID_ARG
