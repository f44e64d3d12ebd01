"""GPU: pt_get_output_noalloc, the readback that stages the packed offsets in
the library's static per-GPU array instead of a per-call allocation, returns
what pt_get_output returns -- on two calls in a row, the second on a smaller
level -- and refuses a level whose offsets exceed the array."""
import ctypes as C

import numpy as np
import pytest

from painttween_amd import Context, _ffi
from painttween_amd import workloads as W
from tests.harness import Session

pytestmark = pytest.mark.gpu


def _noalloc(ctx, level, w, h):
    lib = ctx.lib
    lib.pt_get_output_noalloc.restype = C.c_int
    lib.pt_get_output_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    off = np.zeros((h, w), _ffi.PATCHXF_DTYPE)
    canvas = np.zeros((h, w, 4), np.float32)
    res = np.zeros((h, w), np.float32)
    rc = lib.pt_get_output_noalloc(ctx.handle, level, _ffi.vptr(off), _ffi.vptr(canvas), _ffi.vptr(res))
    return rc, (off, canvas, res)


def _same(a, b, what):
    for name, x, y in zip(("offsets", "canvas", "residual"), a, b):
        assert x.tobytes() == y.tobytes(), "%s: %s differs" % (what, name)


def test_noalloc_readback_equals_get_output_on_two_calls():
    st = W.style_pair(96, 80, seed=91, guides=True)
    fr = W.still_frame(160, 120, seed=92, guides=True)
    with Context.gpu(0) as g:
        s = Session(g, st, [fr], g.default_params(residual_window_size=2, levels=2, lum_shift_enable=1))
        s.run([("style", 1), ("frame", 1, 0), ("randomize",), ("propagate", 4), ("scatter",),
               ("style", 0), ("frame", 0, 0), ("upsample",), ("sync", 0), ("propagate", 2)])
        # level 0 (160 x 120), then the smaller level 1 (80 x 60)
        for level in (0, 1):
            w, h = 160 >> level, 120 >> level
            rc, got = _noalloc(g, level, w, h)
            assert rc == 0, g.lib.pt_last_error(g.handle)
            want = g.get_output(level, w, h)
            _same(got, want, "level %d" % level)
            assert np.any(got[0]["x"] != 0)
        # the array keeps the last call's contents: a repeat of the larger
        # level overwrites every staged element again
        rc, again = _noalloc(g, 0, 160, 120)
        assert rc == 0
        _same(again, g.get_output(0, 160, 120), "level 0 again")


def test_noalloc_readback_refuses_levels_above_the_array():
    limit = _ffi.PT_NOALLOC_SCRATCH_BYTES // 36
    w, h = 2048, 2048
    assert w * h > limit >= (w >> 1) * (h >> 1)
    with Context.gpu(0) as g:
        g.setup(g.default_params(residual_window_size=2, levels=2))
        g.initialize(w, h, 64, 64)
        rc, _ = _noalloc(g, 0, w, h)
        assert rc == _ffi.PT_ERR_UNSUPPORTED
        assert b"PT_NOALLOC_SCRATCH_BYTES" in g.lib.pt_last_error(g.handle)
        rc, got = _noalloc(g, 1, w >> 1, h >> 1)   # a level that fits: zeroed fresh buffers
        assert rc == 0
        _same(got, g.get_output(1, w >> 1, h >> 1), "level 1 of a fresh context")
