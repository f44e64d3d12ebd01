"""Loader for the CPU oracle (oracle/libptoracle.so) -- test infrastructure.

The oracle exports the same C ABI as the HIP library under the ``ptref_``
prefix, so ``OracleContext`` is a ``painttween_amd.Context`` bound to it.
"""
import ctypes as C
import os

import numpy as np

from painttween_amd import _ffi
from painttween_amd.texsynth import Context, PtError

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "libptoracle.so")

_LIB = None


def load_oracle():
    global _LIB
    if _LIB is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError("oracle/libptoracle.so is missing: run __graft_entry__.build()")
        lib = C.CDLL(ORACLE_PATH)
        _ffi.bind(lib, "ptref_")
        lib.ptref_create.restype = C.c_int
        lib.ptref_create.argtypes = [C.POINTER(C.c_void_p), C.c_int]
        lib.ptref_patch_distance.restype = C.c_float
        lib.ptref_patch_distance.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(_ffi.PatchXF)]
        lib.ptref_screen_to_exemplar_angle.restype = C.c_float
        lib.ptref_screen_to_exemplar_angle.argtypes = [C.c_void_p] + [C.c_float] * 4
        lib.ptref_gaussian_weights.restype = C.c_int
        lib.ptref_gaussian_weights.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_int]
        lib.ptref_sincosf.restype = None
        lib.ptref_sincosf.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        lib.ptref_atan2f.restype = C.c_float
        lib.ptref_atan2f.argtypes = [C.c_float, C.c_float]
        for f in ("ptref_logf", "ptref_acosf"):
            getattr(lib, f).restype = C.c_float
            getattr(lib, f).argtypes = [C.c_float]
        for f in ("ptref_log_d", "ptref_exp_d"):
            getattr(lib, f).restype = C.c_double
            getattr(lib, f).argtypes = [C.c_double]
        lib.ptref_rgb_to_lab.restype = None
        lib.ptref_rgb_to_lab.argtypes = [C.c_float, C.c_float, C.c_float, C.POINTER(C.c_float)]
        lib.ptref_luminance_delta.restype = C.c_float
        lib.ptref_luminance_delta.argtypes = [C.c_float] * 6
        lib.ptref_coherence_term.restype = C.c_float
        lib.ptref_coherence_term.argtypes = [C.c_float] * 6 + [C.c_int] * 3 + [C.c_float] * 2
        lib.ptref_tensor_to_angle.restype = C.c_float
        lib.ptref_tensor_to_angle.argtypes = [C.c_float, C.c_float, C.c_float]
        lib.ptref_synthesize_next_frame.restype = C.c_int
        lib.ptref_synthesize_next_frame.argtypes = [C.c_void_p, C.POINTER(_ffi.PtParams), C.c_int, C.c_int, C.c_int,
                                                    C.POINTER(_ffi.FrameGuides), C.c_int, C.c_int,
                                                    C.POINTER(_ffi.StyleGuides), C.c_void_p, C.c_void_p, C.c_void_p]
        lib.ptref_sequence_reset.restype = C.c_int
        lib.ptref_sequence_reset.argtypes = [C.c_void_p]
        lib.ptref_philox.restype = None
        lib.ptref_philox.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        _LIB = lib
    return _LIB


class OracleContext(Context):
    def __init__(self, num_threads=0):
        lib = load_oracle()
        h = C.c_void_p()
        rc = lib.ptref_create(C.byref(h), num_threads)
        if rc != 0:
            raise PtError(rc, "ptref_create failed")
        super().__init__(lib, "ptref_", h)

    def patch_distance(self, x, y, level, cand):
        p = _ffi.PatchXF(*[cand[k] for k in ("x", "y", "theta", "hysteresis", "layer", "scale_u",
                                               "scale_v", "lum_y", "lum_a")])
        return self.lib.ptref_patch_distance(self.handle, x, y, level, C.byref(p))

    def next_frame(self, params, iterations, frame, style):
        """ptref_synthesize_next_frame on host arrays; returns (offsets,
        canvas, residual) of the frame."""
        keep = []

        def arr(a, dt=np.float32):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dt)
            keep.append(a)
            return a

        h, w = frame["color"].shape[:2]
        sh, sw = style["base"].shape[:2]
        fg = _ffi.FrameGuides(_ffi.fptr(arr(frame["color"])), _ffi.fptr(arr(frame.get("orientation"))),
                              _ffi.fptr(arr(frame.get("dist_trans"))), _ffi.fptr(arr(frame.get("scale"))),
                              _ffi.fptr(arr(frame.get("vel_b"))), _ffi.fptr(arr(frame.get("vel_f"))),
                              _ffi.iptr(arr(frame.get("object_id"), np.int32)))
        sg = _ffi.StyleGuides(_ffi.fptr(arr(style["base"])), _ffi.fptr(arr(style["output"])),
                              _ffi.fptr(arr(style.get("orientation"))), _ffi.fptr(arr(style.get("dist_trans"))),
                              _ffi.iptr(arr(style.get("object_ids"), np.int32)))
        off = np.zeros((h, w), _ffi.PATCHXF_DTYPE)
        canvas = np.zeros((h, w, 4), np.float32)
        res = np.zeros((h, w), np.float32)
        self._check(self.lib.ptref_synthesize_next_frame(self.handle, C.byref(params), iterations, w, h, C.byref(fg),
                                                         sw, sh, C.byref(sg), _ffi.vptr(off), _ffi.vptr(canvas),
                                                         _ffi.vptr(res)))
        return off, canvas, res

    def screen_to_exemplar_angle(self, x, y, ex, ey):
        return self.lib.ptref_screen_to_exemplar_angle(self.handle, x, y, ex, ey)

    def gaussian_weights(self, r):
        out = (C.c_float * (2 * r + 1))()
        self._check(self.lib.ptref_gaussian_weights(self.handle, out, 2 * r + 1))
        return list(out)
