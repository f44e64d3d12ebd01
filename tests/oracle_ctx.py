"""Loader for the CPU oracle (oracle/libptoracle.so) -- test infrastructure.

The oracle exports the same C ABI as the HIP library under the ``ptref_``
prefix, so ``OracleContext`` is a ``painttween_amd.Context`` bound to it.
"""
import ctypes as C
import os

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
        lib.ptref_gaussian_weights.restype = C.c_int
        lib.ptref_gaussian_weights.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_int]
        lib.ptref_sincosf.restype = None
        lib.ptref_sincosf.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        lib.ptref_atan2f.restype = C.c_float
        lib.ptref_atan2f.argtypes = [C.c_float, C.c_float]
        lib.ptref_tensor_to_angle.restype = C.c_float
        lib.ptref_tensor_to_angle.argtypes = [C.c_float, C.c_float, C.c_float]
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

    def gaussian_weights(self, r):
        out = (C.c_float * (2 * r + 1))()
        self._check(self.lib.ptref_gaussian_weights(self.handle, out, 2 * r + 1))
        return list(out)
