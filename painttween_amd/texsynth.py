"""Python view of the TexSynth synthesis interface (texSynth_kernel.h:38-93).

``Context`` wraps one context of a library exporting the pt_api.h C ABI and
exposes the reference's operator names (setup, propagate, scatter, upsample,
syncBuffers, advectOffsets, mergeRandomly, ...).  ``Context.gpu()`` opens the
MI355X HIP library.  Errors raise ``PtError`` with the library's message; the
library never exits the process (the reference's checkCUDAError does,
cudaHostUtil.cpp:46-66).
"""
import ctypes as C
import weakref

import numpy as np

from . import _ffi
from ._ffi import PATCHXF_DTYPE, PtParams


class PtError(RuntimeError):
    def __init__(self, code, message):
        super().__init__("error %d: %s" % (code, message))
        self.code = code


def default_params(lib=None, prefix="pt_", **overrides):
    """TsParameters defaults (G/nlParameters.cpp:63-101) with overrides."""
    lib = lib if lib is not None else _ffi.load_library()
    p = PtParams()
    getattr(lib, prefix + "params_default")(C.byref(p))
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise AttributeError("pt_params has no field %r" % k)
        setattr(p, k, v)
    return p


class Pyramid:
    """Device- (or host-) resident ImagePyramid<T> (imagePyramid.h)."""

    CHANNELS = {1: 1, 2: 2, 3: 3, 4: 4, 5: 1}

    def __init__(self, ctx, handle, elem):
        self.ctx, self.handle, self.elem = ctx, handle, elem
        ctx._pyramids.add(self)

    @property
    def levels(self):
        return self.ctx._fn("pyramid_levels")(self.handle)

    def dims(self, level):
        w, h = C.c_int(), C.c_int()
        self.ctx._check(self.ctx._fn("pyramid_dims")(self.handle, level, C.byref(w), C.byref(h)))
        return w.value, h.value

    def download(self, level):
        w, h = self.dims(level)
        ch = self.CHANNELS[self.elem]
        dt = np.int32 if self.elem == _ffi.PT_ELEM_I32 else np.float32
        out = np.empty((h, w, ch) if ch > 1 else (h, w), dt)
        self.ctx._check(self.ctx._fn("pyramid_download")(self.handle, level, _ffi.vptr(out)))
        return out

    def close(self):
        # A pyramid belongs to its context: the context closes it first.
        if self.handle and self.ctx.handle:
            self.ctx._fn("pyramid_destroy")(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One synthesis context (WorkingBuffers + bound textures + parameters)."""

    def __init__(self, lib, prefix, handle):
        self.lib, self.prefix, self.handle = lib, prefix, handle
        self._pyramids = weakref.WeakSet()

    @classmethod
    def gpu(cls, device=0):
        lib = _ffi.load_library()
        h = C.c_void_p()
        rc = lib.pt_create(C.byref(h), device)
        if rc != 0:
            raise PtError(rc, lib.pt_global_last_error().decode())
        return cls(lib, "pt_", h)

    # -- plumbing -------------------------------------------------------
    def _fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def _check(self, rc):
        if rc != 0:
            msg = self._fn("last_error")(self.handle)
            raise PtError(rc, msg.decode() if msg else "")
        return rc

    def call(self, name, *args):
        return self._check(self._fn(name)(self.handle, *args))

    def close(self):
        if self.handle:
            for p in list(self._pyramids):
                p.close()
            self._fn("destroy")(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def default_params(self, **overrides):
        return default_params(self.lib, self.prefix, **overrides)

    def set_tuning(self, key, value):
        """Execution options of the HIP library (pt_set_tuning)."""
        self._check(self.lib.pt_set_tuning(self.handle, key, value))

    # -- TexSynth:: -----------------------------------------------------
    def setup(self, params):
        self.call("setup", C.byref(params))

    def upload_parameters(self, params):
        self.call("upload_parameters", C.byref(params))

    def initialize(self, image_w, image_h, style_w, style_h):
        self.call("initialize", image_w, image_h, style_w, style_h)

    def initialize_exemplars(self, w, h, layers=1):
        self.call("initialize_exemplars", w, h, layers)

    def reset_advected_buffers(self):
        self.call("reset_advected_buffers")

    def flip_anim_buffers(self):
        self.call("flip_anim_buffers")

    def pyramid(self, elem, level0, scale_mode=_ffi.PT_DOWNSAMPLE_SCALE_CONSTANT, default=0.0, shape=None):
        """Build a pyramid from a level-0 float32 array (H, W[, C]) -- int32
        (H, W) for PT_ELEM_I32 -- or, with level0=None and shape=(W, H), from a
        constant default."""
        if level0 is not None:
            level0 = np.ascontiguousarray(level0, np.int32 if elem == _ffi.PT_ELEM_I32 else np.float32)
            h, w = level0.shape[:2]
        else:
            w, h = shape
        out = C.c_void_p()
        self.call("pyramid_create", C.byref(out), elem, w, h, scale_mode, _ffi.vptr(level0), C.c_float(default))
        return Pyramid(self, out, elem)

    def orientation_pyramid(self, tensor_pyr):
        out = C.c_void_p()
        self.call("pyramid_orientation", C.byref(out), tensor_pyr.handle)
        return Pyramid(self, out, _ffi.PT_ELEM_F32)

    def upload_image(self, level, layer, pyr, style_layer=0):
        layer = _ffi.LAYER[layer] if isinstance(layer, str) else layer
        self.call("upload_image", level, layer, pyr.handle, style_layer)

    def create_random_pyramid(self, level, frame, image_color, image_orient, ex_output, ex_orient):
        self.call("create_random_pyramid", level, frame, image_color.handle, image_orient.handle,
                  ex_output.handle, ex_orient.handle)

    def create_random_pyramid_keys(self, level, frame, image_color, image_orient, ex_outputs, ex_orients):
        """createRandomPyramid with keyframes: one exemplar output and
        orientation pyramid per style."""
        n = len(ex_outputs)
        assert len(ex_orients) == n
        outs = (C.c_void_p * n)(*[p.handle.value for p in ex_outputs])
        oris = (C.c_void_p * n)(*[p.handle.value for p in ex_orients])
        self.call("create_random_pyramid_keys", level, frame, image_color.handle, image_orient.handle, n, outs, oris)

    def initialize_cumulative_fields(self, w, h, layers):
        self.call("initialize_cumulative_fields", w, h, layers)

    def offsets_from_frame_to_key_ribbon(self, level, look_forwards):
        self.call("offsets_from_frame_to_key_ribbon", level, int(look_forwards))

    def merge_advected_buffers(self, level, fallback=0):
        self.call("merge_advected_buffers", level, fallback)

    def accumulate_ribbon_field(self, level, forwards, time_step, overwrite):
        self.call("accumulate_ribbon_field", level, int(forwards), time_step, int(overwrite))

    def load_ribbon_field(self, level, ribbons):
        ribbons = np.ascontiguousarray(ribbons, _ffi.RIBBON_DTYPE)
        self.call("load_ribbon_field", level, _ffi.vptr(ribbons))

    def get_ribbon_field(self, level, w, h):
        out = np.zeros((h, w), _ffi.RIBBON_DTYPE)
        self.call("get_ribbon_field", level, _ffi.vptr(out))
        return out

    def randomize_offsets(self, level):
        self.call("randomize_offsets", level)

    def propagate(self, level, jump, key_index=-1):
        self.call("propagate", level, jump, key_index)

    def scatter(self, level, key_index=-1):
        self.call("scatter", level, key_index)

    def upsample(self, fine_level):
        self.call("upsample", fine_level)

    def sync_buffers(self, level, flags=0):
        self.call("sync_buffers", level, flags)

    def update_residual_cache(self, level):
        self.call("update_residual_cache", level)

    def advect_offsets(self, level, time_step, forwards, flags=0):
        self.call("advect_offsets", level, time_step, int(forwards), flags)

    def merge_randomly(self, level, forwards, weight=0.5, force_valid=False):
        self.call("merge_randomly", level, int(forwards), C.c_float(weight), int(force_valid))

    def merge_using_residual(self, level, forwards):
        self.call("merge_using_residual", level, int(forwards))

    def pde_count(self):
        v = C.c_uint64()
        self.call("pde_count", C.byref(v))
        return v.value

    # -- state in/out ---------------------------------------------------
    def get_output(self, level, w, h):
        off = np.zeros((h, w), PATCHXF_DTYPE)
        canvas = np.zeros((h, w, 4), np.float32)
        res = np.zeros((h, w), np.float32)
        self.call("get_output", level, _ffi.vptr(off), _ffi.vptr(canvas), _ffi.vptr(res))
        return off, canvas, res

    def get_advected(self, level, which, w, h):
        off = np.zeros((h, w), PATCHXF_DTYPE)
        base = np.zeros((h, w, 4), np.float32)
        out = np.zeros((h, w, 4), np.float32)
        res = np.zeros((h, w), np.float32)
        ts = np.zeros((h, w), np.int32)
        self.call("get_advected", level, which, _ffi.vptr(off), _ffi.vptr(base), _ffi.vptr(out),
                  _ffi.vptr(res), _ffi.vptr(ts))
        return off, base, out, res, ts

    def load_offsets(self, level, offsets):
        offsets = np.ascontiguousarray(offsets, PATCHXF_DTYPE)
        self.call("load_offsets", level, _ffi.vptr(offsets))

    def load_residual(self, level, residual):
        residual = np.ascontiguousarray(residual, np.float32)
        self.call("load_residual", level, _ffi.vptr(residual))

    # -- one-shot schedule ----------------------------------------------
    def synthesize(self, params, frames, style, scheme=_ffi.PT_S_FMBM_RM, iterations=3, fixup_passes=1,
                   want=("offsets", "canvas", "residual")):
        """Run a schedule (PT_S_FMBM, _FMBM_RM, _FBMO, _TCTF, _IND) over `frames` (list of dicts with
        keys color, orientation, dist_trans, scale, vel_b, vel_f, object_id; arrays are
        (H, W, C) float32, int32 (H, W) object IDs, or None) against `style` (dict: base,
        output, orientation, dist_trans, object_ids).  Returns level-0 results of the
        final pass, one entry per frame."""
        keep = []

        def arr(a, dt=np.float32):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dt)
            keep.append(a)
            return a

        F = len(frames)
        h, w = frames[0]["color"].shape[:2]
        sh, sw = style["base"].shape[:2]
        fg = (_ffi.FrameGuides * F)()
        for i, f in enumerate(frames):
            fg[i] = _ffi.FrameGuides(_ffi.fptr(arr(f["color"])), _ffi.fptr(arr(f.get("orientation"))),
                                     _ffi.fptr(arr(f.get("dist_trans"))), _ffi.fptr(arr(f.get("scale"))),
                                     _ffi.fptr(arr(f.get("vel_b"))), _ffi.fptr(arr(f.get("vel_f"))),
                                     _ffi.iptr(arr(f.get("object_id"), np.int32)))
        sg = _ffi.StyleGuides(_ffi.fptr(arr(style["base"])), _ffi.fptr(arr(style["output"])),
                              _ffi.fptr(arr(style.get("orientation"))), _ffi.fptr(arr(style.get("dist_trans"))),
                              _ffi.iptr(arr(style.get("object_ids"), np.int32)))
        off = np.zeros((F, h, w), PATCHXF_DTYPE) if "offsets" in want else None
        canvas = np.zeros((F, h, w, 4), np.float32) if "canvas" in want else None
        res = np.zeros((F, h, w), np.float32) if "residual" in want else None
        self.call("synthesize", C.byref(params), scheme, iterations, fixup_passes, w, h, F, fg, sw, sh,
                  C.byref(sg), _ffi.vptr(off), _ffi.vptr(canvas), _ffi.vptr(res))
        return off, canvas, res
