"""Batched and serving forms, and streamed frames, over the C ABI.

Guides and outputs live in device memory (torch tensors are used only as
device allocations); the HIP library does all the work:

* :class:`Batch` -- many independent synthesis jobs planned together; the
  same step of every job runs as one multi-job launch, captured in a HIP
  graph (``pt_batch_*``).
* :class:`Server` -- a request loop over streamed sequences with memory
  admission and continuous batching (``pt_server_*``).
* :func:`next_frame` -- the plain per-frame path
  (``pt_synthesize_next_frame``) that the server's results equal.
"""
import ctypes as C

import numpy as np

from . import _ffi
from ._ffi import FrameGuides, StyleGuides, PATCHXF_DTYPE
from .texsynth import PtError


def _torch():
    import torch
    return torch


class DeviceImages:
    """Guide arrays copied to the GPU; keeps the tensors alive."""

    def __init__(self, device=0):
        torch = _torch()
        self.dev = torch.device("cuda:%d" % device)
        self.keep = []

    def ptr(self, a, dt=np.float32, kind=_ffi.FP):
        if a is None:
            return kind()
        torch = _torch()
        t = torch.from_numpy(np.ascontiguousarray(a, dt)).to(self.dev)
        self.keep.append(t)
        return C.cast(C.c_void_p(t.data_ptr()), kind)

    def ids(self, a):
        return self.ptr(a, np.int32, _ffi.IP)

    def frame(self, f):
        return FrameGuides(self.ptr(f["color"]), self.ptr(f.get("orientation")), self.ptr(f.get("dist_trans")),
                           self.ptr(f.get("scale")), self.ptr(f.get("vel_b")), self.ptr(f.get("vel_f")),
                           self.ids(f.get("object_id")))

    def style(self, s):
        return StyleGuides(self.ptr(s["base"]), self.ptr(s["output"]), self.ptr(s.get("orientation")),
                           self.ptr(s.get("dist_trans")), self.ids(s.get("object_ids")))


class DeviceOutputs:
    """Offsets (PatchXF), canvas (RGBA) and residual of `frames` frames of
    w x h pixels in device memory."""

    def __init__(self, w, h, frames=1, device=0):
        torch = _torch()
        dev = torch.device("cuda:%d" % device)
        n = w * h * frames
        self.w, self.h, self.frames = w, h, frames
        self.off = torch.zeros(n * PATCHXF_DTYPE.itemsize, dtype=torch.uint8, device=dev)
        self.canvas = torch.zeros(n * 4, dtype=torch.float32, device=dev)
        self.res = torch.zeros(n, dtype=torch.float32, device=dev)

    def pointers(self):
        return (C.c_void_p(self.off.data_ptr()), C.c_void_p(self.canvas.data_ptr()),
                C.c_void_p(self.res.data_ptr()))

    def numpy(self):
        F, h, w = self.frames, self.h, self.w
        off = self.off.cpu().numpy().view(PATCHXF_DTYPE).reshape(F, h, w)
        return off, self.canvas.cpu().numpy().reshape(F, h, w, 4), self.res.cpu().numpy().reshape(F, h, w)


def next_frame(ctx, params, iterations, frame, style, style_w, style_h, outputs):
    """pt_synthesize_next_frame on `ctx` with device guides (FrameGuides and
    StyleGuides from DeviceImages) into DeviceOutputs."""
    o, c, r = outputs.pointers()
    ctx._check(ctx.lib.pt_synthesize_next_frame(ctx.handle, C.byref(params), iterations, outputs.w, outputs.h,
                                                C.byref(frame), style_w, style_h, C.byref(style), o, c, r))


class Batch:
    """A planned batch of jobs.  `jobs`: list of dicts with keys params
    (PtParams), frames (list of frame dicts), style (dict), and optional
    scheme, iterations, fixup_passes."""

    def __init__(self, jobs, device=0):
        self.lib = _ffi.load_library()
        self.handle = C.c_void_p()
        self.images = DeviceImages(device)
        self.outputs = []
        self._frames = []
        self._styles = []
        arr = (_ffi.PtJob * len(jobs))()
        for i, j in enumerate(jobs):
            h, w = j["frames"][0]["color"].shape[:2]
            sh, sw = j["style"]["base"].shape[:2]
            F = len(j["frames"])
            fg = (FrameGuides * F)()
            for k, f in enumerate(j["frames"]):
                fg[k] = self.images.frame(f)
            sg = self.images.style(j["style"])
            self._frames.append(fg)
            self._styles.append(sg)
            out = DeviceOutputs(w, h, F, device)
            self.outputs.append(out)
            o, c, r = out.pointers()
            arr[i] = _ffi.PtJob(j["params"], j.get("scheme", _ffi.PT_S_FMBM_RM), j.get("iterations", 3),
                                j.get("fixup_passes", 0), w, h, F, fg, sw, sh, C.pointer(sg), o, c, r)
        _torch().cuda.synchronize()
        rc = self.lib.pt_batch_create(C.byref(self.handle), device, len(jobs), arr)
        if rc != 0:
            raise PtError(rc, self.lib.pt_batch_last_error(None).decode())

    def _check(self, rc):
        if rc != 0:
            raise PtError(rc, self.lib.pt_batch_last_error(self.handle).decode())

    def run(self):
        self._check(self.lib.pt_batch_run(self.handle))

    def tune(self, rounds=1):
        self._check(self.lib.pt_batch_tune(self.handle, rounds))

    def set(self, key, value):
        self._check(self.lib.pt_batch_set(self.handle, key, value))

    def stats(self):
        st = _ffi.BatchStats()
        self._check(self.lib.pt_batch_get_stats(self.handle, C.byref(st)))
        shapes = {}
        for op, name in ((0, "propagate"), (1, "scatter")):
            shapes[name] = [(_ffi.SHAPE_NAMES[s] if s >= 0 else "default") for s in st.shape[op]]
        return dict(jobs=st.jobs, ops=st.ops, launches=st.launches, multi_launches=st.multi_launches,
                    ticks=st.ticks, last_run_ms=st.last_run_ms, tune_ms=st.tune_ms, shapes=shapes,
                    xcd_remap=st.xcd_remap, isolate_tiles=st.isolate_tiles, persistent=st.persistent,
                    persistent_grid=st.persistent_grid, persistent_tasks=st.persistent_tasks,
                    persistent_tiles=st.persistent_tiles, persistent_max_chunk=st.persistent_max_chunk,
                    persistent_spanned=st.persistent_spanned)

    def result(self, j):
        return self.outputs[j].numpy()

    def close(self):
        if self.handle:
            self.lib.pt_batch_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """The serving request loop (pt_server_*)."""

    def __init__(self, budget_bytes, max_in_flight, device=0):
        self.lib = _ffi.load_library()
        self.device = device
        self.handle = C.c_void_p()
        self.images = DeviceImages(device)
        self.seqs = {}
        self.outputs = {}
        self._keep = []
        rc = self.lib.pt_server_create(C.byref(self.handle), device, int(budget_bytes), max_in_flight)
        if rc != 0:
            raise PtError(rc, self.lib.pt_server_last_error(None).decode())

    def _check(self, rc):
        if rc != 0:
            raise PtError(rc, self.lib.pt_server_last_error(self.handle).decode())

    def open(self, params, iterations, w, h, style):
        sg = style if isinstance(style, StyleGuides) else self.images.style(style)
        sh, sw = (style["base"].shape[:2]) if isinstance(style, dict) else (None, None)
        return self.open_device(params, iterations, w, h, sg, sw, sh)

    def open_device(self, params, iterations, w, h, sg, sw, sh):
        self._keep.append(sg)
        d = _ffi.SequenceDesc(params, iterations, w, h, sw, sh, C.pointer(sg))
        seq = C.c_int32()
        self._check(self.lib.pt_server_open(self.handle, C.byref(d), C.byref(seq)))
        self.seqs[seq.value] = (w, h)
        return seq.value

    def submit(self, seq, frame, outputs=None):
        fg = frame if isinstance(frame, FrameGuides) else self.images.frame(frame)
        w, h = self.seqs[seq]
        out = outputs if outputs is not None else DeviceOutputs(w, h, 1, self.device)
        o, c, r = out.pointers()
        t = C.c_uint64()
        self._check(self.lib.pt_server_submit(self.handle, seq, C.byref(fg), o, c, r, C.byref(t)))
        self.outputs[t.value] = out
        return t.value

    def poll(self, max_ticks=1, max_done=256):
        done = (C.c_uint64 * max_done)()
        lat = (C.c_double * max_done)()
        n = C.c_int()
        self._check(self.lib.pt_server_poll(self.handle, max_ticks, done, lat, max_done, C.byref(n)))
        return [(done[i], lat[i]) for i in range(n.value)]

    def drain(self):
        self._check(self.lib.pt_server_drain(self.handle))

    def set(self, key, value):
        self._check(self.lib.pt_server_set(self.handle, key, value))

    def close_sequence(self, seq):
        self._check(self.lib.pt_server_close(self.handle, seq))

    def stats(self):
        st = _ffi.ServerStats()
        self._check(self.lib.pt_server_get_stats(self.handle, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def result(self, ticket):
        return self.outputs[ticket].numpy()

    def close(self):
        if self.handle:
            self.lib.pt_server_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
