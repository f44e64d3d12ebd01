"""MI355X (gfx950) implementation of PaintTween's per-frame synthesis hot path.

The compute path is the HIP library ``libpainttween.so`` (kernels in
``csrc/pt_kernels.hip``, runtime and schedule in ``csrc/pt_runtime.hip``)
behind the C ABI of ``include/pt_api.h``.  This package only binds it.
"""
from . import _ffi
from ._ffi import load_library, PtParams, PATCHXF_DTYPE
from .texsynth import Context, Pyramid, PtError, default_params

__all__ = ["_ffi", "load_library", "PtParams", "PATCHXF_DTYPE", "Context", "Pyramid", "PtError",
           "default_params"]
