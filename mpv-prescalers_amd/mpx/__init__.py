"""mpx -- MI355X-native NNEDI3 / RAVU prescalers (drop-in for bjin/mpv-prescalers by name).

Hand-written gfx950 HIP kernels behind a C ABI (include/mpx.h, ``libmpx.so``), driven
from PyTorch-ROCm tensors through ctypes.  The shipped weights are LGPL-3.0 data
extracted from the reference shaders (weights/NOTICE).
"""
from .ops import Context, Plan, default_context, query, resample, upscale, variants  # noqa: F401
from ._abi import MpxError  # noqa: F401
