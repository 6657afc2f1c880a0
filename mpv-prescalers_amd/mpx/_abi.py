"""ctypes binding of libmpx.so (declarations in include/mpx.h).

The library is built in-tree by ``mpx.build``; importing this module without it raises
ImportError -- there is no fallback path.
"""
import ctypes
import os

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmpx.so")
WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "weights")

MPX_OK = 0
MPX_SKIPPED = 1
STATUS = {
    -1: "MPX_ERR_ARG", -2: "MPX_ERR_UNKNOWN_VARIANT", -3: "MPX_ERR_UNAVAILABLE", -4: "MPX_ERR_SHAPE",
    -5: "MPX_ERR_BUDGET", -6: "MPX_ERR_HIP", -7: "MPX_ERR_IO", -8: "MPX_ERR_UNSUPPORTED",
}
U8, U16, F16, F32 = 0, 1, 2, 3
PREC = {"fp32": 0, "fp16": 1, "bf16": 2, "fp8": 3, "fp16x3": 4}
STORE_FP16 = 0x100   # MPX_STORE_FP16, OR'ed into the precision
FILTER = {"bilinear": 0, "catmull_rom": 1, "mitchell": 2, "spline36": 3, "lanczos3": 4}
FAMILY = {0: "nnedi3", 1: "ravu", 2: "ravu-lite", 3: "ravu-3x", 4: "ravu-zoom"}
CHANNELS = {0: "luma", 1: "yuv", 2: "rgb"}

# every exported symbol declared in include/mpx.h
EXPORTS = ("mpx_variant_count", "mpx_variant_name", "mpx_query", "mpx_create", "mpx_destroy", "mpx_ctx_bytes",
           "mpx_plan_create", "mpx_plan_info", "mpx_plan_bucket_count", "mpx_plan_destroy", "mpx_run",
           "mpx_run_ex", "mpx_run_frames", "mpx_run_frames_sized", "mpx_plan_tune", "mpx_plan_set_shape",
           "mpx_plan_set_tile", "mpx_plan_tile", "mpx_resample",
           "mpx_weights_verify", "mpx_run_static", "mpx_queue_create", "mpx_queue_destroy", "mpx_submit", "mpx_poll",
           "mpx_queue_launches", "mpx_queue_hold", "mpx_last_error")
MAX_FRAMES = 16
STATIC_SCRATCH_BYTES = 64 << 20   # MPX_STATIC_SCRATCH_BYTES


class VariantInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "family", "channels", "planes", "radius", "nns", "win_w", "win_h", "ar", "available", "applies",
        "out_w", "out_h", "scale_x", "scale_y", "align")] + [
        ("offset_x", ctypes.c_float), ("offset_y", ctypes.c_float), ("hook", ctypes.c_char * 8)]

    def as_dict(self):
        d = {n: getattr(self, n) for n, _ in self._fields_}
        d["hook"] = self.hook.decode()
        d["family"] = FAMILY[self.family]
        d["channels"] = CHANNELS[self.channels]
        return d


class MpxIo(ctypes.Structure):
    _fields_ = [("in_", ctypes.c_void_p), ("in_pitch", ctypes.c_int64), ("in_cstride", ctypes.c_int64),
                ("in_bstride", ctypes.c_int64), ("out", ctypes.c_void_p), ("out_pitch", ctypes.c_int64),
                ("out_cstride", ctypes.c_int64), ("out_bstride", ctypes.c_int64), ("batch", ctypes.c_int32),
                ("buckets", ctypes.c_void_p)]


class MpxRequest(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("in_", ctypes.c_void_p), ("in_w", ctypes.c_int32), ("in_h", ctypes.c_int32),
                ("in_type", ctypes.c_int32), ("bits", ctypes.c_int32), ("out", ctypes.c_void_p),
                ("out_type", ctypes.c_int32), ("target_w", ctypes.c_int32), ("target_h", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("ready_stream", ctypes.c_void_p), ("deadline_us", ctypes.c_int64)]


MPX_PENDING = 2
STATUS[-9] = "MPX_ERR_DEADLINE"


class MpxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (STATUS.get(code, "?"), code, msg))
        self.code = code


def _load():
    if not os.path.exists(_LIB_PATH):
        raise ImportError("libmpx.so not built at %s: run `python -m mpx.build` (or __graft_entry__.build())"
                          % _LIB_PATH)
    lib = ctypes.CDLL(_LIB_PATH)
    vp, i32, i64, u64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_uint64
    sig = {
        "mpx_variant_count": (i32, []),
        "mpx_variant_name": (ctypes.c_char_p, [i32]),
        "mpx_query": (i32, [ctypes.c_char_p, i32, i32, i32, i32, ctypes.POINTER(VariantInfo)]),
        "mpx_create": (i32, [i32, u64, ctypes.c_char_p, ctypes.POINTER(vp)]),
        "mpx_destroy": (None, [vp]),
        "mpx_ctx_bytes": (u64, [vp]),
        "mpx_plan_create": (i32, [vp, ctypes.c_char_p, i32, i32, i32, i32, i32, i32, i32, i32, i32,
                                  ctypes.POINTER(vp)]),
        "mpx_plan_info": (i32, [vp, ctypes.POINTER(VariantInfo)]),
        "mpx_plan_bucket_count": (i64, [vp]),
        "mpx_plan_destroy": (None, [vp]),
        "mpx_run": (i32, [vp, vp, i64, i64, vp, i64, i64, i32, vp]),
        "mpx_run_ex": (i32, [vp, ctypes.POINTER(MpxIo), vp]),
        "mpx_run_static": (i32, [vp, ctypes.POINTER(MpxIo), vp]),
        "mpx_run_frames": (i32, [vp, i32, ctypes.POINTER(vp), ctypes.POINTER(vp), i64, i64, vp]),
        "mpx_run_frames_sized": (i32, [vp, i32, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(i32),
                                       ctypes.POINTER(i32), ctypes.POINTER(i64), ctypes.POINTER(i64), vp]),
        "mpx_plan_tune": (i32, [vp, ctypes.POINTER(MpxIo), vp, i32, ctypes.POINTER(i32), ctypes.POINTER(i32),
                                ctypes.POINTER(ctypes.c_float)]),
        "mpx_plan_set_shape": (i32, [vp, i32, i32]),
        "mpx_plan_set_tile": (i32, [vp, i32]),
        "mpx_plan_tile": (i32, [vp, ctypes.POINTER(i32), ctypes.POINTER(i32), ctypes.POINTER(i32),
                                ctypes.POINTER(i32)]),
        "mpx_resample": (i32, [ctypes.POINTER(MpxIo), i32, i32, i32, i32, i32, i32, i32, i32, i32,
                               ctypes.c_float, ctypes.c_float, vp]),
        "mpx_weights_verify": (i32, [ctypes.c_char_p, ctypes.c_char_p]),
        "mpx_last_error": (ctypes.c_char_p, []),
        "mpx_queue_create": (i32, [vp, i32, i32, i32, i32, ctypes.POINTER(vp)]),
        "mpx_queue_destroy": (None, [vp]),
        "mpx_submit": (i32, [vp, ctypes.POINTER(MpxRequest), ctypes.POINTER(u64)]),
        "mpx_poll": (i32, [vp, u64, i64, ctypes.POINTER(i64)]),
        "mpx_queue_launches": (u64, [vp]),
        "mpx_queue_hold": (i32, [vp, i32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


LIB = _load()


def check(rc):
    if rc < 0:
        raise MpxError(rc, LIB.mpx_last_error().decode())
    return rc
