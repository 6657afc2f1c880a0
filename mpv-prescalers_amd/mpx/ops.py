"""Python host API over the C ABI: contexts, plans and ``upscale(name, tensor)``.

A variant is chosen by its reference file name without ``.hook`` (the way mpv's
``--glsl-shader`` picks it, README.md:39-42), e.g. ``upscale("ravu-lite-ar-r3", y)``.
Tensors are PyTorch-ROCm device tensors; their memory and stream are handed to the HIP
kernels through ctypes.  Shapes: luma ``(H, W)`` or ``(B, H, W)``; -yuv / -rgb
``(3, H, W)`` or ``(B, 3, H, W)`` (planar).
"""
import ctypes
import threading

import torch

from . import _abi
from ._abi import LIB, MpxError, MpxIo, VariantInfo, check

_DT = {torch.uint8: _abi.U8, torch.uint16: _abi.U16, torch.float16: _abi.F16, torch.float32: _abi.F32}


def variants():
    """All shipped variant names (46), including the three ravu-zoom-ar-r3 forms whose
    weights the reference does not ship (they report available = 0)."""
    return [LIB.mpx_variant_name(i).decode() for i in range(LIB.mpx_variant_count())]


def query(name, in_w, in_h, target_w, target_h):
    """Geometry and gating (the hook's WHEN/WIDTH/HEIGHT/OFFSET) as a dict."""
    info = VariantInfo()
    check(LIB.mpx_query(name.encode(), in_w, in_h, target_w, target_h, ctypes.byref(info)))
    return info.as_dict()


def default_target(name, in_w, in_h):
    """Target size at which a variant applies on both axes: 3x for ravu-3x, else 2x."""
    k = 3 if name.startswith("ravu-3x") else 2
    return in_w * k, in_h * k


class Plan:
    """A variant prepared for fixed sizes / element types (weights resident, scratch for
    ``max_batch`` frames).  Not re-entrant: one plan serves one stream at a time."""

    def __init__(self, ctx, name, in_w, in_h, target_w, target_h, in_dtype, out_dtype, precision="fp32",
                 max_batch=1, bits=10, storage="fp32"):
        self.ctx = ctx
        self.name = name
        self.in_w, self.in_h = in_w, in_h
        self.in_dtype, self.out_dtype = in_dtype, out_dtype
        self.max_batch = max_batch
        self.precision = precision
        if storage not in ("fp32", "fp16"):
            raise ValueError("storage must be 'fp32' or 'fp16'")
        self.storage = storage
        h = ctypes.c_void_p()
        prec = _abi.PREC[precision] | (_abi.STORE_FP16 if storage == "fp16" else 0)
        rc = LIB.mpx_plan_create(ctx._h, name.encode(), in_w, in_h, target_w, target_h, _DT[in_dtype], bits,
                                 _DT[out_dtype], prec, max_batch, ctypes.byref(h))
        check(rc)
        self._h = h
        self.skipped = rc == _abi.MPX_SKIPPED
        info = VariantInfo()
        check(LIB.mpx_plan_info(h, ctypes.byref(info)))
        self.info = info.as_dict()
        self.bucket_count = LIB.mpx_plan_bucket_count(h)

    def close(self):
        if getattr(self, "_h", None):
            LIB.mpx_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def out_shape(self, x):
        lead = tuple(x.shape[:-2])
        return lead + (self.info["out_h"], self.info["out_w"])

    def run(self, x, out=None, stream=None, buckets=False):
        """Upscale ``x`` (device tensor) into ``out`` (allocated if None).  Returns ``out``,
        or ``(out, bucket_rows)`` with ``buckets=True``."""
        io, x, out, bk = self._io(x, out, buckets)
        if stream is None:
            stream = torch.cuda.current_stream(x.device)
        check(LIB.mpx_run_ex(self._h, ctypes.byref(io), ctypes.c_void_p(stream.cuda_stream)))
        return (out, bk) if buckets else out

    def run_static(self, x, out=None, stream=None):
        """``run`` through mpx_run_static: the intermediates go to the library's static
        per-device array (at most 64 MiB per launch) instead of the plan's scratch, and the
        call allocates nothing but ``out`` when it is None.  Calls that overlap on one
        device race on that array."""
        io, x, out, _ = self._io(x, out, False)
        if stream is None:
            stream = torch.cuda.current_stream(x.device)
        check(LIB.mpx_run_static(self._h, ctypes.byref(io), ctypes.c_void_p(stream.cuda_stream)))
        return out

    def tune(self, x, out=None, iters=5, stream=None):
        """Autotune the launch shape on ``x`` / ``out`` (writes ``out``); the plan keeps the
        fastest.  Returns {"grid_per_cu", "split", "tile", "us"}: NNEDI3 MFMA plans tune the
        persistent grid (0 = occupancy maximum) and neuron split, RAVU plans the tile
        shape of their kernels (index into the family's list); plans without a tunable
        shape return zeros."""
        io, x, out, _ = self._io(x, out, False)
        if stream is None:
            stream = torch.cuda.current_stream(x.device)
        g, sp, us = ctypes.c_int(), ctypes.c_int(), ctypes.c_float()
        check(LIB.mpx_plan_tune(self._h, ctypes.byref(io), ctypes.c_void_p(stream.cuda_stream), iters,
                                ctypes.byref(g), ctypes.byref(sp), ctypes.byref(us)))
        return {"grid_per_cu": g.value, "split": sp.value, "tile": self.tile()["tile"], "us": us.value}

    def set_shape(self, grid_per_cu=0, split=0):
        """Set the launch shape directly (e.g. a cached ``tune`` result): blocks per CU
        (0 = occupancy maximum) and, for NNEDI3 nns 64 / 128, the neuron split."""
        check(LIB.mpx_plan_set_shape(self._h, int(grid_per_cu), int(split)))

    def set_tile(self, tile):
        """RAVU plans: launch their kernels on tile shape ``tile`` of the family's list
        (0 = default; results are the same for every shape)."""
        check(LIB.mpx_plan_set_tile(self._h, int(tile)))

    def tile(self):
        """{"tile", "count", "w", "h"}: the current tile shape of a RAVU plan's kernels
        (pixels per block) and the number of candidates (1, size 0 x 0, for NNEDI3)."""
        t, n, w, h = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(LIB.mpx_plan_tile(self._h, ctypes.byref(t), ctypes.byref(n), ctypes.byref(w), ctypes.byref(h)))
        return {"tile": t.value, "count": n.value, "w": w.value, "h": h.value}

    def _check_frame(self, x):
        """One frame of run_frames / run_frames_sized / the serving loop: a contiguous
        device tensor of the plan's element type, (H, W) for a luma plan and (3, H, W)
        for a -yuv / -rgb plan (the kernels read `planes` planes of pitch * height)."""
        planes = self.info["planes"]
        if not x.is_cuda:
            raise ValueError("frame must be a device tensor")
        if x.dtype != self.in_dtype:
            raise TypeError("plan expects %s frames, got %s" % (self.in_dtype, x.dtype))
        if planes == 1 and x.dim() != 2:
            raise ValueError("a luma plan takes (H, W) frames, got shape %s" % (tuple(x.shape),))
        if planes == 3 and (x.dim() != 3 or x.shape[0] != 3):
            raise ValueError("a 3-plane plan takes (3, H, W) frames, got shape %s" % (tuple(x.shape),))
        if not x.is_contiguous():
            raise ValueError("frames must be contiguous")

    def _io(self, x, out, buckets):
        planes = self.info["planes"]
        if x.dtype != self.in_dtype:
            raise TypeError("plan expects %s input, got %s" % (self.in_dtype, x.dtype))
        if x.shape[-2:] != (self.in_h, self.in_w):
            raise ValueError("plan expects %dx%d input, got %s" % (self.in_w, self.in_h, tuple(x.shape)))
        if not x.is_cuda:
            raise ValueError("input must be a device tensor")
        nd = 2 if planes == 1 else 3
        if x.dim() == nd:
            batch = 1
        elif x.dim() == nd + 1:
            batch = x.shape[0]
        else:
            raise ValueError("bad input rank %d" % x.dim())
        if planes == 3 and x.shape[-3] != 3:
            raise ValueError("3-plane variant expects 3 planes")
        if x.stride(-1) != 1:
            x = x.contiguous()
        if out is None:
            out = torch.empty(self.out_shape(x), dtype=self.out_dtype, device=x.device)
        if out.stride(-1) != 1 or out.dtype != self.out_dtype or tuple(out.shape) != self.out_shape(x):
            raise ValueError("bad output tensor")
        bk = None
        if buckets:
            bk = torch.empty((batch, self.bucket_count), dtype=torch.int32, device=x.device)
        io = MpxIo()
        io.in_ = x.data_ptr()
        io.in_pitch = x.stride(-2)
        io.in_cstride = x.stride(-3) if planes == 3 else x.stride(-2) * self.in_h
        io.in_bstride = x.stride(0) if x.dim() == nd + 1 else x.numel()
        io.out = out.data_ptr()
        io.out_pitch = out.stride(-2)
        io.out_cstride = out.stride(-3) if planes == 3 else out.stride(-2) * self.info["out_h"]
        io.out_bstride = out.stride(0) if out.dim() == nd + 1 else out.numel()
        io.batch = batch
        io.buckets = bk.data_ptr() if bk is not None else None
        return io, x, out, bk


    def run_frames(self, xs, outs=None, stream=None):
        """One launch (per pass) over separate same-shape frames (up to 16): each x in
        ``xs`` is a device tensor of the plan's input shape, results go to ``outs``
        (allocated if None).  No copies: the frame addresses travel in the kernel args."""
        n = len(xs)
        if n < 1 or n > min(_abi.MAX_FRAMES, self.max_batch):
            raise ValueError("run_frames takes 1..%d frames" % min(_abi.MAX_FRAMES, self.max_batch))
        for x in xs:
            self._check_frame(x)
            if tuple(x.shape[-2:]) != (self.in_h, self.in_w):
                raise ValueError("frame of shape %s does not match the plan's %dx%d" % (tuple(x.shape), self.in_w,
                                                                                        self.in_h))
        if outs is None:
            outs = [torch.empty(self.out_shape(x), dtype=self.out_dtype, device=x.device) for x in xs]
        if len(outs) != n:
            raise ValueError("need one output per frame")
        pitch_in = xs[0].stride(-2)
        pitch_out = outs[0].stride(-2)
        for x, o in zip(xs, outs):
            if o.dtype != self.out_dtype or tuple(o.shape) != self.out_shape(x) or not o.is_cuda:
                raise ValueError("bad output tensor")
            if x.stride(-2) != pitch_in or o.stride(-2) != pitch_out or not o.is_contiguous():
                raise ValueError("frames must share row pitches")
        ins = (ctypes.c_void_p * n)(*[x.data_ptr() for x in xs])
        ots = (ctypes.c_void_p * n)(*[o.data_ptr() for o in outs])
        if stream is None:
            stream = torch.cuda.current_stream(xs[0].device)
        check(LIB.mpx_run_frames(self._h, n, ins, ots, pitch_in, pitch_out, ctypes.c_void_p(stream.cuda_stream)))
        return outs


    def frame_out_shape(self, x):
        """Output shape of one frame ``x`` of any size within the plan (run_frames_sized)."""
        lead = tuple(x.shape[:-2])
        return lead + (x.shape[-2] * self.info["scale_y"], x.shape[-1] * self.info["scale_x"])

    def run_frames_sized(self, xs, outs=None, stream=None):
        """One launch (per pass) over frames of different sizes (up to 16, each within the
        plan's input size, planes contiguous): every frame is scaled by the plan's factors
        through one tile queue.  Not for ravu-zoom plans."""
        n = len(xs)
        if n < 1 or n > min(_abi.MAX_FRAMES, self.max_batch):
            raise ValueError("run_frames_sized takes 1..%d frames" % min(_abi.MAX_FRAMES, self.max_batch))
        for x in xs:
            self._check_frame(x)
        if outs is None:
            outs = [torch.empty(self.frame_out_shape(x), dtype=self.out_dtype, device=x.device) for x in xs]
        if len(outs) != n:
            raise ValueError("need one output per frame")
        for x, o in zip(xs, outs):
            if not o.is_cuda or not o.is_contiguous() or o.dtype != self.out_dtype:
                raise ValueError("bad output tensor")
            if tuple(o.shape) != self.frame_out_shape(x):
                raise ValueError("bad output shape")
        ins = (ctypes.c_void_p * n)(*[x.data_ptr() for x in xs])
        ots = (ctypes.c_void_p * n)(*[o.data_ptr() for o in outs])
        ws = (ctypes.c_int * n)(*[x.shape[-1] for x in xs])
        hs = (ctypes.c_int * n)(*[x.shape[-2] for x in xs])
        ips = (ctypes.c_int64 * n)(*[x.stride(-2) for x in xs])
        ops_ = (ctypes.c_int64 * n)(*[o.stride(-2) for o in outs])
        if stream is None:
            stream = torch.cuda.current_stream(xs[0].device)
        check(LIB.mpx_run_frames_sized(self._h, n, ins, ots, ws, hs, ips, ops_, ctypes.c_void_p(stream.cuda_stream)))
        return outs


class Context:
    """Device binding + resident weights + HBM budget (bytes, 0 = unlimited)."""

    def __init__(self, device=0, hbm_budget=0, weights_dir=None):
        if isinstance(device, torch.device):
            device = device.index or 0
        self.device = device
        h = ctypes.c_void_p()
        check(LIB.mpx_create(device, hbm_budget, (weights_dir or _abi.WEIGHTS_DIR).encode(), ctypes.byref(h)))
        self._h = h
        self._plans = {}
        self._lock = threading.Lock()

    def bytes(self):
        return LIB.mpx_ctx_bytes(self._h)

    def plan(self, name, in_w, in_h, target=None, in_dtype=torch.float32, out_dtype=torch.float32,
             precision="fp32", max_batch=1, bits=10, storage="fp32"):
        """``storage="fp16"``: intermediates stored in FP16 as mpv's default FBOs store
        them (with ``out_dtype=torch.float16`` the plan reproduces mpv's default storage)."""
        tw, th = target if target is not None else default_target(name, in_w, in_h)
        return Plan(self, name, in_w, in_h, tw, th, in_dtype, out_dtype, precision, max_batch, bits, storage)

    def cached_plan(self, name, in_w, in_h, target, in_dtype, out_dtype, precision, batch, bits, storage="fp32"):
        key = (name, in_w, in_h, target, in_dtype, out_dtype, precision, bits, storage)
        with self._lock:
            p = self._plans.get(key)
            if p is None or p.max_batch < batch:
                if p is not None:
                    p.close()
                p = self.plan(name, in_w, in_h, target, in_dtype, out_dtype, precision, max(batch, 1), bits, storage)
                self._plans[key] = p
            return p

    def close(self):
        for p in self._plans.values():
            p.close()
        self._plans.clear()
        if getattr(self, "_h", None):
            LIB.mpx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_ctx = {}
_ctx_lock = threading.Lock()


def default_context(device=None):
    dev = torch.cuda.current_device() if device is None else (device.index or 0 if isinstance(device, torch.device) else device)
    with _ctx_lock:
        if dev not in _ctx:
            _ctx[dev] = Context(dev)
        return _ctx[dev]


def upscale(name, x, target=None, precision="fp32", out_dtype=None, bits=10, buckets=False, stream=None,
            storage="fp32"):
    """Run variant ``name`` on device tensor ``x``.

    ``target`` = (width, height) the video is displayed at (mpv's OUTPUT size): it gates
    the hook like //!WHEN and, for ravu-zoom, is the output size.  Default: 2x (3x for
    ravu-3x).  ``out_dtype`` defaults to the input dtype for float inputs, float32 else.
    ``storage="fp16"`` stores the intermediates in FP16 as mpv's default FBOs do."""
    planes3 = name.endswith("-yuv") or name.endswith("-rgb")
    in_h, in_w = x.shape[-2:]
    if out_dtype is None:
        out_dtype = x.dtype if x.dtype in (torch.float16, torch.float32) else torch.float32
    batch = 1
    if x.dim() == (4 if planes3 else 3):
        batch = x.shape[0]
    tgt = tuple(target) if target is not None else default_target(name, in_w, in_h)
    plan = default_context(x.device).cached_plan(name, in_w, in_h, tgt, x.dtype, out_dtype, precision, batch, bits,
                                                 storage)
    return plan.run(x, stream=stream, buckets=buckets)


def resample(x, size, offset=(0.0, 0.0), filter="catmull_rom", planes=1, out_dtype=None, bits=10, stream=None):
    """The scaling pass after a prescaler: resample device tensor ``x`` to ``size`` =
    (width, height), sampling at (j + 1/2) * in/out + offset per axis.  ``planes`` = 1:
    ``x`` is (H, W) or (B, H, W); 3: (3, H, W) or (B, 3, H, W).  Pass the prescaler's
    ``info["offset_x"], info["offset_y"]`` to obtain the display-aligned image (what mpv's
    main scaler does with a hook's OFFSET)."""
    return _resample(x, size, offset, filter, out_dtype, bits, stream, planes)


def _resample(x, size, offset, filter, out_dtype, bits, stream, planes):
    if not x.is_cuda:
        raise ValueError("input must be a device tensor")
    if x.stride(-1) != 1:
        x = x.contiguous()
    out_w, out_h = size
    in_h, in_w = x.shape[-2:]
    nd = 2 if planes == 1 else 3
    if x.dim() not in (nd, nd + 1) or (planes == 3 and x.shape[-3] != 3):
        raise ValueError("bad input shape %s for %d plane(s)" % (tuple(x.shape), planes))
    batch = x.shape[0] if x.dim() == nd + 1 else 1
    if out_dtype is None:
        out_dtype = x.dtype if x.dtype in (torch.float16, torch.float32) else torch.float32
    out = torch.empty(tuple(x.shape[:-2]) + (out_h, out_w), dtype=out_dtype, device=x.device)
    io = MpxIo()
    io.in_ = x.data_ptr()
    io.in_pitch = x.stride(-2)
    io.in_cstride = x.stride(-3) if planes == 3 else x.stride(-2) * in_h
    io.in_bstride = x.stride(0) if x.dim() == nd + 1 else x.numel()
    io.out = out.data_ptr()
    io.out_pitch = out.stride(-2)
    io.out_cstride = out.stride(-3) if planes == 3 else out.stride(-2) * out_h
    io.out_bstride = out.stride(0) if x.dim() == nd + 1 else out.numel()
    io.batch = batch
    if stream is None:
        stream = torch.cuda.current_stream(x.device)
    with torch.cuda.device(x.device):
        check(LIB.mpx_resample(ctypes.byref(io), in_w, in_h, _DT[x.dtype], bits, out_w, out_h, _DT[out_dtype], planes,
                               _abi.FILTER[filter], float(offset[0]), float(offset[1]),
                               ctypes.c_void_p(stream.cuda_stream)))
    return out