"""Python face of the library's request queue (mpx_queue_create / mpx_submit / mpx_poll):
continuous batching runs in the library's own threads (no Python on the scheduling path);
this wrapper allocates each request's output tensor, hands the library the frame's and the
output's device pointers with the stream the frame was produced on, and keeps both
tensors alive until the library reports the request finished."""
import ctypes
import threading

import torch

from . import _abi
from ._abi import LIB, MpxError, MpxRequest, check
from .ops import Context, _DT, default_target, query


class Ticket:
    """One submitted request: ``wait()`` returns its output tensor (or raises the
    request's MpxError); ``latency`` is submit -> finish in seconds, measured by the
    library when it saw the request finish."""

    __slots__ = ("queue", "id", "out", "status", "error", "latency", "_lock")

    def __init__(self, queue, tid, out):
        self.queue = queue
        self.id = tid
        self.out = out
        self.status = _abi.MPX_PENDING
        self.error = None
        self.latency = None
        self._lock = threading.Lock()

    def _poll(self, timeout_us):
        with self._lock:
            if self.status != _abi.MPX_PENDING:
                return self.status
            lat = ctypes.c_int64(0)
            rc = LIB.mpx_poll(self.queue._h, self.id, int(timeout_us), ctypes.byref(lat))
            if rc == _abi.MPX_PENDING:
                return rc
            self.status = rc
            self.latency = lat.value * 1e-6
            if rc < 0:
                self.error = MpxError(rc, LIB.mpx_last_error().decode())
            self.queue._finished(self.id)
            return rc

    def done(self):
        return self._poll(0) != _abi.MPX_PENDING

    def wait(self, timeout=None):
        rc = self._poll(-1 if timeout is None else max(1, int(timeout * 1e6)))
        if rc == _abi.MPX_PENDING:
            raise TimeoutError("request not finished")
        if self.error is not None:
            raise self.error
        return self.out


class Queue:
    """Continuous-batching request queue of the library on one device.

    streams: HIP streams of the queue; max_batch: frames per launch (<= 16); max_size:
    (width, height) up to which frames at their variant's default scale share launches
    across sizes.  Plans and weights live in ``ctx`` (its HBM budget; a new Context with
    ``hbm_budget`` when None); outputs are allocated here with torch."""

    def __init__(self, ctx=None, device=0, hbm_budget=16 << 30, streams=4, max_batch=8, max_size=(2560, 1440)):
        self.ctx = ctx if ctx is not None else Context(device, hbm_budget=hbm_budget)
        self._own_ctx = ctx is None
        self.device = torch.device("cuda", self.ctx.device)
        h = ctypes.c_void_p()
        check(LIB.mpx_queue_create(self.ctx._h, streams, max_batch, max_size[0], max_size[1], ctypes.byref(h)))
        self._h = h
        self._live = {}            # ticket id -> (frame, out) until the request finished
        self._lock = threading.Lock()

    @property
    def launches(self):
        return LIB.mpx_queue_launches(self._h)

    def submit(self, name, frame, target=None, precision="fp32", bits=10, out_dtype=None, deadline=None,
               storage="fp32"):
        """Queue one frame ((H, W) luma or (3, H, W), contiguous, on the queue's device);
        returns a Ticket.  The frame is read once the work queued on the current stream
        before this call has run.  deadline: seconds after which a request still queued
        fails with MPX_ERR_DEADLINE instead of running."""
        if frame.dim() < 2:
            raise ValueError("frame must be (H, W) or (3, H, W)")
        in_h, in_w = frame.shape[-2:]
        tgt = tuple(target) if target is not None else default_target(name, in_w, in_h)
        info = query(name, in_w, in_h, tgt[0], tgt[1])
        planes = info["planes"]
        if planes == 1 and frame.dim() != 2:
            raise ValueError("%s takes (H, W) luma frames, got shape %s" % (name, tuple(frame.shape)))
        if planes == 3 and (frame.dim() != 3 or frame.shape[0] != 3):
            raise ValueError("%s takes (3, H, W) frames, got shape %s" % (name, tuple(frame.shape)))
        if not frame.is_cuda or not frame.is_contiguous() or frame.device != self.device:
            raise ValueError("frame must be a contiguous tensor on %s" % self.device)
        if frame.dtype not in _DT:
            raise TypeError("unsupported frame dtype %s" % frame.dtype)
        if out_dtype is None:
            out_dtype = frame.dtype if frame.dtype in (torch.float16, torch.float32) else torch.float32
        out = torch.empty(tuple(frame.shape[:-2]) + (info["out_h"], info["out_w"]), dtype=out_dtype,
                          device=self.device)
        r = MpxRequest()
        r.name = name.encode()
        r.in_ = frame.data_ptr()
        r.in_w, r.in_h = in_w, in_h
        r.in_type = _DT[frame.dtype]
        r.bits = bits
        r.out = out.data_ptr()
        r.out_type = _DT[out_dtype]
        r.target_w, r.target_h = tgt
        r.precision = _abi.PREC[precision] | (_abi.STORE_FP16 if storage == "fp16" else 0)
        r.ready_stream = torch.cuda.current_stream(self.device).cuda_stream
        r.deadline_us = 0 if deadline is None else max(1, int(deadline * 1e6))
        tid = ctypes.c_uint64()
        with self._lock:
            check(LIB.mpx_submit(self._h, ctypes.byref(r), ctypes.byref(tid)))
            self._live[tid.value] = (frame, out)
        return Ticket(self, tid.value, out)

    def hold(self, on=True):
        """Hold the scheduler (submits queue up) or release it: a held set of frames
        is grouped into shared launches when released."""
        check(LIB.mpx_queue_hold(self._h, 1 if on else 0))

    def _finished(self, tid):
        with self._lock:
            self._live.pop(tid, None)

    def close(self):
        """Run everything queued, then free the queue (and its context if it made one)."""
        if getattr(self, "_h", None):
            LIB.mpx_queue_destroy(self._h)
            self._h = None
            with self._lock:
                self._live.clear()
        if self._own_ctx and self.ctx is not None:
            self.ctx.close()
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
