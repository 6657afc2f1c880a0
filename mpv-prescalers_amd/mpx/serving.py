"""Serving loop: continuous batching of upscale requests on one GPU.

Requests (a device frame + variant name + display size) arrive at any time.  A scheduler
thread keeps the weights resident (one Context), admits queued requests in arrival order
while the bytes of in-flight outputs stay within the HBM budget, groups the admitted ones
by plan, and launches each group -- up to 16 frames at independent addresses per launch --
on the least recently used of several HIP streams, so launches of different variants
overlap on the device while new arrivals keep being admitted.  Frames at their variant's
default display size share launches across sizes (``mpx_run_frames_sized``: one tile
queue over frames of different sizes, plans of a fixed capacity); other requests group by
variant and size (``mpx_run_frames``).  A completion thread waits on each launch's event
in order and stamps every request's finish time; latency = finish - arrival.

Each (plan key, stream) pair owns its plan, since a plan's scratch serves one stream at a
time.  All streams use the default priority.
"""
import collections
import threading
import time

import torch

from ._abi import MpxError
from .ops import Context, default_target, query

_BUDGET = -5  # MPX_ERR_BUDGET


class Request:
    __slots__ = ("name", "frame", "target", "precision", "bits", "out_dtype", "t_arrival", "t_admit", "t_done",
                 "out", "bytes", "key", "error", "src_stream", "_done")

    def __init__(self, name, frame, target, precision, bits, out_dtype):
        self.name = name
        self.frame = frame
        self.target = target
        self.precision = precision
        self.bits = bits
        self.out_dtype = out_dtype
        self.t_arrival = time.perf_counter()
        self.t_admit = self.t_done = None
        self.out = None
        self.bytes = 0
        self.key = None
        self.error = None
        self.src_stream = torch.cuda.current_stream(frame.device)  # where the frame was produced
        self._done = threading.Event()

    def wait(self, timeout=None):
        if not self._done.wait(timeout):
            raise TimeoutError("request not finished")
        if self.error is not None:
            raise self.error
        return self.out

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_arrival


class Server:
    """Continuous-batching request loop over one device.

    max_batch: frames per launch (<= 16); streams: concurrent HIP streams;
    hbm_budget: the loop's whole device-memory envelope, split in two: ``admit_bytes``
    (default a quarter) for the outputs of admitted, unfinished requests, and the rest
    for weights + plan scratch (the Context's budget, which evicts idle plans to stay
    inside it).  Weights, scratch and outputs together never exceed hbm_budget."""

    def __init__(self, device=0, hbm_budget=16 << 30, admit_bytes=None, streams=4, max_batch=8,
                 plan_batch=None, mixed_sizes=(2560, 1440)):
        self.device = torch.device("cuda", device) if isinstance(device, int) else device
        self.admit_bytes = int(admit_bytes) if admit_bytes is not None else hbm_budget // 4
        if not 0 < self.admit_bytes < hbm_budget:
            raise ValueError("admit_bytes must be positive and below hbm_budget")
        self.hbm_budget = int(hbm_budget)
        self.ctx_budget = self.hbm_budget - self.admit_bytes
        self.ctx = Context(self.device.index or 0, hbm_budget=self.ctx_budget)
        self.streams = [torch.cuda.Stream(self.device) for _ in range(streams)]
        self.max_batch = max(1, min(16, max_batch))
        self.plan_batch = plan_batch or self.max_batch
        # frames up to this size (width, height) with the variant's default display size
        # share launches across sizes (one plan of this capacity per variant); None: one
        # plan and launch group per size
        self.mixed_sizes = tuple(mixed_sizes) if mixed_sizes else None
        self._plans = collections.OrderedDict()   # (plan key, stream) -> Plan, LRU order
        self._last_use = {}                        # (plan key, stream) -> event of its last launch
        self._queue = collections.deque()
        self._inflight = 0
        self._cv = threading.Condition()
        self._done_q = collections.deque()
        self._done_cv = threading.Condition()
        self._next_stream = 0
        self._stop = False
        self._paused = False
        self.launches = 0
        self._sched = threading.Thread(target=self._schedule, daemon=True)
        self._compl = threading.Thread(target=self._complete, daemon=True)
        self._sched.start()
        self._compl.start()

    # ------------------------------------------------------------------ client side
    def submit(self, name, frame, target=None, precision="fp32", bits=10, out_dtype=None):
        """Queue one frame ((H, W) luma or (3, H, W)); returns a Request (``wait()``).
        A frame whose shape does not fit the variant's planes, or an unknown variant,
        raises here (ValueError / MpxError) and nothing is queued."""
        if frame.dim() < 2:
            raise ValueError("frame must be (H, W) or (3, H, W)")
        in_h, in_w = frame.shape[-2:]
        tgt = tuple(target) if target is not None else default_target(name, in_w, in_h)
        planes = query(name, in_w, in_h, tgt[0], tgt[1])["planes"]
        if planes == 1 and frame.dim() != 2:
            raise ValueError("%s takes (H, W) luma frames, got shape %s" % (name, tuple(frame.shape)))
        if planes == 3 and (frame.dim() != 3 or frame.shape[0] != 3):
            raise ValueError("%s takes (3, H, W) frames, got shape %s" % (name, tuple(frame.shape)))
        if not frame.is_cuda or not frame.is_contiguous():
            raise ValueError("frame must be a contiguous device tensor")
        if out_dtype is None:
            out_dtype = frame.dtype if frame.dtype in (torch.float16, torch.float32) else torch.float32
        r = Request(name, frame, tgt, precision, bits, out_dtype)
        with self._cv:
            self._queue.append(r)
            self._cv.notify()
        return r

    def pause(self):
        """Stop admitting (queued requests wait); for building a burst deterministically."""
        with self._cv:
            self._paused = True

    def resume(self):
        with self._cv:
            self._paused = False
            self._cv.notify()

    def close(self):
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        self._sched.join()
        with self._done_cv:
            self._done_cv.notify_all()
        self._compl.join()
        for p in self._plans.values():
            p.close()
        self._plans.clear()
        self.ctx.close()

    # ------------------------------------------------------------------ scheduler
    def _key(self, r):
        """Plan key of a request: mixed-size launches group by variant (plan of the
        mixed_sizes capacity), the rest by variant and size."""
        in_h, in_w = r.frame.shape[-2:]
        if (self.mixed_sizes and not r.name.startswith("ravu-zoom") and in_w <= self.mixed_sizes[0]
                and in_h <= self.mixed_sizes[1] and r.target == default_target(r.name, in_w, in_h)):
            cw, ch = self.mixed_sizes
            return (r.name, cw, ch, default_target(r.name, cw, ch), r.frame.dtype, r.out_dtype, r.precision, r.bits,
                    True)
        return (r.name, in_w, in_h, r.target, r.frame.dtype, r.out_dtype, r.precision, r.bits, False)

    def _plan(self, key, si):
        """The plan of (key, stream), created on first use.  When the context's HBM budget
        is exhausted, idle plans (their last launch finished) are evicted, oldest first."""
        p = self._plans.get((key, si))
        if p is not None:
            self._plans.move_to_end((key, si))
            return p
        name, in_w, in_h, tgt, in_dt, out_dt, prec, bits, _ = key
        while True:
            try:
                p = self.ctx.plan(name, in_w, in_h, tgt, in_dt, out_dt, prec, self.plan_batch, bits)
                break
            except MpxError as e:
                if e.code != _BUDGET or not self._evict_one():
                    raise
        self._plans[(key, si)] = p
        return p

    def _evict_one(self):
        """Free the least recently used idle plan; if every plan has work in flight, wait
        for the least recently used one to finish.  False when there is no plan."""
        if not self._plans:
            return False
        victim = next((k for k in self._plans if self._last_use.get(k) is None or self._last_use[k].query()),
                      next(iter(self._plans)))
        ev = self._last_use.pop(victim, None)
        if ev is not None:
            ev.synchronize()
        self._plans.pop(victim).close()
        return True

    def _admit_locked(self):
        """Admit queued requests in arrival order while the outputs in flight fit
        admit_bytes; requests whose plan cannot be made, or whose output alone exceeds
        admit_bytes, fail here."""
        admitted = []
        while self._queue:
            r = self._queue[0]
            if r.key is None:
                key = self._key(r)
                try:
                    plan = self._plan(key, 0)
                except Exception as e:  # unknown / unavailable variant, bad size, budget
                    self._queue.popleft()
                    r.error = e
                    r.t_done = time.perf_counter()
                    r._done.set()
                    continue
                r.key = key
                shape = plan.frame_out_shape(r.frame) if key[-1] else plan.out_shape(r.frame)
                r.bytes = int(torch.Size(shape).numel()) * torch.finfo(r.out_dtype).bits // 8
                if r.bytes > self.admit_bytes:
                    self._queue.popleft()
                    r.error = ValueError("output of %d bytes exceeds admit_bytes (%d)" % (r.bytes, self.admit_bytes))
                    r.t_done = time.perf_counter()
                    r._done.set()
                    continue
            if self._inflight + r.bytes > self.admit_bytes:
                break
            self._queue.popleft()
            self._inflight += r.bytes
            r.t_admit = time.perf_counter()
            admitted.append(r)
        return admitted

    def _schedule(self):
        while True:
            with self._cv:
                while True:
                    if self._stop and not self._queue:
                        return
                    admitted = self._admit_locked() if (self._queue and not self._paused) else []
                    if admitted:
                        break
                    # nothing admissible (empty queue, paused, or the outputs in flight fill
                    # admit_bytes): sleep until a submit, resume or completion notifies
                    self._cv.wait()
            groups = collections.OrderedDict()
            for r in admitted:
                groups.setdefault(r.key, []).append(r)
            for key, reqs in groups.items():
                for i in range(0, len(reqs), self.max_batch):
                    self._launch(key, reqs[i:i + self.max_batch])

    def _launch(self, key, reqs):
        si = self._next_stream
        self._next_stream = (si + 1) % len(self.streams)
        st = self.streams[si]
        try:
            plan = self._plan(key, si)
            for src in {r.src_stream for r in reqs}:
                st.wait_stream(src)  # the frames were produced on the submitters' streams
            with torch.cuda.stream(st):
                if key[-1]:   # frames of several sizes, one tile queue
                    outs = [torch.empty(plan.frame_out_shape(r.frame), dtype=r.out_dtype, device=self.device)
                            for r in reqs]
                    plan.run_frames_sized([r.frame for r in reqs], outs, stream=st)
                else:
                    outs = [torch.empty(plan.out_shape(r.frame), dtype=r.out_dtype, device=self.device) for r in reqs]
                    plan.run_frames([r.frame for r in reqs], outs, stream=st)
                ev = torch.cuda.Event()
                ev.record(st)
            self._last_use[(key, si)] = ev
            for r, o in zip(reqs, outs):
                r.out = o
            self.launches += 1
            with self._done_cv:
                self._done_q.append((ev, reqs))
                self._done_cv.notify()
        except Exception as e:  # a failing launch fails its requests, the loop goes on
            for r in reqs:
                r.error = e
                r.t_done = time.perf_counter()
                self._release(r)
                r._done.set()

    def _release(self, r):
        with self._cv:
            self._inflight -= r.bytes
            self._cv.notify()

    def _complete(self):
        while True:
            with self._done_cv:
                while not self._done_q:
                    if self._stop and not self._sched.is_alive():
                        return
                    self._done_cv.wait(0.05)
                ev, reqs = self._done_q.popleft()
            ev.synchronize()
            t = time.perf_counter()
            for r in reqs:
                r.t_done = t
                self._release(r)
                r._done.set()


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    return xs[min(len(xs) - 1, int(round(q / 100.0 * (len(xs) - 1))))]
