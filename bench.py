"""Benchmark: NNEDI3 / RAVU prescalers on one MI355X.

Headline workload: nnedi3-nns32-win8x4 luma 2x by two field-doubling passes,
one 1920x1080 FP16 plane -> 3840x2160, FP16-MFMA form (FP32 accumulation); a step is
one frame through both passes.  Reported: output Mpixel/s, the plain FP32 VALU form's
Mpixel/s and the ratio, p50/p99 per-frame latency, the dominant kernel's roofline
(double_x pass: 2*K*2*nns FLOP per predicted pixel, W x 2H predicted pixels per
launch), secondary variants, the serving loop, and the C FP32 restatement's CPU throughput on
the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--kernel-only]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "mpv-prescalers_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HEADLINE = "nnedi3-nns32-win8x4"
W, H = 1920, 1080
PEAK_FP16 = 2500.0  # TFLOP/s, dense FP16 MFMA (MI355X)
PEAK_HBM = 8000.0   # GB/s
PEAK_LDS = 256 * 128 * 2.4  # GB/s: 128 bytes per clock per CU at 2.4 GHz


RAVU_STEPS = 30   # timed launches per RAVU secondary line (after WARM_S seconds and 5 untimed)

# (variant, input size, target, planes, batch, dtype name) of the RAVU secondary lines
RAVU_SECONDARY = (("ravu-lite-ar-r3", (640, 360), None, 1, 64, "float32"),
                  ("ravu-lite-ar-r3", (1920, 1080), None, 1, 8, "float16"),
                  ("ravu-zoom-ar-r2-rgb", (1280, 720), (2560, 1440), 3, 4, "float16"),
                  ("ravu-r4-yuv", (1920, 1080), None, 3, 2, "float16"),
                  ("ravu-3x-r3", (1920, 1080), None, 1, 2, "float16"),
                  # 1.5x, radius 3 (the r3 anti-ringing table is not shipped)
                  ("ravu-zoom-r3-rgb", (1920, 1080), (2880, 1620), 3, 1, "float16"))


def parity_record(case_prefix):
    """max-abs / PSNR of a benchmark workload at full size against the FP64 oracle, as the
    GPU tests recorded them (profiles/parity_full_size.json), or None."""
    path = os.path.join(ROOT, "profiles", "parity_full_size.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for r in json.load(f)["records"]:
            if r["case"].startswith(case_prefix):
                out = {k: r[k] for k in ("max_abs", "psnr", "tol", "bucket_agree") if k in r}
                out["case"] = r["case"]
                return out
    return None


def zoom_ar3_weights_dir(tmp):
    """A weights directory holding the shipped tables and a supplied ravu-zoom r3
    anti-ringing table (the reference does not ship one): |main r3 table|, the synthetic
    table the GPU tests check against the oracle."""
    import importlib.util
    import numpy as np
    wdir = os.path.join(ROOT, "mpv-prescalers_amd", "weights")
    spec = importlib.util.spec_from_file_location("mpx_extract", os.path.join(wdir, "extract.py"))
    ext = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(ext)
    for f in os.listdir(wdir):
        if f.endswith(".bin"):
            os.symlink(os.path.join(wdir, f), os.path.join(tmp, f))
    ext.write_bin(os.path.join(tmp, "ravu_zoom_lut3_ar.bin"),
                  np.abs(ext.read_bin(os.path.join(wdir, "ravu_zoom_lut3.bin"))).astype(np.float32))
    return tmp


def lds_bytes_per_frame(key, tile):
    """LDS bytes one frame of a RAVU secondary workload moves at tile shape `tile`
    (hardware counters, profiles/counters_ravu_lds.json), or None."""
    path = os.path.join(ROOT, "profiles", "counters_ravu_lds.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        w = json.load(f)["workloads"].get(key)
    return w["lds_bytes_per_frame"] if w and w.get("tile") == list(tile) else None


def tuned_tile(pl, xin, b, torch):
    """Tile shape the autotuner picks for plan `pl` on `xin` (b frames) and its time per
    frame, timed like the line's main figure."""
    pl.tune(xin, iters=3)
    t = pl.tile()
    wall = wall_steps(lambda: pl.run(xin), RAVU_STEPS, 5, torch)
    return {"tile": [t["w"], t["h"]], "us_per_frame": round(wall / (RAVU_STEPS * b) * 1e6, 2)}


def nnedi3_flop_per_px(name):
    nns = int(name.split("nns")[1].split("-")[0])
    k = 8 * int(name.split("win8x")[1])
    return 2 * k * 2 * nns


def timed(fn, steps, warmup, torch):
    """Run fn() warmup+steps times; per-step GPU durations (ms) from events on the
    current stream and the wall time of the timed region."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        starts[i].record()
        fn()
        ends[i].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return [s.elapsed_time(e) for s, e in zip(starts, ends)], wall


WARM_S = 0.3   # seconds of the same work before each timed region


def warm(fn, torch, seconds=WARM_S):
    """Run fn() for `seconds` (synchronising every 8 calls): the GPU idles while the bench
    prepares each workload on the CPU, and a timed region that starts from idle clocks
    measured up to 12 % slower than the same launches at steady state."""
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(8):
            fn()
        torch.cuda.synchronize()


def wall_steps(fn, steps, warmup, torch):
    """Wall time of exactly `steps` back-to-back calls of fn() after `warmup` untimed ones
    (preceded by WARM_S seconds of the same work), synchronised on both sides, with nothing
    else enqueued in the timed region."""
    warm(fn, torch)
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def region_ms(fn, steps, warmup, torch):
    """Average GPU time per call of fn() over `steps` back-to-back calls: one event pair on
    the current stream around the whole timed region, divided by the call count (per-call
    event pairs would also count the host's submission gaps)."""
    warm(fn, torch)
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(steps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / steps


def pct(xs, q):
    xs = sorted(xs)
    return xs[min(len(xs) - 1, int(round(q / 100.0 * (len(xs) - 1))))]


def cpu_baseline(img16, budget_s=12.0):
    """The oracle's C restatement (FP32, OpenMP over all host threads) on whole 1080p
    frames of the same workload, until ~budget_s of CPU time."""
    import numpy as np
    from oracle import cbuild
    x = img16.astype(np.float32)
    cbuild.nnedi3(x[:64, :64], 32, 4)  # load + warm
    n, t0 = 0, time.perf_counter()
    while True:
        cbuild.nnedi3(x, 32, 4)
        n += 1
        dt = time.perf_counter() - t0
        if dt >= budget_s or n >= 40:
            break
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(n * 4 * W * H / dt / 1e6, 3), "unit": "Mpixel/s", "cores": cores, "kind": "port",
            "sample": "%d full 1920x1080 frame(s) of nnedi3-nns32-win8x4, C FP32 OpenMP restatement "
                      "(oracle/c/mpx_cpu.c), %.1f s" % (n, dt)}


SERVE_VARIANTS = ("ravu-lite-ar-r3", "ravu-r4-yuv", "ravu-3x-r3", "nnedi3-nns64-win8x6")
SERVE_SIZES = ((640, 360), (1280, 720), (1366, 768), (1920, 1080), (2560, 1440))


def serving_bench(duration_s=2.0, rate=2000.0, seed=5):
    """Poisson stream of 10-bit FP16 frames, sizes and variants uniform over the sets
    above, through the continuous-batching loop: p50/p99 latency (arrival -> result
    ready) at the offered load, the loop's saturated throughput, and the throughput of
    one frame per launch on one stream for the same requests."""
    import numpy as np
    import torch
    from mpx.frames import make_frame
    from mpx.serving import Server, percentile
    dev = torch.device("cuda:0")
    rng = np.random.Generator(np.random.PCG64(seed))
    pool = {}
    for v in SERVE_VARIANTS:
        planes = 3 if v.endswith("-yuv") else 1
        for (w, h) in SERVE_SIZES:
            pool[(v, w, h)] = [torch.from_numpy(make_frame(w, h, 7 + i, bits=10, channels=planes)
                                                .astype(np.float16)).to(dev) for i in range(2)]
    prec = {v: ("fp16" if v.startswith("nnedi3") else "fp32") for v in SERVE_VARIANTS}
    n = int(duration_s * rate)
    gaps = rng.exponential(1.0 / rate, size=n)
    pick_v = rng.integers(0, len(SERVE_VARIANTS), size=n)
    pick_s = rng.integers(0, len(SERVE_SIZES), size=n)
    reqs = []
    for i in range(n):
        v = SERVE_VARIANTS[pick_v[i]]
        w, h = SERVE_SIZES[pick_s[i]]
        reqs.append((v, pool[(v, w, h)][i % 2]))

    def out_px(v, x):
        k = 3 if v.startswith("ravu-3x") else 2
        return k * k * x.shape[-1] * x.shape[-2] * (3 if v.endswith("-yuv") else 1)
    total_px = sum(out_px(v, x) for v, x in reqs)

    srv = Server(device=0, hbm_budget=16 << 30, streams=4, max_batch=8)
    try:
        for key, frames in pool.items():   # plans + weights resident before timing
            srv.submit(key[0], frames[0], precision=prec[key[0]]).wait(120)
        torch.cuda.synchronize()
        # (a) Poisson arrivals at `rate`
        t0 = time.perf_counter()
        arr = t0 + np.cumsum(gaps)
        handles = []
        for i, (v, x) in enumerate(reqs):
            d = arr[i] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            handles.append(srv.submit(v, x, precision=prec[v]))
        for r in handles:
            r.wait(300)
            r.out = None   # the client consumes each result: its memory returns to the pool
        t_end = max(r.t_done for r in handles)
        lat = [r.latency * 1e3 for r in handles]
        offered = {"p50_ms": round(percentile(lat, 50), 3), "p99_ms": round(percentile(lat, 99), 3),
                   "mpx_s": round(total_px / (t_end - t0) / 1e6, 1), "frames": n,
                   "achieved_fps": round(n / (t_end - t0), 1)}
        # (b) saturated: everything queued at once, at steady state
        sat = total_px / _burst(srv, reqs, prec) / 1e6
        launches = srv.launches
    finally:
        srv.close()
    native = _native_queue(reqs, prec, total_px, gaps)
    # (c) one frame per launch on one stream: the plain request loop (launch, wait for the
    # result, next request), the same launches issued back to back, and the plain request
    # loop with the plain forms (FP32 VALU NNEDI3)
    single_sync = _one_per_launch(pool, reqs, prec, total_px, sync=True)
    single_async = _one_per_launch(pool, reqs, prec, total_px, sync=False)
    plain_sync = _one_per_launch(pool, reqs, {v: "fp32" for v in prec}, total_px, sync=True)
    small = _small_frames(prec)
    return {"workload": "Poisson %.0f frames/s, sizes %s, variants %s, 10-bit fp16 frames" %
                        (rate, "/".join("%dx%d" % s_ for s_ in SERVE_SIZES), "/".join(SERVE_VARIANTS)),
            "offered_load": offered, "saturated_mpx_s": round(sat, 1),
            "one_frame_per_launch_sync_mpx_s": round(single_sync, 1),
            "one_frame_per_launch_async_mpx_s": round(single_async, 1),
            "one_frame_per_launch_plain_fp32_sync_mpx_s": round(plain_sync, 1),
            "speedup_vs_one_frame_per_launch_sync": round(sat / single_sync, 2),
            "speedup_vs_one_frame_per_launch_async": round(sat / single_async, 2),
            "speedup_vs_one_frame_per_launch_plain_fp32_sync": round(sat / plain_sync, 2),
            "launches": launches, "small_frames": small,
            "native_queue": dict(native, speedup_vs_one_frame_per_launch_sync=round(native["saturated_mpx_s"] / single_sync, 2),
                                 speedup_vs_one_frame_per_launch_async=round(native["saturated_mpx_s"] / single_async, 2))}


def _native_queue(reqs, prec, total_px, gaps):
    """The same requests through the library's own queue (mpx_submit / mpx_poll: batching
    in native threads): p50/p99 latency at the Poisson offered load (submit -> finish, as
    the library saw it), then the saturated throughput (everything submitted under a hold,
    released at once, at steady state: the burst runs once untimed first)."""
    import numpy as np
    import torch
    from mpx.queue import Queue
    from mpx.serving import percentile
    q = Queue(hbm_budget=16 << 30, streams=4, max_batch=8)
    try:
        q.hold()   # warm-up burst: every (variant, stream) plan is made before timing
        warm = [q.submit(v, x, precision=prec[v]) for v, x in reqs[:400]]
        q.hold(False)
        for t in warm:
            t.wait(120)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        arr = t0 + np.cumsum(gaps)
        tickets = []
        for i, (v, x) in enumerate(reqs):
            d = arr[i] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets.append((time.perf_counter(), q.submit(v, x, precision=prec[v])))
        for _, t in tickets:
            t.wait(300)
            t.out = None
        t_end = max(ts + t.latency for ts, t in tickets)
        lat = [t.latency * 1e3 for _, t in tickets]
        offered = {"p50_ms": round(percentile(lat, 50), 3), "p99_ms": round(percentile(lat, 99), 3),
                   "mpx_s": round(total_px / (t_end - t0) / 1e6, 1), "frames": len(reqs),
                   "achieved_fps": round(len(reqs) / (t_end - t0), 1)}
        launches0 = q.launches
        for rep in range(2):
            q.hold()
            t0 = time.perf_counter()
            tickets = [(time.perf_counter(), q.submit(v, x, precision=prec[v])) for v, x in reqs]
            q.hold(False)
            for _, t in tickets:
                t.wait(300)
                t.out = None
            dt = max(ts + t.latency for ts, t in tickets) - t0
        launches = q.launches - launches0
    finally:
        q.close()
    return {"offered_load": offered, "saturated_mpx_s": round(total_px / dt / 1e6, 1),
            "saturated_launches_per_burst": launches // 2, "requests_per_burst": len(reqs)}


def _burst(srv, reqs, prec):
    """Seconds for the loop to finish every request queued at once, at steady state: the
    same burst runs once untimed first, which grows the allocator's pool to the burst's
    outputs in flight (admitted up to the loop's output budget; a loop that has been
    serving reuses that memory), then once timed."""
    for rep in range(2):
        t0 = time.perf_counter()
        handles = [srv.submit(v, x, precision=prec[v]) for v, x in reqs]
        for r in handles:
            r.wait(300)
            r.out = None
        dt = max(r.t_done for r in handles) - t0
    return dt


def _one_per_launch(pool, reqs, prec, total_px, sync):
    """Throughput (Mpixel/s) of one launch per request on one stream."""
    import mpx
    import torch
    ctx = mpx.Context(0)
    plans = {}
    for key, frames in pool.items():
        plans[key] = ctx.plan(key[0], key[1], key[2], in_dtype=torch.float16, out_dtype=torch.float16,
                              precision=prec[key[0]])
        plans[key].run(frames[0])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for v, x in reqs:
        plans[(v, x.shape[-1], x.shape[-2])].run(x)
        if sync:
            torch.cuda.synchronize()
    torch.cuda.synchronize()
    mpx_s = total_px / (time.perf_counter() - t0) / 1e6
    for p_ in plans.values():
        p_.close()
    ctx.close()
    return mpx_s


def _small_frames(prec, n=2000, w=640, h=360):
    """All requests 640x360 (the same four variants, round robin), queued at once through
    the loop vs one launch per request (synchronous): where launches dominate, batching
    pays."""
    import numpy as np
    import torch
    from mpx.frames import make_frame
    from mpx.serving import Server
    dev = torch.device("cuda:0")
    pool = {(v, w, h): [torch.from_numpy(make_frame(w, h, 3 + i, bits=10, channels=3 if v.endswith("-yuv") else 1)
                                         .astype(np.float16)).to(dev) for i in range(2)] for v in SERVE_VARIANTS}
    reqs = [(SERVE_VARIANTS[i % len(SERVE_VARIANTS)], pool[(SERVE_VARIANTS[i % len(SERVE_VARIANTS)], w, h)][i % 2])
            for i in range(n)]
    total_px = sum((9 if v.startswith("ravu-3x") else 4) * w * h * (3 if v.endswith("-yuv") else 1) for v, _ in reqs)
    srv = Server(device=0, hbm_budget=16 << 30, streams=4, max_batch=16)
    try:
        for key, frames in pool.items():
            srv.submit(key[0], frames[0], precision=prec[key[0]]).wait(120)
        torch.cuda.synchronize()
        sat = total_px / _burst(srv, reqs, prec) / 1e6
    finally:
        srv.close()
    single = _one_per_launch(pool, reqs, prec, total_px, sync=True)
    return {"workload": "%d requests of %dx%d, variants round robin" % (n, w, h), "saturated_mpx_s": round(sat, 1),
            "one_frame_per_launch_sync_mpx_s": round(single, 1),
            "speedup_vs_one_frame_per_launch_sync": round(sat / single, 2)}


def _double_x(ctx, torch):
    """The double_x pass alone: FP32 W x 2H intermediate in (as inside the two-pass
    pipeline), 2W x 2H FP16 out."""
    dev = torch.device("cuda:0")
    xi = torch.rand((2 * H, W), dtype=torch.float32, device=dev)
    xo = torch.empty((2 * H, 2 * W), dtype=torch.float16, device=dev)
    plan = ctx.plan(HEADLINE, W, 2 * H, target=(2 * W, 2 * H), in_dtype=torch.float32,
                    out_dtype=torch.float16, precision="fp16")
    assert plan.info["scale_x"] == 2 and plan.info["scale_y"] == 1
    plan.tune(xi, xo, iters=5)
    return lambda: plan.run(xi, out=xo)


def dominant_kernel(ctx, steps, torch):
    """Roofline of the dominant kernel, timed with events on the stream it runs on.
    Algorithmic work: 2*K*2*nns FLOP per predicted pixel x W*2H predicted pixels.
    HBM traffic per launch comes from the hardware-counter summary committed beside the
    kernel-trace profiles (FETCH_SIZE + WRITE_SIZE of this launch, calibrated on a plain
    copy), when present."""
    kms = region_ms(_double_x(ctx, torch), max(steps, 20), 5, torch)
    flop = nnedi3_flop_per_px(HEADLINE) * W * 2 * H
    achieved = flop / (kms * 1e-3) / 1e12
    traffic = None
    valu_share = None
    mfma_busy = None
    tf = os.path.join(ROOT, "profiles", "traffic_nnedi3_double_x_nns32_1920x2160.json")
    if os.path.exists(tf):
        with open(tf) as f:
            rec = json.load(f)
        traffic = rec.get("hbm_bytes_per_launch")
        mfma_busy = rec.get("mfma_busy_frac")
        if rec.get("valu_instructions_per_launch"):
            # issue model: 4 cycles per VALU instruction, 16 per transcendental, over the
            # 1024 SIMDs at 2.4 GHz, as a share of the live kernel time
            v, t = rec["valu_instructions_per_launch"], rec["valu_transcendental_instructions_per_launch"]
            valu_share = round(((v - t) * 4 + t * 16) / 1024 / 2.4e9 / (kms * 1e-3), 3)
    return {"bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK_FP16, "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP16, 4), "traffic": traffic,
            "kernel": "nnedi3_mfma double_x (fp16, nns32, win8x4)", "kernel_ms": round(kms, 5),
            "flop_per_launch": flop, "algorithmic_io_bytes": W * 2 * H * 4 + 2 * W * 2 * H * 2,
            "binding_resource": "VALU issue (FP32 normalisation + exp2/rcp epilogue)",
            "valu_issue_share": valu_share, "mfma_busy_frac": mfma_busy}


def kernel_only(ctx, steps, torch, warm_s):
    """The dominant kernel autotuned (7 candidate shapes x 6 runs), run for `warm_s` seconds
    (steady clocks), then `steps` times, then a plain device copy of 256 MiB `steps` times
    (the calibration for the FETCH_SIZE / WRITE_SIZE counters)."""
    fn = _double_x(ctx, torch)
    warm(fn, torch, warm_s)
    for _ in range(steps):
        fn()
    src = torch.empty(64 << 20, dtype=torch.float32, device="cuda:0").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"kernel_launches": steps, "copy_bytes_each_way": src.numel() * 4}), flush=True)
    return 0


def kernel_only_nns256(ctx, steps, torch, warm_s):
    """nnedi3-nns256-win8x6 FP16 MFMA on 8 10-bit 1080p frames: autotuned, run for `warm_s`
    seconds, then `steps` launches (both passes)."""
    import numpy as np
    from mpx.frames import make_plane
    b = 8
    x = torch.from_numpy(np.stack([make_plane(W, H, seed=20 + i, bits=10) for i in range(b)])
                         .astype(np.float16)).to("cuda:0")
    o = torch.empty((b, 2 * H, 2 * W), dtype=torch.float16, device="cuda:0")
    pl = ctx.plan("nnedi3-nns256-win8x6", W, H, in_dtype=torch.float16, out_dtype=torch.float16,
                  precision="fp16", max_batch=b)
    r = pl.tune(x, o, iters=2)
    warm(lambda: pl.run(x, out=o), torch, warm_s)
    for _ in range(steps):
        pl.run(x, out=o)
    torch.cuda.synchronize()
    print(json.dumps({"workload": "nnedi3-nns256-win8x6 fp16 b8 1080p", "launch": r, "launches": steps}), flush=True)
    return 0


def kernel_only_ravu(ctx, steps, torch, warm_s):
    """The RAVU variants of the secondary list, `steps` launches each, at their family's
    default tile shape (the shape the secondary lines' main figures use; a tuner run under
    a profiler would compare distorted timings).  Before each variant's launches a
    one-element fill kernel is launched and another after them: in a kernel trace or
    counter run the dispatches between those two fills are exactly that variant's `steps`
    launches."""
    import numpy as np
    from mpx.frames import make_plane
    tiles = []
    for i, (name, (iw, ih), tgt, planes, b, dt) in enumerate(RAVU_SECONDARY):
        dt = getattr(torch, dt)
        fr = np.stack([make_plane(iw, ih, seed=i) for i in range(b * planes)]).astype(np.float32)
        xin = torch.from_numpy(fr.reshape((b, planes, ih, iw) if planes == 3 else (b, ih, iw))).to(dt).to("cuda:0")
        pl = ctx.plan(name, iw, ih, target=tgt, in_dtype=dt, out_dtype=dt, max_batch=b)
        out = pl.run(xin)
        t = pl.tile()
        tiles.append([t["w"], t["h"]])
        warm(lambda: pl.run(xin, out=out), torch, warm_s)
        torch.full((1,), float(i), device="cuda:0")   # start marker
        for _ in range(steps):
            pl.run(xin, out=out)
        torch.full((1,), -1.0, device="cuda:0")       # end marker
    torch.cuda.synchronize()
    print(json.dumps({"workload": [r[0] for r in RAVU_SECONDARY], "launches_each": steps, "tiles": tiles}),
          flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--kernel-only", action="store_true",
                    help="only the dominant kernel (double_x pass) and a plain copy of known size, for "
                         "hardware-counter runs")
    ap.add_argument("--kernel-warm", type=float, default=0.1,
                    help="with --kernel-only: seconds of the same launches before the counted ones "
                         "(steady clocks for kernel traces; 0 for counter runs)")
    ap.add_argument("--workload", default="headline", choices=("headline", "nns256", "ravu"),
                    help="with --kernel-only: the headline double_x pass (+ copy calibration), "
                         "nnedi3-nns256-win8x6 on 8 10-bit 1080p frames, or the RAVU variants of "
                         "the secondary list")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (frames shard trivially across "
              "GPUs, one process per device); run with --gpus 1", file=sys.stderr)
        return 2

    import numpy as np
    import torch
    import mpx
    from mpx.frames import make_plane

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    ctx = mpx.Context(0)
    img = make_plane(W, H, seed=11, bits=8).astype(np.float16)
    x = torch.from_numpy(img).to(dev)
    out = torch.empty((2 * H, 2 * W), dtype=torch.float16, device=dev)
    steps, warm = args.steps, args.warmup

    if args.kernel_only:
        if args.workload == "nns256":
            return kernel_only_nns256(ctx, args.steps, torch, args.kernel_warm)
        if args.workload == "ravu":
            return kernel_only_ravu(ctx, args.steps, torch, args.kernel_warm)
        return kernel_only(ctx, args.steps, torch, args.kernel_warm)

    plans = {p: ctx.plan(HEADLINE, W, H, in_dtype=torch.float16, out_dtype=torch.float16, precision=p)
             for p in ("fp16", "fp32")}
    launch = plans["fp16"].tune(x, out, iters=5)   # autotuned launch shape, kept by the plan
    res = {}
    for p, plan in plans.items():
        n = steps if p == "fp16" else max(10, steps // 10)
        # throughput: exactly n steps between synchronisations; latency percentiles from a
        # separate pass with an event pair around each step (the event records would
        # otherwise sit between the steps of the timed region)
        wall = wall_steps(lambda: plan.run(x, out=out), n, warm if p == "fp16" else 3, torch)
        times, _ = timed(lambda: plan.run(x, out=out), n, 0, torch)
        res[p] = {"ms_per_step": wall / n * 1e3, "times": times, "mpx_s": 4 * W * H * n / wall / 1e6}

    roofline = dominant_kernel(ctx, max(20, steps // 2), torch)

    secondary = {}
    if not args.no_secondary:
        # nnedi3-nns256-win8x6 FP16-MFMA, 10-bit 1080p frames in a batch of 8
        b = 8
        f10 = torch.from_numpy(np.stack([make_plane(W, H, seed=20 + i, bits=10) for i in range(b)])
                               .astype(np.float16)).to(dev)
        o10 = torch.empty((b, 2 * H, 2 * W), dtype=torch.float16, device=dev)
        accuracy_forms = {}
        for p in ("fp16", "bf16", "fp16x3", "fp8", "fp32"):
            pl = ctx.plan("nnedi3-nns256-win8x6", W, H, in_dtype=torch.float16, out_dtype=torch.float16,
                          precision=p, max_batch=b)
            n = 2 if p == "fp32" else 5
            pl.tune(f10, o10, iters=2)
            wall = wall_steps(lambda: pl.run(f10, out=o10), n, 1, torch)
            fps = b * n / wall
            rec = {"fps": round(fps, 1), "mpx_s": round(fps * 4 * W * H / 1e6, 1),
                   "tflops": round(fps * nnedi3_flop_per_px("nnedi3-nns256-win8x6") * 3 * W * H / 1e12, 1)}
            err = parity_record("nnedi3-nns256-win8x6 %s 1920x1080 10-bit b8" % p)
            if err:
                rec["error_vs_fp64_oracle"] = err
            cf = os.path.join(ROOT, "profiles", "counters_nnedi3_nns256_8x1080p.json")
            if p == "fp16" and os.path.exists(cf):
                with open(cf) as f:
                    rec["mfma_busy_frac"] = json.load(f)["mfma_busy_frac"]   # hardware counters, same workload
            # FP8 (e4m3 weights and window) is slower than FP16 on this kernel and ~40 dB
            # less accurate: reported for its accuracy only
            (accuracy_forms if p == "fp8" else secondary)["nnedi3-nns256-win8x6 %s b8" % p] = rec
        secondary["accuracy_forms"] = accuracy_forms
        for name, (iw, ih), tgt, planes, b, dt in RAVU_SECONDARY:
            dt = getattr(torch, dt)
            fr = np.stack([make_plane(iw, ih, seed=i) for i in range(b * planes)]).astype(np.float32)
            xin = torch.from_numpy(fr.reshape((b, planes, ih, iw) if planes == 3 else (b, ih, iw))).to(dt).to(dev)
            pl = ctx.plan(name, iw, ih, target=tgt, in_dtype=dt, out_dtype=dt, max_batch=b)
            tile = pl.tile()                             # the family's default tile shape
            wall = wall_steps(lambda: pl.run(xin), RAVU_STEPS, 5, torch)
            opx = pl.info["out_w"] * pl.info["out_h"]
            io_bytes = (iw * ih + opx) * planes * (4 if dt == torch.float32 else 2)
            key = "%s %dx%d b%d" % (name, iw, ih, b) if tgt is None or name.endswith("ar-r2-rgb") else \
                "%s %dx%d->%dx%d b%d" % (name, iw, ih, tgt[0], tgt[1], b)
            secondary[key] = {
                "mpx_s": round(b * RAVU_STEPS * opx / wall / 1e6, 1),
                "us_per_frame": round(wall / (RAVU_STEPS * b) * 1e6, 2),
                "hbm_gb_s": round(b * RAVU_STEPS * io_bytes / wall / 1e9, 1),
                "hbm_frac": round(b * RAVU_STEPS * io_bytes / wall / 1e9 / PEAK_HBM, 4),
                "tile": [tile["w"], tile["h"]]}
            lds = lds_bytes_per_frame(key, (tile["w"], tile["h"]))
            if lds:
                # LDS traffic per frame from the hardware counters of the same workload
                secondary[key]["lds_gb_s"] = round(b * RAVU_STEPS * lds / wall / 1e9, 1)
                secondary[key]["lds_frac"] = round(b * RAVU_STEPS * lds / wall / 1e9 / PEAK_LDS, 4)
            secondary[key]["tuned"] = tuned_tile(pl, xin, b, torch)

        # ravu-zoom-ar-r3-rgb at 2x and 1.5x on a supplied (synthetic) anti-ringing table
        import tempfile
        with tempfile.TemporaryDirectory() as tmp:
            zctx = mpx.Context(0, weights_dir=zoom_ar3_weights_dir(tmp))
            for (iw, ih), (tw, th), b in (((1280, 720), (2560, 1440), 4), ((1920, 1080), (2880, 1620), 2)):
                fr = np.stack([make_plane(iw, ih, seed=90 + i, bits=8) for i in range(3 * b)]).astype(np.float16)
                xin = torch.from_numpy(fr.reshape(b, 3, ih, iw)).to(dev)
                pl = zctx.plan("ravu-zoom-ar-r3-rgb", iw, ih, target=(tw, th), in_dtype=torch.float16,
                               out_dtype=torch.float16, max_batch=b)
                tile = pl.tile()
                wall = wall_steps(lambda: pl.run(xin), RAVU_STEPS, 5, torch)
                io_bytes = (iw * ih + tw * th) * 3 * 2
                rec = {"mpx_s": round(b * RAVU_STEPS * tw * th / wall / 1e6, 1),
                       "us_per_frame": round(wall / (RAVU_STEPS * b) * 1e6, 2),
                       "hbm_gb_s": round(b * RAVU_STEPS * io_bytes / wall / 1e9, 1),
                       "hbm_frac": round(b * RAVU_STEPS * io_bytes / wall / 1e9 / PEAK_HBM, 4),
                       "tile": [tile["w"], tile["h"]],
                       "weights": "synthetic AR table (|r3 main table|; the reference ships none)"}
                rec["tuned"] = tuned_tile(pl, xin, b, torch)
                err = parity_record("ravu-zoom-ar-r3-rgb %dx%d->%dx%d" % (iw, ih, tw, th))
                if err:
                    rec["error_vs_fp64_oracle"] = err
                secondary["ravu-zoom-ar-r3-rgb %dx%d->%dx%d b%d" % (iw, ih, tw, th, b)] = rec
                pl.close()
            zctx.close()

    serving = None
    if not args.no_serving:
        serving = serving_bench()

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(img)

    f16 = res["fp16"]
    line = {
        "metric": "upscaled Mpixel/s 1080p->2160p per variant vs plain FP32; p50/p99 frame latency",
        "value": round(f16["mpx_s"], 1),
        "unit": "Mpixel/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warm,
        "ms_per_step": round(f16["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16-mfma/fp32-acc",
        "data": "synthetic (seeded gradient + edges + zone plate + noise, 8-bit quantised)",
        "config": {"workload": "nnedi3-nns32-win8x4 luma 2x, 1920x1080 fp16 plane -> 3840x2160, batch 1",
                   "variant": HEADLINE, "precision": "fp16", "in": [W, H], "out": [2 * W, 2 * H],
                   "launch": {"grid_per_cu": launch["grid_per_cu"], "split": launch["split"]}},
        "plain_fp32": {"value": round(res["fp32"]["mpx_s"], 1), "unit": "Mpixel/s",
                       "ms_per_step": round(res["fp32"]["ms_per_step"], 4)},
        "speedup_vs_plain_fp32": round(f16["mpx_s"] / res["fp32"]["mpx_s"], 2),
        "latency_ms": {"p50": round(pct(f16["times"], 50), 5), "p99": round(pct(f16["times"], 99), 5)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "secondary": secondary,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
