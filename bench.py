"""Benchmark: NNEDI3 / RAVU prescalers on one MI355X.

Headline workload: nnedi3-nns32-win8x4 luma 2x by two field-doubling passes,
one 1920x1080 FP16 plane -> 3840x2160, FP16-MFMA form (FP32 accumulation); a step is
one frame through both passes.  Reported: output Mpixel/s, the plain FP32 VALU form's
Mpixel/s and the ratio, p50/p99 per-frame latency, the dominant kernel's roofline
(double_x pass: 2*K*2*nns FLOP per predicted pixel, W x 2H predicted pixels per
launch), secondary variants, and the FP64 oracle's CPU throughput on the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
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


def pct(xs, q):
    xs = sorted(xs)
    return xs[min(len(xs) - 1, int(round(q / 100.0 * (len(xs) - 1))))]


def cpu_baseline(img16, budget_s=20.0):
    """FP64 NumPy oracle on the host: whole 1080p frames (both passes) until ~budget_s."""
    import numpy as np
    from oracle.nnedi3 import nnedi3
    x = img16.astype(np.float64)
    n, t0 = 0, time.perf_counter()
    while True:
        nnedi3(x, 32, 4)
        n += 1
        dt = time.perf_counter() - t0
        if dt >= budget_s or n >= 3:
            break
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(n * 4 * W * H / dt / 1e6, 3), "unit": "Mpixel/s", "cores": cores, "kind": "port",
            "sample": "%d full 1920x1080 frame(s), nnedi3-nns32-win8x4, FP64 NumPy oracle (BLAS threads=%d), %.1f s"
                      % (n, cores, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
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

    plans = {p: ctx.plan(HEADLINE, W, H, in_dtype=torch.float16, out_dtype=torch.float16, precision=p)
             for p in ("fp16", "fp32")}
    res = {}
    for p, plan in plans.items():
        n = steps if p == "fp16" else max(10, steps // 10)
        times, wall = timed(lambda: plan.run(x, out=out), n, warm if p == "fp16" else 3, torch)
        res[p] = {"ms_per_step": wall / n * 1e3, "times": times, "mpx_s": 4 * W * H * n / wall / 1e6}

    # dominant kernel: the double_x pass alone (FP32 W x 2H intermediate in, as inside the
    # two-pass pipeline), timed with events on the launch stream
    xi = torch.rand((2 * H, W), dtype=torch.float32, device=dev)
    xo = torch.empty((2 * H, 2 * W), dtype=torch.float16, device=dev)
    px_plan = ctx.plan(HEADLINE, W, 2 * H, target=(2 * W, 2 * H), in_dtype=torch.float32,
                       out_dtype=torch.float16, precision="fp16")
    assert px_plan.info["scale_x"] == 2 and px_plan.info["scale_y"] == 1
    ktimes, _ = timed(lambda: px_plan.run(xi, out=xo), max(20, steps // 2), 5, torch)
    kms = sum(ktimes) / len(ktimes)
    flop = nnedi3_flop_per_px(HEADLINE) * W * 2 * H
    achieved = flop / (kms * 1e-3) / 1e12
    roofline = {"bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK_FP16, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP16, 4), "traffic": None,
                "kernel": "nnedi3_mfma double_x (fp16, nns32, win8x4)", "kernel_ms": round(kms, 5),
                "flop_per_launch": flop}

    secondary = {}
    if not args.no_secondary:
        # nnedi3-nns256-win8x6 FP16-MFMA, 10-bit 1080p frames in a batch of 8
        b = 8
        f10 = torch.from_numpy(np.stack([make_plane(W, H, seed=20 + i, bits=10) for i in range(b)])
                               .astype(np.float16)).to(dev)
        o10 = torch.empty((b, 2 * H, 2 * W), dtype=torch.float16, device=dev)
        for p in ("fp16", "fp32"):
            pl = ctx.plan("nnedi3-nns256-win8x6", W, H, in_dtype=torch.float16, out_dtype=torch.float16,
                          precision=p, max_batch=b)
            n = 5 if p == "fp16" else 2
            t, wall = timed(lambda: pl.run(f10, out=o10), n, 1, torch)
            fps = b * n / wall
            secondary["nnedi3-nns256-win8x6 %s b8" % p] = {
                "fps": round(fps, 1), "mpx_s": round(fps * 4 * W * H / 1e6, 1),
                "tflops": round(fps * nnedi3_flop_per_px("nnedi3-nns256-win8x6") * 3 * W * H / 1e12, 1)}
        for name, (iw, ih), tgt, planes, b, dt in (
                ("ravu-lite-ar-r3", (640, 360), None, 1, 64, torch.float32),
                ("ravu-zoom-ar-r2-rgb", (1280, 720), (2560, 1440), 3, 4, torch.float16),
                ("ravu-r4-yuv", (1920, 1080), None, 3, 2, torch.float16),
                ("ravu-3x-r3", (1920, 1080), None, 1, 2, torch.float16)):
            fr = np.stack([make_plane(iw, ih, seed=i) for i in range(b * planes)]).astype(np.float32)
            xin = torch.from_numpy(fr.reshape((b, planes, ih, iw) if planes == 3 else (b, ih, iw))).to(dt).to(dev)
            pl = ctx.plan(name, iw, ih, target=tgt, in_dtype=dt, out_dtype=dt, max_batch=b)
            t, wall = timed(lambda: pl.run(xin), 10, 2, torch)
            opx = pl.info["out_w"] * pl.info["out_h"]
            io_bytes = (iw * ih + opx) * planes * (4 if dt == torch.float32 else 2)
            secondary["%s %dx%d b%d" % (name, iw, ih, b)] = {
                "mpx_s": round(b * 10 * opx / wall / 1e6, 1), "us_per_frame": round(wall / (10 * b) * 1e6, 2),
                "hbm_gb_s": round(b * 10 * io_bytes / wall / 1e9, 1)}

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
                   "variant": HEADLINE, "precision": "fp16", "in": [W, H], "out": [2 * W, 2 * H]},
        "plain_fp32": {"value": round(res["fp32"]["mpx_s"], 1), "unit": "Mpixel/s",
                       "ms_per_step": round(res["fp32"]["ms_per_step"], 4)},
        "speedup_vs_plain_fp32": round(f16["mpx_s"] / res["fp32"]["mpx_s"], 2),
        "latency_ms": {"p50": round(pct(f16["times"], 50), 5), "p99": round(pct(f16["times"], 99), 5)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "secondary": secondary,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
