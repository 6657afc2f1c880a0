"""The library's request queue (mpx_submit / mpx_poll) on the GPU: every result equals
the same request run alone, bit for bit, across variants, sizes and precisions; requests
share launches; a deadline that passes while the queue's streams are full fails the
request without running it; a request whose plan cannot be made fails alone; and a C
program drives the queue through the C ABI alone."""
import os
import shutil
import subprocess

import numpy as np
import pytest
import torch

from conftest import ROOT
from parity import report

pytestmark = pytest.mark.gpu

VARIANTS = ("ravu-lite-ar-r3", "ravu-r4-yuv", "ravu-3x-r3", "nnedi3-nns64-win8x6", "ravu-zoom-r2", "nnedi3-nns32-win8x4")
SIZES = ((640, 360), (1280, 720), (1366, 768), (1920, 1080), (97, 61))


def _frame(v, w, h, seed, bits=10, dtype=np.float16):
    from mpx.frames import make_frame
    planes = 3 if v.endswith("-yuv") else 1
    return torch.from_numpy(make_frame(w, h, seed, bits=bits, channels=planes).astype(dtype)).cuda()


@pytest.mark.timeout(300)
def test_queue_results_equal_alone():
    import mpx
    from mpx.queue import Queue
    q = Queue(hbm_budget=16 << 30, streams=4, max_batch=8)
    reqs = []
    try:
        q.hold()                      # the first set is queued whole, the second while it runs
        for rep in range(2):
            for vi, v in enumerate(VARIANTS):
                for si, (w, h) in enumerate(SIZES):
                    x = _frame(v, w, h, 700 + 31 * rep + 7 * vi + si)
                    prec = "fp16" if v.startswith("nnedi3") else "fp32"
                    tgt = (w * 3 // 2, h * 3 // 2) if v.startswith("ravu-zoom") else None
                    reqs.append((v, x, prec, tgt, q.submit(v, x, target=tgt, precision=prec)))
            if rep == 0:
                q.hold(False)
        results = [(v, x, prec, tgt, t.wait(timeout=120), t.latency) for v, x, prec, tgt, t in reqs]
        launches = q.launches
    finally:
        q.close()
    for v, x, prec, tgt, got, lat in results:
        alone = mpx.upscale(v, x, target=tgt, precision=prec)
        assert torch.equal(got, alone), (v, tuple(x.shape))
        assert lat is not None and lat > 0
    report("native queue, %d requests (%d variants x %d sizes x 2)" % (len(reqs), len(VARIANTS), len(SIZES)),
           launches=launches, equal_alone=True)
    assert launches < len(reqs)


@pytest.mark.timeout(300)
def test_queue_deadline_and_isolated_failures():
    import mpx
    from mpx._abi import MpxError
    from mpx.queue import Queue
    q = Queue(hbm_budget=16 << 30, streams=1, max_batch=1)
    try:
        # three ~20 ms FP32 nns256 frames keep the single stream's two slots full, so the
        # scheduler holds the third; a request submitted now waits in the queue past 1 us
        long = [q.submit("nnedi3-nns256-win8x6", _frame("nnedi3", 1920, 1080, 40 + i), precision="fp32")
                for i in range(3)]
        late = q.submit("ravu-lite-r3", _frame("ravu-lite", 64, 48, 5), deadline=1e-6)
        ok = q.submit("ravu-lite-r3", _frame("ravu-lite", 64, 48, 6))
        with pytest.raises(MpxError) as ei:
            late.wait(timeout=60)
        assert ei.value.code == -9
        assert late.out is not None and ok.wait(timeout=60) is not None
        for t in long:
            t.wait(timeout=60)
    finally:
        q.close()
    # a context too small for a plan: that request fails, a small one still runs
    ctx = mpx.Context(0, hbm_budget=2 << 20)
    q = Queue(ctx=ctx, streams=2, max_batch=4)
    try:
        big = q.submit("ravu-r4-yuv", _frame("ravu-r4-yuv", 640, 360, 7))
        small = q.submit("ravu-lite-r2", _frame("ravu-lite", 32, 24, 8))
        with pytest.raises(MpxError) as ei:
            big.wait(timeout=60)
        assert ei.value.code == -5
        assert torch.equal(small.wait(timeout=60), mpx.upscale("ravu-lite-r2", _frame("ravu-lite", 32, 24, 8)))
    finally:
        q.close()
        ctx.close()


@pytest.mark.timeout(300)
def test_queue_concurrent_submitters():
    """Four host threads submit and wait at the same time (ctypes releases the GIL, so the
    library's submit, scheduler and completion threads all run concurrently): every
    request completes with the result it has alone."""
    import threading
    import mpx
    from mpx.queue import Queue
    names = ("ravu-lite-ar-r3", "ravu-3x-r3", "nnedi3-nns32-win8x4", "ravu-r2-yuv")
    sizes = ((160, 90), (97, 61), (256, 144), (64, 48), (200, 120))
    q = Queue(hbm_budget=4 << 30, streams=3, max_batch=4, max_size=(256, 256))
    results, errors = {}, []

    def worker(t):
        try:
            mine = []
            for i in range(12):
                v = names[(t + i) % len(names)]
                w, h = sizes[(3 * t + i) % len(sizes)]
                x = _frame(v, w, h, 900 + 37 * t + i)
                prec = "fp16" if v.startswith("nnedi3") else "fp32"
                mine.append((v, x, prec, q.submit(v, x, precision=prec)))
            for v, x, prec, tk in mine:
                results[(t, len(results))] = (v, x, prec, tk.wait(timeout=120))
        except Exception as e:  # reported below
            errors.append(repr(e))

    try:
        ts = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
        for th in ts:
            th.start()
        for th in ts:
            th.join(timeout=240)
        assert not any(th.is_alive() for th in ts)
    finally:
        q.close()
    assert not errors, errors
    assert len(results) == 48
    for v, x, prec, got in results.values():
        assert torch.equal(got, mpx.upscale(v, x, precision=prec)), (v, tuple(x.shape))


def test_submit_refuses_bad_requests():
    from mpx._abi import MpxError
    from mpx.queue import Queue
    q = Queue(streams=1)
    try:
        with pytest.raises(MpxError) as ei:
            q.submit("ravu-zoom-ar-r3", _frame("ravu-zoom", 32, 24, 1), target=(48, 36))
        assert ei.value.code == -3
        with pytest.raises(ValueError):
            q.submit("ravu-r3-yuv", _frame("ravu-r3", 32, 24, 1))          # one plane for a 3-plane variant
        with pytest.raises(MpxError) as ei:
            q.submit("ravu-r3", _frame("ravu-r3", 32, 24, 1), precision="fp16")
        assert ei.value.code == -8
    finally:
        q.close()


C_HOST = r'''
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <hip/hip_runtime_api.h>
#include "mpx.h"

#define N 12
int main(void) {
  const char* names[3] = {"ravu-lite-ar-r3", "ravu-3x-r2", "nnedi3-nns16-win8x4"};
  const int ws[4] = {160, 97, 200, 64}, hs[4] = {90, 61, 120, 48};
  mpx_ctx* ctx;
  mpx_queue* q;
  if (mpx_create(0, 0, NULL, &ctx) || mpx_queue_create(ctx, 3, 4, 256, 256, &q)) {
    fprintf(stderr, "%s\n", mpx_last_error());
    return 2;
  }
  void *in[N], *out[N], *ref[N];
  size_t on[N];
  uint64_t t[N];
  for (int i = 0; i < N; ++i) {
    const char* nm = names[i % 3];
    int w = ws[i % 4], h = hs[i % 4], k = (i % 3 == 1) ? 3 : 2;
    size_t n = (size_t)w * h;
    float* host = malloc(n * 4);
    for (size_t j = 0; j < n; ++j) host[j] = (float)((j * 37 + i * 11) % 256) / 255.f;
    on[i] = n * k * k;
    if (hipMalloc(&in[i], n * 4) || hipMalloc(&out[i], on[i] * 4) || hipMalloc(&ref[i], on[i] * 4)) return 3;
    if (hipMemcpy(in[i], host, n * 4, hipMemcpyHostToDevice)) return 4;
    free(host);
    mpx_request r;
    memset(&r, 0, sizeof(r));
    r.name = nm; r.in = in[i]; r.in_w = w; r.in_h = h; r.in_type = MPX_F32; r.out = out[i]; r.out_type = MPX_F32;
    r.precision = (i % 3 == 2) ? MPX_PREC_FP16 : MPX_PREC_FP32;
    if (mpx_submit(q, &r, &t[i])) { fprintf(stderr, "%s\n", mpx_last_error()); return 5; }
    /* the same request alone, through a plan */
    mpx_plan* p;
    if (mpx_plan_create(ctx, nm, w, h, w * k, h * k, MPX_F32, 8, MPX_F32, r.precision, 1, &p) < 0) return 6;
    if (mpx_run(p, in[i], w, n, ref[i], w * k, on[i], 1, NULL)) return 7;
    if (hipDeviceSynchronize()) return 8;
    mpx_plan_destroy(p);
  }
  int bad = 0;
  for (int i = 0; i < N; ++i) {
    int64_t lat = -1;
    int rc = mpx_poll(q, t[i], -1, &lat);
    if (rc != MPX_OK || lat < 0) { fprintf(stderr, "poll %d: %d %s\n", i, rc, mpx_last_error()); return 9; }
    if (mpx_poll(q, t[i], 0, NULL) != MPX_ERR_ARG) return 10;   /* retired */
    float* a = malloc(on[i] * 4);
    float* b = malloc(on[i] * 4);
    hipMemcpy(a, out[i], on[i] * 4, hipMemcpyDeviceToHost);
    hipMemcpy(b, ref[i], on[i] * 4, hipMemcpyDeviceToHost);
    if (memcmp(a, b, on[i] * 4)) ++bad;
    free(a); free(b);
    hipFree(in[i]); hipFree(out[i]); hipFree(ref[i]);
  }
  printf("%d %d %llu\n", N, bad, (unsigned long long)mpx_queue_launches(q));
  mpx_queue_destroy(q);
  mpx_destroy(ctx);
  return 0;
}
'''


@pytest.mark.timeout(300)
def test_c_host_drives_the_queue(tmp_path):
    from mpx import _abi
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    src = tmp_path / "host.c"
    src.write_text(C_HOST)
    exe = tmp_path / "host"
    libdir = os.path.dirname(_abi._LIB_PATH)
    subprocess.run(["gcc", "-std=c99", "-O1", "-D__HIP_PLATFORM_AMD__", "-o", str(exe), str(src),
                    "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include", "-L", libdir, "-lmpx",
                    "-L", "/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath," + libdir, "-Wl,-rpath,/opt/rocm/lib"],
                   check=True)
    print("running the C host program", flush=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr[-3000:])
    n, bad, launches = (int(v) for v in r.stdout.split())
    assert (n, bad) == (12, 0) and launches <= n
