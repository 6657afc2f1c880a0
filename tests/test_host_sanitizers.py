"""The library's host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU).

mpx.build.build_asan() compiles every source with the host pass instrumented
(_build/asan/libmpx_asan.so); a C driver built with clang -fsanitize=address,undefined
then walks the ABI's host-reachable paths: the registry, mpx_query over every name and
hostile sizes, NULL arguments to every entry point, mpx_weights_verify on every shipped
variant and on damaged weight files (truncated, padded, bad magic, huge and wrong
shapes, one flipped bit), and mpx_create without a usable device.  Any out-of-bounds
access, use after free or undefined behaviour aborts the driver with a report."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from conftest import PKG, ROOT

DRIVER = r'''
#include <limits.h>
#include <stdio.h>
#include <string.h>
#include "mpx.h"

static int verify_all(const char* dir) {
  int bad = 0;
  for (int i = 0; i < mpx_variant_count(); ++i) {
    const char* n = mpx_variant_name(i);
    int rc = mpx_weights_verify(dir, n);
    if (rc != MPX_OK && rc != MPX_ERR_UNAVAILABLE && rc != MPX_ERR_IO) return -100 - i;
    if (rc == MPX_ERR_IO) ++bad;
    (void)strlen(mpx_last_error());
  }
  return bad;
}

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  mpx_variant_info info;
  const int sizes[][4] = {{1, 1, 2, 2}, {640, 360, 1280, 720}, {0, 5, 10, 10}, {-3, 5, 10, 10},
                          {INT_MAX, INT_MAX, INT_MAX, INT_MAX}, {32768, 32768, 65536, 65536},
                          {7, 5, 3, 2}, {1919, 1079, 2880, 1620}};
  int n = mpx_variant_count();
  if (n != 46 || mpx_variant_name(-1) != NULL || mpx_variant_name(n) != NULL) return 3;
  for (int i = 0; i < n; ++i)
    for (unsigned s = 0; s < sizeof(sizes) / sizeof(sizes[0]); ++s) {
      int rc = mpx_query(mpx_variant_name(i), sizes[s][0], sizes[s][1], sizes[s][2], sizes[s][3], &info);
      if (rc != MPX_OK && rc != MPX_ERR_SHAPE) return 4;
    }
  if (mpx_query(NULL, 8, 8, 16, 16, &info) != MPX_ERR_UNKNOWN_VARIANT) return 5;
  if (mpx_query("ravu-r3", 8, 8, 16, 16, NULL) != MPX_ERR_ARG) return 6;
  if (mpx_weights_verify(argv[1], NULL) != MPX_ERR_UNKNOWN_VARIANT) return 7;
  /* NULL handles everywhere: status codes, no crash */
  mpx_plan* plan = NULL;
  mpx_io io;
  memset(&io, 0, sizeof(io));
  if (mpx_plan_create(NULL, "ravu-r3", 8, 8, 16, 16, MPX_F32, 8, MPX_F32, 0, 1, &plan) != MPX_ERR_ARG) return 8;
  if (mpx_run(NULL, NULL, 0, 0, NULL, 0, 0, 1, NULL) != MPX_ERR_ARG) return 9;
  if (mpx_run_ex(NULL, &io, NULL) != MPX_ERR_ARG) return 10;
  if (mpx_run_static(NULL, &io, NULL) != MPX_ERR_ARG) return 11;
  if (mpx_run_frames(NULL, 1, NULL, NULL, 8, 16, NULL) != MPX_ERR_ARG) return 12;
  if (mpx_run_frames_sized(NULL, 1, NULL, NULL, NULL, NULL, NULL, NULL, NULL) != MPX_ERR_ARG) return 13;
  if (mpx_resample(NULL, 8, 8, MPX_F32, 8, 16, 16, MPX_F32, 1, 0, 0.f, 0.f, NULL) != MPX_ERR_ARG) return 14;
  if (mpx_resample(&io, 8, 8, MPX_F32, 8, 16, 16, MPX_F32, 1, 0, 0.f, 0.f, NULL) != MPX_ERR_ARG) return 15;
  if (mpx_plan_info(NULL, &info) != MPX_ERR_ARG || mpx_plan_bucket_count(NULL) != 0) return 16;
  if (mpx_plan_tune(NULL, &io, NULL, 1, NULL, NULL, NULL) != MPX_ERR_ARG) return 17;
  if (mpx_plan_set_shape(NULL, 0, 0) != MPX_ERR_ARG || mpx_ctx_bytes(NULL) != 0) return 18;
  if (mpx_plan_set_tile(NULL, 0) != MPX_ERR_ARG || mpx_plan_tile(NULL, NULL, NULL, NULL, NULL) != MPX_ERR_ARG) return 25;
  mpx_plan_destroy(NULL);
  mpx_destroy(NULL);
  mpx_queue* q = NULL;
  uint64_t ticket = 0;
  mpx_request req;
  memset(&req, 0, sizeof(req));
  if (mpx_queue_create(NULL, 2, 4, 64, 64, &q) != MPX_ERR_ARG) return 21;
  if (mpx_submit(NULL, &req, &ticket) != MPX_ERR_ARG || mpx_submit(q, NULL, &ticket) != MPX_ERR_ARG) return 22;
  if (mpx_poll(NULL, 1, 0, NULL) != MPX_ERR_ARG || mpx_queue_launches(NULL) != 0) return 23;
  if (mpx_queue_hold(NULL, 1) != MPX_ERR_ARG) return 24;
  mpx_queue_destroy(NULL);
  /* no usable device in this process: a status, never an abort */
  mpx_ctx* ctx = NULL;
  int crc = mpx_create(0, 0, argv[1], &ctx);
  if (crc == MPX_OK) mpx_destroy(ctx);
  else if (crc != MPX_ERR_HIP && crc != MPX_ERR_ARG) return 19;
  if (mpx_create(-1, 0, NULL, &ctx) >= 0) return 20;
  int good = verify_all(argv[1]);
  int damaged = verify_all(argv[2]);
  printf("%d %d\n", good, damaged);
  return 0;
}
'''


def _write(path, arr, dims=None, magic=b"MPXW"):
    arr = np.ascontiguousarray(arr, dtype="<f4")
    dims = arr.shape if dims is None else dims
    with open(path, "wb") as f:
        f.write(magic + struct.pack("<III", 1, 0, len(dims)) + struct.pack("<%dI" % len(dims), *dims))
        f.write(arr.tobytes())


@pytest.mark.timeout(900)
def test_host_code_under_asan_and_ubsan(tmp_path):
    clang = "/opt/rocm/llvm/bin/clang"
    if not (os.path.exists(clang) and os.path.exists("/opt/rocm/bin/hipcc")):
        pytest.skip("ROCm clang / hipcc not available")
    import importlib.util
    spec = importlib.util.spec_from_file_location("mpx_build", os.path.join(PKG, "mpx", "build.py"))
    b = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(b)
    lib = b.build_asan(verbose=False)
    wdir = os.path.join(PKG, "weights")
    bad = tmp_path / "bad"
    shutil.copytree(wdir, bad, ignore=shutil.ignore_patterns("__pycache__", "*.py"))
    raw = open(os.path.join(wdir, "ravu_lite_lut3.bin"), "rb").read()
    open(bad / "ravu_lite_lut3.bin", "wb").write(raw[:-4])                     # truncated
    open(bad / "ravu_lite_lut2.bin", "wb").write(raw[:9])                      # header cut short
    open(bad / "ravu_lite_lut4.bin", "wb").write(raw + b"\0" * 8)              # padded
    _write(bad / "ravu_3x_lut2.bin", np.zeros(4, np.float32), dims=(0xFFFFFFFF, 0xFFFFFFFF, 4))
    _write(bad / "ravu_3x_lut3.bin", np.zeros((10, 4), np.float32), magic=b"WXPM")
    _write(bad / "ravu_3x_lut4.bin", np.zeros((3, 3, 3, 3, 3, 3, 3, 3, 3), np.float32))  # rank 9
    p = bad / "nnedi3_nns64_win8x6.bin"
    blob = bytearray(open(p, "rb").read())
    blob[len(blob) // 2] ^= 0x01
    open(p, "wb").write(bytes(blob))
    src = tmp_path / "drive.c"
    src.write_text(DRIVER)
    exe = tmp_path / "drive"
    libdir = os.path.dirname(lib)
    subprocess.run([clang, "-std=c99", "-g", "-fsanitize=address,undefined", "-shared-libsan", "-o", str(exe),
                    str(src), "-I", os.path.join(ROOT, "include"), libdir + "/libmpx_asan.so",
                    "-Wl,-rpath," + libdir], check=True)
    env = dict(os.environ)
    rt = [d for d in ("/opt/rocm/llvm/lib/clang/22/lib/x86_64-unknown-linux-gnu",
                      "/opt/rocm/llvm/lib/clang/22/lib/linux") if os.path.isdir(d)]
    env["LD_LIBRARY_PATH"] = ":".join(rt + [env.get("LD_LIBRARY_PATH", "")])
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["HIP_VISIBLE_DEVICES"] = ""   # no device in the process, here and on a GPU machine
    r = subprocess.run([str(exe), wdir, str(bad)], env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-6000:])
    good, damaged = (int(v) for v in r.stdout.split())
    assert good == 0
    # variants reading a damaged file: ravu-lite r2/r3/r4 (+ -ar), ravu-3x r2/r3/r4 in 3
    # channel forms, nnedi3-nns64-win8x6
    assert damaged == 6 + 9 + 1, r.stdout
