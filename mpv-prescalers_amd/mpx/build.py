"""Build libmpx.so (all HIP kernels + the C ABI) for gfx950 with hipcc, in-tree.

Each source under csrc/ is compiled to an object in ``_build/`` (in parallel), then
linked into ``mpx/libmpx.so`` next to this file.  Usage: ``python -m mpx.build`` from
mpv-prescalers_amd/, or ``build()`` from Python.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys
import time

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(ROOT, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libmpx.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON_FLAGS = [
    "-x", "hip", "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    # packed FP32 ops are disabled for every kernel (they have given silent wrong
    # results on this hardware); the host pass ignores the feature with a warning.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unknown-pragmas",
    "-I", CSRC, "-I", os.path.join(os.path.dirname(ROOT), "include"),
]


# Host-side AddressSanitizer + UndefinedBehaviorSanitizer build of the whole library
# (_build/asan/libmpx_asan.so): the C ABI, weight-file parsing and packing instrumented on
# the CPU; device code is built as usual (GPU sanitizers are not used).  Each -fsanitize=
# is given to the host pass only (-Xarch_host); the link is a plain clang++ host link.
ASAN_DIR = os.path.join(BUILD, "asan")
ASAN_LIB = os.path.join(ASAN_DIR, "libmpx_asan.so")
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
       "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=undefined"]
CLANGXX = os.path.join(os.path.dirname(os.path.dirname(HIPCC)), "llvm", "bin", "clang++")


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src, asan=False):
    obj = os.path.join(ASAN_DIR if asan else BUILD, src + ".o")
    cmd = [HIPCC] + COMMON_FLAGS + (SAN if asan else []) + ["-c", os.path.join(CSRC, src), "-o", obj]
    t = time.time()
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("hipcc failed on %s:\n%s\n%s" % (src, " ".join(cmd), p.stderr[-8000:]))
    warn = "\n".join(l for l in p.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    return obj, time.time() - t, warn


def build(verbose=True, jobs=None):
    """Compile every HIP source and link libmpx.so; returns the library path."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 1, len(srcs))
    objs = []
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        for obj, dt, warn in ex.map(_compile, srcs):
            objs.append(obj)
            if verbose:
                print("  compiled %-28s %.1fs" % (os.path.basename(obj)[:-2], dt), flush=True)
            if warn and verbose:
                print(warn, flush=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-ldl"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), p.stderr[-8000:]))
    shutil.move(tmp, LIB)
    if verbose:
        print("  linked", LIB, flush=True)
    return LIB


def build_asan(verbose=True, jobs=None):
    """The sanitizer build (host code instrumented), rebuilt only when a source is newer;
    returns the library path.  Load it only into a process whose executable links the
    sanitizer runtime (e.g. a C program built with clang -fsanitize=address,undefined
    -shared-libsan)."""
    os.makedirs(ASAN_DIR, exist_ok=True)
    srcs = sources()
    newest = max(os.path.getmtime(os.path.join(CSRC, f)) for f in os.listdir(CSRC))
    newest = max(newest, os.path.getmtime(os.path.join(os.path.dirname(ROOT), "include", "mpx.h")))
    if os.path.exists(ASAN_LIB) and os.path.getmtime(ASAN_LIB) > newest:
        return ASAN_LIB
    jobs = jobs or min(8, os.cpu_count() or 1, len(srcs))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = [o for o, _, _ in ex.map(lambda f: _compile(f, asan=True), srcs)]
    tmp = ASAN_LIB + ".tmp"
    cmd = [CLANGXX, "-shared", "-fPIC", "-fsanitize=address,undefined", "-shared-libsan", "-o", tmp] + objs + [
        "-L" + os.path.join(os.path.dirname(os.path.dirname(HIPCC)), "lib"), "-lamdhip64", "-ldl",
        "-Wl,-rpath," + os.path.join(os.path.dirname(os.path.dirname(HIPCC)), "lib")]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("sanitizer link failed:\n%s\n%s" % (" ".join(cmd), p.stderr[-8000:]))
    shutil.move(tmp, ASAN_LIB)
    if verbose:
        print("  linked", ASAN_LIB, flush=True)
    return ASAN_LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
