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


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src):
    obj = os.path.join(BUILD, src + ".o")
    cmd = [HIPCC] + COMMON_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
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


if __name__ == "__main__":
    build()
    sys.exit(0)
