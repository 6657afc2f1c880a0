"""Build the oracle's C restatement (oracle/c/mpx_cpu.c) into oracle/_ref/libmpx_cpu.so
and a ctypes wrapper -- test infrastructure only (CPU baseline + cross-check).

The reference itself (GLSL user shaders) has no C/C++ code that could be compiled here;
the directory name follows the convention for built checkers that travel with the tree.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")
LIB = os.path.join(OUT, "libmpx_cpu.so")
SRC = os.path.join(HERE, "c", "mpx_cpu.c")


def build():
    os.makedirs(OUT, exist_ok=True)
    cmd = ["gcc", "-O3", "-march=x86-64-v2", "-fopenmp", "-fPIC", "-shared", "-o", LIB + ".tmp", SRC, "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
        fp, i = ctypes.c_void_p, ctypes.c_int
        _lib.nnedi3_double_y.argtypes = [fp, i, i, fp, fp, i, i]
        _lib.transpose_f32.argtypes = [fp, i, i, fp]
        _lib.ravu_lite.argtypes = [fp, i, i, fp, fp, i, i, fp]
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def nnedi3(plane, nns, wh):
    """Both doubling passes, FP32: (H, W) -> (2H, 2W)."""
    from .common import load_weights
    wt = np.ascontiguousarray(load_weights("nnedi3_nns%d_win8x%d" % (nns, wh)).astype(np.float32))
    x = np.ascontiguousarray(plane, dtype=np.float32)
    h, w = x.shape
    L = lib()
    y = np.empty((2 * h, w), np.float32)
    L.nnedi3_double_y(_p(x), w, h, _p(y), _p(wt), nns, wh)
    yt = np.empty((w, 2 * h), np.float32)
    L.transpose_f32(_p(y), w, 2 * h, _p(yt))
    z = np.empty((2 * w, 2 * h), np.float32)
    L.nnedi3_double_y(_p(yt), 2 * h, w, _p(z), _p(wt), nns, wh)
    out = np.empty((2 * h, 2 * w), np.float32)
    L.transpose_f32(_p(z), 2 * h, 2 * w, _p(out))
    return out


def ravu_lite(plane, r, ar=False):
    from .common import gaussian_grid, load_lut16
    lut = np.ascontiguousarray(load_lut16("ravu_lite_lut%d" % r).astype(np.float32))
    g = np.ascontiguousarray(gaussian_grid(2 * r - 1)[1], dtype=np.float64)
    x = np.ascontiguousarray(plane, dtype=np.float32)
    h, w = x.shape
    out = np.empty((2 * h, 2 * w), np.float32)
    lib().ravu_lite(_p(x), w, h, _p(out), _p(lut), r, int(ar), _p(g))
    return out


if __name__ == "__main__":
    print(build())
