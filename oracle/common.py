"""Shared oracle pieces: weight files, clamp-to-edge sampling, the structure-tensor hash.

Test infrastructure only (see oracle/__init__.py).
"""
import math
import os
import struct

import numpy as np

EPS = 1.192092896e-7          # ravu-lite-r3.hook:81-83 and every other RAVU hash
PI = 3.141592653589793        # ravu-lite-r3.hook:81,85
BT709 = (0.2126, 0.7152, 0.0722)  # ravu-r3-rgb.hook:21 (color_primary)

WEIGHTS_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                           "mpv-prescalers_amd", "weights")


def load_weights(name, weights_dir=None):
    """Read a MPXW float32 weight file written by the extractor."""
    path = os.path.join(weights_dir or WEIGHTS_DIR, name + ".bin")
    with open(path, "rb") as f:
        data = f.read()
    if data[:4] != b"MPXW":
        raise ValueError("bad weight file %s" % path)
    ver, dt, nd = struct.unpack_from("<III", data, 4)
    dims = struct.unpack_from("<%dI" % nd, data, 16)
    return np.frombuffer(data, "<f4", offset=16 + 4 * nd).reshape(dims).astype(np.float64)


def load_lut16(name, weights_dir=None):
    """LUT as the GPU sees it: float32 hex rounded to FP16 (RNE), shape (rows, width, 4)."""
    lut = load_weights(name, weights_dir).astype(np.float32)
    return lut.astype(np.float16).astype(np.float64)


def tex_off(img, dx, dy):
    """``HOOKED_texOff(vec2(dx, dy))`` for every pixel: clamp-to-edge shift of a plane.

    Works on (..., H, W).  dx, dy are integers."""
    H, W = img.shape[-2:]
    ys = np.clip(np.arange(H) + dy, 0, H - 1)
    xs = np.clip(np.arange(W) + dx, 0, W - 1)
    return img[..., ys[:, None], xs[None, :]]


def luma_of(img, channels):
    """Plane the structure tensor is computed from.

    luma: the plane itself; yuv: ``.x`` (ravu-r3-yuv.hook:22-58 use sample.x);
    rgb: ``dot(rgb, color_primary)`` (ravu-zoom-ar-r2-rgb.hook:24,50-65)."""
    if channels == "luma":
        return img
    if channels == "yuv":
        return img[0]
    if channels == "rgb":
        return BT709[0] * img[0] + BT709[1] * img[1] + BT709[2] * img[2]
    raise ValueError(channels)


# ----------------------------------------------------------------------------- stencil

def gaussian_grid(n):
    """Gradient grid positions and normalised Gaussian weights for an n x n window.

    All positions when n <= 4, else positions 1..n-2 on each axis; weight
    exp(-((X-c)^2+(Y-c)^2)/8) normalised over the grid, c = (n-1)/2.  Computed with
    ``math.exp`` and a left-to-right sum, this reproduces the FP64 literals of
    ravu-lite-r3.hook:50-76 (n=5), ravu-r3.hook:60-107 (n=6), ravu-zoom-r2.hook:50-97
    (n=4) bit for bit."""
    if n <= 4:
        pos = list(range(n))
    else:
        pos = list(range(1, n - 1))
    c = (n - 1) / 2.0
    pts = [(X, Y) for X in pos for Y in pos]
    g = [math.exp(-((X - c) ** 2 + (Y - c) ** 2) / 8.0) for X, Y in pts]
    tot = sum(g)
    return pts, [v / tot for v in g]


def axis_derivative(s, n, p, five_point):
    """Derivative along one axis at position p of the window line ``s`` (indexable by
    position), in the shaders' own formula forms:

    p = 0: ``s[1]-s[0]`` (ravu-zoom-r2.hook:50); p = n-1: ``s[n-1]-s[n-2]``;
    5-point ``(-s[p+2]+8.0*s[p+1]-8.0*s[p-1]+s[p-2])/12.0`` when allowed and p+-2 lie in
    the window (ravu-r3.hook:63); otherwise central ``(s[p+1]-s[p-1])/2.0``
    (ravu-lite-r3.hook:50)."""
    if p == 0:
        return s[1] - s[0]
    if p == n - 1:
        return s[n - 1] - s[n - 2]
    if five_point and p - 2 >= 0 and p + 2 <= n - 1:
        return (-s[p + 2] + 8.0 * s[p + 1] - 8.0 * s[p - 1] + s[p - 2]) / 12.0
    return (s[p + 1] - s[p - 1]) / 2.0


def structure_tensor(samples, n, five_point):
    """abd = sum_g (gx*gx, gx*gy, gy*gy) * G_g over the gradient grid, accumulated in the
    shaders' order (grid X outer, Y inner; ``abd += vec3(...) * G``).

    ``samples[i]`` is the window sample i = X*n + Y (X = column offset, outer), as the
    ``sampleI``/``lumaI`` variables of the shaders."""
    pts, g = gaussian_grid(n)
    a = b = d = 0.0
    for (X, Y), wgt in zip(pts, g):
        col = _Line(samples, lambda k, Y=Y: k * n + Y)
        row = _Line(samples, lambda k, X=X: X * n + k)
        gx = axis_derivative(col, n, X, five_point)
        gy = axis_derivative(row, n, Y, five_point)
        a = a + gx * gx * wgt
        b = b + gx * gy * wgt
        d = d + gy * gy * wgt
    return a, b, d


class _Line:
    def __init__(self, samples, index):
        self.samples = samples
        self.index = index

    def __getitem__(self, k):
        return self.samples[self.index(k)]


STRENGTH_LITE = "lite"    # ravu-lite-r3.hook:86, ravu-zoom-r2.hook:107
STRENGTH_3X = "3x"        # compute/ravu-3x-r3.hook:95
STRENGTH_LOG2 = "log2"    # ravu-r3.hook:117


def hash_abd(a, b, d, rule):
    """Eigen-analysis + bucket of a structure tensor (ravu-lite-r3.hook:77-87).

    Returns (angle, strength, coherence, diag) with float arrays; ``diag`` holds the
    continuous quantities (theta, lambda, mu, L2) used for boundary analysis.
    NaN semantics follow GLSL/IEEE: sqrt of a negative L2 gives NaN, every comparison
    with NaN is false (so coherence 0), ``mix(x, y, bool)`` is a select."""
    with np.errstate(invalid="ignore", divide="ignore"):
        T = a + d
        D = a * d - b * b
        delta = np.sqrt(np.maximum(T * T / 4.0 - D, 0.0))
        L1 = T / 2.0 + delta
        L2 = T / 2.0 - delta
        sqrtL1 = np.sqrt(L1)
        sqrtL2 = np.sqrt(L2)
        at = np.arctan2(L1 - a, b) + PI
        theta = at - PI * np.floor(at / PI)          # GLSL mod(x, y) = x - y*floor(x/y)
        theta = np.where(np.abs(b) < EPS, 0.0, theta)
        lam = sqrtL1
        mu = (sqrtL1 - sqrtL2) / (sqrtL1 + sqrtL2)
        mu = np.where(sqrtL1 + sqrtL2 < EPS, 0.0, mu)
        angle = np.floor(theta * 24.0 / PI)
        if rule == STRENGTH_LITE:
            strength = np.where(lam >= 0.016, np.where(lam >= 0.05, 3.0, 2.0),
                                np.where(lam >= 0.004, 1.0, 0.0))
        elif rule == STRENGTH_3X:
            strength = np.where(lam >= 0.02, 2.0, np.where(lam >= 0.005, 1.0, 0.0))
        elif rule == STRENGTH_LOG2:
            strength = np.clip(np.floor(np.log2(lam * 2000.0 + EPS)), 0.0, 8.0)
        else:
            raise ValueError(rule)
        coherence = np.where(mu >= 0.5, 2.0, np.where(mu >= 0.25, 1.0, 0.0))
    return angle, strength, coherence, {"theta": theta, "lambda": lam, "mu": mu, "L2": L2}


def lut_row(angle, strength, coherence, n_strength, rows):
    """Texel row hit by ``texture(lut, vec2(x, ((angle*S + strength)*3 + coherence + 0.5)/rows))``
    under NEAREST filtering with clamp-to-edge: out-of-range rows (angle = 24, or
    angle = -1 from the mod wrap) clamp to the first/last row."""
    row = (angle * n_strength + strength) * 3.0 + coherence
    row = np.where(np.isnan(row), 0.0, row)
    return np.clip(row, 0, rows - 1).astype(np.int64)
