"""FP64 restatement of the scaling pass that consumes a prescaler's OFFSET -- test
infrastructure only (see oracle/__init__.py).

The reference ships no scaler: mpv applies the accumulated `//!OFFSET` of the hooks in its
main scaling pass (e.g. nnedi3-nns256-win8x6.hook:349-352 `//!OFFSET 0.000000 -0.500000`,
ravu-r3.hook:426-428 `//!OFFSET -0.500000 -0.500000`).  mpv is not in this image, so the
filters are restated from their published definitions (tent; Mitchell-Netravali cubic
with (B, C) = (0, 1/2) Catmull-Rom and (1/3, 1/3) Mitchell; spline36; Lanczos-3) and the
sampling convention from the hook semantics: display pixel j samples the prescaled plane
at t = (j + 1/2) * in/out + offset, taps over texel centres i + 1/2 within the radius,
weights normalised to 1, clamp-to-edge.
Parity with mpv's own scaler is unpinned (mpv absent); this restatement is pinned by
closed forms in tests/test_resample.py (identity, integer shifts, constants, linear
ramps, weight sums).
"""
import numpy as np

RADIUS = {"bilinear": 1, "catmull_rom": 2, "mitchell": 2, "spline36": 3, "lanczos3": 3}


def kernel(name, x):
    x = np.abs(np.asarray(x, dtype=np.float64))
    if name == "bilinear":
        return np.maximum(0.0, 1.0 - x)
    if name in ("catmull_rom", "mitchell"):
        B, C = (0.0, 0.5) if name == "catmull_rom" else (1 / 3, 1 / 3)
        a = ((12 - 9 * B - 6 * C) * x ** 3 + (-18 + 12 * B + 6 * C) * x ** 2 + (6 - 2 * B)) / 6
        b = ((-B - 6 * C) * x ** 3 + (6 * B + 30 * C) * x ** 2 + (-12 * B - 48 * C) * x + (8 * B + 24 * C)) / 6
        return np.where(x < 1, a, np.where(x < 2, b, 0.0))
    if name == "spline36":
        u1, u2 = x - 1, x - 2
        a = ((13 / 11 * x - 453 / 209) * x - 3 / 209) * x + 1
        b = ((-6 / 11 * u1 + 270 / 209) * u1 - 156 / 209) * u1
        c = ((1 / 11 * u2 - 45 / 209) * u2 + 26 / 209) * u2
        return np.where(x < 1, a, np.where(x < 2, b, np.where(x < 3, c, 0.0)))
    if name == "lanczos3":
        return np.where(x < 3, np.sinc(x) * np.sinc(x / 3), 0.0)
    raise ValueError(name)


def axis_matrix(name, in_n, out_n, offset):
    """(out_n, in_n) resampling matrix of one axis."""
    r = RADIUS[name]
    m = np.zeros((out_n, in_n))
    for j in range(out_n):
        p = (j + 0.5) * in_n / out_n + offset - 0.5
        fl = np.floor(p)
        idx = fl - r + 1 + np.arange(2 * r)
        w = kernel(name, p - idx)
        w = w / w.sum()
        for i, wi in zip(np.clip(idx.astype(int), 0, in_n - 1), w):
            m[j, i] += wi
    return m


def resample(plane, out_w, out_h, offset=(0.0, 0.0), name="catmull_rom"):
    """plane (H, W) or (..., H, W) -> (..., out_h, out_w)."""
    p = np.asarray(plane, dtype=np.float64)
    my = axis_matrix(name, p.shape[-2], out_h, offset[1])
    mx = axis_matrix(name, p.shape[-1], out_w, offset[0])
    return np.einsum("ji,...ik,lk->...jl", my, p, mx)
