"""FP64 restatement of NNEDI3 luma doubling (nnedi3-nns256-win8x6.hook:15-706).

Test infrastructure only (see oracle/__init__.py).
"""
import concurrent.futures
import os

import numpy as np

from .common import load_weights, tex_off

VAR_EPS = 1.192092896e-7   # nnedi3-nns16-win8x4.hook:28 `mstd1 >= 1.192092896e-7`


def load_nnedi3(nns, wh):
    """(w1[N,K], w2[N,K], b1[N], b2[N]) in window order k = (dx+3)*wh + (dy-ymin)."""
    t = load_weights("nnedi3_nns%d_win8x%d" % (nns, wh))
    K = 8 * wh
    return t[:, :K], t[:, K:2 * K], t[:, 2 * K], t[:, 2 * K + 1]


def windows_y(plane, wh):
    """(H, W, K) windows of double_y: x offsets -3..4 (outer), y offsets ymin..ymin+wh-1,
    ymin = -(wh/2-1) (nnedi3-nns16-win8x4.hook:39-72 for wh=4; win8x6: -2..3)."""
    ymin = -(wh // 2 - 1)
    return np.stack([tex_off(plane, dx, dy) for dx in range(-3, 5) for dy in range(ymin, ymin + wh)],
                    axis=-1)


def predict(win, w1, w2, b1, b2, chunk=1 << 11):
    """nnedi3() of nnedi3-nns16-win8x4.hook:20-36:
    m = mean, v = E[x^2]-m^2, is = v >= eps ? v^-1/2 : 0, s = v*is;
    e_n = exp(dot(w1_n,x)*is + b1_n); t_n = dot(w2_n,x)*is + b2_n;
    out = clamp(m + 5 * (sum e_n t_n/(1+|t_n|)) / (sum e_n) * s, 0, 1).
    Pixels are independent: chunks of them run on a thread pool (NumPy releases the
    interpreter lock inside its array operations), each chunk computed the same way."""
    shp = win.shape[:-1]
    K = win.shape[-1]
    x = win.reshape(-1, K)
    out = np.empty(x.shape[0])

    def run(s):
        xs = x[s:s + chunk]
        m = xs.sum(axis=1) / K
        v = (xs * xs).sum(axis=1) / K - m * m
        with np.errstate(divide="ignore", invalid="ignore"):
            inv = np.where(v >= VAR_EPS, 1.0 / np.sqrt(v), 0.0)
        sd = v * inv
        s1 = (xs @ w1.T) * inv[:, None] + b1[None, :]
        s2 = (xs @ w2.T) * inv[:, None] + b2[None, :]
        e = np.exp(s1)
        wsum = e.sum(axis=1)
        vsum = (e * (s2 / (1.0 + np.abs(s2)))).sum(axis=1)
        out[s:s + chunk] = np.clip(m + 5.0 * vsum / wsum * sd, 0.0, 1.0)

    starts = range(0, x.shape[0], chunk)
    if x.shape[0] <= 2 * chunk:
        for s in starts:
            run(s)
    else:
        from threadpoolctl import threadpool_limits
        with threadpool_limits(1), concurrent.futures.ThreadPoolExecutor(_threads()) as ex:   # one BLAS thread per chunk
            list(ex.map(run, starts))
    return out.reshape(shp)


def _threads():
    return max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", 0)) or (os.cpu_count() or 1)))


def double_y(plane, nns, wh):
    """double_y + combine_y: output (2H, W), row 2y = input row y, row 2y+1 = the
    prediction between rows y and y+1 (combine_y, nnedi3-nns16-win8x4.hook:90-104).
    Carries OFFSET 0 -0.5."""
    w1, w2, b1, b2 = load_nnedi3(nns, wh)
    pred = predict(windows_y(plane, wh), w1, w2, b1, b2)
    H, W = plane.shape
    out = np.empty((2 * H, W))
    out[0::2] = plane
    out[1::2] = pred
    return out


def double_x(plane, nns, wh):
    """double_x + combine_x: the transpose of double_y (the x-pass weights are the exact
    transpose of the y-pass ones, nnedi3-nns16-win8x4.hook:105-205).  Output (H, 2W)."""
    return double_y(np.ascontiguousarray(plane.T), nns, wh).T


def _f16(a):
    """FP16 storage (round to nearest even) of an FBO-stored texture."""
    return np.asarray(a).astype(np.float16).astype(np.float64)


def nnedi3(plane, nns, wh, axes="yx", fbo16=False):
    """Full hook: double_y (if 'y' in axes) then double_x (if 'x' in axes).
    fbo16: each combine pass's output (the hooked plane: kept rows and the SAVEd
    predictions, nnedi3-nns256-win8x6.hook:18,346-360) is stored in FP16, as mpv's
    default rgba16f FBOs store it, and the next pass reads those values."""
    p = np.asarray(plane, dtype=np.float64)
    if "y" in axes:
        p = double_y(p, nns, wh)
        if fbo16:
            p = _f16(p)
    if "x" in axes:
        p = double_x(p, nns, wh)
        if fbo16:
            p = _f16(p)
    return p
