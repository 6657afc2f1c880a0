"""NumPy emulation of the FP8 (OCP e4m3) NNEDI3 contraction form, used to check the GPU
kernel to a tight tolerance (the FP64 oracle measures the form's own error).

The form: weights (softmax half pre-scaled by log2 e) are stored as e4m3 of w * 2^-e with a
power-of-two scale per neuron and matrix (max|w| * 2^-e in [224, 448)); the normalised
window x' is stored as e4m3 high and low parts of 32 * x'; the contraction sums
w8 * 2^e * (h + l) / 32 exactly (FP32 accumulation on the device) and adds the FP32
biases; exp2 / elliott / blend as in the other forms.
"""
import numpy as np

from oracle.nnedi3 import VAR_EPS, load_nnedi3, windows_y

LOG2E = 1.4426950408889634


def e4m3(v):
    """Round-to-nearest-even to OCP FP8 e4m3 (saturating at 448), returned as float64."""
    v = np.asarray(v, dtype=np.float64)
    a = np.abs(v)
    with np.errstate(divide="ignore"):
        e = np.floor(np.log2(np.where(a > 0, a, 1.0)))
    e = np.maximum(e, -6.0)                 # subnormals share the 2^-6 binade's spacing
    ulp = np.exp2(e - 3.0)
    q = np.round(a / ulp) * ulp             # np.round: half to even
    q = np.minimum(q, 448.0)
    return np.sign(v) * q


def _quant_rows(w):
    mx = np.abs(w).max(axis=1)
    e = np.zeros_like(mx)
    nz = mx > 0
    m, ex = np.frexp(mx[nz] / 448.0)
    e[nz] = ex
    e = e + (np.ldexp(mx, -e.astype(int)) > 448.0)
    sc = np.exp2(e)[:, None]
    return e4m3(w / sc) * sc


def predict_fp8(win, w1, w2, b1, b2):
    w1q = _quant_rows(w1 * LOG2E)
    w2q = _quant_rows(w2)
    b1f = (b1 * LOG2E).astype(np.float32).astype(np.float64)
    b2f = b2.astype(np.float32).astype(np.float64)
    shp = win.shape[:-1]
    K = win.shape[-1]
    x = win.reshape(-1, K)
    m = x.sum(axis=1) / K
    xc = x - m[:, None]
    v = (xc * xc).sum(axis=1) / K
    with np.errstate(divide="ignore"):
        inv = np.where(v >= VAR_EPS, 1.0 / np.sqrt(v), 0.0)
    sd = v * inv
    xs = (xc * inv[:, None]).astype(np.float32).astype(np.float64) * 32.0
    h = e4m3(xs)
    lo = e4m3(xs - h)
    xq = (h + lo) / 32.0
    a1 = xq @ w1q.T + b1f[None, :]
    a2 = xq @ w2q.T + b2f[None, :]
    e = np.exp2(a1)
    vs = (e * (a2 / (1.0 + np.abs(a2)))).sum(axis=1)
    out = np.clip(m + 5.0 * vs / e.sum(axis=1) * sd, 0.0, 1.0)
    return out.reshape(shp)


def nnedi3_fp8(plane, nns, wh, axes="yx"):
    """The doubling passes named in `axes` in the FP8 form (FP64 elsewhere)."""
    w1, w2, b1, b2 = load_nnedi3(nns, wh)

    def dy(p):
        pred = predict_fp8(windows_y(p, wh), w1, w2, b1, b2)
        out = np.empty((2 * p.shape[0], p.shape[1]))
        out[0::2] = p
        out[1::2] = pred
        return out

    p = np.asarray(plane, dtype=np.float64)
    if "y" in axes:
        p = dy(p)
    if "x" in axes:
        p = dy(np.ascontiguousarray(p.T)).T
    return p
