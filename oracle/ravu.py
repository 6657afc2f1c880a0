"""FP64 restatement of the RAVU family (lite, lite-ar, full, 3x, zoom, zoom-ar).

Test infrastructure only (see oracle/__init__.py).  Every function returns the
upscaled planes and the per-site bucket rows so bucket agreement can be measured.
"""
import numpy as np

from .common import (EPS, STRENGTH_3X, STRENGTH_LITE, STRENGTH_LOG2, hash_abd, load_lut16,
                     luma_of, lut_row, structure_tensor, tex_off)


def _as_planes(img, channels):
    img = np.asarray(img, dtype=np.float64)
    if channels == "luma":
        assert img.ndim == 2
        return img[None]
    assert img.ndim == 3 and img.shape[0] == 3
    return img


def _out(planes, channels):
    return planes[0] if channels == "luma" else planes


def _window(planes, channels, lo, n):
    """Window samples i = X*n + Y at offsets (X+lo, Y+lo): list of (C,H,W) arrays and the
    list of hash-plane samples."""
    samp = [tex_off(planes, X + lo, Y + lo) for X in range(n) for Y in range(n)]
    if channels == "luma":
        lum = [s[0] for s in samp]
    else:
        lum = [luma_of(s, channels) for s in samp]
    return samp, lum


def _ar_clamp(res, taps, wgt_pairs, strength=0.8):
    """Anti-ringing of ravu-lite-ar-r3.hook:90-182 (EWA-style, libplacebo):
    hi = sum wg*(0.1+l)^33 / sum wg*(0.1+l)^32 - 0.1,
    lo = 1.1 - sum wg*(1.1-l)^33 / sum wg*(1.1-l)^32,
    res = mix(res, clamp(res, lo, hi), strength) with no final [0,1] clamp."""
    hi = lo = hi2 = lo2 = 0.0
    for l, wg in zip(taps, wgt_pairs):
        cp = (0.1 + l) ** 32
        cm = (1.1 - l) ** 32
        hi = hi + cp * wg
        lo = lo + cm * wg
        hi2 = hi2 + cp * (0.1 + l) * wg
        lo2 = lo2 + cm * (1.1 - l) * wg
    with np.errstate(invalid="ignore", divide="ignore"):
        lo_v = 1.1 - lo2 / lo
        hi_v = hi2 / hi - 0.1
        cl = np.minimum(np.maximum(res, lo_v), hi_v)   # GLSL clamp(x, lo, hi) = min(max(x,lo),hi)
    return res * (1.0 - strength) + cl * strength   # GLSL mix(x, y, a) = x*(1-a) + y*a


# ----------------------------------------------------------------------------- ravu-lite

def _f16(a):
    """FP16 storage (round to nearest even) of an FBO-stored texture."""
    return np.asarray(a).astype(np.float16).astype(np.float64)


def ravu_lite(img, r, ar=False, channels="luma", fbo16=False):
    """RAVU-lite 2x (ravu-lite-r3.hook:22-131, ravu-lite-ar-r3.hook:22-196).

    n = 2r-1 window at offsets -(r-1)..(r-1); LUT texel i holds the 4 sub-pixel
    weights of tap i, tap n^2-1-i uses w.wzyx, the centre tap uses its texel alone
    (ravu-lite-r3.hook:89-115).  res[k] goes to output (2x+dx, 2y+dy), k = 2*dx+dy
    (step2, ravu-lite-r3.hook:127-131).  Non-AR clamps to [0,1] (:116); AR taps are those
    with dx^2+dy^2 <= 4, weights max(0,w), strength 0.8 (ravu-lite-ar-r3.hook:95-182)."""
    assert channels == "luma"
    planes = _as_planes(img, channels)
    n = 2 * r - 1
    lut = load_lut16("ravu_lite_lut%d" % r)             # (288, ntex, 4)
    samp, lum = _window(planes, channels, -(r - 1), n)
    a, b, d = structure_tensor(lum, n, five_point=False)
    angle, strength, coherence, diag = hash_abd(a, b, d, STRENGTH_LITE)
    row = lut_row(angle, strength, coherence, 4, lut.shape[0])
    W = lut[row]                                        # (H, W, ntex, 4)
    c = (n * n - 1) // 2
    res = 0.0
    for i in range(c):
        w = W[..., i, :]
        res = res + samp[i][0][..., None] * w + samp[n * n - 1 - i][0][..., None] * w[..., ::-1]
    res = res + samp[c][0][..., None] * W[..., c, :]
    if ar:
        lo = -(r - 1)
        taps, wgts = [], []
        for i in range(c + 1):
            X, Y = divmod(i, n)
            if (X + lo) ** 2 + (Y + lo) ** 2 > 4:
                continue
            wg = np.maximum(0.0, W[..., i, :])
            taps.append(samp[i][0][..., None])
            wgts.append(wg)
            if i != c:
                taps.append(samp[n * n - 1 - i][0][..., None])
                wgts.append(wg[..., ::-1])
        res = _ar_clamp(res, taps, wgts)
    else:
        res = np.clip(res, 0.0, 1.0)
    if fbo16:   # ravu_lite_int (ravu-lite-r3.hook:19-21) and the step-2 output in FP16
        res = _f16(res)
    H, Wd = planes.shape[-2:]
    out = np.empty((2 * H, 2 * Wd))
    for dx in range(2):
        for dy in range(2):
            out[dy::2, dx::2] = res[..., 2 * dx + dy]
    return out, {"row": row, "diag": diag}


# ----------------------------------------------------------------------------- ravu-3x

def ravu_3x(img, r, channels="luma", fbo16=False):
    """RAVU-3x (compute/ravu-3x-r3.hook:22-163).  LUT texels 2i, 2i+1 hold (w0, w1) of
    tap i; res0 += l_i*w0 + l_j*w1.wzyx, res1 += l_i*w1 + l_j*w0.wzyx (j = n^2-1-i);
    output k = 3*dx+dy: k<4 from res0, k=4 the input pixel, k>4 from res1."""
    planes = _as_planes(img, channels)
    n = 2 * r - 1
    lut = load_lut16("ravu_3x_lut%d" % r)               # (216, 2*ntap, 4)
    samp, lum = _window(planes, channels, -(r - 1), n)
    a, b, d = structure_tensor(lum, n, five_point=False)
    angle, strength, coherence, diag = hash_abd(a, b, d, STRENGTH_3X)
    row = lut_row(angle, strength, coherence, 3, lut.shape[0])
    W = lut[row]
    c = (n * n - 1) // 2
    res0 = 0.0
    res1 = 0.0
    for i in range(c):
        w0 = W[None, ..., 2 * i, :]
        w1 = W[None, ..., 2 * i + 1, :]
        li = samp[i][..., None]
        lj = samp[n * n - 1 - i][..., None]
        res0 = res0 + li * w0 + lj * w1[..., ::-1]
        res1 = res1 + li * w1 + lj * w0[..., ::-1]
    res0 = res0 + samp[c][..., None] * W[None, ..., 2 * c, :]
    res1 = res1 + samp[c][..., None] * W[None, ..., 2 * c + 1, :]
    res0 = np.clip(res0, 0.0, 1.0)
    res1 = np.clip(res1, 0.0, 1.0)
    C, H, Wd = planes.shape
    out = np.empty((C, 3 * H, 3 * Wd))
    for dx in range(3):
        for dy in range(3):
            k = 3 * dx + dy
            if k < 4:
                v = res0[..., k]
            elif k == 4:
                v = samp[c]
            else:
                v = res1[..., k - 5]
            out[:, dy::3, dx::3] = v
    if fbo16:   # imageStore into the hook's FP16 output (compute/ravu-3x-r3.hook:158-162)
        out = _f16(out)
    return _out(out, channels), {"row": row, "diag": diag}


# ----------------------------------------------------------------------------- ravu (full)

def _ravu_taps(samp, W, n):
    """res = sum_{p < n^2/2} (s_p + s_{n^2-1-p}) * w_p, w_p = texel p//4 component p%4
    (ravu-r3.hook:120-145)."""
    res = 0.0
    for p in range(n * n // 2):
        w = W[..., p // 4, p % 4]
        res = res + (samp[p] + samp[n * n - 1 - p]) * w[None]
    return np.clip(res, 0.0, 1.0)


def _lattice_samples(planes, int11, channels, n, step):
    """Samples of step2 (int10, step=2) / step3 (int01, step=3) on the 45-degree lattice:
    sample (X, Y) sits at c + (X-(n-1)/2)*u + (Y-(n-1)/2)*v with u = (1/2,-1/2),
    v = (1/2,1/2); c = (x+1/2, y) for step2, (x, y+1/2) for step3.  Integer positions
    read HOOKED, half-integer positions read ravu_int11 at the floor, each clamped on
    its own texture (ravu-r3.hook:155-191, 289-325)."""
    cx, cy = (0.5, 0.0) if step == 2 else (0.0, 0.5)
    h = (n - 1) / 2.0
    samp = []
    for X in range(n):
        for Y in range(n):
            px = cx + (X - h) * 0.5 + (Y - h) * 0.5
            py = cy - (X - h) * 0.5 + (Y - h) * 0.5
            if float(px).is_integer():
                assert float(py).is_integer()
                samp.append(tex_off(planes, int(px), int(py)))
            else:
                samp.append(tex_off(int11, int(np.floor(px)), int(np.floor(py))))
    if channels == "luma":
        lum = [s[0] for s in samp]
    else:
        lum = [luma_of(s, channels) for s in samp]
    return samp, lum


def ravu(img, r, channels="luma", fbo16=False):
    """RAVU 2x in three predictions + combine (ravu-r3.hook:15-437).

    step1: int11 at (x+1/2, y+1/2) from the n = 2r window at offsets -(r-1)..r;
    step2/3: int10 / int01 on the rotated lattice; step4 interleaves
    out(2x,2y)=in, (2x,2y+1)=int01, (2x+1,2y)=int10, (2x+1,2y+1)=int11 (:426-437).
    The result carries the hook's OFFSET -0.5 -0.5."""
    planes = _as_planes(img, channels)
    n = 2 * r
    lut = load_lut16("ravu_lut%d" % r)                  # (648, n*n/8, 4)
    rows = {}
    samp, lum = _window(planes, channels, -(r - 1), n)
    a, b, d = structure_tensor(lum, n, five_point=True)
    ang, st, co, diag11 = hash_abd(a, b, d, STRENGTH_LOG2)
    rows["int11"] = lut_row(ang, st, co, 9, lut.shape[0])
    int11 = _ravu_taps(samp, lut[rows["int11"]], n)
    if fbo16:   # SAVE ravu_int11 (ravu-r3.hook:19) in an FP16 texture, read by steps 2 and 3
        int11 = _f16(int11)
    outs = {}
    diags = {"int11": diag11}
    for step, key in ((2, "int10"), (3, "int01")):
        samp, lum = _lattice_samples(planes, int11, channels, n, step)
        a, b, d = structure_tensor(lum, n, five_point=True)
        ang, st, co, dg = hash_abd(a, b, d, STRENGTH_LOG2)
        rows[key] = lut_row(ang, st, co, 9, lut.shape[0])
        outs[key] = _ravu_taps(samp, lut[rows[key]], n)
        diags[key] = dg
    C, H, Wd = planes.shape
    out = np.empty((C, 2 * H, 2 * Wd))
    out[:, 0::2, 0::2] = planes
    out[:, 1::2, 0::2] = outs["int01"]
    out[:, 0::2, 1::2] = outs["int10"]
    out[:, 1::2, 1::2] = int11
    if fbo16:   # the combine pass's output FBO (the kept pixels are stored too)
        out = _f16(out)
    return _out(out, channels), {"row": rows, "diag": diags}


# ----------------------------------------------------------------------------- ravu-zoom

def _bilinear_lut(lut, group, tx, ty):
    """FILTER LINEAR fetch of 4 components at texel-space coordinates
    (9*group + tx, ty) where tx, ty are arrays of continuous texel coordinates measured
    from texel centres (so an integer lands exactly on a texel)."""
    rows, width = lut.shape[:2]
    x0 = np.floor(tx)
    y0 = np.floor(ty)
    fx = (tx - x0)[..., None]
    fy = (ty - y0)[..., None]
    x0 = x0.astype(np.int64) + 9 * group
    y0 = y0.astype(np.int64)
    x1 = np.clip(x0 + 1, 0, width - 1)
    y1 = np.clip(y0 + 1, 0, rows - 1)
    x0 = np.clip(x0, 0, width - 1)
    y0 = np.clip(y0, 0, rows - 1)
    v00 = lut[y0, x0]
    v01 = lut[y0, x1]
    v10 = lut[y1, x0]
    v11 = lut[y1, x1]
    return (v00 * (1 - fx) + v01 * fx) * (1 - fy) + (v10 * (1 - fx) + v11 * fx) * fy


def ravu_zoom(img, r, out_w, out_h, ar=False, channels="luma", ar_lut=None, cols=None, rows=None, fbo16=False):
    """RAVU-zoom(-ar) to an arbitrary size (ravu-zoom-r2.hook:22-134,
    ravu-zoom-ar-r2-rgb.hook:25-224).

    pos = (o+1/2)*in/out (OFFSET ALIGN); subpix = fract(pos-1/2); base = pos - subpix.
    Window n = 2r at offsets -(r-1)..r around base.  LUT addressing: group g of 4 taps at
    texel x = 9g + 8f, y = 9*bucket + 8f_y (LUTPOS(f) = (1/2+8f)/9 in normalised coords);
    taps p < n^2/2 use phase f, tap n^2-1-p uses phase 1-f with the same (group, comp).
    Non-AR clamps to [0,1]; AR uses the separate learned `_ar` LUT (all taps, weights as
    stored), per channel, strength 0.8, no final clamp.  Base and phase follow from the
    exact rational pos - 1/2 = ((2o+1)*in - out) / (2*out).
    ``cols`` / ``rows`` (optional index arrays) evaluate only those output columns and
    rows (every output pixel depends on its own position alone): the result is the
    (rows, cols) sub-grid of the full output."""
    planes = _as_planes(img, channels)
    C, H, Wd = planes.shape
    n = 2 * r
    lut = load_lut16("ravu_zoom_lut%d" % r)             # (2592, 9*groups, 4)
    if ar:
        if ar_lut is None:
            ar_lut = load_lut16("ravu_zoom_lut%d_ar" % r)
    ox = np.arange(out_w) if cols is None else np.asarray(cols, dtype=np.int64)
    oy = np.arange(out_h) if rows is None else np.asarray(rows, dtype=np.int64)
    numx = (2 * ox + 1) * Wd - out_w
    numy = (2 * oy + 1) * H - out_h
    bx = numx // (2 * out_w)                            # base texel index (floor)
    by = numy // (2 * out_h)
    fx = (numx - bx * 2 * out_w) / (2.0 * out_w)        # subpix = fract(pos - 0.5)
    fy = (numy - by * 2 * out_h) / (2.0 * out_h)
    lo = -(r - 1)
    samp = []
    for X in range(n):
        for Y in range(n):
            xs = np.clip(bx + X + lo, 0, Wd - 1)
            ys = np.clip(by + Y + lo, 0, H - 1)
            samp.append(planes[:, ys[:, None], xs[None, :]])
    if channels == "luma":
        lum = [s[0] for s in samp]
    else:
        lum = [luma_of(s, channels) for s in samp]
    a, b, d = structure_tensor(lum, n, five_point=True)
    angle, strength, coherence, diag = hash_abd(a, b, d, STRENGTH_LITE)
    # unclamped bucket (angle can be -1 or 24 at the mod wrap): the LINEAR fetch clamps
    # texel rows to the texture, not the bucket index
    bucket = ((angle * 4 + strength) * 3 + coherence).astype(np.int64)   # (out_h, out_w)
    FX = np.broadcast_to(fx[None, :], (len(oy), len(ox)))
    FY = np.broadcast_to(fy[:, None], (len(oy), len(ox)))
    ty = 9 * bucket + 8 * FY
    ty_inv = 9 * bucket + 8 * (1 - FY)
    res = 0.0
    half = n * n // 2
    wts = {}
    for g in range((half + 3) // 4):
        wf = _bilinear_lut(lut, g, 8 * FX, ty)
        wi = _bilinear_lut(lut, g, 8 * (1 - FX), ty_inv)
        if ar:
            af = _bilinear_lut(ar_lut, g, 8 * FX, ty)
            ai = _bilinear_lut(ar_lut, g, 8 * (1 - FX), ty_inv)
        for c in range(4):
            p = 4 * g + c
            if p >= half:
                break
            q = n * n - 1 - p
            res = res + samp[p] * wf[None, ..., c] + samp[q] * wi[None, ..., c]
            if ar:
                wts[p] = af[None, ..., c]
                wts[q] = ai[None, ..., c]
    if ar:
        taps = [samp[p] for p in range(n * n)]
        res = _ar_clamp(res, taps, [wts[p] for p in range(n * n)])
    else:
        res = np.clip(res, 0.0, 1.0)
    if fbo16:   # the hook's output texture in FP16
        res = _f16(res)
    return _out(res, channels), {"row": bucket, "diag": diag}
