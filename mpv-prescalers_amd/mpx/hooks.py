"""mpv's hook semantics around the prescalers: which planes a variant hooks, whether it
runs (WHEN), the geometry of its output and the OFFSET the scaler after it consumes.

A frame is a dict of device planes: ``{"Y", "U", "V"}`` for YUV sources (chroma may be
subsampled) or ``{"RGB"}`` holding a planar (3, H, W) tensor.  Routing follows the hook
points of the reference:

* ``//!HOOK LUMA`` (NNEDI3, ravu, ravu-lite, ravu-3x, ravu-zoom luma forms, e.g.
  ravu-lite-r3.hook:16): the Y plane is replaced; chroma is left to the scaler.  An RGB
  source has no LUMA plane, so the hook does not run.
* ``//!HOOK NATIVE`` with ``LUMA.w 0 >`` in WHEN (the -yuv forms, ravu-r3-yuv.hook:16-20):
  the three planes together, hash from Y.  NATIVE is the merged image after chroma
  scaling (README.md:51-53: the -yuv forms run after --cscale), so subsampled chroma is
  first resampled to the luma size (``cscale`` filter through mpx_resample, chroma siting
  ``chroma_loc``: "left" -- mpv's assumption for 4:2:0 video that does not say, MPEG-2 /
  H.264 siting: horizontally on the even luma columns, vertically between rows -- or
  "center", JPEG / MPEG-1 siting), then the three planes run together.
* ``//!HOOK MAIN`` (the -rgb forms, ravu-r3-rgb.hook:16-21): the RGB planes, hash from
  BT.709 luma.  MAIN is the image after mpv's conversion to RGB, so a YUV source is first
  brought to RGB: chroma resampled to the luma size as for NATIVE, then the YUV -> R'G'B'
  matrix of the source (``matrix`` BT.709 or BT.601, ``levels`` limited "tv" range or
  full "pc" range), values not clipped (mpv's intermediate textures are float).

``WHEN`` is evaluated against the display size (mpv's OUTPUT) on the hooked plane's
size; when it is false the frame passes through unchanged.  ``present`` then resamples
every plane to the display size, consuming the accumulated OFFSET of the hooked planes
(the step mpv's main scaler performs).
"""
import torch

from .ops import query, resample, upscale


def chroma_offset(luma_size, chroma_size, chroma_loc="left"):
    """mpx_resample offset (in chroma pixels) that samples a chroma plane at the luma
    pixel centres when it is resampled to the luma size.  Centre siting: chroma sample k
    covers luma samples s*k .. s*k+s-1 and sits at their centre, offset 0.  Left siting
    (horizontal only): chroma sample k sits on luma sample s*k, i.e. (s-1)/2 luma pixels
    left of the centre, so the luma centre j+1/2 maps to chroma (j+1/2)/s + (s-1)/(2s)."""
    (lw, lh), (cw, ch) = luma_size, chroma_size
    ox = 0.0
    if chroma_loc == "left" and cw < lw:
        s = lw / cw
        ox = (s - 1.0) / (2.0 * s)
    elif chroma_loc not in ("left", "center"):
        raise ValueError("chroma_loc must be 'left' or 'center'")
    return ox, 0.0


def cscale(frame, cscale="bilinear", chroma_loc="left", bits=10):
    """The chroma planes of a YUV frame resampled to the luma size (mpv's --cscale step
    before NATIVE hooks); planes already at the luma size are kept.  Integer planes hold
    codes of ``bits`` bits; when chroma is resampled (a float result), integer planes are
    brought to the same form, code / (2^bits - 1) in float32, as the kernels read them."""
    y = frame["Y"]
    lh, lw = y.shape[-2:]
    fdt = y.dtype if y.dtype in (torch.float16, torch.float32) else torch.float32
    out = dict(frame)
    resampled = False
    for k in ("U", "V"):
        c = frame[k]
        ch, cw = c.shape[-2:]
        if (ch, cw) != (lh, lw):
            out[k] = resample(c, (lw, lh), offset=chroma_offset((lw, lh), (cw, ch), chroma_loc), filter=cscale,
                              out_dtype=fdt, bits=bits)
            resampled = True
    if resampled:
        m = float((1 << bits) - 1)
        for k in ("Y", "U", "V"):
            if out[k].dtype not in (torch.float16, torch.float32):
                out[k] = out[k].to(torch.float32) / m
    return out


# Kr, Kb of the luma equation Y' = Kr R' + (1 - Kr - Kb) G' + Kb B' (ITU-R BT.709 and
# BT.601); the inverse gives R' = Y' + 2 (1 - Kr) Cr, B' = Y' + 2 (1 - Kb) Cb and G' from
# the luma equation.
_MATRICES = {"bt709": (0.2126, 0.0722), "bt601": (0.299, 0.114)}


def yuv_to_rgb(frame, bits=10, matrix="bt709", levels="limited", cscale_filter="bilinear", chroma_loc="left"):
    """A YUV frame as a planar (3, H, W) float32 R'G'B' tensor at the luma size (mpv's
    conversion before MAIN hooks).  Integer planes hold codes of ``bits`` bits; float planes
    hold code / (2^bits - 1).  Limited range: Y' = (code - 16 s) / (219 s) and
    Cb, Cr = (code - 128 s) / (224 s) with s = 2^(bits - 8); full range: Y' = code / m and
    Cb, Cr = (code - 2^(bits-1)) / m with m = 2^bits - 1.  Out-of-range results are kept."""
    if matrix not in _MATRICES:
        raise ValueError("matrix must be 'bt709' or 'bt601'")
    if levels not in ("limited", "full"):
        raise ValueError("levels must be 'limited' or 'full'")
    m = float((1 << bits) - 1)
    merged = cscale(frame, cscale_filter, chroma_loc, bits)

    def codes(t):
        return t.to(torch.float32) * (m if t.dtype in (torch.float16, torch.float32) else 1.0)

    y, cb, cr = (codes(merged[k]) for k in ("Y", "U", "V"))
    if levels == "limited":
        sc = float(1 << (bits - 8))
        y, cb, cr = (y - 16 * sc) / (219 * sc), (cb - 128 * sc) / (224 * sc), (cr - 128 * sc) / (224 * sc)
    else:
        half = float(1 << (bits - 1))
        y, cb, cr = y / m, (cb - half) / m, (cr - half) / m
    kr, kb = _MATRICES[matrix]
    kg = 1.0 - kr - kb
    r = y + 2.0 * (1.0 - kr) * cr
    b = y + 2.0 * (1.0 - kb) * cb
    g = (y - kr * r - kb * b) / kg
    return torch.stack([r, g, b], dim=-3).contiguous()


def apply_hook(name, frame, display, precision="fp32", bits=10, cscale_filter="bilinear", chroma_loc="left",
               matrix="bt709", levels="limited"):
    """Run variant ``name`` on ``frame`` for display size ``display`` = (width, height).
    Returns ``(frame', offsets, ran)``: the new frame dict, a dict plane -> (ox, oy) of
    accumulated OFFSET in the new plane's pixels, and whether the hook ran.  NATIVE
    hooks on subsampled chroma first resample it to the luma size (``cscale_filter``,
    ``chroma_loc``); MAIN hooks on a YUV source first convert it to R'G'B' (``yuv_to_rgb``
    with ``matrix`` and ``levels``), and the returned frame holds the "RGB" plane."""
    out = dict(frame)
    offsets = {k: (0.0, 0.0) for k in frame}
    if "RGB" in frame:
        ref = frame["RGB"]
    elif "Y" in frame:
        ref = frame["Y"]
    else:
        raise ValueError("frame needs a Y or an RGB plane")
    h, w = ref.shape[-2:]
    info = query(name, w, h, display[0], display[1])
    hook = info["hook"]
    tgt = (display[0], display[1])
    if hook == "LUMA":
        if "Y" not in frame:
            return out, offsets, False          # no LUMA plane in an RGB source
        if not info["applies"]:
            return out, offsets, False
        out["Y"] = upscale(name, frame["Y"], target=tgt, precision=precision, bits=bits)
        offsets["Y"] = (info["offset_x"], info["offset_y"])
        return out, offsets, True
    if hook == "NATIVE":
        if "Y" not in frame:                    # LUMA.w 0 > is false for RGB sources
            return out, offsets, False
        if not info["applies"]:
            return out, offsets, False
        merged = cscale(frame, cscale_filter, chroma_loc, bits)
        if len({merged[k].dtype for k in ("Y", "U", "V")}) != 1:
            raise ValueError("%s: the Y, U and V planes must share one element type" % name)
        x = torch.stack([merged["Y"], merged["U"], merged["V"]], dim=-3)
        y = upscale(name, x, target=tgt, precision=precision, bits=bits)
        for i, k in enumerate(("Y", "U", "V")):
            out[k] = y[..., i, :, :]
            offsets[k] = (info["offset_x"], info["offset_y"])
        return out, offsets, True
    if hook == "MAIN":
        if not info["applies"]:
            return out, offsets, False
        if "RGB" in frame:
            rgb = frame["RGB"]
        else:
            rgb = yuv_to_rgb(frame, bits, matrix, levels, cscale_filter, chroma_loc)
            out = {"RGB": rgb}
        out["RGB"] = upscale(name, rgb, target=tgt, precision=precision, bits=bits)
        offsets = {"RGB": (info["offset_x"], info["offset_y"])}
        return out, offsets, True
    raise ValueError("unknown hook point %r" % hook)


def present(frame, offsets, display, filter="catmull_rom"):
    """Resample every plane of ``frame`` to ``display`` = (width, height), consuming each
    plane's OFFSET (chroma planes: centre-sited, offset 0 unless given)."""
    res = {}
    for k, p in frame.items():
        planes = 3 if k == "RGB" else 1
        res[k] = resample(p, display, offset=offsets.get(k, (0.0, 0.0)), filter=filter, planes=planes)
    return res
