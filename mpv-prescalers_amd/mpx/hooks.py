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
  BT.709 luma.  A YUV source would be converted to RGB before MAIN; that conversion is
  the caller's.

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


def cscale(frame, cscale="bilinear", chroma_loc="left"):
    """The chroma planes of a YUV frame resampled to the luma size (mpv's --cscale step
    before NATIVE hooks); planes already at the luma size are kept."""
    y = frame["Y"]
    lh, lw = y.shape[-2:]
    out = dict(frame)
    for k in ("U", "V"):
        c = frame[k]
        ch, cw = c.shape[-2:]
        if (ch, cw) != (lh, lw):
            out[k] = resample(c, (lw, lh), offset=chroma_offset((lw, lh), (cw, ch), chroma_loc), filter=cscale,
                              out_dtype=y.dtype if y.dtype in (torch.float16, torch.float32) else torch.float32)
    return out


def apply_hook(name, frame, display, precision="fp32", bits=10, cscale_filter="bilinear", chroma_loc="left"):
    """Run variant ``name`` on ``frame`` for display size ``display`` = (width, height).
    Returns ``(frame', offsets, ran)``: the new frame dict, a dict plane -> (ox, oy) of
    accumulated OFFSET in the new plane's pixels, and whether the hook ran.  NATIVE
    hooks on subsampled chroma first resample it to the luma size (``cscale_filter``,
    ``chroma_loc``)."""
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
        merged = cscale(frame, cscale_filter, chroma_loc)
        if len({merged[k].dtype for k in ("Y", "U", "V")}) != 1:
            raise ValueError("%s: the Y, U and V planes must share one element type" % name)
        x = torch.stack([merged["Y"], merged["U"], merged["V"]], dim=-3)
        y = upscale(name, x, target=tgt, precision=precision, bits=bits)
        for i, k in enumerate(("Y", "U", "V")):
            out[k] = y[..., i, :, :]
            offsets[k] = (info["offset_x"], info["offset_y"])
        return out, offsets, True
    if hook == "MAIN":
        if "RGB" not in frame:
            raise ValueError("%s hooks MAIN (RGB); convert the YUV source to RGB first" % name)
        if not info["applies"]:
            return out, offsets, False
        out["RGB"] = upscale(name, frame["RGB"], target=tgt, precision=precision, bits=bits)
        offsets["RGB"] = (info["offset_x"], info["offset_y"])
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
