"""Write golden vectors by executing the reference's `.hook` text (hookexec.py).

For every shipped variant (43 names with weights) a small seeded frame is upscaled by
each flavour the reference ships (root, compute/, gather/); the flavours must agree
(they are one math with different load patterns), and the root output (compute/ for
ravu-3x, which exists only there) is stored with its input in ``hooks_golden.npz``:
  <name>/in      float64 input planes, (H, W) or (3, H, W)
  <name>/target  [target_w, target_h] the hook was run for
  <name>/out     float64 output planes
  <name>/flavours  names of the flavours that were executed and agreed
  <name>/out16   float64 output of the root (compute/) flavour with every pass's output
                 texture stored in FP16, as mpv's default rgba16f FBOs do
Run in the container holding the reference:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [HERE, os.path.join(ROOT, "mpv-prescalers_amd")]

import hookexec  # noqa: E402
from mpx.frames import make_frame  # noqa: E402

REF = os.environ.get("MPX_REFERENCE", "/root/reference")
OUT = os.path.join(HERE, "hooks_golden.npz")
W, H = 20, 12           # input frame size of every golden case
FLAVOUR_TOL = 1e-12     # flavours differ only in summation order


def variants():
    names = []
    for nns in (16, 32, 64, 128, 256):
        for wh in (4, 6):
            names.append("nnedi3-nns%d-win8x%d" % (nns, wh))
    for r in (2, 3, 4):
        for ch in ("", "-yuv", "-rgb"):
            names.append("ravu-r%d%s" % (r, ch))
            names.append("ravu-3x-r%d%s" % (r, ch))
    for ar in ("", "-ar"):
        for r in (2, 3, 4):
            names.append("ravu-lite%s-r%d" % (ar, r))
    for r in (2, 3):
        for ch in ("", "-yuv", "-rgb"):
            names.append("ravu-zoom-r%d%s" % (r, ch))
    for ch in ("", "-yuv", "-rgb"):
        names.append("ravu-zoom-ar-r2%s" % ch)   # ravu-zoom-ar-r3*: not shipped
    return names


def target_of(name):
    if name.startswith("ravu-3x"):
        return (3 * W, 3 * H)
    if name.startswith("ravu-zoom"):
        return (W * 3 // 2, H * 3 // 2 + 1)       # 1.5x wide, non-integer ratio tall
    return (2 * W, 2 * H)


def main():
    data, summary = {}, {}
    for i, name in enumerate(variants()):
        planes = 3 if name.endswith(("-yuv", "-rgb")) else 1
        img = make_frame(W, H, seed=1000 + i, bits=8, channels=planes).astype(np.float32).astype(np.float64)
        tgt = target_of(name)
        outs = {}
        for fl in ("", "compute/", "gather/"):
            path = os.path.join(REF, fl + name + ".hook")
            if os.path.exists(path):
                outs[fl or "root"] = hookexec.run_hook(path, img, tgt)
        keys = list(outs)
        base = outs["root"] if "root" in outs else outs[keys[0]]
        for k in keys:
            d = float(np.abs(outs[k] - base).max())
            assert d <= FLAVOUR_TOL, (name, k, d)
        data[name + "/in"] = img
        data[name + "/target"] = np.array(tgt)
        data[name + "/out"] = base
        data[name + "/flavours"] = np.array(keys)
        path = os.path.join(REF, ("" if "root" in outs else keys[0]) + name + ".hook")
        data[name + "/out16"] = hookexec.run_hook(path, img, tgt, fbo16=True)
        summary[name] = keys
        print("%-26s %-28s out %s" % (name, ",".join(keys), base.shape), flush=True)
    np.savez_compressed(OUT, **data)
    with open(os.path.join(HERE, "hooks_golden.json"), "w") as f:
        json.dump({"frame": [W, H], "flavour_tol": FLAVOUR_TOL, "variants": summary}, f, indent=1, sort_keys=True)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
