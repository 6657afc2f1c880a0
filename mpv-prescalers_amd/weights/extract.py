"""Extract the shipped prescaler weights from the reference `.hook` text into binary files.

Runs only where the reference checkout is present (default ``/root/reference``); the
files it writes (``*.bin`` + ``manifest.json`` next to this script) are the data the
library loads at run time.  The weights are LGPL-3.0 data from bjin/mpv-prescalers
(see ``NOTICE`` in this directory).

Two kinds of weights exist in the reference:

* NNEDI3 neuron weights, written inline as ``intBitsToFloat`` integer literals, one
  line per neuron: ``sum1=W(i,w0,w1,w2,w3)+...;sum2=W(...)+...;WS(b1,b2);``
  (``nnedi3-nns256-win8x6.hook:31-33`` define the macros, ``:34-289`` hold the lines).
  They are stored as float32, bit-exact, in the canonical *double_y* window order
  ``k = (dx+3)*Hw + (dy-ymin)`` (dx in [-3,4], dy in [ymin, ymin+Hw-1]) as read from the
  ``samples[i][j] = HOOKED_texOff(vec2(dx, dy))`` lines (``nnedi3-nns16-win8x4.hook:
  39-72``).  The extractor checks that the *double_x* pass holds exactly the transposed
  weights, so one table serves both passes.
* RAVU lookup tables, ``//!TEXTURE`` blocks of little-endian float32 RGBA hex
  (``ravu-lite-r3.hook:132-136``).  They are stored as float32 arrays of shape
  ``(rows, width, 4)`` exactly as written; the GPU uses them FP16-rounded, as the
  ``FORMAT rgba16f`` upload in the reference does.

File format (little-endian): 16-byte header ``b'MPXW'``, u32 version (1), u32 dtype
(0 = float32), u32 ndim, then ndim u32 dims, then the raw float32 data.
"""
import argparse
import glob
import hashlib
import json
import os
import re
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

NNEDI3_NNS = (16, 32, 64, 128, 256)
NNEDI3_WIN = ((8, 4), (8, 6))


def write_bin(path, arr):
    arr = np.ascontiguousarray(arr, dtype=np.float32)
    with open(path, "wb") as f:
        f.write(b"MPXW")
        f.write(struct.pack("<III", 1, 0, arr.ndim))
        f.write(struct.pack("<%dI" % arr.ndim, *arr.shape))
        f.write(arr.tobytes())


def read_bin(path):
    with open(path, "rb") as f:
        data = f.read()
    assert data[:4] == b"MPXW", path
    ver, dt, nd = struct.unpack_from("<III", data, 4)
    assert ver == 1 and dt == 0
    dims = struct.unpack_from("<%dI" % nd, data, 16)
    off = 16 + 4 * nd
    return np.frombuffer(data, dtype="<f4", offset=off).reshape(dims).copy()


def _sha(b):
    return hashlib.sha256(b).hexdigest()


# ----------------------------------------------------------------------------- NNEDI3

_SAMPLE_RE = re.compile(r"samples\[(\d+)\]\[(\d+)\] = HOOKED_texOff\(vec2\(([-0-9.]+), ([-0-9.]+)\)\)\.x;")
_W_RE = re.compile(r"W\((\d+),(-?\d+),(-?\d+),(-?\d+),(-?\d+)\)")
_WS_RE = re.compile(r"WS\((-?\d+),(-?\d+)\)")


def _i2f(v):
    return struct.unpack("<f", struct.pack("<i", int(v)))[0]


def _parse_nnedi3_pass(lines, start):
    """Parse one double_* pass starting at its //!DESC line index.

    Returns (offsets[K] of (dx,dy) by flat sample index, w1[N][K], w2[N][K], b1[N], b2[N],
    first neuron line number (1-based))."""
    i = start
    w1, w2, b1, b2 = [], [], [], []
    first = None
    offsets = {}
    while i < len(lines):
        ln = lines[i]
        if i > start and ln.startswith("//!DESC"):
            break
        if ln.startswith("sum1="):
            if first is None:
                first = i + 1
            parts = ln.split(";")
            s1 = _W_RE.findall(parts[0])
            s2 = _W_RE.findall(parts[1])
            ws = _WS_RE.findall(ln)
            assert len(ws) == 1
            k = 4 * len(s1)
            v1 = np.zeros(k, np.float32)
            v2 = np.zeros(k, np.float32)
            for g in s1:
                idx = int(g[0])
                for j in range(4):
                    v1[4 * idx + j] = _i2f(g[1 + j])
            for g in s2:
                idx = int(g[0])
                for j in range(4):
                    v2[4 * idx + j] = _i2f(g[1 + j])
            w1.append(v1)
            w2.append(v2)
            b1.append(_i2f(ws[0][0]))
            b2.append(_i2f(ws[0][1]))
        m = _SAMPLE_RE.search(ln)
        if m:
            offsets[4 * int(m.group(1)) + int(m.group(2))] = (int(float(m.group(3))), int(float(m.group(4))))
        i += 1
    k = len(offsets)
    offs = [offsets[j] for j in range(k)]
    return (offs, np.array(w1, np.float32), np.array(w2, np.float32),
            np.array(b1, np.float32), np.array(b2, np.float32), first)


def extract_nnedi3(ref, outdir, manifest):
    for nns in NNEDI3_NNS:
        for (ww, wh) in NNEDI3_WIN:
            rel = "nnedi3-nns%d-win%dx%d.hook" % (nns, ww, wh)
            path = os.path.join(ref, rel)
            lines = open(path).read().split("\n")
            descs = [i for i, l in enumerate(lines) if l.startswith("//!DESC")]
            dy_i = [i for i in descs if "double_y" in lines[i]][0]
            dx_i = [i for i in descs if "double_x" in lines[i]][0]
            oy, w1y, w2y, b1y, b2y, first_y = _parse_nnedi3_pass(lines, dy_i)
            ox, w1x, w2x, b1x, b2x, first_x = _parse_nnedi3_pass(lines, dx_i)
            K = ww * wh
            assert w1y.shape == (nns, K) and w1x.shape == (nns, K), (rel, w1y.shape)
            ymin = -(wh // 2 - 1)
            # canonical order for double_y: k = (dx+3)*wh + (dy-ymin)
            can_y = [(dx + 3) * wh + (dy - ymin) for (dx, dy) in oy]
            assert sorted(can_y) == list(range(K)), rel
            # double_x samples the transposed window: canonical index swaps the roles
            can_x = [(dy + 3) * wh + (dx - ymin) for (dx, dy) in ox]
            assert sorted(can_x) == list(range(K)), rel
            W1 = np.zeros((nns, K), np.float32)
            W2 = np.zeros((nns, K), np.float32)
            W1[:, can_y] = w1y
            W2[:, can_y] = w2y
            W1x = np.zeros((nns, K), np.float32)
            W2x = np.zeros((nns, K), np.float32)
            W1x[:, can_x] = w1x
            W2x[:, can_x] = w2x
            # the x pass must be the exact transpose of the y pass (bit-exact)
            assert np.array_equal(W1.view(np.uint32), W1x.view(np.uint32)), rel
            assert np.array_equal(W2.view(np.uint32), W2x.view(np.uint32)), rel
            assert np.array_equal(b1y.view(np.uint32), b1x.view(np.uint32)), rel
            assert np.array_equal(b2y.view(np.uint32), b2x.view(np.uint32)), rel
            table = np.concatenate([W1, W2, b1y[:, None], b2y[:, None]], axis=1)
            name = "nnedi3_nns%d_win%dx%d" % (nns, ww, wh)
            write_bin(os.path.join(outdir, name + ".bin"), table)
            src_lines = lines[first_y - 1:first_y - 1 + nns]
            manifest[name] = {
                "kind": "nnedi3",
                "shape": list(table.shape),
                "layout": "row n = [w1[K], w2[K], b1, b2], k = (dx+3)*%d + (dy%+d)" % (wh, -ymin),
                "source": "%s:%d-%d" % (rel, first_y, first_y + nns - 1),
                "source_sha256": _sha("\n".join(src_lines).encode()),
                "data_sha256": _sha(table.astype("<f4").tobytes()),
                "x_pass_is_transpose": True,
            }


# ----------------------------------------------------------------------------- RAVU LUTs

LUTS = {
    # texture name: (file it is read from, shape (rows, width))
    "ravu_lite_lut2": "ravu-lite-r2.hook",
    "ravu_lite_lut3": "ravu-lite-r3.hook",
    "ravu_lite_lut4": "ravu-lite-r4.hook",
    "ravu_lut2": "ravu-r2.hook",
    "ravu_lut3": "ravu-r3.hook",
    "ravu_lut4": "ravu-r4.hook",
    "ravu_3x_lut2": "compute/ravu-3x-r2.hook",
    "ravu_3x_lut3": "compute/ravu-3x-r3.hook",
    "ravu_3x_lut4": "compute/ravu-3x-r4.hook",
    "ravu_zoom_lut2": "ravu-zoom-r2.hook",
    "ravu_zoom_lut3": "ravu-zoom-r3.hook",
    "ravu_zoom_lut2_ar": "ravu-zoom-ar-r2.hook",
}


def parse_textures(path):
    """Yield (name, width, height, fmt, filter, hex, line_no) for each //!TEXTURE block."""
    lines = open(path).read().split("\n")
    i = 0
    while i < len(lines):
        if lines[i].startswith("//!TEXTURE"):
            name = lines[i].split()[1]
            w, h = (int(v) for v in lines[i + 1].split()[1:3])
            fmt = lines[i + 2].split()[1]
            filt = lines[i + 3].split()[1]
            yield name, w, h, fmt, filt, lines[i + 4].strip(), i + 1
            i += 5
        else:
            i += 1


def extract_luts(ref, outdir, manifest):
    for name, rel in LUTS.items():
        found = None
        for t in parse_textures(os.path.join(ref, rel)):
            if t[0] == name:
                found = t
        assert found is not None, (name, rel)
        _, w, h, fmt, filt, hexs, line = found
        assert fmt == "rgba16f"
        raw = bytes.fromhex(hexs)
        assert len(raw) == w * h * 16, (name, len(raw))
        arr = np.frombuffer(raw, dtype="<f4").reshape(h, w, 4)
        write_bin(os.path.join(outdir, name + ".bin"), arr)
        manifest[name] = {
            "kind": "lut",
            "shape": [h, w, 4],
            "filter": filt,
            "format": fmt,
            "source": "%s:%d-%d" % (rel, line, line + 4),
            "source_sha256": _sha(hexs.encode()),
            "data_sha256": _sha(arr.astype("<f4").tobytes()),
        }


def check_lut_identity(ref):
    """Every file binding the same texture name carries the same hex (all flavours)."""
    seen = {}
    for path in sorted(glob.glob(os.path.join(ref, "*.hook")) + glob.glob(os.path.join(ref, "*", "*.hook"))):
        for name, w, h, fmt, filt, hexs, line in parse_textures(path):
            d = _sha(hexs.encode())
            if name in seen:
                assert seen[name] == d, (name, path)
            seen[name] = d
    return seen


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--out", default=HERE)
    args = ap.parse_args(argv)
    if not os.path.isdir(args.ref):
        print("reference checkout not found at %s" % args.ref, file=sys.stderr)
        return 1
    manifest = {}
    check_lut_identity(args.ref)
    extract_luts(args.ref, args.out, manifest)
    extract_nnedi3(args.ref, args.out, manifest)
    with open(os.path.join(args.out, "manifest.json"), "w") as f:
        json.dump({"format": "MPXW v1 float32", "license": "LGPL-3.0-or-later (bjin/mpv-prescalers)",
                   "weights": manifest}, f, indent=1, sort_keys=True)
    print("wrote %d weight files to %s" % (len(manifest), args.out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
