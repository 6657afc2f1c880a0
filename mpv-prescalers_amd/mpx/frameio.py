"""Frame I/O around the prescalers: YUV4MPEG2 (y4m) and raw planar files.

Frames are planar NumPy arrays: 8-bit as uint8, 9..16-bit as uint16 little-endian
(y4m `C420p10` etc.), planes listed Y, U, V.  ``to_device`` turns a plane into the
integer tensor the C ABI takes directly (MPX_U8 / MPX_U16 with ``bits``), so the
normalisation value/(2^bits-1) happens inside the kernels' loads.
"""
import re

import numpy as np

_CS = {  # y4m colourspace tag -> (chroma subsampling x, y, bits)
    "420jpeg": (2, 2, 8), "420paldv": (2, 2, 8), "420mpeg2": (2, 2, 8), "420": (2, 2, 8), "422": (2, 1, 8),
    "444": (1, 1, 8), "mono": (0, 0, 8),
}


def _parse_cs(tag):
    """(sx, sy, bits) of a colour-space tag: 420/422/444 with optional pNN (ffmpeg's
    C420p10, C444p16 ...), mono and monoNN (ffmpeg's Cmono16)."""
    if tag in _CS:
        return _CS[tag]
    m = re.match(r"^(420|422|444)p(\d+)$", tag) or re.match(r"^(mono)p?(\d+)$", tag)
    if not m or not 8 <= int(m.group(2)) <= 16:
        raise ValueError("unsupported y4m colourspace C%s" % tag)
    sx, sy, _ = _CS[m.group(1)]
    return sx, sy, int(m.group(2))


class Y4MReader:
    """Iterate the frames of a y4m file: each frame is [Y, U, V] (or [Y] for mono)."""

    def __init__(self, path):
        self.f = open(path, "rb")
        header = self.f.readline().decode("ascii").split()
        if not header or header[0] != "YUV4MPEG2":
            raise ValueError("not a YUV4MPEG2 file")
        self.params = {t[0]: t[1:] for t in header[1:]}
        self.width = int(self.params["W"])
        self.height = int(self.params["H"])
        self.cs = self.params.get("C", "420jpeg")
        self.sx, self.sy, self.bits = _parse_cs(self.cs)
        self.dtype = np.uint8 if self.bits <= 8 else np.dtype("<u2")

    def plane_shapes(self):
        shapes = [(self.height, self.width)]
        if self.sx:
            cw, ch = -(-self.width // self.sx), -(-self.height // self.sy)
            shapes += [(ch, cw), (ch, cw)]
        return shapes

    def __iter__(self):
        return self

    def __next__(self):
        line = self.f.readline()
        if not line:
            raise StopIteration
        if not line.startswith(b"FRAME"):
            raise ValueError("bad frame header")
        planes = []
        for shp in self.plane_shapes():
            n = shp[0] * shp[1]
            buf = self.f.read(n * np.dtype(self.dtype).itemsize)
            if len(buf) != n * np.dtype(self.dtype).itemsize:
                raise ValueError("truncated frame")
            planes.append(np.frombuffer(buf, dtype=self.dtype).reshape(shp).copy())
        return planes

    def close(self):
        self.f.close()


class Y4MWriter:
    def __init__(self, path, width, height, cs="420jpeg", fps="30:1"):
        self.f = open(path, "wb")
        self.width, self.height, self.cs = width, height, cs
        self.sx, self.sy, self.bits = _parse_cs(cs)
        self.dtype = np.uint8 if self.bits <= 8 else np.dtype("<u2")
        self.f.write(("YUV4MPEG2 W%d H%d F%s Ip A1:1 C%s\n" % (width, height, fps, cs)).encode("ascii"))

    def write(self, planes):
        self.f.write(b"FRAME\n")
        for p in planes:
            self.f.write(np.ascontiguousarray(p, dtype=self.dtype).tobytes())

    def close(self):
        self.f.close()


def quantise(plane, bits):
    """Float plane in [0, 1] -> integer plane of `bits` (round to nearest, clamp)."""
    q = (1 << bits) - 1
    return np.clip(np.round(np.asarray(plane, dtype=np.float64) * q), 0, q).astype(np.uint8 if bits <= 8 else np.uint16)


def read_raw_planar(path, width, height, planes=1, bits=8, frame=0):
    """One frame of a raw planar file (all planes full size)."""
    dt = np.uint8 if bits <= 8 else np.dtype("<u2")
    n = width * height * planes
    with open(path, "rb") as f:
        f.seek(frame * n * np.dtype(dt).itemsize)
        a = np.frombuffer(f.read(n * np.dtype(dt).itemsize), dtype=dt)
    if a.size != n:
        raise ValueError("truncated raw frame")
    return a.reshape((planes, height, width)) if planes > 1 else a.reshape((height, width))


def to_device(plane, device="cuda"):
    """Integer plane -> device tensor in the dtype the C ABI reads (uint8 / uint16)."""
    import torch
    a = np.ascontiguousarray(plane)
    if a.dtype == np.uint8:
        return torch.from_numpy(a).to(device)
    return torch.from_numpy(a.astype(np.int32)).to(torch.uint16).to(device)
