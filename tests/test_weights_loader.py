"""Weight-file checks of the C library (host only, no GPU): every shipped table passes;
a wrongly shaped supplied ravu-zoom-ar-r3 table, a truncated file, a damaged shipped
file and a header that claims a huge shape all fail with MPX_ERR_IO before anything is
uploaded (mpx_plan_create runs the same checks first)."""
import ctypes
import os
import shutil
import struct

import numpy as np
import pytest

from conftest import PKG

WDIR = os.path.join(PKG, "weights")
IO, UNAVAILABLE = -7, -3


def _verify(d, name):
    from mpx._abi import LIB
    rc = LIB.mpx_weights_verify(d.encode() if d else None, name.encode())
    return rc, LIB.mpx_last_error().decode()


def _write(path, arr, dims=None):
    arr = np.ascontiguousarray(arr, dtype="<f4")
    dims = arr.shape if dims is None else dims
    with open(path, "wb") as f:
        f.write(b"MPXW" + struct.pack("<III", 1, 0, len(dims)) + struct.pack("<%dI" % len(dims), *dims))
        f.write(arr.tobytes())


def _copy_weights(tmp_path):
    d = tmp_path / "w"
    shutil.copytree(WDIR, d, ignore=shutil.ignore_patterns("__pycache__", "*.py"))
    return str(d)


def test_every_shipped_variant_verifies():
    import mpx
    for n in mpx.variants():
        rc, msg = _verify(None, n)
        if n.startswith("ravu-zoom-ar-r3"):
            assert rc == UNAVAILABLE, (n, rc, msg)
        else:
            assert rc == 0, (n, rc, msg)


def test_supplied_ar_table_shape(tmp_path):
    d = _copy_weights(tmp_path)
    rng = np.random.default_rng(0)
    good = rng.random((2592, 45, 4), dtype=np.float32)
    _write(os.path.join(d, "ravu_zoom_lut3_ar.bin"), good)
    assert _verify(d, "ravu-zoom-ar-r3-rgb")[0] == 0      # no shipped hash: shape + size only
    _write(os.path.join(d, "ravu_zoom_lut3_ar.bin"), good[:, :18])   # an r2-shaped table
    rc, msg = _verify(d, "ravu-zoom-ar-r3-rgb")
    assert rc == IO and "shape 2592x18x4, expected 2592x45x4" in msg, msg
    _write(os.path.join(d, "ravu_zoom_lut3_ar.bin"), good.reshape(-1, 4))   # wrong rank
    assert _verify(d, "ravu-zoom-ar-r3")[0] == IO


def test_truncated_and_padded_files(tmp_path):
    d = _copy_weights(tmp_path)
    p = os.path.join(d, "ravu_lite_lut3.bin")
    data = open(p, "rb").read()
    open(p, "wb").write(data[:-4])
    rc, msg = _verify(d, "ravu-lite-ar-r3")
    assert rc == IO and "size" in msg, msg
    open(p, "wb").write(data + b"\0\0\0\0")
    assert _verify(d, "ravu-lite-r3")[0] == IO
    open(p, "wb").write(data[:10])
    assert _verify(d, "ravu-lite-r3")[0] == IO


def test_damaged_shipped_file(tmp_path):
    d = _copy_weights(tmp_path)
    for name, variant in (("nnedi3_nns64_win8x6", "nnedi3-nns64-win8x6"), ("ravu_zoom_lut2_ar", "ravu-zoom-ar-r2-yuv"),
                          ("ravu_lut4", "ravu-r4-rgb")):
        p = os.path.join(d, name + ".bin")
        b = bytearray(open(p, "rb").read())
        b[len(b) // 2] ^= 0x01          # one bit of one weight
        open(p, "wb").write(bytes(b))
        rc, msg = _verify(d, variant)
        assert rc == IO and "SHA-256" in msg, (variant, msg)


def test_huge_header_is_rejected_before_reading(tmp_path):
    d = _copy_weights(tmp_path)
    _write(os.path.join(d, "ravu_3x_lut2.bin"), np.zeros(4, np.float32), dims=(0xFFFFFFFF, 0xFFFFFFFF, 4))
    rc, msg = _verify(d, "ravu-3x-r2")
    assert rc == IO and "shape" in msg, msg


def test_missing_file(tmp_path):
    d = _copy_weights(tmp_path)
    os.remove(os.path.join(d, "ravu_3x_lut4.bin"))
    rc, msg = _verify(d, "ravu-3x-r4-yuv")
    assert rc == IO and "cannot open" in msg, msg


def test_manifest_header_is_current():
    """csrc/weights_manifest.h (compiled into the library) is generated from manifest.json."""
    from mpx import gen_tables
    with open(gen_tables.OUT_MANIFEST) as f:
        assert f.read() == gen_tables.render_manifest()
