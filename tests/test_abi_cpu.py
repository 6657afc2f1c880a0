"""C-ABI library: loads, exports every symbol of include/mpx.h, and its host-only
queries reproduce the reference's //!WHEN / //!WIDTH / //!HEIGHT / //!OFFSET rules.
No GPU compute here."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "mpx.h")


def test_library_exports_every_declared_symbol():
    from mpx import _abi
    decl = re.findall(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b(mpx_\w+)\s*\(", open(HEADER).read(), re.M)
    assert set(decl) == set(_abi.EXPORTS), set(decl) ^ set(_abi.EXPORTS)
    lib = ctypes.CDLL(_abi._LIB_PATH)
    for name in decl:
        assert hasattr(lib, name), name


def test_registry_names_match_reference_files(reference_dir):
    import mpx
    names = set(mpx.variants())
    files = {f[:-5] for f in os.listdir(reference_dir) if f.endswith(".hook")}
    files |= {f[:-5] for f in os.listdir(os.path.join(reference_dir, "compute")) if f.endswith(".hook")}
    missing = {l.strip().split("/")[-1][:-5] for l in open(os.path.join(reference_dir, ".MISSING_LARGE_BLOBS"))
               if l.strip()}
    assert names == files | missing
    assert len(names) == 46


def test_registry_availability():
    import mpx
    for n in mpx.variants():
        info = mpx.query(n, 100, 100, 200, 200)
        assert info["available"] == (0 if n.startswith("ravu-zoom-ar-r3") else 1), n


@pytest.mark.parametrize("name,planes,hook", [
    ("ravu-lite-r3", 1, "LUMA"), ("ravu-r4-yuv", 3, "NATIVE"), ("ravu-zoom-ar-r2-rgb", 3, "MAIN"),
    ("nnedi3-nns32-win8x4", 1, "LUMA"), ("ravu-3x-r3-rgb", 3, "MAIN")])
def test_channel_forms(name, planes, hook):
    import mpx
    info = mpx.query(name, 64, 48, 128, 96)
    assert info["planes"] == planes and info["hook"] == hook


def test_when_rules():
    import mpx
    # ravu-lite: both axes, ratio < 0.707106 (ravu-lite-r3.hook:20)
    q = mpx.query("ravu-lite-r3", 640, 360, 1280, 720)
    assert q["applies"] == 1 and (q["out_w"], q["out_h"]) == (1280, 720)
    assert (q["offset_x"], q["offset_y"]) == (0.0, 0.0)
    q = mpx.query("ravu-lite-r3", 640, 360, 1280, 480)   # 360/480 = 0.75 -> no
    assert q["applies"] == 0 and (q["out_w"], q["out_h"]) == (640, 360)
    # ravu: offset -0.5 -0.5 (ravu-r3.hook:424)
    q = mpx.query("ravu-r3", 640, 360, 1280, 720)
    assert (q["offset_x"], q["offset_y"]) == (-0.5, -0.5)
    # nnedi3: per-axis (nnedi3-nns16-win8x4.hook:18, :109)
    q = mpx.query("nnedi3-nns16-win8x4", 640, 360, 640, 720)
    assert (q["out_w"], q["out_h"], q["offset_x"], q["offset_y"]) == (640, 720, 0.0, -0.5)
    q = mpx.query("nnedi3-nns16-win8x4", 640, 360, 1280, 720)
    assert (q["out_w"], q["out_h"], q["offset_x"], q["offset_y"]) == (1280, 720, -0.5, -0.5)
    # ravu-3x: ratio < 0.471404 (compute/ravu-3x-r3.hook:21)
    assert mpx.query("ravu-3x-r3", 100, 100, 200, 200)["applies"] == 0
    q = mpx.query("ravu-3x-r3", 100, 100, 300, 300)
    assert q["applies"] == 1 and q["out_w"] == 300
    # ravu-zoom: in < out on both axes, output = target, OFFSET ALIGN
    q = mpx.query("ravu-zoom-r3", 1920, 1080, 2880, 1620)
    assert q["applies"] == 1 and (q["out_w"], q["out_h"], q["align"]) == (2880, 1620, 1)
    assert mpx.query("ravu-zoom-r3", 1920, 1080, 1920, 1620)["applies"] == 0


def test_errors():
    import mpx
    from mpx._abi import MpxError
    with pytest.raises(MpxError) as e:
        mpx.query("ravu-lite-r9", 10, 10, 20, 20)
    assert e.value.code == -2
    with pytest.raises(MpxError):
        mpx.query("ravu-lite-r3", 0, 10, 20, 20)


def test_header_compiles_as_c_and_links(tmp_path):
    """include/mpx.h is plain C: a C program compiled with gcc against it links to
    libmpx.so and answers the host-only queries (what a cgo / FFI binding relies on)."""
    import shutil
    import subprocess
    from mpx import _abi
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    src = tmp_path / "probe.c"
    src.write_text(r'''
#include <stdio.h>
#include <string.h>
#include "mpx.h"
int main(void) {
  mpx_variant_info info;
  int n = mpx_variant_count();
  if (mpx_query("nnedi3-nns256-win8x6", 1920, 1080, 3840, 2160, &info) != MPX_OK) return 2;
  if (mpx_query("no-such-variant", 8, 8, 16, 16, &info) != MPX_ERR_UNKNOWN_VARIANT) return 3;
  if (strstr(mpx_last_error(), "no-such-variant") == NULL) return 4;
  printf("%d %d %d %.2f %.2f %s\n", n, info.out_w, info.out_h, info.offset_x, info.offset_y, info.hook);
  return 0;
}
''')
    exe = tmp_path / "probe"
    libdir = os.path.dirname(_abi._LIB_PATH)
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-o", str(exe), str(src), "-I", os.path.join(ROOT, "include"),
                    "-L", libdir, "-lmpx", "-Wl,-rpath," + libdir], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert out == ["46", "3840", "2160", "-0.50", "-0.50", "LUMA"]
