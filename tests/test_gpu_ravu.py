"""RAVU (full), RAVU-3x and RAVU-zoom(-ar) on the GPU vs the FP64 oracle, in the luma,
-yuv and -rgb forms (ravu-r3*.hook, compute/ravu-3x-r3*.hook, ravu-zoom*-r2*.hook).

Bar: LUT-row (bucket) agreement >= 99.99 % of sites; on outputs whose sites (and, for the
int10/int01 predictions of full RAVU, the int11 sites they read) agree, max-abs <= 2e-6
(5e-6 for zoom's bilinear LUT fetch)."""
import numpy as np
import pytest
import torch
from scipy.ndimage import binary_dilation

from parity import compare

pytestmark = pytest.mark.gpu


def _frame(w, h, seed, planes, bits=8):
    from mpx.frames import make_frame
    return make_frame(w, h, seed, bits, channels=planes).astype(np.float32)


def _suffix(ch):
    return "" if ch == "luma" else "-" + ch


@pytest.mark.parametrize("r", [2, 3, 4])
@pytest.mark.parametrize("ch", ["luma", "yuv", "rgb"])
def test_ravu_3x(r, ch):
    import mpx
    from oracle.ravu import ravu_3x
    planes = 1 if ch == "luma" else 3
    img = _frame(320, 180, seed=30 + r, planes=planes)
    name = "ravu-3x-r%d%s" % (r, _suffix(ch))
    out, rows = mpx.upscale(name, torch.from_numpy(img).cuda(), buckets=True)
    torch.cuda.synchronize()
    ref, aux = ravu_3x(img.astype(np.float64), r, channels=ch)
    agree_site = rows.cpu().numpy().reshape(img.shape[-2:]) == aux["row"]

    def to_out(a):
        m = np.repeat(np.repeat(a, 3, axis=0), 3, axis=1)
        return m if planes == 1 else np.broadcast_to(m, (3,) + m.shape)
    f = compare(name + " 320x180", out.cpu().numpy(), ref, rows.cpu().numpy(), aux["row"],
                site_shape=img.shape[-2:], site_to_out=to_out)
    assert f["bucket_agree"] >= 0.9999
    assert f["max_abs_agreeing"] <= 2e-6
    # the centre of every 3x3 block is the input pixel, bit-exact
    o = out.cpu().numpy()
    assert np.array_equal(o[..., 1::3, 1::3], img)
    del agree_site


@pytest.mark.parametrize("r", [2, 3, 4])
@pytest.mark.parametrize("ch", ["luma", "yuv", "rgb"])
def test_ravu_full(r, ch):
    import mpx
    from oracle.ravu import ravu
    planes = 1 if ch == "luma" else 3
    img = _frame(256, 144, seed=50 + r, planes=planes)
    name = "ravu-r%d%s" % (r, _suffix(ch))
    out, rows = mpx.upscale(name, torch.from_numpy(img).cuda(), buckets=True)
    torch.cuda.synchronize()
    ref, aux = ravu(img.astype(np.float64), r, channels=ch)
    H, W = img.shape[-2:]
    g = rows.cpu().numpy().reshape(3, H, W)
    rr = np.stack([aux["row"]["int11"], aux["row"]["int10"], aux["row"]["int01"]])
    ok11 = g[0] == rr[0]
    bad11 = binary_dilation(~ok11, structure=np.ones((2 * r + 1, 2 * r + 1), bool))
    ok10 = (g[1] == rr[1]) & ~bad11
    ok01 = (g[2] == rr[2]) & ~bad11

    def to_out(_):
        m = np.zeros((2 * H, 2 * W), bool)
        m[0::2, 0::2] = True
        m[0::2, 1::2] = ok10
        m[1::2, 0::2] = ok01
        m[1::2, 1::2] = ok11
        return m if planes == 1 else np.broadcast_to(m, (3,) + m.shape)
    f = compare(name + " 256x144", out.cpu().numpy(), ref, g, rr, site_shape=g.shape, site_to_out=to_out)
    assert f["bucket_agree"] >= 0.9999
    assert f["max_abs_agreeing"] <= 2e-6
    o = out.cpu().numpy()
    assert np.array_equal(o[..., 0::2, 0::2], img)   # combine: (2x, 2y) is the input


ZOOM_CASES = [(2, False, "luma", (160, 90), (320, 180)), (2, True, "luma", (160, 90), (240, 135)),
              (3, False, "luma", (160, 90), (400, 225)), (2, True, "rgb", (128, 72), (256, 144)),
              (2, True, "yuv", (120, 80), (180, 120)), (3, False, "rgb", (96, 64), (144, 96)),
              (2, False, "yuv", (97, 61), (301, 173))]


@pytest.mark.parametrize("r,ar,ch,ins,outs", ZOOM_CASES)
def test_ravu_zoom(r, ar, ch, ins, outs):
    import mpx
    from oracle.ravu import ravu_zoom
    planes = 1 if ch == "luma" else 3
    img = _frame(ins[0], ins[1], seed=70 + r, planes=planes)
    name = ("ravu-zoom-ar-r%d%s" if ar else "ravu-zoom-r%d%s") % (r, _suffix(ch))
    out, rows = mpx.upscale(name, torch.from_numpy(img).cuda(), target=outs, buckets=True)
    torch.cuda.synchronize()
    ref, aux = ravu_zoom(img.astype(np.float64), r, outs[0], outs[1], ar=ar, channels=ch)
    g = rows.cpu().numpy().reshape(outs[1], outs[0])

    def to_out(a):
        return a if planes == 1 else np.broadcast_to(a, (3,) + a.shape)
    f = compare("%s %dx%d->%dx%d" % (name, ins[0], ins[1], outs[0], outs[1]), out.cpu().numpy(), ref, g,
                aux["row"], site_shape=g.shape, site_to_out=to_out)
    assert f["bucket_agree"] >= 0.9999
    assert f["max_abs_agreeing"] <= 5e-6


def test_ravu_zoom_phase_zero_passthrough():
    """At an exact 3x zoom every output with x = 1 (mod 3) and y = 1 (mod 3) has phase 0,
    where both LUTs are the unit vector on the base tap: the input pixel (exactly without
    AR; with AR the clamp bounds are (0.1+l)^33/(0.1+l)^32-0.1 = l up to FP32 rounding)."""
    import mpx
    img = _frame(64, 48, seed=3, planes=1)
    for name in ("ravu-zoom-r2", "ravu-zoom-r3", "ravu-zoom-ar-r2"):
        out = mpx.upscale(name, torch.from_numpy(img).cuda(), target=(192, 144)).cpu().numpy()
        if "-ar-" in name:
            assert np.abs(out[1::3, 1::3] - img).max() <= 1e-6, name
        else:
            assert np.array_equal(out[1::3, 1::3], img), name


def test_ravu_zoom_ar_r3_unavailable():
    import mpx
    from mpx._abi import MpxError
    x = torch.zeros((16, 16), device="cuda")
    with pytest.raises(MpxError) as e:
        mpx.upscale("ravu-zoom-ar-r3", x, target=(32, 32))
    assert e.value.code == -3


def test_ravu_zoom_ar_r3_with_supplied_table(tmp_path):
    """ravu-zoom-ar-r3's AR table is not shipped; with a user-supplied table in the weights
    directory the variant runs.  A synthetic non-negative table (|main r3 table|) checks
    the r3 AR path against the oracle using the same table."""
    import os
    import mpx
    from oracle.common import WEIGHTS_DIR, load_weights
    from oracle.ravu import ravu_zoom
    import importlib.util
    spec = importlib.util.spec_from_file_location("extract", os.path.join(WEIGHTS_DIR, "extract.py"))
    ext = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(ext)
    for f in os.listdir(WEIGHTS_DIR):
        if f.endswith(".bin"):
            os.symlink(os.path.join(WEIGHTS_DIR, f), str(tmp_path / f))
    table = np.abs(load_weights("ravu_zoom_lut3")).astype(np.float32)
    ext.write_bin(str(tmp_path / "ravu_zoom_lut3_ar.bin"), table)
    ctx = mpx.Context(0, weights_dir=str(tmp_path))
    img = _frame(80, 48, seed=9, planes=1)
    plan = ctx.plan("ravu-zoom-ar-r3", 80, 48, target=(120, 72))
    out = plan.run(torch.from_numpy(img).cuda()).cpu().numpy()
    ref, _ = ravu_zoom(img.astype(np.float64), 3, 120, 72, ar=True,
                       ar_lut=table.astype(np.float16).astype(np.float64))
    assert np.abs(out - ref).max() <= 5e-5


@pytest.mark.parametrize("name,target", [("ravu-r4-yuv", None), ("ravu-r3-rgb", None), ("ravu-3x-r3-yuv", None),
                                         ("ravu-zoom-ar-r2-rgb", (250, 141)), ("ravu-lite-ar-r4", None)])
@pytest.mark.parametrize("bits", [10, 16])
def test_ravu_integer_input_equals_float(name, target, bits):
    """Integer planes in 16-bit containers (value / (2^bits - 1), as mpv's sampler
    normalises them): each output equals the run on the FP32 quotients, bit for bit."""
    import mpx
    planes = 1 if name.endswith(("-r4", "-r3")) else 3
    full = 2 ** bits - 1
    u16 = np.stack([np.round(_frame(100, 57, seed=60 + s, planes=planes, bits=bits) * full) for s in range(2)])
    x = torch.from_numpy(u16.astype(np.int32)).to(torch.uint16).cuda()
    out = mpx.upscale(name, x, target=target, out_dtype=torch.float32, bits=bits)
    for i in range(2):
        xf = torch.from_numpy(u16[i].astype(np.float32) / np.float32(full)).cuda()
        single = mpx.upscale(name, xf, target=target)
        assert torch.equal(out[i], single), (name, bits, i)


@pytest.mark.parametrize("r,ar,ch,ins,outs", [(2, True, "rgb", (128, 72), (256, 144)),
                                              (2, False, "luma", (120, 80), (180, 120)),
                                              (3, False, "yuv", (96, 64), (192, 128))])
def test_ravu_zoom_weight_table_equals_per_pixel(r, ar, ch, ins, outs):
    """Where the output phase repeats with a short period (2x: 2, 1.5x: 3), a plan takes
    its bilinear LUT weights from a table interpolated once at plan creation.  A context
    whose HBM budget holds the LUTs but not that table interpolates per output pixel
    instead.  Both give the same output, bit for bit."""
    import mpx
    planes = 1 if ch == "luma" else 3
    img = _frame(ins[0], ins[1], seed=70 + r, planes=planes)
    name = "ravu-zoom%s-r%d%s" % ("-ar" if ar else "", r, _suffix(ch))
    x = torch.from_numpy(img).cuda()
    groups = ((2 * r) ** 2 // 2 + 3) // 4
    lut_bytes = 2592 * 9 * groups * 4 * 2 * (2 if ar else 1)   # FP16 LUT(s) resident in the context
    outs_ = []
    for budget in (16 << 30, lut_bytes + (32 << 10)):       # the smallest table is 74 KB
        ctx = mpx.Context(0, hbm_budget=budget)
        plan = ctx.plan(name, ins[0], ins[1], target=outs)
        outs_.append(plan.run(x).cpu())
        plan.close()
        ctx.close()
    assert torch.equal(outs_[0], outs_[1])
