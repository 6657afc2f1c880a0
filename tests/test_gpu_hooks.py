"""Hook-point routing, WHEN gating and presentation (mpx.hooks) on the GPU."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _yuv(w, h, sub=2, seed=0):
    from mpx.frames import make_plane
    y = torch.from_numpy(make_plane(w, h, seed, bits=10).astype(np.float32)).cuda()
    cw, ch = -(-w // sub), -(-h // sub)
    u = torch.from_numpy(make_plane(cw, ch, seed + 1, bits=10).astype(np.float32)).cuda()
    v = torch.from_numpy(make_plane(cw, ch, seed + 2, bits=10).astype(np.float32)).cuda()
    return {"Y": y, "U": u, "V": v}


def test_luma_hook_on_420():
    import mpx
    from mpx.hooks import apply_hook, present
    fr = _yuv(64, 36)
    out, offs, ran = apply_hook("ravu-r3", fr, (128, 72))
    assert ran and out["U"] is fr["U"] and out["V"] is fr["V"]
    assert torch.equal(out["Y"], mpx.upscale("ravu-r3", fr["Y"]))
    assert offs["Y"] == (-0.5, -0.5) and offs["U"] == (0.0, 0.0)
    shown = present(out, offs, (128, 72))
    assert all(tuple(p.shape) == (72, 128) for p in shown.values())


def test_when_false_passes_through():
    from mpx.hooks import apply_hook
    fr = _yuv(64, 36)
    out, offs, ran = apply_hook("nnedi3-nns32-win8x4", fr, (80, 45))   # 1.25x: WHEN false
    assert not ran and out["Y"] is fr["Y"]


def test_native_and_main_routing():
    import mpx
    from mpx.hooks import apply_hook
    fr444 = _yuv(48, 32, sub=1, seed=5)
    out, offs, ran = apply_hook("ravu-r2-yuv", fr444, (96, 64))
    want = mpx.upscale("ravu-r2-yuv", torch.stack([fr444["Y"], fr444["U"], fr444["V"]]))
    assert ran and torch.equal(out["V"], want[2])
    rgb = {"RGB": torch.stack([fr444["Y"], fr444["U"], fr444["V"]])}
    out, offs, ran = apply_hook("ravu-zoom-ar-r2-rgb", rgb, (72, 48))
    assert ran and tuple(out["RGB"].shape) == (3, 48, 72)
    out, offs, ran = apply_hook("ravu-lite-r2", rgb, (96, 64))             # no LUMA in RGB
    assert not ran


@pytest.mark.parametrize("loc", ["left", "center"])
def test_native_on_420_after_cscale_vs_oracle(loc):
    """4:2:0 into a NATIVE hook: chroma resampled to the luma size (bilinear, the given
    siting), then the three planes through ravu-r2-yuv -- against the FP64 oracle's
    resampler and RAVU on the same planes."""
    from mpx.hooks import apply_hook, chroma_offset
    from oracle.ravu import ravu
    from oracle.resample import resample
    w, h = 50, 34
    fr = _yuv(w, h, seed=9)
    out, offs, ran = apply_hook("ravu-r2-yuv", fr, (2 * w, 2 * h), chroma_loc=loc)
    assert ran and offs["U"] == (-0.5, -0.5)
    planes = [fr["Y"].cpu().numpy().astype(np.float64)]
    for k in ("U", "V"):
        c = fr[k].cpu().numpy().astype(np.float64)
        planes.append(resample(c, w, h, offset=chroma_offset((w, h), c.shape[::-1], loc), name="bilinear"))
    ref, _ = ravu(np.stack(planes), 2, channels="yuv")
    got = np.stack([out[k].cpu().numpy() for k in ("Y", "U", "V")])
    err = float(np.abs(got - ref).max())
    print("ravu-r2-yuv on 4:2:0 (%s siting) vs FP64 oracle: max-abs %.3g" % (loc, err))
    assert err <= 5e-5


def test_main_hook_on_yuv_420_converts_to_rgb():
    """A YUV 4:2:0 source into a MAIN (-rgb) hook: chroma to the luma size, the BT.709
    limited-range matrix, then the RGB planes through the variant -- the same as running
    the variant on yuv_to_rgb's planes; the result replaces the frame's planes."""
    import mpx
    from mpx.hooks import apply_hook, yuv_to_rgb
    fr = _yuv(56, 40, seed=21)
    out, offs, ran = apply_hook("ravu-r2-rgb", fr, (112, 80), bits=10)
    assert ran and set(out) == {"RGB"} and offs == {"RGB": (-0.5, -0.5)}
    rgb = yuv_to_rgb(fr, bits=10)
    assert tuple(rgb.shape) == (3, 40, 56) and rgb.is_cuda
    assert torch.equal(out["RGB"], mpx.upscale("ravu-r2-rgb", rgb))


def test_native_on_integer_420():
    """Integer 4:2:0 planes into a NATIVE hook: resampled chroma is float, so the luma
    plane is brought to code / (2^bits - 1) as well; the result equals the same frame given
    as normalised floats."""
    from mpx.hooks import apply_hook
    f = _yuv(48, 32, seed=13)
    fi = {k: torch.round(v * 1023).to(torch.uint16) for k, v in f.items()}   # 10-bit codes
    ff = {k: v.to(torch.float32) / 1023 for k, v in fi.items()}
    a, _, ran_a = apply_hook("ravu-r2-yuv", fi, (96, 64), bits=10)
    b, _, ran_b = apply_hook("ravu-r2-yuv", ff, (96, 64), bits=10)
    assert ran_a and ran_b
    for k in ("Y", "U", "V"):
        assert torch.allclose(a[k], b[k], atol=1e-6), k
