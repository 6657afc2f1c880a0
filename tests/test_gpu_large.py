"""Frames far larger than the other tests use: sample offsets beyond 2^24 elements, planes
2^25 elements apart and the 24-bit row multiply at rows past 4,000.  Each variant's output
over a crop's interior equals, bit for bit, the output of the same variant run on the
cropped frame (the per-pixel arithmetic does not depend on where a pixel sits; only the
clamped border of the crop differs, so a margin of 16 input pixels is left out)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

M = 16  # input pixels left out at the crop's edges


def _frame(w, h, planes, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    shape = (planes, h, w) if planes == 3 else (h, w)
    # smooth structure plus noise, quantised to 10 bits, as FP16
    y = torch.linspace(0, 1, h, device="cuda").view(-1, 1)
    x = torch.linspace(0, 1, w, device="cuda").view(1, -1)
    base = 0.5 + 0.3 * torch.sin(40 * x + 17 * y) * torch.cos(23 * y - 9 * x)
    noise = torch.rand(shape, generator=g, device="cuda") * 0.2
    return (torch.round((base + noise).clamp(0, 1) * 1023) / 1023).to(torch.float16)


@pytest.mark.parametrize("name,w,h,prec,crop", [
    ("ravu-lite-r3", 16384, 2048, "fp32", (15000, 1500, 384, 200)),
    ("ravu-r3-yuv", 8192, 2048, "fp32", (7000, 1000, 256, 128)),
    ("ravu-3x-r2-rgb", 8192, 1536, "fp32", (6000, 1400, 256, 100)),
    ("nnedi3-nns16-win8x4", 8192, 4096, "fp16", (7500, 3900, 256, 160)),
])
def test_large_frame_equals_crop(name, w, h, prec, crop):
    import mpx
    planes = 3 if name.endswith(("-yuv", "-rgb")) else 1
    k = 3 if name.startswith("ravu-3x") else 2
    x = _frame(w, h, planes, 5)
    print("large frame", name, w, h, flush=True)
    big = mpx.upscale(name, x, precision=prec)
    c0, r0, cw, ch = crop
    xc = x[..., r0:r0 + ch, c0:c0 + cw].contiguous()
    small = mpx.upscale(name, xc, precision=prec)
    a = big[..., k * (r0 + M):k * (r0 + ch - M), k * (c0 + M):k * (c0 + cw - M)]
    b = small[..., k * M:k * (ch - M), k * M:k * (cw - M)]
    assert a.shape == b.shape
    assert torch.equal(a, b), (name, (a.float() - b.float()).abs().max().item())


@pytest.mark.parametrize("name", ["nnedi3-nns16-win8x4", "nnedi3-nns32-win8x4", "nnedi3-nns64-win8x4",
                                  "nnedi3-nns256-win8x4", "nnedi3-nns32-win8x6", "nnedi3-nns64-win8x6",
                                  "nnedi3-nns256-win8x6"])
def test_tall_staged_tiles_equal_short(name):
    """The one-product FP16 NNEDI3 forms stage 16-row tiles where a launch has enough of them
    (a 1920x1080 frame, both passes) and 8-row tiles otherwise (a 256x128 crop): the two
    give bit-identical outputs over the crop's interior."""
    import mpx
    x = _frame(1920, 1080, 1, 9)
    print("tall and short staged tiles", name, flush=True)
    big = mpx.upscale(name, x, precision="fp16")
    c0, r0, cw, ch = 800, 400, 256, 128
    small = mpx.upscale(name, x[r0:r0 + ch, c0:c0 + cw].contiguous(), precision="fp16")
    a = big[2 * (r0 + M):2 * (r0 + ch - M), 2 * (c0 + M):2 * (c0 + cw - M)]
    b = small[2 * M:2 * (ch - M), 2 * M:2 * (cw - M)]
    assert a.shape == b.shape
    assert torch.equal(a, b), (name, (a.float() - b.float()).abs().max().item())
