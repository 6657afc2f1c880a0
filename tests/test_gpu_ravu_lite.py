"""RAVU-lite(-ar) on the GPU vs the FP64 oracle (ravu-lite-r3.hook, ravu-lite-ar-r3.hook).

Bar: bucket rows agree on >= 99.99 % of sites (the HIP hash re-decides in FP64 near
bucket boundaries) and, on pixels whose site agrees, max-abs <= 2e-6 (FP32 taps over
FP16-rounded weights vs FP64)."""
import numpy as np
import pytest
import torch

from parity import compare

pytestmark = pytest.mark.gpu

TOL = 2e-6


def _frame(w, h, seed, bits=8):
    from mpx.frames import make_plane
    return make_plane(w, h, seed, bits).astype(np.float32)


def _site_to_out_2x(agree):
    return np.repeat(np.repeat(agree, 2, axis=0), 2, axis=1)


@pytest.mark.parametrize("r", [2, 3, 4])
@pytest.mark.parametrize("ar", [False, True])
def test_lite_parity_640x360(r, ar):
    import mpx
    from oracle.ravu import ravu_lite
    name = ("ravu-lite-ar-r%d" if ar else "ravu-lite-r%d") % r
    img = _frame(640, 360, seed=100 + r)
    x = torch.from_numpy(img).cuda()
    out, rows = mpx.upscale(name, x, buckets=True)
    torch.cuda.synchronize()
    ref, aux = ravu_lite(img.astype(np.float64), r, ar=ar)
    f = compare(name + " 640x360 f32", out.cpu().numpy(), ref, rows.cpu().numpy(), aux["row"],
                site_shape=img.shape, site_to_out=_site_to_out_2x)
    assert f["bucket_agree"] >= 0.9999
    assert f["max_abs_agreeing"] <= TOL


@pytest.mark.parametrize("shape", [(1, 1), (3, 2), (5, 7), (67, 5), (130, 9)])
def test_lite_ragged_and_tiny(shape):
    """Frames smaller than the halo and widths that are not tile multiples: clamp-to-edge
    everywhere, no out-of-range writes (the output buffer carries guard bands)."""
    import mpx
    from oracle.ravu import ravu_lite
    w, h = shape
    img = _frame(w, h, seed=7)
    x = torch.from_numpy(img).cuda()
    plan = mpx.default_context().plan("ravu-lite-ar-r3", w, h, in_dtype=torch.float32, out_dtype=torch.float32)
    guard = torch.full((2 * h + 2, 2 * w + 8), -7.0, device="cuda")
    view = guard[1:1 + 2 * h, 4:4 + 2 * w]
    plan.run(x, out=view)
    torch.cuda.synchronize()
    g = guard.cpu().numpy()
    ref, _ = ravu_lite(img.astype(np.float64), 3, ar=True)
    assert np.abs(g[1:1 + 2 * h, 4:4 + 2 * w] - ref).max() <= 1e-5
    g[1:1 + 2 * h, 4:4 + 2 * w] = -7.0
    assert (g == -7.0).all(), "write outside the output view"


def test_lite_batch_fp16_u8():
    """Batched launch, 8-bit integer input and FP16 output: every frame equals its own
    single-frame FP32 run rounded to FP16."""
    import mpx
    frames = np.stack([_frame(96, 64, seed=s) for s in range(4)])
    u8 = np.round(frames * 255).astype(np.uint8)
    x8 = torch.from_numpy(u8).cuda()
    out16 = mpx.upscale("ravu-lite-r3", x8, out_dtype=torch.float16)
    for i in range(4):
        single = mpx.upscale("ravu-lite-r3", torch.from_numpy(u8[i].astype(np.float32) / 255.0).cuda())
        # 8-bit unorm load is value/255 in FP32, as torch's float32 division
        ref16 = single.half().float()
        assert torch.equal(out16[i].float(), ref16), i


def test_lite_skipped_plan_copies():
    """When //!WHEN is false the hook is skipped: the output is the input."""
    import mpx
    img = _frame(64, 48, seed=3)
    x = torch.from_numpy(img).cuda()
    out = mpx.upscale("ravu-lite-r3", x, target=(64, 60))
    assert out.shape == x.shape and torch.equal(out, x)
