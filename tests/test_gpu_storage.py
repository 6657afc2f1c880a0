"""FP16 intermediate storage (MPX_STORE_FP16), mpv's default rgba16f FBOs, against the
oracle's FP16-storage mode (itself held to the reference's passes executed with FP16
storage, tests/test_oracle_golden.py).

With FP16 output, a plan storing its intermediates in FP16 follows the FP16-storage
oracle: NNEDI3 within a few FP16 steps (the FP32 predictions round like the FP64 ones
except within ~1e-7 of a rounding boundary), full RAVU with >= 99.9 % of the pixels within
2^-10 (a rounding flip of int11 can move a step-2/3 bucket).  The same plans without the
flag follow the FP32-storage oracle instead, and single-pass variants (ravu-lite) are
unaffected by the flag."""
import numpy as np
import pytest
import torch

from parity import report

pytestmark = pytest.mark.gpu

W, H = 160, 96


def _oracle(name, img, fbo16):
    from test_oracle_golden import run_oracle
    return run_oracle(name, img, (2 * W, 2 * H), fbo16=fbo16)


@pytest.mark.parametrize("name,prec", [("nnedi3-nns32-win8x4", "fp32"), ("nnedi3-nns64-win8x6", "fp16"),
                                       ("ravu-r3", "fp32"), ("ravu-r4-yuv", "fp32"), ("ravu-lite-ar-r3", "fp32")])
def test_fp16_storage_follows_mpv_default(name, prec):
    import mpx
    from mpx.frames import make_frame
    ch = 3 if name.endswith("-yuv") else 1
    img = make_frame(W, H, seed=31, bits=8, channels=ch).astype(np.float16)
    x = torch.from_numpy(img).cuda()
    ctx = mpx.Context(0)
    outs = {}
    for storage in ("fp32", "fp16"):
        pl = ctx.plan(name, W, H, in_dtype=torch.float16, out_dtype=torch.float16, precision=prec, storage=storage)
        outs[storage] = pl.run(x).float().cpu().numpy().astype(np.float64)
        pl.close()
    torch.cuda.synchronize()
    ref16 = _oracle(name, img.astype(np.float64), True)
    ref32 = _oracle(name, img.astype(np.float64), False).astype(np.float16).astype(np.float64)
    e16 = np.abs(outs["fp16"] - ref16)
    e32 = np.abs(outs["fp32"] - ref32)
    gap = np.abs(ref16 - ref32)
    within = float(np.mean(e16 <= 2.0 ** -10))
    report("%s %s FP16 storage vs FP16-storage oracle %dx%d" % (name, prec, W, H), max_abs=float(e16.max()),
           within_2m10=within, storage_gap_max=float(gap.max()), fp32_storage_vs_oracle_max=float(e32.max()))
    if name.startswith("ravu-lite"):
        assert np.array_equal(outs["fp16"], outs["fp32"])     # single pass: nothing stored between passes
        assert gap.max() == 0.0
    else:
        # the storage changes the result, and the flag moves the output to mpv's default
        assert gap.max() > 0.0 and not np.array_equal(outs["fp16"], outs["fp32"])
        assert float(e16.mean()) < float(np.abs(outs["fp32"] - ref16).mean())
    if name.startswith("nnedi3"):
        assert e16.max() <= 4e-3
    assert within >= 0.999
    assert float(np.mean(e32 <= 2.0 ** -10)) >= 0.999
