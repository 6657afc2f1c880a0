"""How far FP32 intermediates (this library's default) are from mpv's default FP16
storage of a hook's textures, measured with the oracle's two storage modes on seeded
8-bit frames (both outputs rounded to FP16, so only the intermediates differ).  The
FP16-storage mode is pinned by the reference's passes executed with FP16 storage
(tests/test_oracle_golden.py).  CPU only.

Measured here (192x108, 3 seeds): single-pass variants (ravu-lite, ravu-3x, ravu-zoom)
store nothing between passes, gap 0; NNEDI3 stores the double_y plane, gap about one FP16
step (~1e-3, 79 dB); full RAVU stores int11, whose FP16 rounding moves step-2/3 buckets
of a few per cent of the pixels (gap up to ~0.1, 72-77 dB)."""
import numpy as np
import pytest

from parity import psnr, report

W, H = 192, 108


@pytest.mark.parametrize("name,single_pass", [
    ("nnedi3-nns32-win8x4", False), ("nnedi3-nns64-win8x6", False), ("ravu-r3", False), ("ravu-r4-yuv", False),
    ("ravu-lite-ar-r3", True), ("ravu-3x-r3", True), ("ravu-zoom-ar-r2-rgb", True)])
def test_fp16_storage_gap(name, single_pass):
    from mpx.frames import make_frame
    from test_oracle_golden import run_oracle
    ch = 3 if name.endswith(("-yuv", "-rgb")) else 1
    worst, d_all = 0.0, []
    for seed in range(3):
        img = make_frame(W, H, seed=50 + seed, bits=8, channels=ch).astype(np.float64)
        tgt = (3 * W, 3 * H) if "-3x-" in name else ((W * 3 // 2, H * 3 // 2) if "zoom" in name else (2 * W, 2 * H))
        a = run_oracle(name, img, tgt).astype(np.float16).astype(np.float64)
        b = run_oracle(name, img, tgt, fbo16=True)
        d_all.append(a - b)
        worst = max(worst, float(np.abs(a - b).max()))
    d = np.concatenate([x.reshape(-1) for x in d_all])
    report("FP16-storage gap %s %dx%d x3 (FP32 vs FP16 intermediates, FP16 out)" % (name, W, H), max_abs=worst,
           psnr=psnr(d), differing=float(np.mean(d != 0)))
    if single_pass:
        assert worst == 0.0
    elif name.startswith("nnedi3"):
        assert 0.0 < worst <= 2.0 ** -9          # a step or two of FP16 near 1
    else:
        assert worst > 2.0 ** -9                 # bucket moves: a visible difference
