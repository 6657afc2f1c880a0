"""The offset-consuming scaler on the GPU against the FP64 restatement (oracle/resample.py),
and display alignment of a prescaler's output through it."""
import numpy as np
import pytest
import torch

from parity import compare

pytestmark = pytest.mark.gpu

FILTERS = ["bilinear", "catmull_rom", "mitchell", "spline36", "lanczos3"]


@pytest.mark.parametrize("name", FILTERS)
@pytest.mark.parametrize("sizes,offset", [(((50, 30), (75, 45)), (0.0, 0.0)), (((64, 36), (100, 70)), (-0.5, -0.5)),
                                          (((40, 40), (37, 23)), (0.25, -0.75))])
def test_resample_matches_oracle(name, sizes, offset):
    import mpx
    from mpx.frames import make_plane
    from oracle.resample import resample
    (w, h), (ow, oh) = sizes
    img = make_plane(w, h, 3, bits=10).astype(np.float32)
    got = mpx.resample(torch.from_numpy(img).cuda(), (ow, oh), offset=offset, filter=name).cpu().numpy()
    f = compare("resample %s %dx%d->%dx%d off %s" % (name, w, h, ow, oh, offset), got,
                resample(img.astype(np.float64), ow, oh, offset, name))
    assert f["max_abs"] <= 2e-6


def test_resample_u16_planes_batch():
    import mpx
    from oracle.resample import resample
    rng = np.random.default_rng(4)
    x = rng.integers(0, 1024, size=(2, 3, 20, 30)).astype(np.uint16)
    t = torch.from_numpy(x.astype(np.int32)).to(torch.uint16).cuda()
    got = mpx.resample(t, (45, 31), offset=(-0.5, 0.0), filter="spline36", planes=3, bits=10).cpu().numpy()
    want = resample(x / 1023.0, 45, 31, (-0.5, 0.0), "spline36")
    assert np.abs(got - want).max() <= 2e-6


def test_nnedi3_display_alignment():
    """A linear ramp doubled by NNEDI3 carries OFFSET (-0.5, -0.5); resampled with that
    offset it lands on the ramp evaluated at the display pixel centres (up to NNEDI3's own
    deviation from a ramp), and without it half a pixel away."""
    import mpx
    w, h = 64, 48
    x = (np.arange(w) + 0.5) / w
    y = (np.arange(h) + 0.5) / h
    img = (0.05 + 0.6 * x[None, :] + 0.3 * y[:, None]).astype(np.float32)
    info = mpx.query("nnedi3-nns32-win8x4", w, h, 2 * w, 2 * h)
    assert (info["offset_x"], info["offset_y"]) == (-0.5, -0.5)
    pre = mpx.upscale("nnedi3-nns32-win8x4", torch.from_numpy(img).cuda(), precision="fp16")
    ow, oh = 2 * w, 2 * h
    jx = (np.arange(ow) + 0.5) / ow
    jy = (np.arange(oh) + 0.5) / oh
    ideal = 0.05 + 0.6 * jx[None, :] + 0.3 * jy[:, None]
    inner = (slice(8, oh - 8), slice(8, ow - 8))
    aligned = mpx.resample(pre, (ow, oh), offset=(info["offset_x"], info["offset_y"]),
                           filter="catmull_rom").cpu().numpy()
    raw = mpx.resample(pre, (ow, oh), filter="catmull_rom").cpu().numpy()
    err_aligned = np.abs(aligned - ideal)[inner].max()
    err_raw = np.abs(raw - ideal)[inner].max()
    # half a display pixel of the ramp: 0.5 * (0.6 / 128 + 0.3 / 96) = 3.9e-3
    assert err_raw > 3e-3, err_raw
    assert err_aligned < 0.25 * err_raw, (err_aligned, err_raw)
