"""The resampling oracle against closed forms (CPU)."""
import numpy as np
import pytest

from oracle.resample import RADIUS, axis_matrix, kernel, resample

FILTERS = sorted(RADIUS)


@pytest.mark.parametrize("name", ["bilinear", "catmull_rom", "spline36", "lanczos3"])
def test_identity_and_integer_shift(name):
    """Interpolating filters (kernel(0) = 1, kernel(k) = 0 at other integers) reproduce
    the plane at scale 1 and shift it by integer offsets (Mitchell is not interpolating)."""
    rng = np.random.default_rng(1)
    img = rng.random((9, 13))
    assert np.allclose(resample(img, 13, 9, name=name), img, atol=1e-12)
    shifted = resample(img, 13, 9, offset=(1.0, 0.0), name=name)
    assert np.allclose(shifted[:, :-1], img[:, 1:], atol=1e-12)
    assert np.allclose(shifted[:, -1], img[:, -1], atol=1e-12)   # clamp-to-edge


@pytest.mark.parametrize("name", FILTERS)
def test_constant_and_weight_sums(name):
    for in_n, out_n in ((7, 15), (16, 24), (40, 17)):
        m = axis_matrix(name, in_n, out_n, 0.25)
        assert np.allclose(m.sum(axis=1), 1.0, atol=1e-12)
    assert np.allclose(resample(np.full((5, 6), 0.3), 17, 11, offset=(-0.5, 0.5), name=name), 0.3, atol=1e-12)


@pytest.mark.parametrize("name", ["bilinear", "catmull_rom", "mitchell"])
def test_linear_ramp_reproduced(name):
    """Tent and the (B, C) cubics with B + 2C = 1 reproduce linear functions: away from the
    clamped border, sampling a ramp gives the ramp at the sample point."""
    w, h, out_w = 24, 4, 36
    x = np.arange(w) + 0.5
    img = np.tile(0.1 + 0.02 * x, (h, 1))
    off = -0.5
    out = resample(img, out_w, h, offset=(off, 0.0), name=name)
    t = (np.arange(out_w) + 0.5) * w / out_w + off
    want = 0.1 + 0.02 * t
    inner = slice(6, out_w - 6)
    assert np.allclose(out[0, inner], want[inner], atol=1e-12)


def test_kernel_values():
    assert kernel("catmull_rom", 0.0) == 1.0 and kernel("catmull_rom", 1.0) == 0.0
    assert abs(kernel("mitchell", 0.0) - 8 / 9) < 1e-15
    assert abs(kernel("spline36", 0.0) - 1.0) < 1e-15 and abs(kernel("spline36", 1.0)) < 1e-15
    assert abs(kernel("lanczos3", 0.0) - 1.0) < 1e-15 and abs(kernel("lanczos3", 2.0)) < 1e-15


@pytest.mark.parametrize("loc,sub", [("left", 2), ("center", 2), ("left", 4), ("center", 3)])
def test_chroma_siting_offset_reproduces_a_ramp(loc, sub):
    """A chroma plane whose samples hold their own luma x position (left siting: chroma k
    on luma sample sub*k; centre siting: at the centre of luma samples sub*k..sub*k+sub-1,
    luma position sub*k + (sub-1)/2) resampled bilinearly to the luma width with the
    siting's offset gives back the luma positions 0, 1, 2, ... away from the clamped
    edges: the offset places every chroma sample where its siting says."""
    from mpx.hooks import chroma_offset
    from oracle.resample import resample
    cw, lw = 12, 12 * sub
    k = np.arange(cw, dtype=np.float64)
    pos = sub * k if loc == "left" else sub * k + (sub - 1) / 2.0
    c = np.tile(pos, (3, 1))
    off = chroma_offset((lw, 3), (cw, 3), loc)
    assert off[1] == 0.0
    up = resample(c, lw, 3, offset=off, name="bilinear")[1]
    j = np.arange(lw, dtype=np.float64)
    inner = (j >= pos[0]) & (j <= pos[-1])
    assert np.allclose(up[inner], j[inner], atol=1e-12)
