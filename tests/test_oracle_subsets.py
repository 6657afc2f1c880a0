"""The oracle's partial evaluations that the full-size GPU parity tests use: an NNEDI3
output region computed on an input crop, and a ravu-zoom output sub-grid, equal the
full-frame oracle exactly (CPU only)."""
import numpy as np
import pytest

from parity import nnedi3_crop_ref


@pytest.mark.parametrize("wh", [4, 6])
def test_nnedi3_crop_equals_full(wh):
    from mpx.frames import make_plane
    from oracle.nnedi3 import nnedi3
    img = make_plane(90, 70, 5, 10)
    full = nnedi3(img, 16, wh)
    for (ox0, oy0, ow, oh) in [(0, 0, 40, 36), (70, 50, 40, 40), (140, 100, 40, 40), (33, 77, 21, 19)]:
        ref = nnedi3_crop_ref(img, 16, wh, ox0, oy0, ow, oh)
        assert np.array_equal(ref, full[oy0:oy0 + oh, ox0:ox0 + ow]), (ox0, oy0)


def test_nnedi3_crop_margin_is_needed():
    """With a margin too small, interior crops differ from the full oracle: the check above
    has teeth."""
    from mpx.frames import make_plane
    from oracle.nnedi3 import nnedi3
    img = make_plane(90, 70, 5, 10)
    full = nnedi3(img, 16, 6)
    ref = nnedi3_crop_ref(img, 16, 6, 70, 50, 40, 40, margin=2)
    assert not np.array_equal(ref, full[50:90, 70:110])


@pytest.mark.parametrize("ar,ch", [(True, "rgb"), (False, "yuv"), (True, "luma")])
def test_zoom_subgrid_equals_full(ar, ch):
    from mpx.frames import make_frame
    from oracle.ravu import ravu_zoom
    planes = 1 if ch == "luma" else 3
    img = make_frame(40, 30, 3, 8, channels=planes)
    full, aux = ravu_zoom(img, 2, 100, 70, ar=ar, channels=ch)
    cols, rows = [0, 5, 17, 98, 99], [0, 33, 69]
    sub, aux2 = ravu_zoom(img, 2, 100, 70, ar=ar, channels=ch, cols=cols, rows=rows)
    assert np.array_equal(full[..., rows, :][..., cols], sub)
    assert np.array_equal(aux["row"][np.ix_(rows, cols)], aux2["row"])
