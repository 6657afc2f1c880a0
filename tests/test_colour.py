"""YUV -> R'G'B' conversion before MAIN hooks (mpx.hooks.yuv_to_rgb), on the CPU with
4:4:4 planes: known colours, and a round trip through the forward transform written from
its definition (Y' = Kr R' + Kg G' + Kb B', Cb = (B' - Y') / (2 (1 - Kb)),
Cr = (R' - Y') / (2 (1 - Kr))) for BT.709 / BT.601 at limited and full range."""
import numpy as np
import pytest
import torch

K = {"bt709": (0.2126, 0.0722), "bt601": (0.299, 0.114)}


def _forward(rgb, matrix, levels, bits):
    kr, kb = K[matrix]
    r, g, b = rgb
    y = kr * r + (1 - kr - kb) * g + kb * b
    cb = (b - y) / (2 * (1 - kb))
    cr = (r - y) / (2 * (1 - kr))
    m = (1 << bits) - 1
    if levels == "limited":
        s = 1 << (bits - 8)
        codes = (16 * s + 219 * s * y, 128 * s + 224 * s * cb, 128 * s + 224 * s * cr)
    else:
        codes = (m * y, (1 << (bits - 1)) + m * cb, (1 << (bits - 1)) + m * cr)
    return codes, m


@pytest.mark.parametrize("matrix", ["bt709", "bt601"])
@pytest.mark.parametrize("levels", ["limited", "full"])
@pytest.mark.parametrize("bits", [8, 10])
def test_round_trip(matrix, levels, bits):
    from mpx.hooks import yuv_to_rgb
    rng = np.random.default_rng(3)
    rgb = rng.random((3, 7, 11))
    codes, m = _forward(rgb, matrix, levels, bits)
    frame = {k: torch.from_numpy((c / m).astype(np.float32)) for k, c in zip("YUV", codes)}
    got = yuv_to_rgb(frame, bits=bits, matrix=matrix, levels=levels).numpy()
    assert got.shape == (3, 7, 11)
    assert np.abs(got - rgb).max() <= 2e-6


def test_known_colours_and_integer_planes():
    from mpx.hooks import yuv_to_rgb
    # 8-bit limited range: black, white, BT.709 red at its exact (fractional) codes
    y = torch.tensor([[16.0, 235.0, 16 + 219 * 0.2126]])
    u = torch.tensor([[128.0, 128.0, 128 - 112 * 0.2126 / (1 - 0.0722)]])
    v = torch.tensor([[128.0, 128.0, 240.0]])
    got = yuv_to_rgb({"Y": y / 255, "U": u / 255, "V": v / 255}, bits=8).numpy()
    want = np.array([[[0, 1, 1]], [[0, 1, 0]], [[0, 1, 0]]], dtype=np.float64)
    assert np.abs(got - want).max() <= 2e-6
    # integer codes give the same result as their normalised floats
    yi, ui, vi = (torch.tensor([[64, 940, 502]], dtype=torch.int16) for _ in range(3))
    fi = {"Y": yi, "U": ui, "V": vi}
    ff = {k: t.to(torch.float32) / 1023 for k, t in fi.items()}
    assert torch.allclose(yuv_to_rgb(fi, bits=10), yuv_to_rgb(ff, bits=10), atol=1e-6)
    with pytest.raises(ValueError):
        yuv_to_rgb(ff, matrix="bt2020")
