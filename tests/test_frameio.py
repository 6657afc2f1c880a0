"""y4m / raw planar frame I/O (CPU)."""
import numpy as np
import pytest


@pytest.mark.parametrize("cs,bits", [("420jpeg", 8), ("420p10", 10), ("444", 8), ("mono", 8)])
def test_y4m_round_trip(tmp_path, cs, bits):
    from mpx.frameio import Y4MReader, Y4MWriter, quantise
    from mpx.frames import make_plane
    w, h = 37, 22
    path = str(tmp_path / "a.y4m")
    wr = Y4MWriter(path, w, h, cs=cs)
    frames = []
    for i in range(3):
        y = quantise(make_plane(w, h, i, bits=bits), bits)
        planes = [y]
        if cs != "mono":
            sx = 1 if cs == "444" else 2
            cw, ch = -(-w // sx), -(-h // sx)
            planes += [quantise(make_plane(cw, ch, 10 + i, bits), bits), quantise(make_plane(cw, ch, 20 + i, bits), bits)]
        wr.write(planes)
        frames.append(planes)
    wr.close()
    rd = Y4MReader(path)
    assert (rd.width, rd.height, rd.bits) == (w, h, bits)
    got = list(rd)
    rd.close()
    assert len(got) == 3
    for a, b in zip(got, frames):
        for pa, pb in zip(a, b):
            assert pa.dtype == (np.uint8 if bits == 8 else np.uint16)
            assert np.array_equal(pa, pb)


def test_raw_planar(tmp_path):
    from mpx.frameio import read_raw_planar
    a = (np.arange(2 * 3 * 5 * 4) % 1024).astype("<u2").reshape(2, 3, 5, 4)
    p = tmp_path / "r.raw"
    p.write_bytes(a.tobytes())
    assert np.array_equal(read_raw_planar(str(p), 4, 5, planes=3, bits=10, frame=1), a[1])


def test_quantise_matches_generator_levels():
    from mpx.frameio import quantise
    from mpx.frames import make_plane
    f = make_plane(16, 9, 1, bits=10)
    q = quantise(f, 10)
    assert np.array_equal(q.astype(np.float64) / 1023, f)
