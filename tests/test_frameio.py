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


def test_y4m_byte_fixtures():
    """Hand-built YUV4MPEG2 streams (tests/golden/y4m/make_y4m.py writes them with struct
    from the format's rules): header tags in any order with comments, FRAME lines with
    parameters, little-endian 16-bit samples, 4:2:0 / 4:2:2 / 4:4:4 / mono16, odd sizes,
    and a stream with no C tag (4:2:0 8-bit)."""
    import json
    import os
    from mpx.frameio import Y4MReader
    d = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "y4m")
    exp = json.load(open(os.path.join(d, "expected.json")))
    assert len(exp) == 5
    for name, e in exp.items():
        rd = Y4MReader(os.path.join(d, name))
        assert (rd.width, rd.height, rd.bits) == (e["width"], e["height"], e["bits"]), name
        got = list(rd)
        rd.close()
        assert len(got) == len(e["frames"]), name
        for gf, ef in zip(got, e["frames"]):
            assert len(gf) == len(ef), name
            for gp, ep in zip(gf, ef):
                assert gp.dtype == (np.uint8 if e["bits"] <= 8 else np.uint16)
                assert np.array_equal(gp, np.array(ep)), name


def test_y4m_writer_bytes(tmp_path):
    """The writer's bytes: header line, FRAME line, Y then U then V, 16-bit little-endian."""
    import struct
    from mpx.frameio import Y4MWriter
    p = str(tmp_path / "w.y4m")
    wr = Y4MWriter(p, 3, 2, cs="420p10", fps="25:1")
    y = np.array([[0, 1, 1023], [512, 2, 3]], np.uint16)
    u = np.array([[7, 8]], np.uint16)
    v = np.array([[9, 1000]], np.uint16)
    wr.write([y, u, v])
    wr.close()
    want = b"YUV4MPEG2 W3 H2 F25:1 Ip A1:1 C420p10\nFRAME\n" + struct.pack("<10H", 0, 1, 1023, 512, 2, 3, 7, 8, 9, 1000)
    assert open(p, "rb").read() == want


@pytest.mark.parametrize("tag", ["411", "420p7", "mono17", "444alpha"])
def test_y4m_rejects_unknown_colourspace(tmp_path, tag):
    from mpx.frameio import Y4MReader
    p = tmp_path / "x.y4m"
    p.write_bytes(("YUV4MPEG2 W2 H2 C%s\nFRAME\n" % tag).encode() + bytes(16))
    with pytest.raises(ValueError):
        Y4MReader(str(p))
