"""Byte-level YUV4MPEG2 fixtures, written from the format's rules with struct (not with
mpx.frameio), with their expected samples in expected.json.

A stream is one header line -- "YUV4MPEG2" and space-separated tagged parameters (W width,
H height, F rate, I interlacing, A aspect, C colour space, X comment), "\\n" -- then per
frame "FRAME", optional tagged parameters, "\\n" and the planes Y, U, V (or Y alone for
mono) row by row, one byte per sample up to 8 bits, two bytes little-endian above.
Chroma planes of 4:2:0 are ceil(W/2) x ceil(H/2), of 4:2:2 ceil(W/2) x H.

    python tests/golden/y4m/make_y4m.py
"""
import json
import os
import struct

HERE = os.path.dirname(os.path.abspath(__file__))


def samples(n, seed, top):
    v, out = seed, []
    for _ in range(n):
        v = (v * 1103515245 + 12345) % (1 << 31)
        out.append(v % (top + 1))
    return out


CASES = [
    # name, header after "YUV4MPEG2 ", W, H, chroma (sx, sy) or None, bits, frame lines
    ("c420p10.y4m", "W5 H3 F25:1 It A0:0 C420p10 XYSCSS=420P10", 5, 3, (2, 2), 10, ["FRAME", "FRAME Ixyz Xnote"]),
    ("c422.y4m", "W4 H2 F30000:1001 Ip A1:1 C422 XCOLORRANGE=LIMITED", 4, 2, (2, 1), 8, ["FRAME"]),
    ("mono16.y4m", "W3 H2 F24:1 Ip A1:1 Cmono16", 3, 2, None, 16, ["FRAME", "FRAME"]),
    ("c444p12.y4m", "W2 H2 F1:1 Ib C444p12", 2, 2, (1, 1), 12, ["FRAME Ib"]),
    ("nocs.y4m", "W3 H3 F25:1", 3, 3, (2, 2), 8, ["FRAME"]),   # no C tag: 4:2:0, 8-bit
]


def main():
    expected = {}
    for name, hdr, w, h, ch, bits, frames in CASES:
        body = ("YUV4MPEG2 " + hdr + "\n").encode("ascii")
        fr = []
        for i, line in enumerate(frames):
            body += (line + "\n").encode("ascii")
            shapes = [(h, w)]
            if ch:
                shapes += [(-(-h // ch[1]), -(-w // ch[0]))] * 2
            planes = []
            for k, (ph, pw) in enumerate(shapes):
                vals = samples(ph * pw, 7 + 31 * i + k, (1 << bits) - 1)
                body += struct.pack("<%d%s" % (len(vals), "B" if bits <= 8 else "H"), *vals)
                planes.append([vals[r * pw:(r + 1) * pw] for r in range(ph)])
            fr.append(planes)
        with open(os.path.join(HERE, name), "wb") as f:
            f.write(body)
        expected[name] = {"width": w, "height": h, "bits": bits, "frames": fr}
    with open(os.path.join(HERE, "expected.json"), "w") as f:
        json.dump(expected, f)
    print("wrote", ", ".join(c[0] for c in CASES))


if __name__ == "__main__":
    main()
