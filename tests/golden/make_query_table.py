"""Expected geometry and gating of every shipped variant, from the reference's own
`//!WHEN`, `//!WIDTH`, `//!HEIGHT` and `//!OFFSET` directives.

For each `.hook` file (root flavour; ravu-3x from compute/, its only flavour) and each
(input size, display size) pair of a grid, the passes are walked as mpv's user-shader
engine walks them:
  * a pass runs when its WHEN expression is non-zero.  The expressions are RPN over
    HOOKED.w/h (the hooked plane as left by earlier passes), OUTPUT.w/h (the display
    size) and LUMA.w (non-zero: the -yuv forms hook NATIVE of a frame that has a luma
    plane), evaluated in float32 as mpv does (e.g. nnedi3-nns32-win8x4.hook:19,
    ravu-r4-yuv.hook:20, compute/ravu-3x-r3.hook:21, ravu-zoom-ar-r2-rgb.hook:23);
  * a pass with SAVE writes a named texture and leaves the hooked plane alone; a pass
    without it replaces the hooked plane with its output, sized by WIDTH/HEIGHT (RPN,
    default: the hooked plane's size);
  * OFFSET ox oy is a shift in the pixels of the pass's output; a later pass that scales
    an axis by s scales the shift accumulated so far by s (the offset is part of the
    plane's texture transform, composed pass by pass).  OFFSET ALIGN marks an output
    aligned to the display (ravu-zoom).
The result per case: whether any pass replaced the hooked plane ("applies"), the final
size, the accumulated offset and the ALIGN flag -- what mpx_query reports.

Writes tests/golden/query_golden.json.  Run here, where the reference checkout exists:
    python tests/golden/make_query_table.py [/root/reference]
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from hookexec import parse_hook_file  # noqa: E402

F = np.float32

# (in_w, in_h, target_w, target_h): 2x and 3x targets, targets just either side of the
# 0.707106 and 0.471404 ratio thresholds on each axis separately, shrinking and equal
# targets, odd and tiny sizes
SIZES = [
    (640, 360, 1280, 720), (640, 360, 1920, 1080), (1920, 1080, 3840, 2160), (1920, 1080, 2880, 1620),
    (1280, 720, 2560, 1440), (1366, 768, 1920, 1080), (1000, 1000, 1414, 1414), (1000, 1000, 1415, 1415),
    (1000, 1000, 1415, 1414), (1000, 1000, 1414, 1415), (1000, 1000, 2121, 2121), (1000, 1000, 2122, 2122),
    (1000, 1000, 2122, 2121), (1000, 1000, 3000, 1000), (1000, 1000, 1000, 3000), (1000, 1000, 1000, 1000),
    (1000, 1000, 999, 2000), (1000, 1000, 2000, 999), (1920, 1080, 1280, 720), (1, 1, 2, 2), (1, 1, 3, 3),
    (3, 2, 5, 3), (7, 5, 16, 11), (707106, 1000, 1000000, 1415), (33, 17, 47, 25), (33, 17, 70, 37),
    (4096, 2160, 8192, 4320), (720, 576, 1920, 1080), (853, 480, 1707, 961), (16, 9, 34, 19),
]


def rpn32(expr, env):
    """mpv's size / condition expressions: RPN over float32 (+ - * / < > ! =)."""
    st = []
    for tok in expr.split():
        if tok in ("+", "-", "*", "/", "<", ">", "="):
            b, a = st.pop(), st.pop()
            if tok == "+": st.append(F(a + b))
            elif tok == "-": st.append(F(a - b))
            elif tok == "*": st.append(F(a * b))
            elif tok == "/": st.append(F(a / b))
            elif tok == "<": st.append(F(1.0 if a < b else 0.0))
            elif tok == ">": st.append(F(1.0 if a > b else 0.0))
            else: st.append(F(1.0 if a == b else 0.0))
        elif tok == "!":
            st.append(F(0.0 if st.pop() else 1.0))
        elif tok in env:
            st.append(F(env[tok]))
        else:
            st.append(F(float(tok)))
    assert len(st) == 1, expr
    return st[0]


def evaluate(passes, in_w, in_h, tw, th):
    hw, hh = in_w, in_h
    off = [0.0, 0.0]
    align = 0
    applies = 0
    for p in passes:
        env = {"HOOKED.w": hw, "HOOKED.h": hh, "OUTPUT.w": tw, "OUTPUT.h": th, "LUMA.w": in_w, "LUMA.h": in_h}
        if "WHEN" in p and not rpn32(p["WHEN"], env):
            continue
        if "SAVE" in p:
            continue
        nw = int(rpn32(p["WIDTH"], env)) if "WIDTH" in p else hw
        nh = int(rpn32(p["HEIGHT"], env)) if "HEIGHT" in p else hh
        off = [off[0] * nw / hw, off[1] * nh / hh]
        o = p.get("OFFSET")
        if o == "ALIGN":
            align = 1
            off = [0.0, 0.0]
        elif o:
            ox, oy = (float(v) for v in o.split())
            off = [off[0] + ox, off[1] + oy]
        hw, hh = nw, nh
        applies = 1
    return {"applies": applies, "out_w": hw, "out_h": hh, "offset_x": off[0], "offset_y": off[1], "align": align}


def hook_files(ref):
    out = {}
    for f in sorted(os.listdir(ref)):
        if f.endswith(".hook"):
            out[f[:-5]] = os.path.join(ref, f)
    for f in sorted(os.listdir(os.path.join(ref, "compute"))):
        if f.startswith("ravu-3x") and f.endswith(".hook"):
            out[f[:-5]] = os.path.join(ref, "compute", f)
    return out


def main(ref):
    rows = []
    sources = {}
    for name, path in sorted(hook_files(ref).items()):
        passes, _ = parse_hook_file(path)
        lines = [p["line"] for p in passes]
        sources[name] = "%s (directive blocks at lines %s)" % (os.path.relpath(path, ref), ",".join(map(str, lines)))
        for (iw, ih, tw, th) in SIZES:
            r = evaluate(passes, iw, ih, tw, th)
            r.update(name=name, in_w=iw, in_h=ih, target_w=tw, target_h=th)
            rows.append(r)
    cols = ["name", "in_w", "in_h", "target_w", "target_h", "applies", "out_w", "out_h", "offset_x", "offset_y",
            "align"]
    doc = {"about": "expected mpx_query results from the reference's //!WHEN/WIDTH/HEIGHT/OFFSET "
                    "directives (tests/golden/make_query_table.py)",
           "sources": sources, "columns": cols, "rows": [[r[c] for c in cols] for r in rows]}
    path = os.path.join(HERE, "query_golden.json")
    with open(path, "w") as f:
        f.write(json.dumps({k: v for k, v in doc.items() if k != "rows"}, sort_keys=True)[:-1])
        f.write(', "rows": [\n' + ",\n".join(json.dumps(r) for r in doc["rows"]) + "\n]}\n")
    print("wrote %s: %d variants x %d size cases" % (path, len(sources), len(SIZES)))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
