"""mpx_query against the reference's own //!WHEN / WIDTH / HEIGHT / OFFSET directives.

tests/golden/query_golden.json holds, for every variant the reference ships a file for
(43) and 30 (input, display) size pairs (2x / 3x targets, targets on either side of the
0.707106 and 0.471404 thresholds on each axis, shrinking and equal targets, tiny and odd
sizes), what mpv's pass walk gives: whether the hook replaces the plane, its final size,
the accumulated OFFSET and the ALIGN flag (make_query_table.py evaluates the directive
RPN in float32 from the hook text).  The three ravu-zoom-ar-r3 forms have no file (the
reference lists them in .MISSING_LARGE_BLOBS); every ravu-zoom file carries the same
directive block, so they are held to their ravu-zoom-ar-r2 counterparts."""
import json
import os

import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "query_golden.json")


@pytest.fixture(scope="module")
def table():
    d = json.load(open(GOLDEN))
    return [dict(zip(d["columns"], r)) for r in d["rows"]]


def test_table_covers_every_shipped_file(table):
    import mpx
    names = {r["name"] for r in table}
    shipped = set(mpx.variants()) - {"ravu-zoom-ar-r3", "ravu-zoom-ar-r3-rgb", "ravu-zoom-ar-r3-yuv"}
    assert names == shipped


def test_query_matches_directives(table):
    import mpx
    bad = []
    for r in table:
        q = mpx.query(r["name"], r["in_w"], r["in_h"], r["target_w"], r["target_h"])
        got = (q["applies"], q["out_w"], q["out_h"], q["offset_x"], q["offset_y"], q["align"])
        want = (r["applies"], r["out_w"], r["out_h"], r["offset_x"], r["offset_y"], r["align"])
        if got != want:
            bad.append((r["name"], r["in_w"], r["in_h"], r["target_w"], r["target_h"], got, want))
    assert not bad, bad[:10]


def test_zoom_ar_r3_follows_the_zoom_directives(table):
    import mpx
    for r in table:
        if r["name"].startswith("ravu-zoom-ar-r2"):
            name = r["name"].replace("-r2", "-r3")
            a = mpx.query(name, r["in_w"], r["in_h"], r["target_w"], r["target_h"])
            b = mpx.query(r["name"], r["in_w"], r["in_h"], r["target_w"], r["target_h"])
            for k in ("applies", "out_w", "out_h", "offset_x", "offset_y", "align", "hook"):
                assert a[k] == b[k], (name, k)
