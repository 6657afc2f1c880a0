"""Every tile shape of every RAVU family gives the same output, bit for bit, as the
default shape: luma / yuv / rgb forms, ragged sizes, batches, frames of different sizes in
one launch, and zoom's weight-table and per-pixel paths.  (The per-pixel arithmetic does
not depend on how pixels are grouped into blocks; the default shape's output is held to
the FP64 oracle by the family tests.)"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _frames(planes, sizes, seed, dtype=np.float16):
    from mpx.frames import make_frame
    return [torch.from_numpy(make_frame(w, h, seed + i, bits=10, channels=planes).astype(dtype)).cuda()
            for i, (w, h) in enumerate(sizes)]


def _each_tile(plan, run):
    n = plan.tile()["count"]
    assert n >= 2
    outs = []
    for t in range(n):
        plan.set_tile(t)
        outs.append(run())
    plan.set_tile(0)
    return outs


NAMES = ["ravu-lite-r2", "ravu-lite-ar-r3", "ravu-lite-r4", "ravu-lite-ar-r4",
         "ravu-3x-r2", "ravu-3x-r3-yuv", "ravu-3x-r4-rgb",
         "ravu-r2", "ravu-r3-yuv", "ravu-r4-rgb", "ravu-r4"]


@pytest.mark.parametrize("name", NAMES)
def test_tiles_equal_batch_and_sized(name):
    import mpx
    planes = 3 if name.endswith(("-yuv", "-rgb")) else 1
    ctx = mpx.Context(0)
    w, h = 197, 61                       # ragged in both axes for every shape
    xs = _frames(planes, [(w, h)] * 3, 40)
    x = torch.stack(xs)
    plan = ctx.plan(name, w, h, in_dtype=torch.float16, out_dtype=torch.float16, max_batch=3)
    outs = _each_tile(plan, lambda: plan.run(x).clone())
    for t, o in enumerate(outs[1:], 1):
        assert torch.equal(o, outs[0]), (name, "batch", t)
    # frames of different sizes in one launch (per-frame tile prefix per shape)
    sizes = [(197, 61), (64, 36), (130, 97), (17, 9)]
    fs = _frames(planes, sizes, 70)
    splan = ctx.plan(name, 256, 128, in_dtype=torch.float16, out_dtype=torch.float16, max_batch=4)
    souts = _each_tile(splan, lambda: [o.clone() for o in splan.run_frames_sized(fs)])
    for t, os_ in enumerate(souts[1:], 1):
        for i, (a, b) in enumerate(zip(os_, souts[0])):
            assert torch.equal(a, b), (name, "sized", t, sizes[i])


@pytest.mark.parametrize("name,src,dst", [("ravu-zoom-r2", (97, 61), (194, 122)),       # 2x: weight table
                                          ("ravu-zoom-ar-r2-rgb", (97, 61), (150, 100)),  # per-pixel weights
                                          ("ravu-zoom-r3-yuv", (120, 68), (180, 102)),   # 1.5x table
                                          ("ravu-zoom-ar-r2", (64, 48), (200, 131))])
def test_zoom_tiles_equal(name, src, dst):
    import mpx
    planes = 3 if name.endswith(("-yuv", "-rgb")) else 1
    x = torch.stack(_frames(planes, [src] * 2, 90))
    plan = mpx.Context(0).plan(name, src[0], src[1], target=dst, in_dtype=torch.float16,
                               out_dtype=torch.float16, max_batch=2)
    outs = _each_tile(plan, lambda: plan.run(x).clone())
    for t, o in enumerate(outs[1:], 1):
        assert torch.equal(o, outs[0]), (name, t)
