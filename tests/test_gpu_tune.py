"""Launch-shape autotuner: picks a candidate, keeps it in the plan, and the tuned plan
still meets the parity bar."""
import numpy as np
import pytest
import torch

from parity import compare

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nns,wh", [(32, 4), (64, 6), (128, 6)])
def test_tune_nnedi3(nns, wh):
    import mpx
    from mpx.frames import make_plane
    from oracle.nnedi3 import nnedi3
    img = make_plane(160, 96, 7, bits=10).astype(np.float32)
    x = torch.from_numpy(img).cuda()
    plan = mpx.Context(0).plan("nnedi3-nns%d-win8x%d" % (nns, wh), 160, 96, precision="fp16")
    res = plan.tune(x, iters=3)
    assert res["grid_per_cu"] in (0, 1, 2, 3, 4, 6, 8) and res["us"] > 0
    assert res["split"] in ((0, 1) if nns in (64, 128) else (0,))
    f = compare("nnedi3-nns%d-win8x%d fp16 tuned %s" % (nns, wh, res), plan.run(x).cpu().numpy(),
                nnedi3(img.astype(np.float64), nns, wh))
    assert f["max_abs"] <= 2.0 ** -10


def test_tune_other_plans_untouched():
    """NNEDI3 FP32 plans have no tunable shape: tune runs nothing and reports zeros."""
    import mpx
    x = torch.rand((32, 48), device="cuda")
    plan = mpx.Context(0).plan("nnedi3-nns16-win8x4", 48, 32)
    before = plan.run(x).clone()
    assert plan.tune(x) == {"grid_per_cu": 0, "split": 0, "tile": 0, "us": 0.0}
    assert plan.tile() == {"tile": 0, "count": 1, "w": 0, "h": 0}
    assert torch.equal(plan.run(x), before)


@pytest.mark.parametrize("name", ["ravu-lite-ar-r3", "ravu-3x-r3-yuv", "ravu-r3-rgb", "ravu-zoom-ar-r2"])
def test_tune_ravu_tile(name):
    """RAVU plans tune their tile shape: the choice is kept and the result is unchanged."""
    import mpx
    from mpx.frames import make_frame
    planes = 3 if name.endswith(("-yuv", "-rgb")) else 1
    x = torch.from_numpy(make_frame(320, 180, 3, bits=10, channels=planes).astype(np.float16)).cuda()
    plan = mpx.Context(0).plan(name, 320, 180, in_dtype=torch.float16, out_dtype=torch.float16,
                               target=(480, 270) if "zoom" in name else None)
    before = plan.run(x).clone()
    res = plan.tune(x, iters=3)
    t = plan.tile()
    assert res["grid_per_cu"] == 0 and res["split"] == 0 and res["us"] > 0
    assert res["tile"] == t["tile"] and 0 <= t["tile"] < t["count"] and t["count"] >= 2 and t["w"] * t["h"] >= 128
    assert torch.equal(plan.run(x), before)


@pytest.mark.parametrize("nns,wh,split,grid", [(64, 6, 0, 0), (64, 6, 1, 0), (64, 4, 1, 2), (128, 6, 0, 3),
                                               (128, 6, 1, 0)])
def test_set_shape_nnedi3(nns, wh, split, grid):
    """Every neuron split / grid the tuner can pick meets the parity bar when set directly."""
    import mpx
    from mpx.frames import make_plane
    from oracle.nnedi3 import nnedi3
    img = make_plane(136, 72, 11, bits=10).astype(np.float32)
    x = torch.from_numpy(img).cuda()
    plan = mpx.Context(0).plan("nnedi3-nns%d-win8x%d" % (nns, wh), 136, 72, precision="fp16")
    plan.set_shape(grid, split)
    f = compare("nnedi3-nns%d-win8x%d fp16 split %d grid %d" % (nns, wh, split, grid), plan.run(x).cpu().numpy(),
                nnedi3(img.astype(np.float64), nns, wh))
    assert f["max_abs"] <= 2.0 ** -10


def test_set_shape_rejects():
    import mpx
    from mpx._abi import MpxError
    ctx = mpx.Context(0)
    nn32 = ctx.plan("nnedi3-nns32-win8x4", 64, 32, precision="fp16")
    with pytest.raises(MpxError):
        nn32.set_shape(0, 1)          # nns 32 has no neuron split
    with pytest.raises(MpxError):
        nn32.set_shape(65, 0)
    nn32.set_shape(4, 0)
    lite = ctx.plan("ravu-lite-r3", 64, 32)
    lite.set_shape(0, 0)
    with pytest.raises(MpxError):
        lite.set_shape(2, 0)          # no grid / split to set
    with pytest.raises(MpxError):
        lite.set_tile(lite.tile()["count"])
    with pytest.raises(MpxError):
        lite.set_tile(-1)
    with pytest.raises(MpxError):
        nn32.set_tile(1)              # NNEDI3 plans have one tile shape
    nn32.set_tile(0)
