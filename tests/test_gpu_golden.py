"""Every shipped variant on the GPU against the golden vectors made by executing the
reference's own `.hook` text (tests/golden/hooks_golden.npz): RAVU family in FP32
(max-abs <= 2e-5), NNEDI3 plain FP32 (<= 2e-5) and FP16-MFMA (<= 2^-10)."""
import os

import numpy as np
import pytest
import torch

from conftest import ROOT
from parity import compare

pytestmark = pytest.mark.gpu

_G = np.load(os.path.join(ROOT, "tests", "golden", "hooks_golden.npz"))
NAMES = sorted({k.split("/")[0] for k in _G.files})


@pytest.mark.parametrize("name", NAMES)
def test_gpu_matches_reference_text(name):
    import mpx
    img = _G[name + "/in"].astype(np.float32)
    tgt = tuple(int(v) for v in _G[name + "/target"])
    want = _G[name + "/out"]
    x = torch.from_numpy(img).cuda()
    precs = ("fp32", "fp16") if name.startswith("nnedi3") else ("fp32",)
    for prec in precs:
        out = mpx.upscale(name, x, target=tgt, precision=prec).cpu().numpy()
        f = compare("golden %s %s" % (name, prec), out, want)
        assert f["max_abs"] <= (2.0 ** -10 if prec == "fp16" else 2e-5), (name, prec)
