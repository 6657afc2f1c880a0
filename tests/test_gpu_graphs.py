"""HIP-graph variant chains: every replay equals the eager runs of the same steps, with
eager work and synchronisation between replays and new input contents each time."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _planes(w, h, n, seed):
    from mpx.frames import make_frame
    return torch.from_numpy(make_frame(w, h, seed, bits=10, channels=n).astype(np.float16)).cuda()


def test_chain_replays_match_eager():
    import mpx
    from mpx.graphs import Chain
    ctx = mpx.default_context()
    specs = [("nnedi3-nns64-win8x6", 1, (96, 64), None, "fp16"), ("ravu-lite-ar-r3", 1, (80, 48), None, "fp32"),
             ("ravu-zoom-ar-r2-rgb", 3, (64, 40), (96, 60), "fp32"), ("ravu-r3-yuv", 3, (48, 32), None, "fp32"),
             ("ravu-3x-r2", 1, (40, 24), None, "fp32")]
    steps = []
    for name, planes, (w, h), tgt, prec in specs:
        x = _planes(w, h, planes, 1)
        plan = ctx.plan(name, w, h, tgt, in_dtype=torch.float16, out_dtype=torch.float16, precision=prec)
        steps.append((plan, x, torch.empty(plan.out_shape(x), dtype=torch.float16, device="cuda")))
    chain = Chain(steps)
    for rep in range(4):
        for i, (plan, x, o) in enumerate(steps):   # new contents in place
            x.copy_(_planes(x.shape[-1], x.shape[-2], 3 if x.dim() == 3 else 1, 10 * rep + i + 2))
        torch.cuda.synchronize()
        chain.replay()
        torch.cuda.synchronize()
        got = [o.clone() for o in chain.outputs()]
        eager = [plan.run(x) for plan, x, _ in steps]
        torch.cuda.synchronize()
        for (name, *_), g, e in zip(specs, got, eager):
            assert torch.equal(g, e), (rep, name)
