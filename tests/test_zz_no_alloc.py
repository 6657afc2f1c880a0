"""mpx_run_static (intermediates in the library's static per-device array) against
mpx_run_ex (intermediates in the plan's scratch): bit-identical outputs on two calls in a
row, the second smaller than the first, for every family with intermediates (NNEDI3 with
both axes doubled, RAVU full luma and yuv) and one without (ravu-lite); and the size
limit rejected before anything is enqueued."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

# (variant, precision, [(w, h, batch) first call, second call (smaller)])
CASES = [
    ("nnedi3-nns32-win8x4", "fp16", [(200, 120, 3), (96, 64, 1)]),
    ("nnedi3-nns64-win8x6", "fp32", [(160, 90, 2), (48, 40, 1)]),
    ("ravu-r3-yuv", "fp32", [(192, 128, 2), (80, 48, 1)]),
    ("ravu-r4", "fp32", [(256, 144, 2), (64, 40, 1)]),
    ("ravu-lite-ar-r3", "fp32", [(128, 96, 2), (40, 24, 1)]),
]


def _frames(name, w, h, batch, seed):
    from mpx.frames import make_frame
    planes = 3 if name.endswith("-yuv") else 1
    fr = [make_frame(w, h, seed + i, bits=10, channels=planes) for i in range(batch)]
    return torch.from_numpy(np.stack(fr).astype(np.float16)).cuda()


@pytest.mark.parametrize("name,prec,calls", CASES, ids=[c[0] + "-" + c[1] for c in CASES])
def test_run_static_equals_run_ex(name, prec, calls):
    import mpx
    from mpx import _abi
    ctx = mpx.Context(0)
    for k, (w, h, batch) in enumerate(calls):
        print("mpx_run_static %s %s %dx%d batch %d (call %d)" % (name, prec, w, h, batch, k + 1), flush=True)
        plan = ctx.plan(name, w, h, in_dtype=torch.float16, out_dtype=torch.float32, precision=prec,
                        max_batch=batch)
        x = _frames(name, w, h, batch, 1000 * k + 7)
        ref = plan.run(x)
        out = torch.full_like(ref, float("nan"))
        io, _, _, _ = plan._io(x, out, False)
        stream = torch.cuda.current_stream()
        _abi.check(_abi.LIB.mpx_run_static(plan._h, ctypes.byref(io), ctypes.c_void_p(stream.cuda_stream)))
        torch.cuda.synchronize()
        assert torch.equal(out, ref), (name, w, h, batch)
        plan.close()
    ctx.close()


def test_run_static_rejects_oversize_batch():
    import mpx
    from mpx import _abi
    ctx = mpx.Context(0)
    # one 1920x1080 NNEDI3 2x frame needs 1920*1080*2*4 B = 16.6 MB of intermediates: 4 fit
    # in the 64 MiB array, 5 do not
    w, h = 1920, 1080
    need = w * h * 2 * 4
    assert 4 * need <= _abi.STATIC_SCRATCH_BYTES < 5 * need
    plan = ctx.plan("nnedi3-nns32-win8x4", w, h, in_dtype=torch.float16, out_dtype=torch.float32,
                    precision="fp16", max_batch=5)
    x = _frames("nnedi3-nns32-win8x4", w, h, 5, 11)
    out = torch.full((5, 2 * h, 2 * w), -1.0, device="cuda")
    io, _, _, _ = plan._io(x, out, False)
    stream = torch.cuda.current_stream()
    rc = _abi.LIB.mpx_run_static(plan._h, ctypes.byref(io), ctypes.c_void_p(stream.cuda_stream))
    assert rc == -4, rc
    assert b"static scratch" in _abi.LIB.mpx_last_error()
    torch.cuda.synchronize()
    assert bool((out == -1.0).all())          # nothing was enqueued
    io.batch = 4                              # the largest batch that fits runs
    _abi.check(_abi.LIB.mpx_run_static(plan._h, ctypes.byref(io), ctypes.c_void_p(stream.cuda_stream)))
    ref = plan.run(x[:4])
    torch.cuda.synchronize()
    assert torch.equal(out[:4], ref)
    assert bool((out[4] == -1.0).all())
    plan.close()
    ctx.close()
