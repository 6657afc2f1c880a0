"""Frames of different sizes in one launch (mpx_run_frames_sized): every frame equals the
same variant run on that frame alone, bit for bit."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SIZES = [(200, 120), (37, 23), (128, 64), (5, 3), (199, 119), (64, 4)]


def _frames(planes, seed):
    from mpx.frames import make_frame
    return [torch.from_numpy(make_frame(w, h, seed + i, bits=10, channels=planes).astype(np.float16)).cuda()
            for i, (w, h) in enumerate(SIZES)]


@pytest.mark.parametrize("name,prec", [("nnedi3-nns32-win8x4", "fp16"), ("nnedi3-nns64-win8x6", "fp32"),
                                       ("nnedi3-nns128-win8x6", "fp16"), ("ravu-lite-ar-r3", "fp32"),
                                       ("ravu-r3-yuv", "fp32"), ("ravu-r4", "fp32"), ("ravu-3x-r3-rgb", "fp32"),
                                       ("ravu-3x-r2", "fp32")])
def test_sized_launch_equals_single_frames(name, prec):
    import mpx
    planes = 3 if name.endswith(("-yuv", "-rgb")) else 1
    frames = _frames(planes, 40)
    plan = mpx.Context(0).plan(name, 200, 120, in_dtype=torch.float16, out_dtype=torch.float16, precision=prec,
                               max_batch=len(frames))
    outs = plan.run_frames_sized(frames)
    for (w, h), x, o in zip(SIZES, frames, outs):
        alone = mpx.upscale(name, x, precision=prec)
        assert o.shape == alone.shape, (w, h)
        assert torch.equal(o, alone), (name, w, h)


def test_sized_launch_rejections():
    import mpx
    from mpx import MpxError
    ctx = mpx.Context(0)
    zoom = ctx.plan("ravu-zoom-r2", 64, 48, target=(96, 72), max_batch=2)
    x = torch.rand((48, 64), device="cuda")
    with pytest.raises(MpxError):
        zoom.run_frames_sized([x])
    lite = ctx.plan("ravu-lite-r2", 64, 48, max_batch=2)
    with pytest.raises(MpxError):
        lite.run_frames_sized([torch.rand((49, 64), device="cuda")])   # taller than the plan


@pytest.mark.parametrize("name,planes", [("ravu-r3-yuv", 3), ("ravu-lite-r3", 1), ("nnedi3-nns32-win8x4", 1)])
def test_frame_planes_checked_before_launch(name, planes):
    """run_frames / run_frames_sized / Server.submit reject a frame whose planes do not
    match the plan (a luma frame sent to a 3-plane plan would make the kernels read and
    write two planes past the tensors): ValueError, and nothing is launched."""
    import mpx
    from mpx.serving import Server
    ctx = mpx.Context(0)
    plan = ctx.plan(name, 64, 48, in_dtype=torch.float16, out_dtype=torch.float16, max_batch=4)
    wrong = [torch.rand((48, 64), device="cuda", dtype=torch.float16)] if planes == 3 else \
        [torch.rand((3, 48, 64), device="cuda", dtype=torch.float16), torch.rand((2, 48, 64), device="cuda",
                                                                                  dtype=torch.float16)]
    for x in wrong:
        with pytest.raises(ValueError):
            plan.run_frames([x])
        with pytest.raises(ValueError):
            plan.run_frames_sized([x])
    srv = Server(device=0, hbm_budget=1 << 30, streams=1, max_batch=2)
    try:
        for x in wrong:
            with pytest.raises(ValueError):
                srv.submit(name, x)
        assert srv.launches == 0
    finally:
        srv.close()


def test_plan_refuses_another_stream_while_its_intermediates_are_busy():
    """A plan with intermediates (NNEDI3 doubling both axes) runs long work on stream A; a
    run of the same plan on stream B before A finishes is refused (MPX_ERR_ARG) without
    enqueuing anything; after A is waited for, B runs and equals A's result."""
    import mpx
    from mpx.frames import make_plane
    w, h, b = 1920, 1080, 4
    ctx = mpx.Context(0)
    plan = ctx.plan("nnedi3-nns256-win8x6", w, h, in_dtype=torch.float16, out_dtype=torch.float32,
                    precision="fp32", max_batch=b)
    x = torch.from_numpy(np.stack([make_plane(w, h, seed=70 + i, bits=10) for i in range(b)])
                         .astype(np.float16)).cuda()
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    oa = torch.empty((b, 2 * h, 2 * w), device="cuda")
    ob = torch.full((1, 2 * h, 2 * w), -1.0, device="cuda")
    torch.cuda.synchronize()
    print("long FP32 nns256 run on stream A, then the same plan on stream B", flush=True)
    plan.run(x, out=oa, stream=sa)            # ~17 ms of work: still in flight below
    with pytest.raises(mpx.MpxError) as ei:
        plan.run(x[:1], out=ob, stream=sb)
    assert ei.value.code == -1 and "another stream" in str(ei.value)
    sa.synchronize()
    sb.synchronize()
    assert bool((ob == -1.0).all())           # the refused run enqueued nothing
    plan.run(x[:1], out=ob, stream=sb)        # A is done: B may take the plan
    sb.synchronize()
    assert torch.equal(ob[0], oa[0])
    plan.close()
    ctx.close()


def test_ravu_plan_refuses_another_stream_while_its_int11_plane_is_busy():
    """Full RAVU keeps its int11 plane in the plan's scratch, and its last kernel (steps 2
    and 3) completes the plan's event.  With slow FP32 NNEDI3 work queued ahead of it on
    stream A, the RAVU run on A is still in flight when the same plan is called on stream B:
    refused, nothing enqueued; after A finishes, B runs and equals A's result."""
    import mpx
    from mpx.frames import make_frame, make_plane
    w, h = 1920, 1080
    ctx = mpx.Context(0)
    slow = ctx.plan("nnedi3-nns256-win8x6", w, h, in_dtype=torch.float16, out_dtype=torch.float32,
                    precision="fp32", max_batch=2)
    plan = ctx.plan("ravu-r4-yuv", w, h, in_dtype=torch.float16, out_dtype=torch.float32, max_batch=2)
    xs = torch.from_numpy(np.stack([make_plane(w, h, seed=80 + i, bits=10) for i in range(2)])
                          .astype(np.float16)).cuda()
    os_ = torch.empty((2, 2 * h, 2 * w), device="cuda")
    x = torch.from_numpy(np.stack([make_frame(w, h, 90 + i, bits=10, channels=3) for i in range(2)])
                         .astype(np.float16)).cuda()
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    oa = torch.empty((2, 3, 2 * h, 2 * w), device="cuda")
    ob = torch.full((1, 3, 2 * h, 2 * w), -1.0, device="cuda")
    torch.cuda.synchronize()
    print("FP32 nns256 then ravu-r4-yuv on stream A, then the RAVU plan on stream B", flush=True)
    slow.run(xs, out=os_, stream=sa)         # ~8 ms of work ahead of the RAVU run on A
    plan.run(x, out=oa, stream=sa)
    with pytest.raises(mpx.MpxError) as ei:
        plan.run(x[:1], out=ob, stream=sb)
    assert ei.value.code == -1 and "another stream" in str(ei.value)
    sa.synchronize()
    sb.synchronize()
    assert bool((ob == -1.0).all())           # the refused run enqueued nothing
    plan.run(x[:1], out=ob, stream=sb)        # A is done: B may take the plan
    sb.synchronize()
    assert torch.equal(ob[0], oa[0])
    plan.close()
    slow.close()
    ctx.close()


def test_frame_span_limits_refused_before_launch():
    """The staging loaders address samples with 32-bit element offsets: mpx_run refuses a
    pitch of 2^24 elements or an input frame spanning more than 2^32 elements with
    MPX_ERR_SHAPE before anything is launched (the buffers here are far smaller than the
    strides claim, so a launch would read out of bounds)."""
    import ctypes
    import mpx
    from mpx._abi import LIB
    ctx = mpx.Context(0)
    x = torch.zeros((600, 64), device="cuda")
    o = torch.zeros((1200, 128), device="cuda")
    plan = ctx.plan("ravu-lite-r2", 64, 600)
    for pitch in (1 << 24, 1 << 23):     # pitch limit; 600 rows x 2^23 > 2^32 elements
        rc = LIB.mpx_run(plan._h, ctypes.c_void_p(x.data_ptr()), pitch, pitch * 600, ctypes.c_void_p(o.data_ptr()),
                         128, 128 * 1200, 1, None)
        assert rc == -4, (pitch, rc)
    torch.cuda.synchronize()
    assert torch.equal(plan.run(x), mpx.upscale("ravu-lite-r2", x))
