"""Serving loop (continuous batching) on the GPU: every result equals the same request
run alone, batching through mpx_run_frames matches per-frame runs, and admission keeps
working under a budget smaller than the offered work."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

MIX = [("ravu-lite-ar-r3", 1, (160, 90)), ("ravu-r4-yuv", 3, (128, 72)), ("ravu-3x-r3", 1, (96, 64)),
       ("nnedi3-nns64-win8x6", 1, (160, 90)), ("ravu-lite-ar-r3", 1, (200, 120)), ("ravu-zoom-ar-r2-rgb", 3, (96, 54))]


def _frame(w, h, planes, seed):
    from mpx.frames import make_frame
    f = make_frame(w, h, seed, bits=10, channels=planes).astype(np.float16)
    return torch.from_numpy(f).cuda()


def _alone(name, x, prec, target=None):
    import mpx
    return mpx.upscale(name, x, target=target, precision=prec)


def test_run_frames_equals_single_runs():
    import mpx
    frames = [_frame(128, 72, 1, s) for s in range(5)]
    plan = mpx.default_context().plan("nnedi3-nns32-win8x4", 128, 72, in_dtype=torch.float16,
                                      out_dtype=torch.float16, precision="fp16", max_batch=8)
    outs = plan.run_frames(frames)
    for f, o in zip(frames, outs):
        assert torch.equal(o, plan.run(f))
    plan3 = mpx.default_context().plan("ravu-r3-rgb", 64, 48, in_dtype=torch.float16, out_dtype=torch.float16,
                                       max_batch=4)
    fr3 = [_frame(64, 48, 3, 10 + s) for s in range(3)]
    for f, o in zip(fr3, plan3.run_frames(fr3)):
        assert torch.equal(o, plan3.run(f))


def test_mix_results_equal_alone():
    from mpx.serving import Server
    srv = Server(device=0, hbm_budget=4 << 30, streams=3, max_batch=4)
    try:
        srv.pause()   # queue the first burst so the scheduler sees several per plan
        reqs = []
        for i in range(40):
            name, planes, (w, h) = MIX[i % len(MIX)]
            prec = "fp16" if name.startswith("nnedi3") else "fp32"
            tgt = (w * 3 // 2, h * 3 // 2) if "zoom" in name else None
            x = _frame(w, h, planes, 100 + i)
            reqs.append((name, x, prec, tgt, srv.submit(name, x, target=tgt, precision=prec)))
            if i == 23:
                srv.resume()
        for name, x, prec, tgt, r in reqs:
            got = r.wait(timeout=60)
            assert r.latency is not None and r.latency > 0
            assert torch.equal(got, _alone(name, x, prec, tgt)), name
        assert srv.launches < len(reqs)  # requests were batched
    finally:
        srv.close()


def test_admission_under_small_budget():
    """One envelope: the context gets hbm_budget - admit_bytes, outputs of in-flight
    requests may not exceed admit_bytes (requests queue and all complete), and a request
    whose output alone exceeds admit_bytes fails without stopping the loop."""
    from mpx.serving import Server
    one = 2 * 2 * 160 * 2 * 90 * 2  # one 320x180 fp16 output ... twice
    srv = Server(device=0, hbm_budget=2 << 30, admit_bytes=one, streams=2, max_batch=8)
    try:
        assert srv.ctx_budget + srv.admit_bytes == 2 << 30
        rs = [srv.submit("ravu-lite-r3", _frame(160, 90, 1, s)) for s in range(12)]
        big = srv.submit("ravu-lite-r3", _frame(320, 180, 1, 99))      # 4x one output
        after = srv.submit("ravu-lite-r3", _frame(160, 90, 1, 100))
        for r in rs:
            r.wait(timeout=60)
        with pytest.raises(ValueError):
            big.wait(timeout=60)
        assert after.wait(timeout=60).shape == (180, 320)
        assert srv._inflight == 0
        assert 0 < srv.ctx.bytes() <= srv.ctx_budget
    finally:
        srv.close()
    with pytest.raises(ValueError):
        Server(device=0, hbm_budget=1 << 20, admit_bytes=1 << 20)


def test_bad_request_fails_alone():
    """A request that cannot run (weights not shipped) fails without stopping the loop."""
    from mpx.serving import Server
    srv = Server(device=0, hbm_budget=1 << 30, streams=1)
    try:
        bad = srv.submit("ravu-zoom-ar-r3", _frame(32, 32, 1, 1), target=(48, 48))
        good = srv.submit("ravu-lite-r2", _frame(32, 32, 1, 2))
        with pytest.raises(Exception):
            bad.wait(timeout=30)
        assert good.wait(timeout=30).shape == (64, 64)
    finally:
        srv.close()


def test_plan_eviction_under_context_budget():
    """A context budget that holds one plan's scratch at a time: plans of other sizes are
    made by evicting idle ones, and every request completes with the result it has alone."""
    from mpx.serving import Server
    import mpx
    sizes = [(256, 256), (200, 200), (160, 160), (256, 256)]
    budget = 64 * 1024 + 256 * 256 * 4 + 32 * 1024   # weights + the largest int11 scratch + slack
    srv = Server(device=0, hbm_budget=budget + (1 << 30), admit_bytes=1 << 30, streams=1, max_batch=1,
                 mixed_sizes=None)
    assert srv.ctx_budget == budget
    try:
        frames = [_frame(w, h, 1, 30 + i) for i, (w, h) in enumerate(sizes)]
        rs = [srv.submit("ravu-r2", f) for f in frames]
        for f, r in zip(frames, rs):
            assert torch.equal(r.wait(timeout=60), mpx.upscale("ravu-r2", f))
    finally:
        srv.close()


def test_mixed_sizes_share_launches():
    """Frames of different sizes of one variant go out in shared launches (one tile queue
    over all of them) and each equals its result alone."""
    from mpx.serving import Server
    import mpx
    sizes = [(160, 90), (128, 72), (200, 120), (64, 36), (96, 54), (150, 100)]
    srv = Server(device=0, hbm_budget=4 << 30, streams=1, max_batch=8, mixed_sizes=(256, 256))
    try:
        srv.pause()
        frames = [_frame(w, h, 1, 70 + i) for i, (w, h) in enumerate(sizes)]
        rs = [srv.submit("ravu-lite-ar-r3", f) for f in frames]
        before = srv.launches
        srv.resume()
        for f, r in zip(frames, rs):
            assert torch.equal(r.wait(timeout=60), mpx.upscale("ravu-lite-ar-r3", f))
        assert srv.launches - before == 1
    finally:
        srv.close()

