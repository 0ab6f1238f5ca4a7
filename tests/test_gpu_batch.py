"""GPU checks of batched multi-view rendering (include/arbor_batch.h): every
view of a batch equals its single-view render bit for bit (Hit, Linear,
Guide, Filtered, History, Pixels); the hipGraph-captured batch equals the
eager batch on every replay, with eager work between replays; the autotuner
changes speed, not results."""
import numpy as np
import pytest

from gpu_util import require_gpu, to_dev, to_host

pytestmark = pytest.mark.gpu

# batch_variant 0-4: fused trace + shade; 5-9: split (7, 8 and 9 read the
# tree records from device memory over the SAH hierarchy; 9 takes three
# interior steps per gathered round)
BATCH_VARIANTS = 10

SIZES = [(256, 256), (320, 180), (200, 120), (64, 64), (333, 97), (128, 72), (96, 200), (40, 24)]


@pytest.fixture(scope="module")
def px():
    require_gpu()
    import pyarbor

    return pyarbor


@pytest.fixture(scope="module")
def scene(px):
    return px.DeviceScene().generate(7)


def _views(px, n):
    cams, _ = px.batch_walk_views(3, 7, n)
    return cams, SIZES[:n]


def _single(px, scene, C, w, h):
    fb = px.FrameBuffers(w, h)
    Cd = to_dev(C)
    px.trace(scene, Cd, fb)
    px.shade(scene, Cd, fb)
    px.filter_lighting(Cd, fb)
    px.resolve(Cd, fb)
    return {k: to_host(getattr(fb, k)) for k in ("Hit", "Linear", "Guide", "Filtered", "History")} | {
        "Pixels": to_host(fb.Pixels, np.uint32)}


def _check_batch(px, b, refs, tag):
    for i, ref in enumerate(refs):
        for name, want in ref.items():
            got = b.download(i, name)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (tag, i, name)


@pytest.mark.parametrize("shadows", [0, 1])
def test_batch_equals_single_views(px, scene, shadows):
    """Every batch_variant (fused trace+shade, or split trace then shade), with
    sun visibility off and on (four views each), against single-view renders."""
    print("case: batch of 8 uneven views vs single-view renders, shadows", shadows, flush=True)
    cams, sizes = _views(px, 8)
    cams = cams.copy()
    cams[:4, 22] = float(shadows)
    refs = [_single(px, scene, cams[i], w, h) for i, (w, h) in enumerate(sizes)]
    b = px.Batch(16, sum(w * h for w, h in sizes) + 1000)
    b.set_views(cams, sizes)
    assert b.pixels() == sum(w * h for w, h in sizes)
    old = px.tune_get("batch_variant")
    try:
        for v in range(BATCH_VARIANTS):
            px.tune_set("batch_variant", v)
            b.set_views(cams, sizes)
            b.render(scene)
            _check_batch(px, b, refs, "variant %d" % v)
    finally:
        px.tune_set("batch_variant", old)
    b.close()


def test_batch_graph_replays(px, scene):
    print("case: hipGraph-captured batch, 3 replays with eager work between", flush=True)
    import torch

    cams, sizes = _views(px, 5)
    refs = [_single(px, scene, cams[i], w, h) for i, (w, h) in enumerate(sizes)]
    b = px.Batch(8, sum(w * h for w, h in sizes))
    b.set_views(cams, sizes)
    stream = torch.cuda.Stream()
    b.capture(scene, stream)
    for r in range(3):
        b.launch(stream)
        stream.synchronize()
        _check_batch(px, b, refs, "replay %d" % r)
        _single(px, scene, cams[0], 48, 40)  # unrelated eager work between replays
    b.close()


def test_batch_autotune_keeps_results(px, scene):
    print("case: batch autotuner", flush=True)
    cams, sizes = _views(px, 4)
    b = px.Batch(4, sum(w * h for w, h in sizes))
    b.set_views(cams, sizes)
    b.render(scene)
    before = [b.download(i, "Hit") for i in range(4)]
    old = px.tune_get("batch_variant")
    res = b.autotune(scene, reps=1)
    print("autotune:", res)
    assert 0 <= res["best"] < len(res["ms"]) and px.tune_get("batch_variant") == res["best"]
    b.render(scene)
    for i in range(4):
        assert np.array_equal(b.download(i, "Hit").view(np.uint32), before[i].view(np.uint32))
    px.tune_set("batch_variant", old)
    b.close()


def test_trace_autotune_keeps_results(px, scene, oracle):
    print("case: single-view trace autotuner", flush=True)
    C = oracle.init_camera(7)
    fb = px.FrameBuffers(160, 96)
    Cd = to_dev(C)
    px.trace(scene, Cd, fb)
    before = to_host(fb.Hit)
    old = px.tune_get("trace_variant")
    res = px.autotune_trace(scene, Cd, fb, reps=1)
    print("autotune:", res)
    px.trace(scene, Cd, fb)
    assert np.array_equal(to_host(fb.Hit).view(np.uint32), before.view(np.uint32))
    px.tune_set("trace_variant", old)


def test_batched_views_workload_full_size(px, scene, oracle):
    """The batched-views workload at its size: 64 cameras at seeded positions
    inside the 324-tree forest, resolutions drawn from {256x256, 640x360,
    1280x720, 1920x1080, 2560x1440} (98,180,096 pixels), 1 spp, sun visibility
    off, over the SAH hierarchy, rendered by two replays of one captured
    hipGraph. Every view's Hit, Linear, Guide, Filtered, History and Pixels
    equal its single-view render (arbor_tracePixelsAccel, shadePixels,
    filterLighting, resolvePixels) bit for bit; 20,000 sampled pixels of a
    1920x1080 view and of a 2560x1440 view meet the trace bar against the
    FP32 and FP64 oracle."""
    import torch

    from gpu_util import check_hits

    print("case: 64 batched views at native resolutions", flush=True)
    cams, sizes = px.batch_walk_views(3, 7, 64)
    total = sum(w * h for w, h in sizes)
    assert total == 98_180_096 and len(set(sizes)) == 5, (total, set(sizes))
    accel = px.TraceAccel(scene)
    b = px.Batch(64, total)
    b.set_views(cams, sizes)
    b.set_accel(accel)
    stream = torch.cuda.Stream()
    b.capture(scene, stream)
    for _ in range(2):
        b.launch(stream)
    stream.synchronize()
    lib = px.lib()
    fbs = {}
    per = {"Hit": 16, "Linear": 16, "Guide": 16, "Filtered": 16, "History": 16, "Pixels": 4}
    for i, (w, h) in enumerate(sizes):
        fb = fbs.setdefault((w, h), px.FrameBuffers(w, h))
        Cd = to_dev(cams[i])
        fb.History.zero_()
        px.trace_accel(scene, accel, Cd, fb)
        px.shade(scene, Cd, fb)
        px.filter_lighting(Cd, fb)
        px.resolve(Cd, fb)
        torch.cuda.synchronize()
        for name, bpp in per.items():
            ptr = lib.arbor_batch_buffer(b.h, i, name.encode())
            got = px._wrap_device(ptr, w * h * bpp)
            want = getattr(fb, name).view(torch.uint8).reshape(-1)
            assert torch.equal(got, want), (i, (w, h), name)
    print("64 views equal their single-view renders;", len(fbs), "sizes")
    ref = oracle.generate(7)
    for size in ((1920, 1080), (2560, 1440)):
        i = next(k for k, s in enumerate(sizes) if s == size)
        W, H = size
        hit = b.download(i, "Hit").reshape(-1, 4)
        rng = np.random.default_rng(4000 + W)
        n = 20000
        xy = np.stack([rng.integers(0, W, n), rng.integers(0, H, n)], axis=1).astype(np.int32)
        gpu = hit[xy[:, 1] * W + xy[:, 0]].reshape(-1)
        C = cams[i]
        h32 = oracle.trace_pixels(ref, C, xy, W, H)
        h64 = oracle.trace_pixels(ref, C, xy, W, H, fp64=True)
        check_hits("batch view %d %dx%d (sampled)" % (i, W, H), gpu, h32, h64, (oracle, ref, C, W, H, xy))
    b.close()


def test_batch_rejects_mixed_worlds_and_launch_count(px, scene):
    """All views of a batch see one forest: a view whose instancing seed C[31]
    differs is rejected. A batched render (default split variant) is four
    kernel launches whatever the number of views: trace, sun visibility,
    shading, and filter + resolve over every view's pixels (profiler count)."""
    import torch
    from torch.profiler import ProfilerActivity, profile

    print("case: batch world-seed check and launch count", flush=True)
    cams, sizes = _views(px, 8)
    b = px.Batch(8, sum(w * h for w, h in sizes))
    bad = cams.copy()
    bad[5, 31] = 8.0
    with pytest.raises(px.ArborError):
        b.set_views(bad, sizes)
    b.set_views(cams, sizes)
    b.render(scene)
    torch.cuda.synchronize()
    with profile(activities=[ProfilerActivity.CUDA]) as prof:
        b.render(scene)
        torch.cuda.synchronize()
    names = [e.name for e in prof.events() if e.device_type == torch.autograd.DeviceType.CUDA
             and "memcpy" not in e.name.lower() and "memset" not in e.name.lower()]
    print(len(names), "kernels:", names)
    assert px.tune_get("batch_variant") in (5, 6, 7, 8, 9) and len(names) == 4
    # over the SAH hierarchy with the default variant: the record-tree trace
    old = px.tune_get("batch_variant")
    try:
        px.tune_set("batch_variant", 9)
        acc = px.TraceAccel(scene)  # held until the batch is closed
        b.set_accel(acc)
        b.render(scene)
        torch.cuda.synchronize()
        with profile(activities=[ProfilerActivity.CUDA]) as prof:
            b.render(scene)
            torch.cuda.synchronize()
        names = [e.name for e in prof.events() if e.device_type == torch.autograd.DeviceType.CUDA
                 and "memcpy" not in e.name.lower() and "memset" not in e.name.lower()]
        print(len(names), "kernels over the hierarchy:", names)
        assert len(names) == 4 and "k_traceViewsRec" in names[0] and "k_sunVisibilityViewsRec" in names[1]
    finally:
        px.tune_set("batch_variant", old)
    b.close()
