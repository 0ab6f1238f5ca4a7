"""GPU checks of batched multi-view rendering (include/arbor_batch.h): every
view of a batch equals its single-view render bit for bit (Hit, Linear,
Guide, Filtered, History, Pixels); the hipGraph-captured batch equals the
eager batch on every replay, with eager work between replays; the autotuner
changes speed, not results."""
import numpy as np
import pytest

from gpu_util import require_gpu, to_dev, to_host

pytestmark = pytest.mark.gpu

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
        for v in range(7):
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
