"""GPU parity at the edges of the view sizes the C-ABI accepts
(ARBOR_MAX_DIM = 8192 in include/arbor.h; the sizes below and above it are
rejected in test_abi.py):

* tiny and ragged views, whose 8x8 tiles are partial on both axes (down to
  a single pixel): tracePixels against the FP32/FP64 oracle (the trace bar
  of gpu_util.check_hits from 10,000 pixels up; below, every id as the FP32
  oracle's outside near-ties and t within 1e-3 relative), the SAH walk
  (unscheduled and tile-scheduled) bit-identical to the heap walk,
  shadePixels with sun visibility on within 1e-3 of the FP64 run (at most 1%
  of the pixels, and at least one, beyond it where the FP32 oracle is
  within it), filterLighting and History bit-exact, Pixels within one 8-bit
  step, and the engine's five-launch frame equal to the entry points;
* the largest view, 8192x8192 (67 M rays): the SAH walk against the heap
  walk over the whole frame (differences only where two objects are hit at
  nearly the same t, the walk-order-dependent case), 20,000 seeded pixels
  against the oracle, and the size-independent properties of Hit.
"""
import numpy as np
import pytest

from gpu_util import check_hits, near_tie, require_gpu, to_dev, to_host

pytestmark = pytest.mark.gpu

MAX_DIM = 8192  # ARBOR_MAX_DIM


@pytest.fixture(scope="module")
def px():
    require_gpu()
    import pyarbor

    return pyarbor


@pytest.fixture(scope="module")
def scene(px, oracle):
    dev = px.DeviceScene().generate(7)
    return dev, oracle.generate(7), px.TraceAccel(dev)


def _camera(oracle, w, h, view, shadows):
    C = oracle.init_camera(7)
    I = np.zeros(32, np.float32)
    I[5] = view
    oracle.update_camera(C, I, 1.0 / 60.0, w, h)
    C[22] = shadows
    C[13] = 0.610865  # sun at 35 degrees
    return C


def _order_dependent(a, b):
    """Pixels where two Hit buffers differ; each must be two object hits at
    nearly the same t."""
    diff = np.any(a.view(np.uint32) != b.view(np.uint32), axis=1)
    for k in np.nonzero(diff)[0]:
        assert a[k, 1] >= 0 and b[k, 1] >= 0, (k, a[k], b[k])
        assert abs(a[k, 0] - b[k, 0]) <= 1e-3 * a[k, 0], (k, a[k], b[k])
    return int(diff.sum())


@pytest.mark.parametrize("w,h,view", [(1, 1, 1), (7, 5, 3), (9, 17, 1), (333, 187, 3)])
def test_ragged_views(px, oracle, scene, w, h, view):
    print("case: ragged view %dx%d preset %d, sun visibility on" % (w, h, view), flush=True)
    dev, ref, acc = scene
    C = _camera(oracle, w, h, view, 1.0)
    Cd = to_dev(C)
    fb = px.FrameBuffers(w, h)
    px.trace(dev, Cd, fb)
    hit = to_host(fb.Hit)
    h32 = oracle.trace(ref, C, w, h)
    h64 = oracle.trace64(ref, C, w, h)
    if w * h >= 10000:
        check_hits("%dx%d" % (w, h), hit, h32, h64, (oracle, ref, C, w, h, None))
    else:
        # too few pixels for the bar's fractions: every id as the FP32
        # oracle's outside near-ties, t within 1e-3 relative where they agree
        g, r = hit.reshape(-1, 4), h32.reshape(-1, 4)
        same = (g[:, 1] == r[:, 1]) & (g[:, 2] == r[:, 2])
        assert (same | near_tie(h32, h64)).all(), (g, r)
        assert (np.abs(g[same, 0] - r[same, 0]) <= 1e-3 * r[same, 0]).all(), (g, r)
    for sched in (None, px.TileSchedule(w, h)):
        fa = px.FrameBuffers(w, h)
        px.trace_accel(dev, acc, Cd, fa, sched)
        assert np.array_equal(to_host(fa.Hit).view(np.uint32), hit.view(np.uint32)), sched is None
    # shading on the oracle's hits (fused shadow rays), then filter and resolve
    fb.Hit.copy_(to_dev(h32))
    px.shade(dev, Cd, fb)
    L, G = to_host(fb.Linear), to_host(fb.Guide)
    L64, _ = oracle.shade64(ref, C, h32, w, h)
    L32, _ = oracle.shade(ref, C, h32, w, h)
    err = np.abs(L.reshape(-1, 4)[:, :3] - L64.reshape(-1, 4)[:, :3]).max(axis=1)
    err32 = np.abs(L32.reshape(-1, 4)[:, :3] - L64.reshape(-1, 4)[:, :3]).max(axis=1)
    bad = (err > 1e-3) & (err32 <= 1e-3)
    print("shade frac<=1e-3 %.5f (fp32 oracle %.5f), beyond where the oracle is within: %d"
          % ((err <= 1e-3).mean(), (err32 <= 1e-3).mean(), int(bad.sum())))
    assert bad.sum() <= max(1, w * h // 100)
    px.filter_lighting(Cd, fb)
    F = to_host(fb.Filtered)
    assert np.array_equal(F.view(np.uint32), oracle.filter(L, G, h32, C, w, h).view(np.uint32))
    px.resolve(Cd, fb)
    hist = np.zeros_like(F)
    want = oracle.resolve(F, hist, C, w, h)
    pix = to_host(fb.Pixels, np.uint32)
    assert np.abs(pix.view(np.uint8).astype(int) - want.view(np.uint8).astype(int)).max() <= 1
    assert np.array_equal(to_host(fb.History).view(np.uint32), hist.view(np.uint32))


@pytest.mark.parametrize("w,h", [(1, 1), (7, 5), (333, 187)])
def test_engine_ragged_views(px, scene, w, h):
    """The engine's frame (cooperative prologue, SAH trace, visibility,
    shading, filter + resolve) at ragged sizes: its Hit is the heap walk's
    for the engine's own camera, its Linear the shading entry point's."""
    print("case: engine frame %dx%d" % (w, h), flush=True)
    dev, _, _ = scene
    eng = px.Engine(w, h, 7)
    try:
        inp = np.zeros(32, np.float32)
        for f in range(3):
            inp[0] = 7.0 * f
            eng.set_input(inp)
            eng.frame()
        eng.sync()
        C = eng.download("C", 64)
        hit = eng.download("Hit", w * h * 4).reshape(-1, 4)
        lin = eng.download("Linear", w * h * 4)
    finally:
        eng.close()
    fb = px.FrameBuffers(w, h)
    Cd = to_dev(C)
    px.trace(dev, Cd, fb)
    assert _order_dependent(to_host(fb.Hit).reshape(-1, 4), hit) == 0
    px.shade(dev, Cd, fb)
    assert np.array_equal(to_host(fb.Linear).view(np.uint32), lin.view(np.uint32))


def test_max_view(px, oracle, scene):
    print("case: %dx%d walk frame 0, SAH walk vs heap walk" % (MAX_DIM, MAX_DIM), flush=True)
    dev, ref, acc = scene
    W = H = MAX_DIM
    C = oracle.init_camera(7)
    I = px.walk_input(7, 0)
    I[10] = 1.0  # sun visibility off, as the benchmark walk runs
    oracle.update_camera(C, I, 1.0 / 60.0, W, H)
    Cd = to_dev(C)
    fb = px.FrameBuffers(W, H)
    px.trace(dev, Cd, fb)
    heap = to_host(fb.Hit).reshape(-1, 4)
    px.trace_accel(dev, acc, Cd, fb, px.TileSchedule(W, H))
    sah = to_host(fb.Hit).reshape(-1, 4)
    del fb
    n = _order_dependent(heap, sah)
    print("pixels differing:", n, "of", W * H)
    assert n <= 64  # at most 4 per 1080p frame elsewhere; 32x the pixels here
    rng = np.random.default_rng(8192)
    k = 20000
    xy = np.stack([rng.integers(0, W, k), rng.integers(0, H, k)], axis=1).astype(np.int32)
    gpu = sah[xy[:, 1] * W + xy[:, 0]].reshape(-1)
    h32 = oracle.trace_pixels(ref, C, xy, W, H)
    h64 = oracle.trace_pixels(ref, C, xy, W, H, fp64=True)
    check_hits("%d^2 (sampled)" % W, gpu, h32, h64, (oracle, ref, C, W, H, xy))
    t, ids, tree = sah[:, 0], sah[:, 1], sah[:, 2]
    assert (t > 0).all() and (t <= np.float32(1e5)).all()
    assert ((ids == -1) | ((ids >= 0) & (ids < 29412))).all()
    assert np.array_equal(ids >= 0, tree >= 0)
    w_want = (np.float32(C[14]) * t) / np.float32(H)
    assert np.array_equal(sah[:, 3].view(np.uint32), w_want.astype(np.float32).view(np.uint32))
