"""The SAH trace hierarchy (arbor_buildTraceAccel, built on the device) and
the trace and sun-visibility passes over it.

* Build: the device builder's words equal the oracle's host builder's (the
  same binned SAH, costs, partition and median fallback) bit for bit on four
  scenes and on leaf sets built to defeat the heuristic, at the walk's depth
  bound 32 and at lower bounds (where the median splits take over).
* Structure: deterministic bytes; every primitive of the heap BVH appears once
  with its heap leaf box; every interior child box is the exact union of its
  children's; depth within the walk's 32-entry stack; the slot table maps
  each primitive to its heap leaf (the tie order).
* Results: the Accel passes equal the heap passes (arbor_tracePixels,
  arbor_sunVisibility) bit for bit on whole 1920x1080 walk frames, except for
  rays whose reported hit depends on the walk order in the reference itself:
  a swaying leaf hit before its rest-pose leaf box is entered. Such pixels
  carry two valid object hits of the reference's tests at nearly the same t;
  at most 4 per frame are allowed (measured: 0 or 1 per 2 million rays).
* Launch shapes (accel_variant, accel_vis_variant) and the tile schedule
  change nothing.
"""
import numpy as np
import pytest

# accel_variant 0-9 and 13-31: tile-per-wave SAH kernels (13, 14 and 16-31 read
# the tree records from the schedule); 10-12: the persistent form; 32-34: the
# stage-split form (start, queue rounds, finish)
ACCEL_VARIANTS = 35
# accel_vis_variant 0-21 (10-21 read the tree records from the schedule)
ACCEL_VIS_VARIANTS = 28

from accel_check import LeafSet, check_accel_structure, skewed_leaf_set
from gpu_util import require_gpu, to_dev, to_host

pytestmark = pytest.mark.gpu



@pytest.fixture(scope="module")
def px():
    require_gpu()
    import pyarbor

    return pyarbor


@pytest.fixture(scope="module")
def scene(px):
    dev = px.DeviceScene().generate(7)
    return dev, px.TraceAccel(dev)


def _walk_camera(px, oracle, frame, w, h, shadows):
    C = oracle.init_camera(7)
    for f in range(frame + 1):
        I = px.walk_input(7, f)
        if f == 0:
            I[10] = 1.0
        oracle.update_camera(C, I, 1.0 / 60.0, w, h)
    C[22] = shadows
    return C


def test_accel_structure(px, scene):
    print("case: accel structure, seed 7", flush=True)
    dev, acc = scene
    words = acc.words()
    again = px.TraceAccel(dev).words()
    assert np.array_equal(words, again), "the build is not deterministic"
    check_accel_structure(words, to_host(dev.Nodes), to_host(dev.Order))


def _host_leaf_set(dev):
    s = LeafSet()
    s.Nodes = to_host(dev.Nodes)
    s.Order = to_host(dev.Order).view(np.uint32)
    return s


@pytest.mark.parametrize("seed", [7, 1, 42, 1788])
def test_device_build_equals_oracle_builder(px, oracle, seed):
    """The device build of a scene's hierarchy equals the oracle's host
    build word for word, at depth bounds 32 (arbor_buildTraceAccel) and 15
    (arbor_buildTraceAccelLimit: median splits over the upper levels)."""
    dev = px.DeviceScene().generate(seed)
    host = _host_leaf_set(dev)
    for limit in (None, 15):
        print("case: device SAH build, seed", seed, "depth limit", limit or 32, flush=True)
        got = px.TraceAccel(dev, depth_limit=limit).words()
        want = oracle.build_sah(host, depth_limit=limit)
        bad = np.nonzero(got != want)[0]
        assert bad.size == 0, (limit, bad[:8], got[bad[:8]], want[bad[:8]])
        depth = check_accel_structure(got, host.Nodes, host.Order)
        assert depth <= (limit or 32)


@pytest.mark.parametrize("shape", ["identical", "exponential", "nested", "full"])
def test_device_build_on_skewed_leaf_sets(px, oracle, shape):
    """Leaf sets built to defeat the heuristic (identical boxes, centres
    spaced exponentially, nested boxes, every slot used): the device build
    equals the oracle's at depth bounds 32 and 16 and keeps every leaf
    within the bound."""
    s = skewed_leaf_set(shape)

    class Dev:
        Nodes = to_dev(s.Nodes)
        Order = to_dev(s.Order)

    for limit in (None, 16):
        print("case: device SAH build,", shape, "leaf set, depth limit", limit or 32, flush=True)
        got = px.TraceAccel(Dev, depth_limit=limit).words()
        want = oracle.build_sah(s, depth_limit=limit)
        bad = np.nonzero(got != want)[0]
        assert bad.size == 0, (limit, bad[:8], got[bad[:8]], want[bad[:8]])
        assert check_accel_structure(got, s.Nodes, s.Order) <= (limit or 32)


@pytest.mark.parametrize("seed", [1, 42, 1788])
def test_accel_other_scenes(px, oracle, seed):
    """Other scenes: the hierarchy is valid and the Accel trace and sun
    visibility equal the heap passes on a 640x360 view of the scene's own
    forest (walk frame 0 of its seed; shadows on for the visibility pass)."""
    print("case: accel seed", seed, "640x360", flush=True)
    import torch

    dev = px.DeviceScene().generate(seed)
    acc = px.TraceAccel(dev)
    check_accel_structure(acc.words(), to_host(dev.Nodes), to_host(dev.Order))
    W, H = 640, 360
    C = oracle.init_camera(seed)
    I = px.walk_input(seed, 0)
    oracle.update_camera(C, I, 1.0 / 60.0, W, H)
    C[22] = 1.0
    Cd = to_dev(C)
    fa, fb = px.FrameBuffers(W, H), px.FrameBuffers(W, H)
    px.trace(dev, Cd, fa)
    px.trace_accel(dev, acc, Cd, fb)
    n = _order_dependent(to_host(fa.Hit).reshape(-1, 4), to_host(fb.Hit).reshape(-1, 4))
    va = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
    vb = torch.zeros_like(va)
    px.sun_visibility(dev, Cd, fa, va)
    px.sun_visibility_accel(dev, acc, Cd, fa, vb)
    nv = int(np.any(to_host(va).reshape(-1, 4).view(np.uint32) != to_host(vb).reshape(-1, 4).view(np.uint32),
                    axis=1).sum())
    print("Hit pixels differing", n, "Vis pixels differing", nv)
    assert n <= 2 and nv <= 2


def _order_dependent(a, b):
    """Pixels where two Hit buffers differ; each must be two object hits at
    nearly the same t (the walk-order-dependent case)."""
    diff = np.any(a.view(np.uint32) != b.view(np.uint32), axis=1)
    for k in np.nonzero(diff)[0]:
        assert a[k, 1] >= 0 and b[k, 1] >= 0, (k, a[k], b[k])
        assert abs(a[k, 0] - b[k, 0]) <= 1e-3 * a[k, 0], (k, a[k], b[k])
    return int(diff.sum())


@pytest.mark.parametrize("frame", [0, 57])
def test_accel_trace_equals_heap_trace(px, oracle, scene, frame):
    print("case: accel trace 1920x1080 walk frame", frame, flush=True)
    dev, acc = scene
    W, H = 1920, 1080
    C = to_dev(_walk_camera(px, oracle, frame, W, H, 0.0))
    fa, fb = px.FrameBuffers(W, H), px.FrameBuffers(W, H)
    px.trace(dev, C, fa)
    px.trace_accel(dev, acc, C, fb)
    a, b = to_host(fa.Hit).reshape(-1, 4), to_host(fb.Hit).reshape(-1, 4)
    n = _order_dependent(a, b)
    print("pixels differing:", n)
    assert n <= 4


def test_accel_launch_shapes_and_schedule_change_nothing(px, oracle, scene):
    print("case: accel variants and schedule, 640x360", flush=True)
    dev, acc = scene
    W, H = 640, 360
    C = to_dev(_walk_camera(px, oracle, 3, W, H, 1.0))
    lib = px.lib()
    old_t, old_v = lib.arbor_tune_get(b"accel_variant"), lib.arbor_tune_get(b"accel_vis_variant")
    try:
        ref_fb = px.FrameBuffers(W, H)
        px.trace_accel(dev, acc, C, ref_fb)
        want = to_host(ref_fb.Hit)
        sched = px.TileSchedule(W, H)
        for v in range(ACCEL_VARIANTS):
            assert lib.arbor_tune_set(b"accel_variant", v) == 0
            for s in (None, sched, sched):  # unscheduled, first and second scheduled call
                fb = px.FrameBuffers(W, H)
                px.trace_accel(dev, acc, C, fb, s)
                assert np.array_equal(to_host(fb.Hit).view(np.uint32), want.view(np.uint32)), (v, s is None)
        import torch

        vis = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
        assert lib.arbor_tune_set(b"accel_vis_variant", 0) == 0
        px.sun_visibility_accel(dev, acc, C, ref_fb, vis)
        want_vis = to_host(vis).view(np.uint32).copy()
        vsched = px.TileSchedule(W, H)
        for v in range(ACCEL_VIS_VARIANTS):
            assert lib.arbor_tune_set(b"accel_vis_variant", v) == 0
            for s in (None, vsched):
                vis.zero_()
                px.sun_visibility_accel(dev, acc, C, ref_fb, vis, s)
                assert np.array_equal(to_host(vis).view(np.uint32), want_vis), (v, s is None)
    finally:
        lib.arbor_tune_set(b"accel_variant", old_t)
        lib.arbor_tune_set(b"accel_vis_variant", old_v)


def test_accel_visibility_equals_heap_visibility(px, oracle, scene):
    print("case: accel sun visibility 1920x1080 walk frame 0, shadows on", flush=True)
    import torch

    dev, acc = scene
    W, H = 1920, 1080
    C = to_dev(_walk_camera(px, oracle, 0, W, H, 1.0))
    fb = px.FrameBuffers(W, H)
    px.trace(dev, C, fb)
    va = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
    vb = torch.zeros_like(va)
    px.sun_visibility(dev, C, fb, va)
    px.sun_visibility_accel(dev, acc, C, fb, vb)
    a, b = to_host(va).reshape(-1, 4), to_host(vb).reshape(-1, 4)
    diff = np.any(a.view(np.uint32) != b.view(np.uint32), axis=1)
    print("pixels differing:", int(diff.sum()), "shadowed fraction %.3f" % (a[:, 0] < 1).mean())
    assert diff.sum() <= 4


def test_engine_accel_on_and_off(px):
    print("case: engine 320x192 seed 7, accel on vs off, 3 frames with shadows", flush=True)
    w, h = 320, 192
    out = []
    for on in (True, False):
        eng = px.Engine(w, h, 7)
        eng.set_accel(on)
        inp = np.zeros(32, np.float32)
        for f in range(3):
            inp[0] = 0.3 * f
            eng.set_input(inp)
            eng.frame()
        eng.sync()
        out.append(eng.download("Hit", w * h * 4).reshape(-1, 4))
    n = _order_dependent(out[0], out[1])
    print("pixels differing:", n)
    assert n <= 2


def test_batch_accel_equals_batch_heap(px, scene):
    """The batched views over the SAH hierarchy (every batch_variant, shadows
    on) give the heap batch's Hit, Linear and Guide, up to the
    walk-order-dependent pixels."""
    print("case: batch of 6 walk views, accel vs heap, batch_variant 0..9, shadows on", flush=True)
    dev, acc = scene
    cams, sizes = px.batch_walk_views(3, 7, 6)
    sizes = [(min(w, 640), min(h, 360)) for w, h in sizes]
    cams = np.array(cams, np.float32)
    cams[:, 22] = 1.0
    lib = px.lib()
    old = lib.arbor_tune_get(b"batch_variant")
    try:
        for v in range(10):
            assert lib.arbor_tune_set(b"batch_variant", v) == 0
            out = []
            for use in (None, acc):
                b = px.Batch(len(sizes), sum(w * h for w, h in sizes))
                b.set_views(cams, sizes)
                b.set_accel(use)
                b.render(dev, px.BATCH_TRACE_SHADE)
                out.append([b.download(i, "Hit") for i in range(len(sizes))] +
                           [b.download(i, "Linear") for i in range(len(sizes))])
                b.close()
            n_hit = sum(_order_dependent(a.reshape(-1, 4), c.reshape(-1, 4)) for a, c in zip(out[0][:6], out[1][:6]))
            lin = [np.any(a.reshape(-1, 4).view(np.uint32) != c.reshape(-1, 4).view(np.uint32), axis=1)
                   for a, c in zip(out[0][6:], out[1][6:])]
            print("batch_variant", v, "Hit pixels differing", n_hit, "Linear", int(sum(x.sum() for x in lin)))
            assert n_hit <= 2 and sum(x.sum() for x in lin) <= 2 + 2 * n_hit
    finally:
        lib.arbor_tune_set(b"batch_variant", old)


def test_accel_autotune(px, oracle, scene):
    """The SAH-walk autotuner times every accel_variant, keeps the fastest,
    and the trace it selects gives the same Hit."""
    print("case: accel autotune 640x360", flush=True)
    dev, acc = scene
    W, H = 640, 360
    C = to_dev(_walk_camera(px, oracle, 3, W, H, 0.0))
    lib = px.lib()
    old = lib.arbor_tune_get(b"accel_variant")
    try:
        ref = px.FrameBuffers(W, H)
        px.trace_accel(dev, acc, C, ref)
        fb = px.FrameBuffers(W, H)
        r = px.autotune_trace_accel(dev, acc, C, fb, reps=1)
        assert len(r["ms"]) == ACCEL_VARIANTS and r["best"] == int(np.argmin(r["ms"]))
        assert lib.arbor_tune_get(b"accel_variant") == r["best"]
        px.trace_accel(dev, acc, C, fb)
        assert np.array_equal(to_host(fb.Hit).view(np.uint32), to_host(ref.Hit).view(np.uint32))
        print("autotune:", r)
    finally:
        lib.arbor_tune_set(b"accel_variant", old)
