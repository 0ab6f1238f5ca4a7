"""GPU parity of the per-frame path (tracePixels, traceTreePixels,
shadePixels, filterLighting, resolvePixels, the engine frame) against the
oracle's FP32 restatement and its FP64 run.

The reference algorithm is ill-conditioned in FP32 (its own FP32 and FP64
runs disagree on a few percent of pixels: the 7-step leaf Newton solve and
the 3e-4 acceptance threshold). The bar, per view:
  * hit ids agree with the FP32 oracle except on near-tie rays (rays where
    the FP32 and FP64 oracle runs themselves disagree), with at most 2e-4 of
    the view's pixels disagreeing outside them, and each of those within
    TIE_EPS_PX (1e-3 pixel) of a tie: the GPU's (id, tree) is what the FP32
    or FP64 oracle answers for a ray through a point 1e-3 pixel away;
  * on rays where the GPU and FP64 ids agree, t relative error <= 1e-5 for
    at least the fraction the FP32 oracle itself reaches, minus 0.5%;
  * radiance (Linear) within 1e-3 absolute of the FP64 run on the same
    hits for >= 99% of pixels;
  * filterLighting bit-exact on identical inputs, History bit-exact, Pixels
    within one 8-bit step (the device powf against glibc's).
"""
import numpy as np
import pytest

from gpu_util import check_hits, require_gpu, to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def px():
    require_gpu()
    import pyarbor

    return pyarbor


@pytest.fixture(scope="module")
def scene7(px, oracle):
    return px.DeviceScene().generate(7), oracle.generate(7)


def _camera(oracle, seed, view, w, h, shadows=0):
    C = oracle.init_camera(seed)
    I = np.zeros(32, np.float32)
    I[5] = view
    oracle.update_camera(C, I, 1.0 / 60.0, w, h)
    C[22] = shadows
    return C


def _check_hits(tag, gpu, h32, h64, tie=None):
    return check_hits(tag, gpu, h32, h64, tie)


def test_single_tree_c1(px, oracle):
    """Config 1: seed 1, one tree, 256x256 primary rays, camera 8 m from the
    trunk at eye height 1.72 m, shadows off."""
    print("case: traceTreePixels seed 1 256x256", flush=True)
    dev = px.DeviceScene().generate(1)
    ref = oracle.generate(1)
    C = oracle.init_camera(1)
    C[0:3] = [0.0, 1.72, 8.0]
    C[4:7] = [0.0, 1.72, 7.0]
    C[22] = 0
    fb = px.FrameBuffers(256, 256)
    Cd = to_dev(C)
    px.trace_tree(dev, Cd, fb)
    gpu = to_host(fb.Hit)
    h32 = oracle.trace_tree(ref, C, 256, 256)
    h64 = oracle.trace_tree64(ref, C, 256, 256)
    _check_hits("C1", gpu, h32, h64, (oracle, ref, C, 256, 256, None, True))
    assert (gpu[1::4] >= 0).mean() > 0.2


@pytest.mark.parametrize("view", [1, 3, 5])
def test_forest_trace(px, oracle, scene7, view):
    print("case: tracePixels seed 7 view", view, "320x180", flush=True)
    dev, ref = scene7
    w, h = 320, 180
    C = _camera(oracle, 7, view, w, h)
    fb = px.FrameBuffers(w, h)
    px.trace(dev, to_dev(C), fb)
    gpu = to_host(fb.Hit)
    h32 = oracle.trace(ref, C, w, h)
    h64 = oracle.trace64(ref, C, w, h)
    _check_hits("forest v%d" % view, gpu, h32, h64, (oracle, ref, C, w, h, None))


@pytest.mark.parametrize("frame", [0, 57])
def test_forest_trace_full_size(px, oracle, scene7, frame):
    """The benchmark's view size: a 1920x1080 walk frame (seed 7, 324 trees)
    traced on the GPU. 20,000 seeded pixels are checked against the FP32 and
    FP64 oracle with the reduced views' bars, and the whole frame against
    size-independent properties: every t in (0, 1e5], objects nearer than
    1e5, ids and trees in range and present together, Hit.w = C[14] * t /
    height."""
    print("case: 1920x1080 walk frame", frame, flush=True)
    dev, ref = scene7
    W, H = 1920, 1080
    C = oracle.init_camera(7)
    for f in range(frame + 1):
        I = px.walk_input(7, f)
        if f == 0:
            I[10] = 1.0  # sun visibility off, as the benchmark walk runs
        oracle.update_camera(C, I, 1.0 / 60.0, W, H)
    fb = px.FrameBuffers(W, H)
    px.trace(dev, to_dev(C), fb)
    hit = to_host(fb.Hit).reshape(-1, 4)
    rng = np.random.default_rng(1000 + frame)
    n = 20000
    xy = np.stack([rng.integers(0, W, n), rng.integers(0, H, n)], axis=1).astype(np.int32)
    gpu = hit[xy[:, 1] * W + xy[:, 0]].reshape(-1)
    h32 = oracle.trace_pixels(ref, C, xy, W, H)
    h64 = oracle.trace_pixels(ref, C, xy, W, H, fp64=True)
    _check_hits("1080p frame %d (sampled)" % frame, gpu, h32, h64, (oracle, ref, C, W, H, xy))
    t, ids, tree = hit[:, 0], hit[:, 1], hit[:, 2]
    assert (t > 0).all()
    sky = ids < 0
    assert ((ids == -1) | ((ids >= 0) & (ids < 29412))).all()
    assert ((tree == -1) | ((tree >= 0) & (tree < 324))).all()
    assert np.array_equal(ids >= 0, tree >= 0)
    assert (t <= np.float32(1e5)).all() and (t[~sky] < np.float32(1e5)).all()  # sky and ground carry no id
    w_want = (np.float32(C[14]) * t.astype(np.float32)) / np.float32(H)
    assert np.array_equal(hit[:, 3].view(np.uint32), w_want.astype(np.float32).view(np.uint32))
    print("objects %.3f ground %.3f sky %.3f" % ((ids >= 0).mean(), ((ids < 0) & (t < 1e5)).mean(), (t == 1e5).mean()))


def test_trace_variants_identical(px, oracle, scene7):
    """Every trace_variant / grid_cap setting gives bit-identical Hit (they
    change only how lanes are scheduled)."""
    print("case: tracePixels variants seed 7 view 2 240x136", flush=True)
    dev, _ = scene7
    w, h = 240, 136
    Cd = to_dev(_camera(oracle, 7, 2, w, h))
    keys = ("trace_variant", "grid_cap", "sched_tiles", "sched_leaf_min", "sched_newton_min")
    old = {k: px.tune_get(k) for k in keys}
    settings = [dict(trace_variant=v, grid_cap=65535) for v in range(10)]
    settings += [dict(trace_variant=1, grid_cap=7), dict(trace_variant=0, grid_cap=5),
                 dict(trace_variant=8, sched_tiles=1, sched_leaf_min=64, sched_newton_min=1),
                 dict(trace_variant=9, sched_tiles=37, sched_leaf_min=1, sched_newton_min=64)]
    outs = []
    try:
        for st in settings:
            for k, v in st.items():
                px.tune_set(k, v)
            fb = px.FrameBuffers(w, h)
            px.trace(dev, Cd, fb)
            outs.append(to_host(fb.Hit))
    finally:
        for k, v in old.items():
            px.tune_set(k, v)
    for o in outs[1:]:
        assert np.array_equal(o.view(np.uint32), outs[0].view(np.uint32))


def test_scheduled_trace_identical(px, oracle, scene7):
    """tracePixelsScheduled over a moving camera (walk frames, then a preset
    jump) gives Hit bit-identical to tracePixels on every frame; its order is
    a permutation of the tiles and every tile gets a recorded cost."""
    print("case: scheduled trace, 6 frames 328x184", flush=True)
    dev, _ = scene7
    w, h = 328, 184  # ragged: 41 x 23 tiles
    sched = px.TileSchedule(w, h)
    a, b = px.FrameBuffers(w, h), px.FrameBuffers(w, h)
    C = oracle.init_camera(7)
    tiles = 41 * 23
    for f in range(6):
        I = px.walk_input(7, f)
        if f == 4:
            I[5] = 3  # jump to a view preset: no overlap with the previous view
        oracle.update_camera(C, I, 1.0 / 60.0, w, h)
        Cd = to_dev(C)
        px.trace(dev, Cd, a)
        px.trace_scheduled(dev, Cd, b, sched)
        assert np.array_equal(to_host(a.Hit).view(np.uint32), to_host(b.Hit).view(np.uint32)), f
        order = sched.order()
        assert np.array_equal(np.sort(order), np.arange(tiles, dtype=np.uint32)), f
        assert (sched.cost() >= 1).all()


def test_visibility_pass_identical(px, oracle, scene7):
    """sunVisibility (plain and tile-scheduled) followed by shadePixelsVis gives
    Linear and Guide bit-identical to the fused shadePixels: sun visibility on,
    breeze on, over moving frames and jittered samples; with visibility off
    the pass writes nothing and shading reads nothing (Vis filled with NaN)."""
    import torch

    print("case: visibility pass, 5 frames 328x184, wind", flush=True)
    dev, _ = scene7
    w, h = 328, 184
    sched = px.TileSchedule(w, h)
    a, b = px.FrameBuffers(w, h), px.FrameBuffers(w, h)
    vis = torch.zeros(w * h * 4, dtype=torch.float32, device="cuda")
    C = oracle.init_camera(7)
    for f in range(5):
        oracle.update_camera(C, px.walk_input(7, f), 1.0 / 60.0, w, h)
        c = C.copy()
        c[22], c[11], c[8], c[13] = float(f != 2), 0.6, 0.25 * f, 0.610865
        c[9], c[10] = float(f % 3), float(f % 3 == 0)
        Cd = to_dev(c)
        px.trace(dev, Cd, a)
        b.Hit.copy_(a.Hit)
        px.shade(dev, Cd, a)
        if f == 2:
            vis.fill_(float("nan"))
        if f % 2:
            px.sun_visibility(dev, Cd, b, vis)
        else:
            px.sun_visibility(dev, Cd, b, vis, sched)
        px.shade_vis(dev, Cd, b, vis)
        for name in ("Linear", "Guide"):
            want = to_host(getattr(a, name)).view(np.uint32)
            got = to_host(getattr(b, name)).view(np.uint32)
            assert np.array_equal(want, got), (f, name, int((want != got).sum()))
        if f == 2:
            assert torch.isnan(vis).all()
        else:
            v = to_host(vis).reshape(-1, 4)
            assert (v[:, 3] == 1).all() and (v[:, :3] >= 0).all() and (v[:, :3] <= 1).all()
            print("frame %d mean visibility %.4f" % (f, float(v[:, :3].mean())))


def test_visibility_pass_full_size(px, oracle, scene7):
    """The shadowed 4K workload's setting at its full size (3840x2160, seed 7,
    sun at 35 deg, breeze on, sample 1): the tile-scheduled visibility pass
    and shading from it give Linear and Guide bit-identical to the fused
    shadePixels on every pixel."""
    import torch

    print("case: visibility pass 3840x2160 vs fused shading", flush=True)
    dev, _ = scene7
    W, H = 3840, 2160
    C = oracle.init_camera(7)
    I = np.zeros(32, np.float32)
    I[7] = 1.0  # breeze on
    oracle.update_camera(C, I, 1.0 / 60.0, W, H)
    C[13], C[9], C[10] = 0.610865, 1.0, 0.0
    assert C[22] == 1.0 and C[11] > 0.0
    Cd = to_dev(C)
    a, b = px.FrameBuffers(W, H), px.FrameBuffers(W, H)
    px.trace(dev, Cd, a)
    b.Hit.copy_(a.Hit)
    px.shade(dev, Cd, a)
    vis = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
    px.sun_visibility(dev, Cd, b, vis, px.TileSchedule(W, H))
    px.shade_vis(dev, Cd, b, vis)
    for name in ("Linear", "Guide"):
        assert torch.equal(getattr(a, name).view(torch.int32), getattr(b, name).view(torch.int32)), name
    v = vis.view(-1, 4)[:, :3]
    print("shadowed (visibility < 1) fraction %.3f" % float((v < 1).any(dim=1).float().mean()))


def test_visibility_variants_identical(px, oracle, scene7):
    """Every vis_variant (launch shape, walk options, register budget) gives
    the same Vis bit for bit."""
    import torch

    print("case: visibility variants seed 7 view 3 240x136, wind", flush=True)
    dev, _ = scene7
    w, h = 240, 136
    C = _camera(oracle, 7, 3, w, h, shadows=1)
    C[11], C[8] = 0.6, 0.5
    Cd = to_dev(C)
    fb = px.FrameBuffers(w, h)
    px.trace(dev, Cd, fb)
    old = px.tune_get("vis_variant")
    outs = []
    try:
        for v in range(8):
            px.tune_set("vis_variant", v)
            vis = torch.zeros(w * h * 4, dtype=torch.float32, device="cuda")
            px.sun_visibility(dev, Cd, fb, vis)
            outs.append(to_host(vis))
    finally:
        px.tune_set("vis_variant", old)
    assert (outs[0].reshape(-1, 4)[:, :3] < 1).mean() > 0.05  # some pixels in shadow
    for o in outs[1:]:
        assert np.array_equal(o.view(np.uint32), outs[0].view(np.uint32))


@pytest.mark.parametrize("view,shadows", [(1, 0), (3, 1), (4, 0)])
def test_shade_filter_resolve(px, oracle, scene7, view, shadows):
    """Shading on identical hits (the oracle's FP32 Hit uploaded) isolates the
    material/lighting arithmetic; filter and resolve then run on the GPU's own
    Linear/Guide and are compared with the oracle on the same inputs."""
    print("case: shade/filter/resolve view", view, "shadows", shadows, flush=True)
    dev, ref = scene7
    w, h = 160, 96
    C = _camera(oracle, 7, view, w, h, shadows)
    Cd = to_dev(C)
    H = oracle.trace(ref, C, w, h)
    fb = px.FrameBuffers(w, h)
    fb.Hit.copy_(to_dev(H))
    px.shade(dev, Cd, fb)
    L = to_host(fb.Linear)
    G = to_host(fb.Guide)
    L64, G64 = oracle.shade64(ref, C, H, w, h)
    L32, _ = oracle.shade(ref, C, H, w, h)
    err = np.abs(L.reshape(-1, 4)[:, :3] - L64.reshape(-1, 4)[:, :3]).max(axis=1)
    err32 = np.abs(L32.reshape(-1, 4)[:, :3] - L64.reshape(-1, 4)[:, :3]).max(axis=1)
    print("shade |gpu-fp64|: p50 %.3g p99 %.3g max %.3g frac<=1e-3 %.5f; fp32 oracle frac<=1e-3 %.5f" % (
        np.median(err), np.percentile(err, 99), err.max(), (err <= 1e-3).mean(), (err32 <= 1e-3).mean()))
    assert (err <= 1e-3).mean() >= min(0.99, (err32 <= 1e-3).mean() - 0.005)
    gerr = np.abs(G - G64).max()
    print("guide max |gpu-fp64| %.3g" % gerr)
    assert np.isfinite(L).all() and (L >= 0).all()
    # filter on the GPU's own Linear/Guide vs the oracle's filter on those inputs
    px.filter_lighting(Cd, fb)
    F = to_host(fb.Filtered)
    F32 = oracle.filter(L, G, H, C, w, h)
    # same operands, same operation order, the tap weights from the host's expf
    assert np.array_equal(F.view(np.uint32), F32.view(np.uint32)), int((F != F32).sum())
    # resolve: first sample (C[10]=1) then an accumulation step
    px.resolve(Cd, fb)
    pix = to_host(fb.Pixels, np.uint32)
    hist = np.zeros_like(F)
    want = oracle.resolve(F, hist, C, w, h)
    dpx = np.abs(pix.view(np.uint8).astype(int) - want.view(np.uint8).astype(int)).max()
    print("resolve max channel diff", dpx)
    assert dpx <= 1
    assert np.array_equal(to_host(fb.History).view(np.uint32), hist.view(np.uint32))


@pytest.mark.parametrize("view", [1, 3])
def test_wind_shadows_four_spp(px, oracle, scene7, view):
    """The shadowed, windy configuration at a reduced size: breeze on
    (C[11]=0.6) at wind time C[8]=0.75 s, sun elevation 35 deg, tree-to-tree
    sun visibility on, four jittered samples per pixel (C[9]=s, C[10]=(s==0)),
    each traced, shaded, filtered and accumulated by resolvePixels."""
    print("case: wind+shadows 4 spp view", view, "192x108", flush=True)
    dev, ref = scene7
    w, h = 192, 108
    C = _camera(oracle, 7, view, w, h, shadows=1)
    C[11], C[8], C[13] = 0.6, 0.75, 0.610865
    fb = px.FrameBuffers(w, h)
    hist = np.zeros(w * h * 4, np.float32)
    for s in range(4):
        C[9], C[10] = float(s), float(s == 0)
        Cd = to_dev(C)
        px.trace(dev, Cd, fb)
        h32 = oracle.trace(ref, C, w, h)
        h64 = oracle.trace64(ref, C, w, h)
        _check_hits("wind v%d s%d" % (view, s), to_host(fb.Hit), h32, h64, (oracle, ref, C, w, h, None))
        # shade, filter and resolve on the oracle's hits (isolates the lighting arithmetic)
        fb.Hit.copy_(to_dev(h32))
        px.shade(dev, Cd, fb)
        L, G = to_host(fb.Linear), to_host(fb.Guide)
        L64, _ = oracle.shade64(ref, C, h32, w, h)
        L32, G32 = oracle.shade(ref, C, h32, w, h)
        err = np.abs(L.reshape(-1, 4)[:, :3] - L64.reshape(-1, 4)[:, :3]).max(axis=1)
        err32 = np.abs(L32.reshape(-1, 4)[:, :3] - L64.reshape(-1, 4)[:, :3]).max(axis=1)
        print("s%d shade frac<=1e-3 gpu %.5f fp32 oracle %.5f" % (s, (err <= 1e-3).mean(), (err32 <= 1e-3).mean()))
        assert (err <= 1e-3).mean() >= min(0.99, (err32 <= 1e-3).mean() - 0.005)
        px.filter_lighting(Cd, fb)
        F = to_host(fb.Filtered)
        assert np.array_equal(F.view(np.uint32), oracle.filter(L, G, h32, C, w, h).view(np.uint32))
        px.resolve(Cd, fb)
        want = oracle.resolve(F, hist, C, w, h)
        pix = to_host(fb.Pixels, np.uint32)
        assert np.abs(pix.view(np.uint8).astype(int) - want.view(np.uint8).astype(int)).max() <= 1
        assert np.array_equal(to_host(fb.History).view(np.uint32), hist.view(np.uint32))
        fb.History.copy_(to_dev(hist))  # both sides accumulate from the same history
    assert (hist[3::4] == 4.0).all()  # four samples accumulated


def test_engine_render_mode_and_ppm(px, oracle, tmp_path):
    """The harness's render mode through the engine (native_harness.cpp:60-63:
    view preset 2, sun visibility on, 4 samples accumulated by resolve) against
    the oracle's same loop; the PPM carries the engine's Pixels."""
    print("case: engine render view 2, 4 samples, 320x192", flush=True)
    w, h = 320, 192
    eng = px.Engine(w, h, 7)
    eng.render(2, 4, light=True)
    got = eng.download("Pixels", w * h, np.uint32)
    ref = oracle.generate(7)
    C = oracle.init_camera(7)
    I = np.zeros(32, np.float32)
    I[5] = 2
    oracle.update_camera(C, I, 1.0 / 60.0, w, h)
    C[22] = 1
    hist = np.zeros(w * h * 4, np.float32)
    for s in range(4):
        C[9], C[10] = float(s), float(s == 0)
        H = oracle.trace(ref, C, w, h)
        L, G = oracle.shade(ref, C, H, w, h)
        F = oracle.filter(L, G, H, C, w, h)
        want = oracle.resolve(F, hist, C, w, h)
    d = np.abs(got.view(np.uint8).astype(int) - want.view(np.uint8).astype(int)).reshape(-1, 4).max(axis=1)
    print("render pixels: frac<=1 %.5f frac<=2 %.5f max %d, %d pixels > 1" % (
        (d <= 1).mean(), (d <= 2).mean(), d.max(), int((d > 1).sum())))
    # four samples, each with its own near-tie hits and shadow rays: the
    # pixels off by more than one step are those a differing hit or
    # visibility reaches through the filter (measured 12 of 61,440)
    assert (d <= 1).mean() >= 0.999
    path = tmp_path / "render.ppm"
    eng.save_ppm(path)
    data = path.read_bytes()
    head = b"P6\n%d %d\n255\n" % (w, h)
    assert data[:len(head)] == head and len(data) == len(head) + 3 * w * h
    rgb = np.frombuffer(data[len(head):], np.uint8).reshape(-1, 3)
    assert np.array_equal(rgb, got.view(np.uint8).reshape(-1, 4)[:, :3])
    eng.close()


def test_engine_record_replay(px, tmp_path):
    """Frames recorded on one engine (host and device inputs, a breeze toggle,
    a multi-sample frame) replay on a fresh engine of the same size and seed
    to the same camera state and bit-identical Pixels; partial replays and a
    file that is not a recording are handled."""
    import torch

    print("case: engine record/replay 320x192 seed 7", flush=True)
    w, h = 320, 192
    rec = tmp_path / "walk.rec"
    a = px.Engine(w, h, 7)
    a.record(rec)
    for f in range(5):
        I = px.walk_input(7, f)
        if f == 1:
            I[7] = 1.0  # breeze on
        if f % 2:
            a.frame(1.0 / 60.0, torch.from_numpy(I).cuda())
        else:
            a.set_input(I)
            a.frame(1.0 / 60.0)
    a.set_input(px.walk_input(7, 5))
    a.frame_samples(3, 1.0 / 60.0)
    a.record(None)
    want_c = a.download("C", 64).view(np.uint32).copy()
    want_px = a.download("Pixels", w * h, np.uint32).copy()
    a.frame(1.0 / 60.0)  # not recorded
    assert rec.stat().st_size == 24 + 6 * 136
    b = px.Engine(w, h, 7)
    assert b.replay(rec) == 6
    assert np.array_equal(b.download("C", 64).view(np.uint32), want_c)
    assert np.array_equal(b.download("Pixels", w * h, np.uint32), want_px)
    c = px.Engine(w, h, 7)
    assert c.replay(rec, 4, 1) == 1 and c.replay(rec, 10) == 0
    bad = tmp_path / "bad.rec"
    bad.write_bytes(b"not a recording")
    with pytest.raises(px.ArborError):
        c.replay(bad)
    for e in (a, b, c):
        e.close()


def _dilate(mask, w, h, r=2):
    """Pixels within r (Chebyshev) of a set pixel: the filter's 5x5 reach."""
    m = mask.reshape(h, w)
    out = np.zeros_like(m)
    for dy in range(-r, r + 1):
        for dx in range(-r, r + 1):
            ys0, ys1 = max(0, dy), h + min(0, dy)
            xs0, xs1 = max(0, dx), w + min(0, dx)
            out[ys0 - dy:ys1 - dy, xs0 - dx:xs1 - dx] |= m[ys0:ys1, xs0:xs1]
    return out.reshape(-1)


def test_engine_frame_pipeline(px, oracle):
    """The engine's frame (updateCamera, the scheduled trace over the SAH
    hierarchy, the sun-visibility pass, shading, filter, resolve) against the
    oracle pipeline on the same camera path. The last frame's Hit meets the
    trace bar (ties within TIE_EPS_PX); its Linear meets the shading bar on
    the pixels whose hit agrees; and Pixels are within one 8-bit step of the
    oracle's everywhere outside the filter's 5x5 reach of a pixel whose hit
    or radiance differs (the only pixels whose inputs differ)."""
    print("case: engine frame 320x192 seed 7", flush=True)
    w, h = 320, 192
    eng = px.Engine(w, h, 7)
    assert eng.size() == (w, h)
    ref = oracle.generate(7)
    C = oracle.init_camera(7)
    hist = np.zeros(w * h * 4, np.float32)
    for f in range(3):
        I = px.walk_input(7, f)
        if f == 0:
            I[10] = 1.0  # toggle sun visibility off (updateCamera I[10])
        eng.set_input(I)
        eng.frame(1.0 / 60.0)
        oracle.update_camera(C, I, 1.0 / 60.0, w, h)
        H = oracle.trace(ref, C, w, h)
        L, G = oracle.shade(ref, C, H, w, h)
        F = oracle.filter(L, G, H, C, w, h)
        want = oracle.resolve(F, hist, C, w, h)
    Cg = eng.download("C", 64)
    assert np.array_equal(Cg.view(np.uint32), C.view(np.uint32))
    hit = eng.download("Hit", w * h * 4)
    check_hits("engine frame", hit, H, oracle.trace64(ref, C, w, h), (oracle, ref, C, w, h, None))
    same = np.all(hit.reshape(-1, 4)[:, :3] == H.reshape(-1, 4)[:, :3], axis=1)
    lin = eng.download("Linear", w * h * 4).reshape(-1, 4)[:, :3]
    L64, _ = oracle.shade64(ref, C, H, w, h)
    err = np.abs(lin - L64.reshape(-1, 4)[:, :3]).max(axis=1)
    err32 = np.abs(L.reshape(-1, 4)[:, :3] - L64.reshape(-1, 4)[:, :3]).max(axis=1)
    print("engine shade frac<=1e-3 %.5f (fp32 oracle %.5f)" % ((err[same] <= 1e-3).mean(), (err32 <= 1e-3).mean()))
    assert (err[same] <= 1e-3).mean() >= min(0.99, (err32 <= 1e-3).mean() - 0.005)
    got = eng.download("Pixels", w * h, np.uint32)
    d = np.abs(got.view(np.uint8).astype(int) - want.view(np.uint8).astype(int)).reshape(-1, 4).max(axis=1)
    differs = ~same | (np.abs(lin - L.reshape(-1, 4)[:, :3]).max(axis=1) > 1e-3)
    reach = _dilate(differs, w, h)
    print("engine pixels: max step %d; outside reach max %d; %d pixels in reach of %d differing inputs" % (
        d.max(), d[~reach].max() if (~reach).any() else 0, int(reach.sum()), int(differs.sum())))
    assert d[~reach].max() <= 1
    # the differing inputs: the hit mismatches (trace bar) and the radiance
    # differences beyond 1e-3 (shading bar); measured 180 of 61,440 pixels
    assert differs.mean() <= 0.005
    info = eng.inspect()
    assert info["sceneBytes"] < 7 * 1048576  # source.test.mjs:6
    assert info["frames"] == 3
    assert info["accel"] is True
    px_bytes = w * h * (6 * 16 + 4)
    assert info["allocatedBytes"] > info["sceneBytes"] + px_bytes + px.lib().arbor_trace_accel_bytes()
    eng.close()


def test_engine_frame_is_five_launches(px):
    """The reference engine issues five dispatches per frame (src/engine.js:48-50,
    asserted at tests/source.test.mjs:6). The engine's frame is five kernel
    launches too, counted by the profiler over one frame on a device input
    (no host copy): the prologue (updateCamera and the two tile orders), the
    trace, the visibility pass, shading, and filter + resolve. A 3-sample
    frame is five launches per sample."""
    import torch
    from torch.profiler import ProfilerActivity, profile

    print("case: launches per engine frame 320x192", flush=True)
    w, h = 320, 192
    eng = px.Engine(w, h, 7)
    dev_in = torch.from_numpy(px.walk_input(7, 0)).cuda()
    eng.frame(1.0 / 60.0, dev_in)  # warm: every kernel loaded
    eng.frame_samples(3, 1.0 / 60.0, dev_in)
    eng.sync()
    torch.cuda.synchronize()
    counts = []
    for samples in (1, 3):
        with profile(activities=[ProfilerActivity.CUDA]) as prof:
            if samples == 1:
                eng.frame(1.0 / 60.0, dev_in)
            else:
                eng.frame_samples(samples, 1.0 / 60.0, dev_in)
            eng.sync()
            torch.cuda.synchronize()
        names = [e.name for e in prof.events() if e.device_type == torch.autograd.DeviceType.CUDA
                 and "memcpy" not in e.name.lower() and "memset" not in e.name.lower()]
        print(samples, "sample(s):", len(names), "kernels:", names)
        counts.append(len(names))
    eng.close()
    assert counts == [5, 15]


def test_engine_passes_equal_entry_points(px):
    """The engine's five launches give every buffer the bits of the public
    entry points chained one by one on the same camera: arbor_tracePixelsAccel,
    arbor_sunVisibilityAccel, arbor_shadePixelsVis, arbor_filterLighting,
    arbor_resolvePixels (sun visibility on, History accumulating over a still
    camera)."""
    import torch

    print("case: engine passes vs entry points 328x184", flush=True)
    w, h = 328, 184
    eng = px.Engine(w, h, 7)
    still = np.zeros(32, np.float32)
    eng.set_input(px.walk_input(7, 3))
    eng.frame(1.0 / 60.0)
    eng.set_input(still)
    eng.frame(1.0 / 60.0)
    before = eng.download("History", w * h * 4).copy()
    eng.frame(1.0 / 60.0)
    C = eng.download("C", 64).copy()
    assert C[22] == 1.0 and C[10] == 0.0 and C[9] == 2.0  # visibility on, third sample accumulating
    dev = px.DeviceScene().generate(7)
    acc = px.TraceAccel(dev)
    fb = px.FrameBuffers(w, h)
    fb.History.copy_(to_dev(before))
    Cd = to_dev(C)
    px.trace_accel(dev, acc, Cd, fb)
    vis = torch.zeros(w * h * 4, dtype=torch.float32, device="cuda")
    px.sun_visibility_accel(dev, acc, Cd, fb, vis)
    px.shade_vis(dev, Cd, fb, vis)
    px.filter_lighting(Cd, fb)
    px.resolve(Cd, fb)
    for name in ("Hit", "Linear", "Guide", "Filtered", "History"):
        want = to_host(getattr(fb, name)).view(np.uint32)
        got = eng.download(name, w * h * 4).view(np.uint32)
        assert np.array_equal(got, want), (name, int((got != want).sum()))
    assert np.array_equal(eng.download("Pixels", w * h, np.uint32), to_host(fb.Pixels, np.uint32))
    assert (to_host(vis).reshape(-1, 4)[:, :3] < 1).mean() > 0.01  # some pixels in shadow
    eng.close()
