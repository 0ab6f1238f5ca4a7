"""The allocation-free trace entry point, arbor_tracePixelsAccelNoAlloc,
against the entry point it varies, arbor_tracePixelsAccel with a
caller-owned tile schedule: Hit bit for bit on two calls in a row, the second
with a smaller view (the library's schedule array then holds the first
call's tile times and order); views beyond the array's 512 x 512 tiles are
rejected with ARBOR_ERR_ARG before any launch."""
import ctypes as ct

import numpy as np
import pytest

from gpu_util import require_gpu, to_dev, to_host

pytestmark = pytest.mark.gpu


def test_no_alloc_trace_equals_scheduled_trace(oracle):
    require_gpu()
    import pyarbor as px

    scene = px.DeviceScene().generate(7)
    acc = px.TraceAccel(scene)
    lib = px.lib()
    assert lib.arbor_no_alloc_scratch_bytes() == 2560 + 324 * 48 + 12 * 512 * 512
    C = oracle.init_camera(7)
    for (w, h), frames in (((640, 360), 4), ((328, 184), 3)):
        print("case: no-alloc trace %dx%d after %d walk frames" % (w, h, frames), flush=True)
        for f in range(frames):
            oracle.update_camera(C, px.walk_input(7, f), 1.0 / 60.0, w, h)
        Cd = to_dev(C)
        a, b = px.FrameBuffers(w, h), px.FrameBuffers(w, h)
        px.trace_accel(scene, acc, Cd, a, px.TileSchedule(w, h))
        b.Hit.fill_(float("nan"))
        px.call("arbor_tracePixelsAccelNoAlloc", scene.P, scene.Nodes, scene.Order, acc.buf, Cd, b.Hit, w, h)
        want, got = to_host(a.Hit), to_host(b.Hit)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), int((got != want).sum())
        assert (want.reshape(-1, 4)[:, 1] >= 0).mean() > 0.05
    big = px.FrameBuffers(8, 8)
    args = [px._ptr(t) for t in (scene.P, scene.Nodes, scene.Order, acc.buf, Cd, big.Hit)]
    assert lib.arbor_tracePixelsAccelNoAlloc(*args, 4104, 4096, px._stream()) == -2  # 513 x 512 tiles
    assert lib.arbor_tracePixelsAccelNoAlloc(*args[:5], None, 64, 64, px._stream()) == -1
