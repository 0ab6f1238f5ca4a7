"""GPU parity on the reference harness's own frames of the shadowed, windy
forest: the golden cases written by the compiled reference (breeze on at two
wind times, a progressive 4-sample sequence with jittered rays and sun-disk
samples, the sun at 35 degrees with visibility on and off, resolves over a
non-zero History, sample numbers 37 and 200, the three debug views), each
run through the C-ABI entry points.

Per case:
  * tracePixels on the case's camera meets the trace bar against the FP32
    and FP64 oracle (ties within TIE_EPS_PX), and the SAH walk's Hit equals
    the heap walk's;
  * shadePixels on the reference's Hit meets the shading bar (Linear within
    1e-3 of the FP64 run), and the visibility pass plus shadePixelsVis equals
    the fused shadePixels bit for bit;
  * filterLighting on the reference's Linear, Guide and Hit reproduces the
    reference's Filtered bit for bit (its digest);
  * resolvePixels on the reference's Filtered and the History the reference
    saw reproduces the reference's History bit for bit (its digest) and its
    Pixels within one 8-bit step (the device powf against glibc's).
The oracle buffers used as inputs are first checked against the reference's
digests, so every input is the reference's own.
"""
import numpy as np
import pytest

from gpu_util import check_hits, require_gpu, to_dev, to_host
from oracle_lib import C3_CASES, digest, replay_c3

pytestmark = pytest.mark.gpu

NAMES = [c["name"] for c in C3_CASES]


@pytest.fixture(scope="module")
def px():
    require_gpu()
    import pyarbor

    return pyarbor


@pytest.fixture(scope="module")
def setup(px, oracle, golden):
    fr = golden[0]["frames"]
    w, h = fr["width"], fr["height"]
    ref = oracle.generate(7)
    frames = replay_c3(oracle, ref, w, h)
    dev = px.DeviceScene().generate(7)
    acc = px.TraceAccel(dev)
    return dict(ref=ref, frames=frames, dev=dev, acc=acc, w=w, h=h)


@pytest.mark.parametrize("name", NAMES)
def test_golden_case_on_gpu(px, oracle, golden, setup, name):
    import torch

    print("case: golden shadowed/windy frame", name, flush=True)
    g = golden[0]["frames_c3"][name]
    f = setup["frames"][name]
    ref, dev, w, h = setup["ref"], setup["dev"], setup["w"], setup["h"]
    C = f["C"]
    assert np.array_equal(C, np.array(g["C"], np.float32))
    for buf in ("Hit", "Linear", "Guide", "Filtered", "Pixels"):
        assert digest(f[buf]) == g[buf], buf  # the oracle's inputs are the reference's
    Cd = to_dev(C)

    # trace on the case's camera
    fb = px.FrameBuffers(w, h)
    px.trace(dev, Cd, fb)
    hit = to_host(fb.Hit)
    check_hits("golden " + name, hit, f["Hit"], oracle.trace64(ref, C, w, h), (oracle, ref, C, w, h, None))
    fa = px.FrameBuffers(w, h)
    px.trace_accel(dev, setup["acc"], Cd, fa)
    assert np.array_equal(to_host(fa.Hit).view(np.uint32), hit.view(np.uint32))

    # shading on the reference's hits
    fb.Hit.copy_(to_dev(f["Hit"]))
    px.shade(dev, Cd, fb)
    L, G = to_host(fb.Linear), to_host(fb.Guide)
    L64, _ = oracle.shade64(ref, C, f["Hit"], w, h)
    err = np.abs(L.reshape(-1, 4)[:, :3] - L64.reshape(-1, 4)[:, :3]).max(axis=1)
    err32 = np.abs(f["Linear"].reshape(-1, 4)[:, :3] - L64.reshape(-1, 4)[:, :3]).max(axis=1)
    print("%s shade frac<=1e-3 gpu %.5f reference %.5f; guide max |gpu-ref| %.3g" % (
        name, (err <= 1e-3).mean(), (err32 <= 1e-3).mean(), float(np.abs(G - f["Guide"]).max())))
    assert (err <= 1e-3).mean() >= min(0.99, (err32 <= 1e-3).mean() - 0.005)
    assert np.isfinite(L).all() and (L >= 0).all()
    vis = torch.zeros(w * h * 4, dtype=torch.float32, device="cuda")
    fv = px.FrameBuffers(w, h)
    fv.Hit.copy_(fb.Hit)
    px.sun_visibility(dev, Cd, fv, vis)
    px.shade_vis(dev, Cd, fv, vis)
    assert np.array_equal(to_host(fv.Linear).view(np.uint32), L.view(np.uint32))
    assert np.array_equal(to_host(fv.Guide).view(np.uint32), G.view(np.uint32))

    # filter on the reference's Linear, Guide and Hit
    fb.Linear.copy_(to_dev(f["Linear"]))
    fb.Guide.copy_(to_dev(f["Guide"]))
    px.filter_lighting(Cd, fb)
    F = to_host(fb.Filtered)
    assert digest(F) == g["Filtered"], int((F != f["Filtered"]).sum())

    # resolve on the reference's Filtered and the History it resolved over
    fb.Filtered.copy_(to_dev(f["Filtered"]))
    fb.History.copy_(to_dev(f["before"]))
    px.resolve(Cd, fb)
    assert digest(to_host(fb.History)) == g["History"]
    pix = to_host(fb.Pixels, np.uint32)
    d = np.abs(pix.view(np.uint8).astype(int) - f["Pixels"].view(np.uint8).astype(int))
    print("%s pixels: %d of %d bit-exact, max step %d" % (name, int((pix == f["Pixels"]).sum()), pix.size, d.max()))
    assert d.max() <= 1
