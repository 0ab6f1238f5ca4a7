"""GPU checks of the lean denoiser network (levels = 0 in
include/arbor_denoise.h: compact 8-channel features, one 3x3 convolution of
16 or 32 channels with ReLU, a linear 1x1 output layer):

* the compact features the shading pass writes (arbor_shadePixelsFeatures)
  equal, bit for bit, the ones the runner computes from shadePixels' Linear
  and Guide and the Hit;
* forward against a plain PyTorch fp32 restatement (CPU) of the graph fed the
  runner's bf16 input and weights, each layer rounded to bf16 as the runner
  stores it: relative L2 <= 2e-2; one training step's weight and bias
  gradients against autograd of the same L1 loss: cosine >= 0.99, and the
  1x1 layer's taps other than the centre stay exactly 0;
* the one-pass kernel (arbor_denoiser_denoise, and arbor_denoiser_denoise_
  features from shading's features) equals the layered runner bit for bit;
* the engine with the lean network attached equals the entry points chained
  by hand bit for bit and keeps five launches per frame;
* the committed lean checkpoint on a 1920x1080 tracer frame scores a higher
  PSNR than filterLighting against the 64-sample render.
"""
import os
import sys

import numpy as np
import pytest

from gpu_util import require_gpu, to_dev, to_host

pytestmark = pytest.mark.gpu
N, H, W, w_img, h_img = 2, 64, 96, 90, 61


@pytest.fixture(scope="module")
def env():
    torch = require_gpu()
    import pyarbor

    sys.path.insert(0, os.path.dirname(pyarbor.__file__))
    return torch, pyarbor


def _frames(torch, seed):
    g = torch.Generator().manual_seed(seed)
    n = w_img * h_img
    lin = torch.rand(N, n, 4, generator=g) * 3.0
    gui = torch.rand(N, n, 4, generator=g)
    hit = torch.rand(N, n, 4, generator=g) * 60.0
    hit[:, ::7, 0] = 1e5  # sky
    img = lin.reshape(N, h_img, w_img, 4).permute(0, 3, 1, 2)
    tgt = torch.nn.functional.avg_pool2d(img, 5, 1, 2, count_include_pad=False).permute(0, 2, 3, 1).reshape(N, n, 4)
    return [t.contiguous().cuda() for t in (lin, gui, hit, tgt)]


def _bf16(dn, name, count, torch):
    x = torch.from_numpy(dn.tensor(name, count, np.uint16).astype(np.int16).view(np.int16))
    return x.view(torch.bfloat16).float()


def _weights(dn, torch, width):
    """fp32 weights as torch layouts: conv [width, 8, 3, 3], 1x1 [16, width]
    (the stored centre tap), and the full stored 1x1 [9][16][width]."""
    w0 = torch.from_numpy(dn.tensor("weight:0", 9 * width * 8, np.float32)).reshape(3, 3, width, 8)
    w1 = torch.from_numpy(dn.tensor("weight:1", 9 * 16 * width, np.float32)).reshape(9, 16, width)
    b0 = torch.from_numpy(dn.tensor("bias:0", width, np.float32))
    b1 = torch.from_numpy(dn.tensor("bias:1", 16, np.float32))
    return w0.permute(2, 3, 0, 1).contiguous(), w1[4].contiguous(), b0, b1, w1


def _ref(torch, x, w0, w1, b0, b1):
    F = torch.nn.functional

    def bfr(t):
        return t.to(torch.bfloat16).float()

    a = bfr(F.relu(F.conv2d(x, bfr(w0), b0, padding=1)))
    y = (a.permute(0, 2, 3, 1) @ bfr(w1).t().contiguous()).permute(0, 3, 1, 2)
    return bfr(y + b1[None, :, None, None])


def test_shading_features_equal_runner_features(env):
    """arbor_shadePixelsFeatures writes the compact features arbor_denoiser_
    set_input computes for the lean network from shadePixels' Linear and
    Guide and the Hit: bit for bit, sun visibility off and on."""
    torch, px = env
    print("case: compact features from shading, 200x120, seed 7", flush=True)
    w, h = 200, 120
    scene = px.DeviceScene().generate(7)
    acc = px.TraceAccel(scene)
    dn = px.Denoiser(1, 128, 208, train=False, seed=1, levels=0, widths=[16])
    cams, _ = px.batch_walk_views(5, 7, 2)
    for k, light in ((0, 0.0), (1, 1.0)):
        cam = cams[k].copy()
        cam[22] = light
        Cd = to_dev(cam)
        fb = px.FrameBuffers(w, h)
        px.trace_accel(scene, acc, Cd, fb)
        px.shade(scene, Cd, fb)
        feat = torch.zeros(w * h * 8, dtype=torch.bfloat16, device="cuda")
        px.shade_features(scene, Cd, fb, feat)
        dn.set_input(0, fb.Linear, fb.Guide, None, fb.Hit, w, h)
        torch.cuda.synchronize()
        slot = torch.from_numpy(dn.tensor("input", 128 * 208 * 8, np.uint16)).reshape(128, 208, 8)
        got = to_host(feat.view(torch.int16)).view(np.uint16).reshape(h, w, 8)
        assert np.array_equal(got, slot[:h, :w].numpy()), (light, int((got != slot[:h, :w].numpy()).sum()))
        assert (slot[h:].numpy() == 0).all() and (slot[:, w:].numpy() == 0).all()
        assert (got[..., 7] == 0).all() and got[..., :7].any()
    dn.close()


@pytest.mark.parametrize("width", [16, 32])
def test_lean_network_forward_and_gradients(env, width, tmp_path):
    torch, px = env
    print("case: lean network of %d channels, batch 2, 64x96 slots" % width, flush=True)
    dn = px.Denoiser(N, H, W, train=True, seed=4, levels=0, widths=[width])
    assert dn.shape() == (0, [width])
    assert dn.parameters() == 9 * 8 * width + width + 9 * width * 16 + 16
    lin, gui, hit, tgt = _frames(torch, 5)
    for i in range(N):
        dn.set_input(i, lin[i], gui[i], None, hit[i], w_img, h_img)
        dn.set_target(i, tgt[i], w_img, h_img)
    dn.forward(N)
    x = _bf16(dn, "input", N * H * W * 8, torch).reshape(N, H, W, 8).permute(0, 3, 1, 2)
    w0, w1, b0, b1, w1full = _weights(dn, torch, width)
    assert (w1full[torch.arange(9) != 4] == 0).all()
    ref = _ref(torch, x, w0, w1, b0, b1)
    got = _bf16(dn, "output", N * H * W * 16, torch).reshape(N, H, W, 16).permute(0, 3, 1, 2)
    rel = ((got[:, :3] - ref[:, :3]).norm() / ref[:, :3].norm()).item()
    print("forward relative L2 error %.3g" % rel)
    assert rel <= 2e-2
    params = [t.clone().requires_grad_(True) for t in (w0, w1, b0, b1)]
    out = _ref(torch, x, *params)
    t = torch.from_numpy(dn.tensor("target", N * H * W * 4, np.float32)).reshape(N, H, W, 4).permute(0, 3, 1, 2)
    loss = ((out[:, :3] - t[:, :3]).abs() * t[:, 3:4]).sum() / (N * H * W * 3)
    loss.backward()
    dn.train_step(N, 1e-3)
    assert abs(dn.loss() - loss.item()) <= 0.02 * loss.item()
    gw0 = torch.from_numpy(dn.tensor("wgrad:0", 9 * width * 8, np.float32)).reshape(3, 3, width, 8).permute(2, 3, 0, 1)
    gw1 = torch.from_numpy(dn.tensor("wgrad:1", 9 * 16 * width, np.float32)).reshape(9, 16, width)
    assert (gw1[torch.arange(9) != 4] == 0).all()
    gb0 = torch.from_numpy(dn.tensor("bgrad:0", width, np.float32))
    gb1 = torch.from_numpy(dn.tensor("bgrad:1", 16, np.float32))
    for name, a, b in (("w0", gw0, params[0].grad), ("w1", gw1[4], params[1].grad), ("b0", gb0, params[2].grad),
                       ("b1", gb1, params[3].grad)):
        nb = b.norm().item()
        if nb == 0:
            assert a.norm().item() == 0, name
            continue
        cos = (a * b).sum().item() / (a.norm().item() * nb)
        print("%s grad cos %.5f" % (name, cos))
        assert cos >= 0.99, name
    # the step keeps the 1x1 layer's other taps at 0
    _, _, _, _, w1after = _weights(dn, torch, width)
    assert (w1after[torch.arange(9) != 4] == 0).all() and not torch.equal(w1after[4], w1full[4])
    path = str(tmp_path / "lean.bin")
    dn.save(path)
    assert px.checkpoint_shape(path) == (0, [width])
    inf = px.Denoiser.from_checkpoint(path, N, H, W)
    for i in range(N):
        inf.set_input(i, lin[i], gui[i], None, hit[i], w_img, h_img)
    inf.forward(N)
    dn.forward(N)
    assert torch.equal(_bf16(inf, "output", N * H * W * 16, torch), _bf16(dn, "output", N * H * W * 16, torch))
    # a lean checkpoint does not load into a U-Net and the other way round
    unet = px.Denoiser(N, H, W, train=False, seed=1, levels=1, widths=[width])
    with pytest.raises(px.ArborError):
        unet.load(path)
    unet.save(str(tmp_path / "unet.bin"))
    with pytest.raises(px.ArborError):
        inf.load(str(tmp_path / "unet.bin"))
    for d in (inf, unet, dn):
        d.close()


@pytest.mark.parametrize("width", [16, 32])
def test_lean_one_pass_equals_layered(env, width):
    """The lean network's one-pass kernel equals the layered runner
    (set_input, forward, get_output) bit for bit: a seeded network with random
    biases on random frames in a zero-padded slot (90x61 in 96x64), through
    arbor_denoiser_denoise and through arbor_denoiser_denoise_features from
    the features the runner staged; and on a 1920x1080 tracer frame from the
    features the shading pass wrote."""
    torch, px = env
    import dn_train

    lin, gui, hit, _ = _frames(torch, 11)
    dn = px.Denoiser(1, H, W, train=False, seed=9, levels=0, widths=[width])
    g = torch.Generator(device="cuda").manual_seed(5)
    for li, co in ((0, width), (1, 16)):  # random biases: a bias mix-up shows
        b = px._wrap_device(dn.tensor_ptr("bias:%d" % li), co * 4).view(torch.float32)
        b.copy_(torch.randn(co, device="cuda", generator=g) * 0.3)
    for k in range(N):
        print("case: lean one pass, width %d, random frame %d, %dx%d in %dx%d" % (width, k, w_img, h_img, W, H),
              flush=True)
        a = torch.zeros(w_img * h_img * 4, dtype=torch.float32, device="cuda")
        b, c = torch.full_like(a, -1.0), torch.full_like(a, -2.0)
        dn.set_input(0, lin[k], gui[k], None, hit[k], w_img, h_img)
        dn.forward(1)
        dn.output(0, a, w_img, h_img)
        dn.denoise(lin[k], gui[k], None, hit[k], w_img, h_img, b)
        # the staged slot as a w x h feature image
        slot = px._wrap_device(dn.tensor_ptr("input"), H * W * 16).view(torch.bfloat16).reshape(H, W, 8)
        feat = slot[:h_img, :w_img].contiguous()
        dn.denoise_features(feat, w_img, h_img, c)
        torch.cuda.synchronize()
        assert torch.equal(a.view(torch.int32), b.view(torch.int32)), (k, int((a != b).sum()))
        assert torch.equal(a.view(torch.int32), c.view(torch.int32)), (k, int((a != c).sum()))
    dn.close()

    print("case: lean one pass, width %d, 1920x1080 tracer frame from shading's features" % width, flush=True)
    Wf, Hf = 1920, 1080
    scene = px.DeviceScene().generate(7)
    cams, _ = px.batch_walk_views(11 + 777777, 7, 2)
    cam = cams[1].copy()
    cam[22] = 1.0
    fb, _, Cd = dn_train.render_inputs(px, torch, scene, cam, Wf, Hf)
    feat = torch.zeros(Wf * Hf * 8, dtype=torch.bfloat16, device="cuda")
    px.shade_features(scene, Cd, fb, feat)
    dn = px.Denoiser(1, 1088, 1920, train=False, seed=3, levels=0, widths=[width])
    a = torch.zeros(Wf * Hf * 4, dtype=torch.float32, device="cuda")
    b = torch.full_like(a, -1.0)
    dn.set_input(0, fb.Linear, fb.Guide, None, fb.Hit, Wf, Hf)
    dn.forward(1)
    dn.output(0, a, Wf, Hf)
    dn.denoise_features(feat, Wf, Hf, b)
    torch.cuda.synchronize()
    assert torch.equal(a.view(torch.int32), b.view(torch.int32)), int((a != b).sum())
    dn.close()


def test_engine_with_lean_network(env):
    """arbor_engine_set_denoiser with the lean network: the frame equals the
    entry points chained on the same camera (trace, sun visibility, the
    shading pass's compact features, the one-pass network into Filtered,
    resolvePixels) bit for bit, in five launches; a frame larger than the
    slot falls back to filterLighting as the plain engine renders it."""
    torch, px = env
    from torch.profiler import ProfilerActivity, profile

    print("case: engine with the lean network, 328x184 in a 336x192 slot", flush=True)
    w, h = 328, 184
    dn = px.Denoiser(1, 192, 336, train=False, seed=5, levels=0, widths=[16])
    eng, plain = px.Engine(w, h, 7), px.Engine(w, h, 7)
    eng.set_denoiser(dn)
    for e in (eng, plain):
        e.set_input(px.walk_input(7, 3))
        e.frame(1.0 / 60.0)
        e.set_input(np.zeros(32, np.float32))
    before = eng.download("History", w * h * 4).copy()
    dev_in = torch.zeros(32, dtype=torch.float32, device="cuda")
    eng.sync()
    torch.cuda.synchronize()
    with profile(activities=[ProfilerActivity.CUDA]) as prof:
        eng.frame(1.0 / 60.0, dev_in)
        eng.sync()
        torch.cuda.synchronize()
    names = [e.name for e in prof.events() if e.device_type == torch.autograd.DeviceType.CUDA
             and "memcpy" not in e.name.lower() and "memset" not in e.name.lower()]
    print(len(names), "kernels:", names)
    assert len(names) == 5 and any("denoiseLean" in n for n in names)
    assert eng.inspect()["denoiser"] == "network"
    C = eng.download("C", 64).copy()
    assert C[22] == 1.0
    dev = px.DeviceScene().generate(7)
    acc = px.TraceAccel(dev)
    fb = px.FrameBuffers(w, h)
    fb.History.copy_(to_dev(before))
    Cd = to_dev(C)
    px.trace_accel(dev, acc, Cd, fb)
    vis = torch.zeros(w * h * 4, dtype=torch.float32, device="cuda")
    px.sun_visibility_accel(dev, acc, Cd, fb, vis)
    feat = torch.zeros(w * h * 8, dtype=torch.bfloat16, device="cuda")
    px.shade_features(dev, Cd, fb, feat)
    dn.denoise_features(feat, w, h, fb.Filtered)
    px.resolve(Cd, fb)
    assert np.array_equal(eng.download("Normal", w * h * 4).view(np.uint32),
                          to_host(feat.view(torch.float32)).view(np.uint32))
    for name in ("Hit", "Filtered", "History"):
        want = to_host(getattr(fb, name)).view(np.uint32)
        got = eng.download(name, w * h * 4).view(np.uint32)
        assert np.array_equal(got, want), (name, int((got != want).sum()))
    assert np.array_equal(eng.download("Pixels", w * h, np.uint32), to_host(fb.Pixels, np.uint32))
    plain.frame(1.0 / 60.0)
    assert not np.array_equal(eng.download("Filtered", w * h * 4), plain.download("Filtered", w * h * 4))

    print("case: engine frame larger than the lean network's slot", flush=True)
    W2, H2 = 384, 200
    for e in (eng, plain):
        e.resize(W2, H2)
        e.set_input(px.walk_input(7, 4))
        e.frame(1.0 / 60.0)
    assert eng.inspect()["denoiser"] == "filterLighting"
    for name in ("Filtered", "History"):
        assert np.array_equal(eng.download(name, W2 * H2 * 4).view(np.uint32),
                              plain.download(name, W2 * H2 * 4).view(np.uint32)), name
    eng.close()
    plain.close()
    dn.close()


def test_lean_checkpoint_on_1080p_tracer_frame(env):
    """The committed lean checkpoint on a 1920x1080 tracer frame (seed-7
    forest, a held-out camera, sun visibility on, one sample), run from the
    shading pass's compact features: after resolvePixels' display mapping it
    scores a higher PSNR than filterLighting against the 64-sample render."""
    torch, px = env
    import dn_train

    print("case: lean checkpoint on a 1920x1080 tracer frame", flush=True)
    Wf, Hf = 1920, 1080
    path = os.path.join(os.path.dirname(px.__file__), "weights", "denoiser_forest_lean.bin")
    scene = px.DeviceScene().generate(7)
    cams, _ = px.batch_walk_views(11 + 777777, 7, 2)
    cam = cams[0].copy()
    cam[22] = 1.0
    fb, _, Cd = dn_train.render_inputs(px, torch, scene, cam, Wf, Hf)
    target = dn_train.render_targets_batched(px, torch, scene, cam[None, :], Wf, Hf, 64)[0]
    feat = torch.zeros(Wf * Hf * 8, dtype=torch.bfloat16, device="cuda")
    px.shade_features(scene, Cd, fb, feat)
    dn = px.Denoiser.from_checkpoint(path, 1, 1088, 1920)
    assert dn.shape()[0] == 0
    out = torch.zeros(Wf * Hf * 4, dtype=torch.float32, device="cuda")
    dn.denoise_features(feat, Wf, Hf, out)
    refimg = dn_train.display(px, torch, target, Cd, Wf, Hf)
    scores = {}
    for k, buf in (("1 spp", fb.Linear), ("filterLighting", fb.Filtered), ("lean", out)):
        img = dn_train.display(px, torch, buf, Cd, Wf, Hf)
        scores[k] = dn_train.psnr(img, refimg)
    print("PSNR against 64 samples:", {k: round(v, 3) for k, v in scores.items()})
    assert scores["lean"] > scores["filterLighting"] > scores["1 spp"]
    dn.close()
