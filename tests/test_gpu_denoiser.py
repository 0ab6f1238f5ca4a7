"""GPU checks of the denoiser network runner (include/arbor_denoise.h) against
a plain PyTorch fp32 restatement of the same graph (CPU), fed the runner's own
bf16 input tensor and weights, with each layer's output rounded to bf16 as the
runner stores it:

* forward: relative L2 error of the output <= 2e-2;
* one training step: every layer's weight and bias gradient against autograd
  of the same L1 loss (the runner keeps bf16 gradients between layers):
  cosine >= 0.99 and relative norm error <= 0.1;
* training reduces the loss on a small fixed problem; save/load round trip
  reproduces the outputs bit for bit.
"""
import numpy as np
import pytest

from gpu_util import require_gpu

pytestmark = pytest.mark.gpu

def net_layers(widths):
    """(cin, cout, level, relu) of every layer, as the runner builds an
    L-level network: encoder pairs, decoder pairs (deepest first), output."""
    L = len(widths)
    out = []
    for k in range(L):
        out += [(16 if k == 0 else widths[k - 1], widths[k], k, 1), (widths[k], widths[k], k, 1)]
    for k in range(L - 2, -1, -1):
        out += [(widths[k + 1] + widths[k], widths[k], k, 1), (widths[k], widths[k], k, 1)]
    return out + [(widths[0], 16, 0, 0)]


# (cin, cout, level, relu) of L0..L18 of the default network
LAYERS = net_layers([32, 48, 64, 96, 128])
assert LAYERS[10] == (224, 96, 3, 1) and LAYERS[16] == (80, 32, 0, 1) and len(LAYERS) == 19
N, H, W, w_img, h_img = 2, 64, 96, 90, 61


@pytest.fixture(scope="module")
def env():
    torch = require_gpu()
    import pyarbor

    return torch, pyarbor


def _frames(torch, seed):
    g = torch.Generator().manual_seed(seed)
    n = w_img * h_img
    lin = torch.rand(N, n, 4, generator=g) * 3.0
    gui = torch.rand(N, n, 4, generator=g)
    nor = torch.nn.functional.normalize(torch.randn(N, n, 4, generator=g)[..., :3], dim=-1)
    nor = torch.cat([nor, torch.zeros(N, n, 1)], -1)
    hit = torch.rand(N, n, 4, generator=g) * 60.0
    hit[:, ::7, 0] = 1e5  # sky
    # target: a smoothed version of the radiance
    img = lin.reshape(N, h_img, w_img, 4).permute(0, 3, 1, 2)
    tgt = torch.nn.functional.avg_pool2d(img, 5, 1, 2, count_include_pad=False).permute(0, 2, 3, 1).reshape(N, n, 4)
    return [t.contiguous().cuda() for t in (lin, gui, nor, hit, tgt)]


def _load(dn, torch, layers=LAYERS):
    Ws, Bs = [], []
    for i, (ci, co, _, _) in enumerate(layers):
        w = torch.from_numpy(dn.tensor("weight:%d" % i, 9 * co * ci, np.float32))
        b = torch.from_numpy(dn.tensor("bias:%d" % i, co, np.float32))
        Ws.append(w.reshape(3, 3, co, ci).permute(2, 3, 0, 1).contiguous())
        Bs.append(b)
    return Ws, Bs


def _ref_forward(torch, x, Ws, Bs, layers=LAYERS):
    F = torch.nn.functional
    L = (len(layers) + 1) // 4

    def bfr(t):
        return t.to(torch.bfloat16).float()

    def conv(i, t):
        y = F.conv2d(t, bfr(Ws[i]), Bs[i], padding=1)
        return bfr(F.relu(y) if layers[i][3] else y)

    def up(t):
        return t.repeat_interleave(2, 2).repeat_interleave(2, 3)

    skips, t = [], x
    for k in range(L):
        a = conv(2 * k, t)
        s = conv(2 * k + 1, a)
        if k < L - 1:
            skips.append(s)
            t = bfr(F.avg_pool2d(s, 2))
        else:
            t = s
    for j, k in enumerate(range(L - 2, -1, -1)):
        c = torch.cat([up(t), skips[k]], 1)
        t = conv(2 * L + 1 + 2 * j, conv(2 * L + 2 * j, c))
    return conv(4 * L - 2, t)


def _input(dn, torch):
    x = torch.from_numpy(dn.tensor("input", N * H * W * 16, np.uint16).astype(np.int16).view(np.int16))
    return x.view(torch.bfloat16).float().reshape(N, H, W, 16).permute(0, 3, 1, 2)


def _output(dn, torch):
    y = torch.from_numpy(dn.tensor("output", N * H * W * 16, np.uint16).astype(np.int16).view(np.int16))
    return y.view(torch.bfloat16).float().reshape(N, H, W, 16).permute(0, 3, 1, 2)


def test_forward_and_gradients_match_torch(env):
    torch, px = env
    print("case: denoiser forward + gradients, batch 2, 64x96 slots", flush=True)
    dn = px.Denoiser(N, H, W, train=True, seed=3)
    assert dn.parameters() == 1021904
    lin, gui, nor, hit, tgt = _frames(torch, 1)
    for i in range(N):
        dn.set_input(i, lin[i], gui[i], nor[i], hit[i], w_img, h_img)
        dn.set_target(i, tgt[i], w_img, h_img)
    dn.forward(N)
    x = _input(dn, torch)
    assert x[:, 10:].abs().max() == 0 and x[:, :, h_img:].abs().max() == 0 and x[:, :, :, w_img:].abs().max() == 0
    Ws, Bs = _load(dn, torch)
    ref = _ref_forward(torch, x, Ws, Bs)
    got = _output(dn, torch)
    rel = ((got[:, :3] - ref[:, :3]).norm() / ref[:, :3].norm()).item()
    print("forward relative L2 error %.3g" % rel)
    assert rel <= 2e-2
    assert got[:, 3:].abs().max() == 0
    # gradients of one step (taken before the step's weight update)
    Ws_g = [w.clone().requires_grad_(True) for w in Ws]
    Bs_g = [b.clone().requires_grad_(True) for b in Bs]
    out = _ref_forward(torch, x, Ws_g, Bs_g)
    t = torch.from_numpy(dn.tensor("target", N * H * W * 4, np.float32)).reshape(N, H, W, 4).permute(0, 3, 1, 2)
    loss = ((out[:, :3] - t[:, :3]).abs() * t[:, 3:4]).sum() / (N * H * W * 3)
    loss.backward()
    dn.train_step(N, 1e-3)
    print("loss runner %.5f torch %.5f" % (dn.loss(), loss.item()))
    assert abs(dn.loss() - loss.item()) <= 0.02 * loss.item()
    for i, (ci, co, _, _) in enumerate(LAYERS):
        gw = torch.from_numpy(dn.tensor("wgrad:%d" % i, 9 * co * ci, np.float32)).reshape(3, 3, co, ci).permute(
            2, 3, 0, 1)
        gb = torch.from_numpy(dn.tensor("bgrad:%d" % i, co, np.float32))
        for name, a, b in (("w", gw, Ws_g[i].grad), ("b", gb, Bs_g[i].grad)):
            nb = b.norm().item()
            if nb == 0:
                assert a.norm().item() == 0
                continue
            cos = (a * b).sum().item() / (a.norm().item() * nb)
            relerr = (a - b).norm().item() / nb
            print("L%d %s grad cos %.5f relerr %.4f" % (i, name, cos, relerr))
            assert cos >= 0.99 and relerr <= 0.1, (i, name)
    dn.close()


def test_training_reduces_loss_and_checkpoint_roundtrip(env, tmp_path):
    torch, px = env
    print("case: denoiser training 60 steps + save/load", flush=True)
    dn = px.Denoiser(N, H, W, train=True, seed=5)
    lin, gui, nor, hit, tgt = _frames(torch, 2)
    for i in range(N):
        dn.set_input(i, lin[i], gui[i], nor[i], hit[i], w_img, h_img)
        dn.set_target(i, tgt[i], w_img, h_img)
    losses = []
    for step in range(60):
        dn.train_step(N, 2e-3)
        losses.append(dn.loss())
    print("loss", ["%.4f" % v for v in losses[::10]], "final %.4f" % losses[-1])
    assert losses[-1] < 0.5 * losses[0]
    path = tmp_path / "dn.bin"
    dn.save(path)
    dn.forward(N)
    want = _output(dn, torch)
    inf = px.Denoiser(N, H, W, train=False, seed=99)
    inf.load(path)
    for i in range(N):
        inf.set_input(i, lin[i], gui[i], nor[i], hit[i], w_img, h_img)
    inf.forward(N)
    assert torch.equal(_output(inf, torch), want)
    out = torch.zeros(w_img * h_img * 4, dtype=torch.float32, device="cuda")
    inf.output(1, out, w_img, h_img)
    o = out.cpu().reshape(-1, 4)
    assert torch.allclose(o[:, :3], torch.expm1(want[1, :3, :h_img, :w_img].permute(1, 2, 0).reshape(-1, 3)).clamp(
        min=0), rtol=1e-6, atol=1e-6)
    assert (o[:, 3] == 1).all()
    inf.close()
    dn.close()


def test_checkpoint_on_1080p_tracer_frame(env):
    """The learned-denoiser workload at its size: the committed checkpoint on a
    1920x1080 frame the tracer renders (seed-7 forest, a held-out camera, sun
    visibility on, one sample). Its output matches the fp32 PyTorch
    restatement of the same graph (relative L2 <= 2e-2 on the runner's own
    bf16 input), and after resolvePixels' display mapping it scores a higher
    PSNR and SSIM than filterLighting against the 64-sample render."""
    import os
    import sys

    torch, px = env
    sys.path.insert(0, os.path.dirname(px.__file__))
    import dn_train

    print("case: denoiser checkpoint on a 1920x1080 tracer frame", flush=True)
    W, Hh = 1920, 1080
    scene = px.DeviceScene().generate(7)
    cams, _ = px.batch_walk_views(11 + 777777, 7, 2)
    cam = cams[0].copy()
    cam[22] = 1.0
    fb, normal, Cd = dn_train.render_inputs(px, torch, scene, cam, W, Hh)
    target = dn_train.render_targets_batched(px, torch, scene, cam[None, :], W, Hh, 64)[0]
    inf = px.Denoiser(1, 1088, 1920, train=False, seed=0)
    inf.load(os.path.join(os.path.dirname(px.__file__), "weights", "denoiser_forest.bin"))
    out = torch.zeros(W * Hh * 4, dtype=torch.float32, device="cuda")
    inf.set_input(0, fb.Linear, fb.Guide, normal, fb.Hit, W, Hh)
    inf.forward(1)
    inf.output(0, out, W, Hh)
    torch.cuda.synchronize()
    x = torch.from_numpy(inf.tensor("input", 1088 * 1920 * 16, np.uint16).astype(np.int16).view(np.int16))
    x = x.view(torch.bfloat16).float().reshape(1, 1088, 1920, 16).permute(0, 3, 1, 2)
    y = torch.from_numpy(inf.tensor("output", 1088 * 1920 * 16, np.uint16).astype(np.int16).view(np.int16))
    y = y.view(torch.bfloat16).float().reshape(1, 1088, 1920, 16).permute(0, 3, 1, 2)
    Ws, Bs = _load(inf, torch)
    with torch.no_grad():
        ref = _ref_forward(torch, x, Ws, Bs)
    rel = ((y[:, :3] - ref[:, :3]).norm() / ref[:, :3].norm()).item()
    print("forward relative L2 error at 1088x1920: %.4g" % rel)
    assert rel <= 2e-2
    refimg = dn_train.display(px, torch, target, Cd, W, Hh)
    scores = {}
    for k, buf in (("1 spp", fb.Linear), ("filterLighting", fb.Filtered), ("learned", out)):
        img = dn_train.display(px, torch, buf, Cd, W, Hh)
        scores[k] = (dn_train.psnr(img, refimg), dn_train.ssim(img, refimg))
    print("PSNR / SSIM against 64 samples:", {k: (round(a, 3), round(b, 4)) for k, (a, b) in scores.items()})
    assert scores["learned"][0] > scores["filterLighting"][0] > scores["1 spp"][0]
    assert scores["learned"][1] > scores["filterLighting"][1]
    inf.close()


def test_bad_checkpoint_leaves_network_unchanged(env, tmp_path):
    """A truncated checkpoint, or one with bytes after its last layer, is
    rejected before any weight changes: the network's output stays bit for
    bit what it was."""
    torch, px = env
    print("case: denoiser truncated / over-long checkpoint", flush=True)
    dn = px.Denoiser(N, H, W, train=True, seed=7)
    lin, gui, nor, hit, _ = _frames(torch, 3)
    for i in range(N):
        dn.set_input(i, lin[i], gui[i], nor[i], hit[i], w_img, h_img)
    good = tmp_path / "good.bin"
    dn.save(good)
    dn.forward(N)
    want = _output(dn, torch)
    data = good.read_bytes()
    other = px.Denoiser(N, H, W, train=True, seed=8)
    other.save(tmp_path / "other.bin")
    odata = (tmp_path / "other.bin").read_bytes()
    for name, blob in (("trunc", odata[: len(odata) * 3 // 4]), ("long", odata + b"\0"), ("short", odata[:10])):
        path = tmp_path / (name + ".bin")
        path.write_bytes(blob)
        with pytest.raises(px.ArborError):
            dn.load(path)
        dn.forward(N)
        assert torch.equal(_output(dn, torch), want), name
    assert len(data) == len(odata)
    dn.load(tmp_path / "other.bin")
    dn.forward(N)
    assert not torch.equal(_output(dn, torch), want)
    dn.close()
    other.close()


@pytest.mark.parametrize("widths", [[16, 32], [16]])
def test_small_network_forward_and_gradients(env, widths):
    """A network of other levels and widths (arbor_denoiser_create_net): the
    same forward and gradient bars against the PyTorch restatement of its
    graph, and its shape read back from the runner and from a checkpoint."""
    torch, px = env
    print("case: denoiser %d levels %s, batch 2, 64x96 slots" % (len(widths), widths), flush=True)
    layers = net_layers(widths)
    dn = px.Denoiser(N, H, W, train=True, seed=4, levels=len(widths), widths=widths)
    assert dn.shape() == (len(widths), widths)
    assert dn.parameters() == sum(9 * ci * co + co for ci, co, _, _ in layers)
    lin, gui, nor, hit, tgt = _frames(torch, 5)
    for i in range(N):
        dn.set_input(i, lin[i], gui[i], nor[i], hit[i], w_img, h_img)
        dn.set_target(i, tgt[i], w_img, h_img)
    dn.forward(N)
    x = _input(dn, torch)
    Ws, Bs = _load(dn, torch, layers)
    ref = _ref_forward(torch, x, Ws, Bs, layers)
    got = _output(dn, torch)
    rel = ((got[:, :3] - ref[:, :3]).norm() / ref[:, :3].norm()).item()
    print("forward relative L2 error %.3g" % rel)
    assert rel <= 2e-2
    Ws_g = [w.clone().requires_grad_(True) for w in Ws]
    Bs_g = [b.clone().requires_grad_(True) for b in Bs]
    out = _ref_forward(torch, x, Ws_g, Bs_g, layers)
    t = torch.from_numpy(dn.tensor("target", N * H * W * 4, np.float32)).reshape(N, H, W, 4).permute(0, 3, 1, 2)
    loss = ((out[:, :3] - t[:, :3]).abs() * t[:, 3:4]).sum() / (N * H * W * 3)
    loss.backward()
    dn.train_step(N, 1e-3)
    assert abs(dn.loss() - loss.item()) <= 0.02 * loss.item()
    for i, (ci, co, _, _) in enumerate(layers):
        gw = torch.from_numpy(dn.tensor("wgrad:%d" % i, 9 * co * ci, np.float32)).reshape(3, 3, co, ci).permute(
            2, 3, 0, 1)
        gb = torch.from_numpy(dn.tensor("bgrad:%d" % i, co, np.float32))
        for name, a, b in (("w", gw, Ws_g[i].grad), ("b", gb, Bs_g[i].grad)):
            nb = b.norm().item()
            if nb == 0:
                assert a.norm().item() == 0
                continue
            cos = (a * b).sum().item() / (a.norm().item() * nb)
            print("L%d %s grad cos %.5f" % (i, name, cos))
            assert cos >= 0.99, (i, name)
    import tempfile

    with tempfile.TemporaryDirectory() as tmp:
        path = tmp + "/small.bin"
        dn.save(path)
        assert px.checkpoint_shape(path) == (len(widths), widths)
        inf = px.Denoiser.from_checkpoint(path, N, H, W)
        for i in range(N):
            inf.set_input(i, lin[i], gui[i], nor[i], hit[i], w_img, h_img)
        inf.forward(N)
        dn.forward(N)
        assert torch.equal(_output(inf, torch), _output(dn, torch))
        inf.close()
    dn.close()


@pytest.mark.parametrize("levels,widths", [(2, [16, 32]), (1, [16])])
def test_engine_denoises_with_network_and_falls_back(env, levels, widths):
    """arbor_engine_set_denoiser: a frame that fits the network's slot equals
    the entry points chained one by one on the same camera (trace, sun
    visibility, shadePixelsVis, the normal of shadePixelsAOV, the network's
    input / forward / output into Filtered, resolvePixels), bit for bit; a
    frame larger than the slot, and every frame after detaching, equals the
    filterLighting engine's bit for bit. The 1-level 16-channel network runs
    as the one-pass kernel with the resolve fused."""
    torch, px = env
    from gpu_util import to_dev, to_host

    print("case: engine with a %d-level network attached, 328x184 in a 336x192 slot" % levels, flush=True)
    w, h = 328, 184
    dn = px.Denoiser(1, 192, 336, train=False, seed=5, levels=levels, widths=widths)
    assert dn.slot() == (1, 192, 336)
    eng, plain = px.Engine(w, h, 7), px.Engine(w, h, 7)
    assert not eng.buffer_ptr("Normal")
    eng.set_denoiser(dn)
    assert eng.buffer_ptr("Normal")
    for e in (eng, plain):
        e.set_input(px.walk_input(7, 3))
        e.frame(1.0 / 60.0)
        e.set_input(np.zeros(32, np.float32))
    before = eng.download("History", w * h * 4).copy()
    eng.frame(1.0 / 60.0)
    plain.frame(1.0 / 60.0)
    assert eng.inspect()["denoiser"] == "network" and plain.inspect()["denoiser"] == "filterLighting"
    C = eng.download("C", 64).copy()
    assert C[22] == 1.0 and C[9] == 1.0
    dev = px.DeviceScene().generate(7)
    acc = px.TraceAccel(dev)
    fb, aov = px.FrameBuffers(w, h), px.FrameBuffers(w, h)
    fb.History.copy_(to_dev(before))
    Cd = to_dev(C)
    px.trace_accel(dev, acc, Cd, fb)
    vis = torch.zeros(w * h * 4, dtype=torch.float32, device="cuda")
    normal = torch.zeros(w * h * 4, dtype=torch.float32, device="cuda")
    px.sun_visibility_accel(dev, acc, Cd, fb, vis)
    px.shade_vis(dev, Cd, fb, vis)
    aov.Hit.copy_(fb.Hit)
    px.shade_aov(dev, Cd, aov, normal)
    dn.set_input(0, fb.Linear, fb.Guide, normal, fb.Hit, w, h)
    dn.forward(1)
    dn.output(0, fb.Filtered, w, h)
    px.resolve(Cd, fb)
    for name in ("Hit", "Linear", "Guide", "Filtered", "History"):
        want = to_host(getattr(fb, name)).view(np.uint32)
        got = eng.download(name, w * h * 4).view(np.uint32)
        assert np.array_equal(got, want), (name, int((got != want).sum()))
    assert np.array_equal(eng.download("Normal", w * h * 4).view(np.uint32), to_host(normal).view(np.uint32))
    assert np.array_equal(eng.download("Pixels", w * h, np.uint32), to_host(fb.Pixels, np.uint32))
    # the network's image is not the filter's
    assert not np.array_equal(eng.download("Filtered", w * h * 4), plain.download("Filtered", w * h * 4))

    print("case: engine frame larger than the slot, then detached", flush=True)
    W2, H2 = 384, 200
    for e in (eng, plain):
        e.resize(W2, H2)
        e.set_input(px.walk_input(7, 4))
    for step in range(2):
        if step == 1:
            eng.set_denoiser(None)
            assert not eng.buffer_ptr("Normal")
        eng.frame(1.0 / 60.0)
        plain.frame(1.0 / 60.0)
        assert eng.inspect()["denoiser"] == "filterLighting"
        for name in ("Filtered", "History"):
            assert np.array_equal(eng.download(name, W2 * H2 * 4).view(np.uint32),
                                  plain.download(name, W2 * H2 * 4).view(np.uint32)), (step, name)
        assert np.array_equal(eng.download("Pixels", W2 * H2, np.uint32), plain.download("Pixels", W2 * H2, np.uint32))
    eng.close()
    plain.close()
    dn.close()


def test_one_level_network_in_one_pass_equals_layered(env):
    """arbor_denoiser_denoise runs a 1-level 16-channel network as one fused
    kernel (features, three convolutions and output, intermediates in LDS).
    Its output equals the layered runner's (set_input, forward, get_output)
    bit for bit: the committed small checkpoint and a seeded He-initialised
    network with random biases, on random frames in a zero-padded slot (90x61
    in 96x64: image and slot borders inside the tiles) and on a 1920x1080
    tracer frame in a 1088x1920 slot. A 2-level network takes the layered
    path through the same call."""
    import os
    import sys

    torch, px = env
    sys.path.insert(0, os.path.dirname(px.__file__))
    import dn_train

    small = os.path.join(os.path.dirname(px.__file__), "weights", "denoiser_forest_small.bin")
    lin, gui, nor, hit, _ = _frames(torch, 11)
    nets = [("checkpoint", px.Denoiser.from_checkpoint(small, 1, H, W)),
            ("seeded", px.Denoiser(1, H, W, train=False, seed=9, levels=1, widths=[16])),
            ("2-level", px.Denoiser(1, H, W, train=False, seed=9, levels=2, widths=[16, 32]))]
    g = torch.Generator(device="cuda").manual_seed(5)
    for li in range(3):  # random biases: a bias mix-up shows
        b = px._wrap_device(nets[1][1].tensor_ptr("bias:%d" % li), 16 * 4).view(torch.float32)
        b.copy_(torch.randn(16, device="cuda", generator=g) * 0.3)
    for name, dn in nets:
        for k in range(N):
            print("case: one-pass denoise, %s network, random frame %d, %dx%d in %dx%d" % (name, k, w_img, h_img, W, H),
                  flush=True)
            a = torch.zeros(w_img * h_img * 4, dtype=torch.float32, device="cuda")
            b = torch.full_like(a, -1.0)
            dn.set_input(0, lin[k], gui[k], nor[k], hit[k], w_img, h_img)
            dn.forward(1)
            dn.output(0, a, w_img, h_img)
            dn.denoise(lin[k], gui[k], nor[k], hit[k], w_img, h_img, b)
            torch.cuda.synchronize()
            assert torch.equal(a.view(torch.int32), b.view(torch.int32)), (name, k, int((a != b).sum()))
        dn.close()

    print("case: one-pass denoise, checkpoint, 1920x1080 tracer frame", flush=True)
    Wf, Hf = 1920, 1080
    scene = px.DeviceScene().generate(7)
    cams, _ = px.batch_walk_views(11 + 777777, 7, 2)
    cam = cams[1].copy()
    cam[22] = 1.0
    fb, normal, _ = dn_train.render_inputs(px, torch, scene, cam, Wf, Hf)
    dn = px.Denoiser.from_checkpoint(small, 1, 1088, 1920)
    a = torch.zeros(Wf * Hf * 4, dtype=torch.float32, device="cuda")
    b = torch.full_like(a, -1.0)
    dn.set_input(0, fb.Linear, fb.Guide, normal, fb.Hit, Wf, Hf)
    dn.forward(1)
    dn.output(0, a, Wf, Hf)
    dn.denoise(fb.Linear, fb.Guide, normal, fb.Hit, Wf, Hf, b)
    torch.cuda.synchronize()
    assert torch.equal(a.view(torch.int32), b.view(torch.int32)), int((a != b).sum())
    dn.close()


def test_engine_frame_with_one_pass_network_is_five_launches(env):
    """With a 1-level 16-channel network attached the engine's frame stays
    five kernel launches: the network and the resolve are one kernel."""
    torch, px = env
    from torch.profiler import ProfilerActivity, profile

    print("case: launches per engine frame with the one-pass network, 320x192", flush=True)
    eng = px.Engine(320, 192, 7)
    dn = px.Denoiser(1, 192, 320, train=False, seed=5, levels=1, widths=[16])
    eng.set_denoiser(dn)
    dev_in = torch.from_numpy(px.walk_input(7, 0)).cuda()
    eng.frame(1.0 / 60.0, dev_in)
    eng.sync()
    torch.cuda.synchronize()
    with profile(activities=[ProfilerActivity.CUDA]) as prof:
        eng.frame(1.0 / 60.0, dev_in)
        eng.sync()
        torch.cuda.synchronize()
    names = [e.name for e in prof.events() if e.device_type == torch.autograd.DeviceType.CUDA
             and "memcpy" not in e.name.lower() and "memset" not in e.name.lower()]
    print(len(names), "kernels:", names)
    assert len(names) == 5 and any("denoise1" in n for n in names)
    assert eng.inspect()["denoiser"] == "network"
    eng.close()
    dn.close()
