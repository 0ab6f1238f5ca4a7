"""GPU numerics of the denoiser kernels (include/arbor_denoise.h) against
plain PyTorch fp32 references of the same ops, computed on the CPU from the
same bf16-rounded operands:

* conv3x3 (MFMA implicit GEMM) within bf16 output rounding of the fp32
  convolution (|err| <= 2^-8 |ref| + 2^-12 max|ref|), over ragged sizes and
  every supported output width, 8 input channels (four taps per
  instruction) included;
* conv1x1 within the same bound, 16 and 32 input channels;
* weight / bias gradients and the flipped-weight input gradient against
  autograd (fp32 sums of exact bf16 products: rtol 1e-4 of the largest);
* pooling, upsampling, concat/split, ReLU gradient, add: exact.
"""
import numpy as np
import pytest

from gpu_util import require_gpu

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    torch = require_gpu()
    import pyarbor

    return torch, pyarbor


def _bf(torch, a):
    return a.to(torch.bfloat16)


def _ours_w(torch, w):
    """torch [co, ci, 3, 3] -> ours [tap][co][ci], tap = 3*ky + kx."""
    return w.permute(2, 3, 0, 1).reshape(9, w.shape[0], w.shape[1]).contiguous()


def _nhwc(torch, x):
    return x.permute(0, 2, 3, 1).contiguous()


@pytest.mark.parametrize("n,h,w,cin,cout,relu", [
    (2, 10, 12, 16, 16, 1), (1, 7, 9, 48, 32, 0), (2, 16, 16, 32, 48, 1), (1, 5, 33, 224, 96, 1),
    (1, 8, 8, 160, 64, 0), (1, 6, 10, 80, 128, 1), (1, 3, 3, 8, 80, 0), (2, 9, 11, 112, 112, 1),
    (1, 4, 6, 96, 224, 1), (2, 17, 21, 8, 16, 1), (1, 9, 40, 8, 32, 1)])
def test_conv3x3_forward(env, n, h, w, cin, cout, relu):
    torch, px = env
    print("case: conv3x3 n%d %dx%d %d->%d relu%d" % (n, h, w, cin, cout, relu), flush=True)
    g = torch.Generator().manual_seed(cin * 1000 + cout)
    x = _bf(torch, torch.randn(n, cin, h, w, generator=g)).float()
    wt = _bf(torch, torch.randn(cout, cin, 3, 3, generator=g) * (1.0 / (3 * cin ** 0.5))).float()
    b = torch.randn(cout, generator=g) * 0.1
    ref = torch.nn.functional.conv2d(x, wt, b, padding=1)
    if relu:
        ref = torch.relu(ref)
    X = _nhwc(torch, x).to(torch.bfloat16).cuda()
    W = _ours_w(torch, wt).to(torch.bfloat16).cuda()
    B = b.float().cuda()
    Y = torch.zeros(n, h, w, cout, dtype=torch.bfloat16, device="cuda")
    px.call("arbor_dn_conv3x3", X, n, h, w, cin, W, B, relu, Y, cout)
    got = Y.float().cpu().permute(0, 3, 1, 2)
    err = (got - ref).abs()
    bound = ref.abs() * 2.0 ** -8 + ref.abs().max() * 2.0 ** -12
    print("max err %.3g (bound at max %.3g)" % (err.max(), bound.max()))
    assert (err <= bound).all()


@pytest.mark.parametrize("pixels,cin,cout,relu", [(1000, 16, 16, 0), (77, 32, 16, 1), (4099, 16, 48, 1),
                                                 (513, 32, 32, 0)])
def test_conv1x1_forward(env, pixels, cin, cout, relu):
    torch, px = env
    print("case: conv1x1 %d pixels %d->%d relu%d" % (pixels, cin, cout, relu), flush=True)
    g = torch.Generator().manual_seed(pixels + cin)
    x = _bf(torch, torch.randn(pixels, cin, generator=g)).float()
    wt = _bf(torch, torch.randn(cout, cin, generator=g) * (1.0 / cin ** 0.5)).float()
    b = torch.randn(cout, generator=g) * 0.1
    ref = x @ wt.t() + b
    if relu:
        ref = torch.relu(ref)
    Y = torch.zeros(pixels, cout, dtype=torch.bfloat16, device="cuda")
    px.call("arbor_dn_conv1x1", x.to(torch.bfloat16).cuda(), pixels, cin, wt.to(torch.bfloat16).cuda(), b.cuda(),
            relu, Y, cout)
    got = Y.float().cpu()
    err = (got - ref).abs()
    bound = ref.abs() * 2.0 ** -8 + ref.abs().max() * 2.0 ** -12
    print("max err %.3g (bound at max %.3g)" % (err.max(), bound.max()))
    assert (err <= bound).all()


@pytest.mark.parametrize("n,h,w,cin,cout", [(2, 10, 12, 16, 32), (1, 64, 70, 48, 96), (3, 17, 9, 224, 48)])
def test_conv3x3_gradients(env, n, h, w, cin, cout):
    torch, px = env
    print("case: conv3x3 gradients n%d %dx%d %d->%d" % (n, h, w, cin, cout), flush=True)
    g = torch.Generator().manual_seed(7 + cin + cout)
    x = _bf(torch, torch.randn(n, cin, h, w, generator=g)).float().requires_grad_(True)
    wt = _bf(torch, torch.randn(cout, cin, 3, 3, generator=g) * 0.1).float().requires_grad_(True)
    b = torch.zeros(cout, requires_grad=True)
    gy = _bf(torch, torch.randn(n, cout, h, w, generator=g)).float()
    y = torch.nn.functional.conv2d(x, wt, b, padding=1)
    (y * gy).sum().backward()
    X = _nhwc(torch, x.detach()).to(torch.bfloat16).cuda()
    W = _ours_w(torch, wt.detach()).to(torch.bfloat16).cuda()
    dY = _nhwc(torch, gy).to(torch.bfloat16).cuda()
    nparts = px.lib().arbor_dn_wgrad_partial_floats(n, h, w, cin, cout)
    part = torch.zeros(nparts, dtype=torch.float32, device="cuda")
    dW = torch.zeros(9, cout, cin, dtype=torch.float32, device="cuda")
    dB = torch.zeros(cout, dtype=torch.float32, device="cuda")
    px.call("arbor_dn_conv3x3_wgrad", X, dY, n, h, w, cin, cout, part, nparts, dW, dB)
    want_w = _ours_w(torch, wt.grad)
    assert (dW.cpu() - want_w).abs().max() <= 1e-4 * want_w.abs().max()
    assert (dB.cpu() - b.grad).abs().max() <= 1e-4 * b.grad.abs().max()
    # input gradient: the convolution of dY with the flipped, transposed weights
    Wf = torch.zeros(9, cin, cout, dtype=torch.bfloat16, device="cuda")
    px.call("arbor_dn_flip_weights", W, cin, cout, Wf)
    assert torch.equal(Wf.float().cpu(), W.float().cpu().flip(0).transpose(1, 2))
    dX = torch.zeros(n, h, w, cin, dtype=torch.bfloat16, device="cuda")
    px.call("arbor_dn_conv3x3", dY, n, h, w, cout, Wf, None, 0, dX, cin)  # cin = 224: two launches
    got = dX.float().cpu().permute(0, 3, 1, 2)
    ref = x.grad
    assert ((got - ref).abs() <= ref.abs() * 2.0 ** -8 + ref.abs().max() * 2.0 ** -12).all()


@pytest.mark.parametrize("n,h,w,cc,cs,cout", [(1, 10, 12, 128, 96, 96), (2, 8, 18, 48, 32, 32), (1, 34, 16, 64, 48, 48)])
def test_conv3x3_cat_equals_materialised_concat(env, n, h, w, cc, cs, cout):
    """The fused [up2(Xc), X] convolution gives the same bits as up2 + concat + conv."""
    torch, px = env
    print("case: conv3x3_cat n%d %dx%d %d+%d->%d" % (n, h, w, cc, cs, cout), flush=True)
    g = torch.Generator().manual_seed(cc + cs)
    xc = _bf(torch, torch.randn(n, h // 2, w // 2, cc, generator=g)).cuda()
    xs = _bf(torch, torch.randn(n, h, w, cs, generator=g)).cuda()
    wt = _bf(torch, torch.randn(9, cout, cc + cs, generator=g) * 0.05).cuda()
    b = (torch.randn(cout, generator=g) * 0.1).cuda()
    up = torch.zeros(n, h, w, cc, dtype=torch.bfloat16, device="cuda")
    px.call("arbor_dn_up2", xc, n, h // 2, w // 2, cc, up)
    cat = torch.zeros(n, h, w, cc + cs, dtype=torch.bfloat16, device="cuda")
    px.call("arbor_dn_concat", up, cc, xs, cs, n * h * w, cat)
    want = torch.zeros(n, h, w, cout, dtype=torch.bfloat16, device="cuda")
    px.call("arbor_dn_conv3x3", cat, n, h, w, cc + cs, wt, b, 1, want, cout)
    got = torch.zeros_like(want)
    px.call("arbor_dn_conv3x3_cat", xc, cc, xs, cs, n, h, w, wt, b, 1, got, cout)
    assert torch.equal(got.cpu(), want.cpu())


def test_resampling_concat_and_elementwise(env):
    torch, px = env
    print("case: pool/up/concat/split/relu/add", flush=True)
    g = torch.Generator().manual_seed(3)
    n, h, w, c = 2, 6, 10, 24
    x = _bf(torch, torch.randn(n, h, w, c, generator=g))
    X = x.cuda()
    Y = torch.zeros(n, h // 2, w // 2, c, dtype=torch.bfloat16, device="cuda")
    px.call("arbor_dn_pool2", X, n, h, w, c, Y)
    ref = _bf(torch, torch.nn.functional.avg_pool2d(x.float().permute(0, 3, 1, 2), 2).permute(0, 2, 3, 1))
    assert torch.equal(Y.cpu(), ref)
    dX = torch.zeros_like(X)
    px.call("arbor_dn_pool2_bwd", Y, n, h, w, c, dX)
    ref = _bf(torch, Y.float().cpu().repeat_interleave(2, 1).repeat_interleave(2, 2) * 0.25)
    assert torch.equal(dX.cpu(), ref)
    U = torch.zeros(n, 2 * h, 2 * w, c, dtype=torch.bfloat16, device="cuda")
    px.call("arbor_dn_up2", X, n, h, w, c, U)
    assert torch.equal(U.cpu(), x.repeat_interleave(2, 1).repeat_interleave(2, 2))
    dU = torch.zeros_like(X)
    px.call("arbor_dn_up2_bwd", U, n, h, w, c, dU)
    assert torch.equal(dU.cpu(), _bf(torch, x.float() * 4))
    b = _bf(torch, torch.randn(n, h, w, 16, generator=g))
    cat = torch.zeros(n, h, w, c + 16, dtype=torch.bfloat16, device="cuda")
    px.call("arbor_dn_concat", X, c, b.cuda(), 16, n * h * w, cat)
    assert torch.equal(cat.cpu(), torch.cat([x, b], dim=3))
    A2 = torch.zeros_like(X)
    B2 = torch.zeros(n, h, w, 16, dtype=torch.bfloat16, device="cuda")
    px.call("arbor_dn_split", cat, c, 16, n * h * w, A2, B2)
    assert torch.equal(A2.cpu(), x) and torch.equal(B2.cpu(), b)
    G = X.clone()
    act = torch.relu(X)
    px.call("arbor_dn_relu_bwd", G, act, G.numel())
    assert torch.equal(G.cpu(), torch.where(act.cpu().float() > 0, x, torch.zeros_like(x)))
    S = X.clone()
    px.call("arbor_dn_add", S, X, S.numel())
    assert torch.equal(S.cpu(), _bf(torch, x.float() * 2))
