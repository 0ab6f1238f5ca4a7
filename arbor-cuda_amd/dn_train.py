"""Training and evaluation driver of the learned denoiser (include/arbor_denoise.h).

Data: the tracer renders its own training pairs on the GPU. Each crop is a
camera at a seeded position inside the 324-tree forest (seed 7), sun
visibility on, so one sample per pixel carries the soft-shadow noise that
filterLighting exists to remove.
  * input: sample 0's Linear, Guide, shading Normal (shadePixelsAOV) and Hit;
  * target: the mean of `spp` jittered samples of Linear (C[9] = s), accumulated
    by resolvePixels, i.e. the unfiltered converged image.
Targets come from the batched-view renderer (all crops traced and shaded in one
launch per sample).

Evaluation: held-out 1920x1080 frames (the walk's first camera and seeded
views not used for training). Each is compared with its `spp`-sample target
after the display mapping of resolvePixels (exposure, ACES, gamma, vignette),
by PSNR and SSIM:
  * the raw sample;
  * filterLighting applied to it;
  * the network's output.
The times of filterLighting and of the network (features, forward, output) at
1080p are measured with HIP events.

Several networks can be trained on the same data and evaluated on the same
held-out frames: --nets "5:32,48,64,96,128;2:16,32;0:16" (levels:widths, ';'
between networks; levels 0 is the lean network, whose time is that of its
one-pass kernel from the compact features the shading pass writes); the i-th
network is saved as <out> with _i before the extension when there is more
than one.

Usage: python arbor-cuda_amd/dn_train.py --crops 4096 --size 512 --spp 64 --steps 6000
       --out denoiser.bin --report report.json
(the committed checkpoint's configuration; about 12 minutes on one MI355X)
"""
import argparse
import ctypes as ct
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
MAX_DIM = 8192  # ARBOR_MAX_DIM


def _log(*a):
    print(*a, flush=True)


def render_inputs(px, torch, scene, C, w, h):
    """Sample 0 of camera C: Linear, Guide, Normal, Hit and filterLighting's output."""
    fb = px.FrameBuffers(w, h)
    normal = torch.zeros(w * h * 4, dtype=torch.float32, device="cuda")
    c = C.copy()
    c[9], c[10] = 0.0, 1.0
    Cd = torch.from_numpy(c).cuda()
    px.trace(scene, Cd, fb)
    px.shade_aov(scene, Cd, fb, normal)
    px.filter_lighting(Cd, fb)
    return fb, normal, Cd


def render_targets_batched(px, torch, scene, cams, w, h, spp):
    """Mean of `spp` samples of Linear for every camera (rows of cams), rendered
    as one batch of views per sample; returns [n, w*h*4] float32 on the GPU."""
    lib = px.lib()
    n = len(cams)
    out = torch.zeros(n, w * h * 4, dtype=torch.float32, device="cuda")
    rows = MAX_DIM // h  # views per resolve call (one tall image)
    for v0 in range(0, n, 256):
        part = cams[v0:v0 + 256].copy()
        k = len(part)
        b = px.Batch(k, k * w * h)
        hist = out[v0:v0 + k]
        pixels = torch.zeros(k * w * h, dtype=torch.int32, device="cuda")
        for s in range(spp):
            part[:, 9], part[:, 10] = float(s), float(s == 0)
            b.set_views(part, [(w, h)] * k)
            b.render(scene, px.BATCH_TRACE_SHADE)
            lin = lib.arbor_batch_buffer(b.h, 0, b"Linear")
            c0 = lib.arbor_batch_buffer(b.h, 0, b"C")
            for r0 in range(0, k, rows):
                rk = min(rows, k - r0)
                off = r0 * w * h * 16
                px.check(lib.arbor_resolvePixels(ct.c_void_p(lin + off), ct.c_void_p(hist[r0].data_ptr()),
                                                 ct.c_void_p(pixels[r0 * w * h].data_ptr()), ct.c_void_p(c0),
                                                 w, rk * h, px._stream()), "arbor_resolvePixels")
        torch.cuda.synchronize()
        b.close()
        if n > 1:
            _log("targets: views %d-%d of %d done" % (v0, v0 + k - 1, n))
    return out


def display(px, torch, radiance, Cd, w, h):
    """resolvePixels' display mapping of one radiance image -> float RGB [0,1] (h, w, 3)."""
    c = Cd.clone()
    c[10] = 1.0
    hist = torch.zeros(w * h * 4, dtype=torch.float32, device="cuda")
    pix = torch.zeros(w * h, dtype=torch.int32, device="cuda")
    px.call("arbor_resolvePixels", radiance, hist, pix, c, w, h)
    torch.cuda.synchronize()
    rgba = pix.cpu().numpy().view(np.uint8).reshape(h, w, 4)
    return rgba[:, :, :3].astype(np.float64) / 255.0


def psnr(a, b):
    mse = float(np.mean((a - b) ** 2))
    return 99.0 if mse == 0 else 10.0 * math.log10(1.0 / mse)


def ssim(a, b):
    """Mean SSIM over the three channels: Gaussian window sigma 1.5, K1 0.01,
    K2 0.03, data range 1 (Wang et al. 2004)."""
    from scipy.ndimage import gaussian_filter

    c1, c2 = 0.01 ** 2, 0.03 ** 2
    vals = []
    for ch in range(3):
        x, y = a[:, :, ch], b[:, :, ch]
        mx, my = gaussian_filter(x, 1.5, truncate=3.5), gaussian_filter(y, 1.5, truncate=3.5)
        sxx = gaussian_filter(x * x, 1.5, truncate=3.5) - mx * mx
        syy = gaussian_filter(y * y, 1.5, truncate=3.5) - my * my
        sxy = gaussian_filter(x * y, 1.5, truncate=3.5) - mx * my
        m = ((2 * mx * my + c1) * (2 * sxy + c2)) / ((mx * mx + my * my + c1) * (sxx + syy + c2))
        vals.append(float(m.mean()))
    return float(np.mean(vals))


def event_ms(torch, fn, reps=20):
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--crops", type=int, default=4096)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--views-seed", type=int, default=11)
    ap.add_argument("--eval-views", type=int, default=2)
    ap.add_argument("--eval-only", default="", help="skip training: evaluate this checkpoint")
    ap.add_argument("--nets", default="5:32,48,64,96,128",
                    help="networks to train, levels:widths separated by ';' (default: the 5-level network)")
    ap.add_argument("--out", default="denoiser.bin")
    ap.add_argument("--report", default="denoiser_report.json")
    args = ap.parse_args()

    import torch

    import pyarbor as px

    torch.manual_seed(args.seed)
    rng = np.random.default_rng(args.seed)
    scene = px.DeviceScene().generate(args.seed)
    w = h = args.size
    report = {"config": vars(args)}

    # ---- data ----
    if not args.eval_only:
        t0 = time.time()
        cams, _ = px.batch_walk_views(args.views_seed, args.seed, min(args.crops, 256))
        all_cams = [cams]
        for k in range(256, args.crops, 256):
            c, _ = px.batch_walk_views(args.views_seed + 1000 * (k // 256), args.seed, min(256, args.crops - k))
            all_cams.append(c)
        cams = np.concatenate(all_cams)[:args.crops].copy()
        cams[:, 22] = 1.0  # sun visibility on
        n = len(cams)
        hw4 = w * h * 4
        L0 = torch.zeros(n, hw4, device="cuda")
        G0 = torch.zeros_like(L0)
        N0 = torch.zeros_like(L0)
        H0 = torch.zeros_like(L0)
        for i in range(n):
            fb, normal, _ = render_inputs(px, torch, scene, cams[i], w, h)
            L0[i].copy_(fb.Linear)
            G0[i].copy_(fb.Guide)
            N0[i].copy_(normal)
            H0[i].copy_(fb.Hit)
            if (i + 1) % 512 == 0:
                _log("inputs: %d of %d crops" % (i + 1, n))
        torch.cuda.synchronize()
        t1 = time.time()
        _log("inputs: %d crops %dx%d in %.1f s" % (n, w, h, t1 - t0))
        T = render_targets_batched(px, torch, scene, cams, w, h, args.spp)
        t2 = time.time()
        rays = n * w * h * args.spp
        _log("targets: %d x %d spp in %.1f s (%.1f Mrays/s)" % (n, args.spp, t2 - t1, rays / (t2 - t1) / 1e6))
        report["data"] = {"crops": n, "size": [w, h], "spp": args.spp, "input_s": round(t1 - t0, 2),
                          "target_s": round(t2 - t1, 2), "target_mrays_per_s": round(rays / (t2 - t1) / 1e6, 1)}

        # ---- training, one network after another on the same data ----
        nets = [(int(a), [int(x) for x in b.split(",")]) for a, b in (n.split(":") for n in args.nets.split(";"))]
        root, ext = os.path.splitext(args.out)
        ckpts, report["train"] = [], []
        for ni, (levels, widths) in enumerate(nets):
            path = args.out if len(nets) == 1 else "%s_%d%s" % (root, ni, ext)
            dn = px.Denoiser(args.batch, h, w, train=True, seed=args.seed, levels=levels, widths=widths)
            losses = []
            t3 = time.time()
            for step in range(args.steps):
                lr = args.lr * (0.05 + 0.95 * 0.5 * (1 + math.cos(math.pi * step / args.steps)))
                idx = rng.choice(n, args.batch, replace=False)
                for j, i in enumerate(idx):
                    dn.set_input(j, L0[i], G0[i], N0[i], H0[i], w, h)
                    dn.set_target(j, T[i], w, h)
                dn.train_step(args.batch, lr)
                if step % 100 == 0 or step == args.steps - 1:
                    losses.append((step, dn.loss()))
                    _log("net %d step %5d loss %.5f lr %.2e  %.1f s" % (ni, step, losses[-1][1], lr, time.time() - t3))
            torch.cuda.synchronize()
            t4 = time.time()
            dn.save(path)
            report["train"].append({"levels": levels, "widths": widths, "checkpoint": os.path.basename(path),
                                    "steps": args.steps, "batch": args.batch, "seconds": round(t4 - t3, 1),
                                    "ms_per_step": round((t4 - t3) / args.steps * 1e3, 2), "loss": losses,
                                    "parameters": dn.parameters()})
            dn.close()
            ckpts.append(path)
        del L0, G0, N0, H0, T
        torch.cuda.empty_cache()
    else:
        ckpts = args.eval_only.split(";")

    # ---- evaluation at 1920x1080 ----
    W, H = 1920, 1080
    nets = [px.Denoiser.from_checkpoint(c, 1, 1088, 1920) for c in ckpts]
    eng = px.Engine(W, H, args.seed)
    I = px.walk_input(args.seed, 0)
    eng.set_input(I)
    eng.frame(1.0 / 60.0)
    walk_cam = eng.download("C", 64)
    eng.close()
    eval_cams = [("walk frame 0", walk_cam)]
    vc, _ = px.batch_walk_views(args.views_seed + 777777, args.seed, args.eval_views)
    for i in range(args.eval_views):
        eval_cams.append(("held-out view %d" % i, vc[i]))
    results = []
    for name, cam in eval_cams:
        cam = cam.copy()
        cam[22] = 1.0
        fb, normal, Cd = render_inputs(px, torch, scene, cam, W, H)
        target = render_targets_batched(px, torch, scene, cam[None, :], W, H, args.spp)[0]
        ref = display(px, torch, target, Cd, W, H)
        r = {"camera": name}
        for k, buf in (("1 spp", fb.Linear), ("filterLighting", fb.Filtered)):
            img = display(px, torch, buf, Cd, W, H)
            r[k] = {"psnr": round(psnr(img, ref), 3), "ssim": round(ssim(img, ref), 4)}
        r["filter_ms"] = round(event_ms(torch, lambda: px.filter_lighting(Cd, fb)), 4)
        r["learned"] = []
        feat = torch.zeros(W * H * 8, dtype=torch.bfloat16, device="cuda")
        px.shade_features(scene, Cd, fb, feat)  # the lean network's input, as the shading pass writes it
        for ni, inf in enumerate(nets):
            out = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")

            def run_net():
                if inf.shape()[0] == 0:
                    inf.denoise_features(feat, W, H, out)
                    return
                inf.set_input(0, fb.Linear, fb.Guide, normal, fb.Hit, W, H)
                inf.forward(1)
                inf.output(0, out, W, H)

            run_net()
            img = display(px, torch, out, Cd, W, H)
            levels, widths = inf.shape()
            r["learned"].append({"checkpoint": os.path.basename(ckpts[ni]), "levels": levels, "widths": widths,
                                 "psnr": round(psnr(img, ref), 3), "ssim": round(ssim(img, ref), 4),
                                 "denoiser_ms": round(event_ms(torch, run_net, reps=10), 4)})
        _log(json.dumps(r))
        results.append(r)
    for inf in nets:
        inf.close()
    report["eval"] = results
    with open(args.report, "w") as f:
        json.dump(report, f, indent=1)
    _log("report written to", args.report)


if __name__ == "__main__":
    main()
