"""Throughput of the ARBOR render path on one MI355X.

Workload (the full-forest configuration): seed 7, 18x18 = 324 instanced trees
sharing one resident BVH, 1920x1080, 1 spp primary rays, shade, filterLighting
and resolve, tree-to-tree sun visibility off, breeze off, along a seeded
120-frame WASD walk (dt = 1/60 s). A step is one frame of the C++ engine:
updateCamera -> tracePixels -> shadePixels -> filterLighting -> resolvePixels
in five kernel launches, inputs resident in HBM (the walk's 32-float input
rows are uploaded before the timed region). The timed steps are walk frames
0..steps-1; the whole walk's per-frame GPU time is reported beside them.

Prints one JSON line. `value` is end-to-end primary Mrays/s (W*H per frame /
frame time); the trace kernel's own rate, per-stage ms, its roofline fraction
and the reference CPU harness timed on this host are reported beside it.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "arbor-cuda_amd"))

# Algorithmic FP32 operations per primary ray of the engine's trace on this
# workload, averaged over every 8th frame of the 120-frame walk at every 8th
# pixel: `python oracle/count_trace_ops.py` (the oracle's counting
# instantiation, oracle/flops.cpp: each +,-,*,/,sqrt,min,max,abs,floor and libm
# call counts 1, with the GPU's walk over the SAH hierarchy and its tie-order
# box tests, rejection tests and early exits, including the leaf Newton
# solve's fixed-point and 2-cycle exits, the skipped band-weighted terms whose
# band weight is 0, and the stored leaf basis with the breeze off). The heap
# BVH walk of arbor_tracePixels does 8785.9 (`--heap`) for the same Hit.
TRACE_FLOP_PER_RAY = 5030.3
TRACE_FLOP_PER_RAY_HEAP = 8785.9
# HBM bytes per scheduled SAH-walk trace launch on this workload (walk frame
# 0, the default launch shape): rocprofv3 --pmc FETCH_SIZE 32,511 KiB x 2.0 +
# WRITE_SIZE 33,352 KiB x 1.0, the factors being what the same counters read
# for a 256 MiB device copy (131,093 KiB and 262,144 KiB);
# profiles/tracePixels_forest1080p_pmc.md.
TRACE_HBM_BYTES = (32511.0 * 2.0 + 33352.0 * 1.0) * 1024
FP32_PEAK_TFLOPS = 157.3
# One non-packed FP32 VALU op per lane per clock: 256 CUs x 64 lanes x 2.4 GHz.
FP32_SINGLE_ISSUE_TOPS = 39.32
HBM_PEAK_GBS = 8000.0
# Frames of the seeded walk (the full-forest workload's camera path).
WALK_FRAMES = 120


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=120)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--profile-trace", type=int, default=0, metavar="N",
                    help="profiling driver: N scheduled SAH-walk tracePixels launches of walk frame 0 (the "
                         "engine's trace call: tile-order kernels + trace kernel), then a 256 MiB device copy "
                         "(the calibration kernel for FETCH_SIZE/WRITE_SIZE); prints no JSON")
    ap.add_argument("--accel-variant", type=int, default=-1, metavar="V",
                    help="SAH-walk trace launch shape (the tuning key accel_variant) for this run; -1 keeps the "
                         "library default")
    ap.add_argument("--accel-vis-variant", type=int, default=-1, metavar="V",
                    help="SAH-walk sun-visibility launch shape (the tuning key accel_vis_variant) for this run; -1 "
                         "keeps the library default")
    ap.add_argument("--batch-variant", type=int, default=-1, metavar="V",
                    help="batched-views launch shape (the tuning key batch_variant) for this run; -1 keeps the "
                         "library default")
    ap.add_argument("--views", type=int, default=0, metavar="N",
                    help="batched-views workload instead: N seeded walk views of uneven resolution rendered by one "
                         "hipGraph-captured batch per step (about 0.3 s per step at N=64; use --steps 10)")
    ap.add_argument("--views-seed", type=int, default=3)
    ap.add_argument("--shadows4k", action="store_true",
                    help="sun-visibility workload instead: 3840x2160, 4 jittered samples per frame, tree-to-tree "
                         "sun visibility on at 35 deg elevation, breeze on, static initial camera (use --steps 10)")
    ap.add_argument("--profile-denoiser", type=int, default=0, metavar="N",
                    help="profiling driver: N learned-denoiser passes (features, forward, output; for the lean "
                         "network its one pass from compact features) at 1920x1080 on fixed random inputs with "
                         "the network of --denoiser (default: the 5-level checkpoint); prints no JSON")
    ap.add_argument("--denoiser", nargs="?", const=",".join(DENOISER_NETS), default="", metavar="CHECKPOINTS",
                    help="learned-denoiser workload instead: held-out 1920x1080 frames (sun visibility on), "
                         "1-sample inputs against 64-sample targets; filterLighting vs each network of the "
                         "comma-separated checkpoints (default: the lean network, the 1-level U-Net, then the "
                         "5-level one); the first one's ms is the value")
    return ap.parse_args()


WEIGHTS = os.path.join(HERE, "arbor-cuda_amd", "weights")
DENOISER_WEIGHTS = os.path.join(WEIGHTS, "denoiser_forest.bin")
# The smallest U-Net at or above filterLighting's PSNR on the held-out frames.
DENOISER_SMALL = os.path.join(WEIGHTS, "denoiser_forest_small.bin")
# The lean network (one 3x3 convolution and a 1x1 output, compact features
# written by the shading pass): at or above filterLighting's PSNR, one pass.
DENOISER_LEAN = os.path.join(WEIGHTS, "denoiser_forest_lean.bin")
DENOISER_NETS = (DENOISER_LEAN, DENOISER_SMALL, DENOISER_WEIGHTS)
# BF16 dense MFMA peak of MI355X (TFLOP/s).
BF16_PEAK_TFLOPS = 2500.0


def denoiser_layers(widths):
    """(cin, cout, level, cat) of every convolution of the U-Net of `widths`
    channels per level, in the runner's order (encoder pairs, decoder pairs
    deepest first, output layer); cat = channels that come upsampled from the
    level below (0 for a plain convolution). cin counts the 10 used input
    features, cout of the output layer its 3 used channels."""
    L = len(widths)
    out = []
    for k in range(L):
        out += [(10 if k == 0 else widths[k - 1], widths[k], k, 0), (widths[k], widths[k], k, 0)]
    for k in range(L - 2, -1, -1):
        out += [(widths[k + 1] + widths[k], widths[k], k, widths[k + 1]), (widths[k], widths[k], k, 0)]
    return out + [(widths[0], 3, 0, 0)]


def denoiser_flops(height, width, widths=(32, 48, 64, 96, 128)):
    """Algorithmic FLOPs of one forward pass at height x width (multiples of
    16): 2 * 9 * cin * cout per output pixel of every convolution."""
    return sum(2.0 * 9 * ci * co * (height >> lv) * (width >> lv) for ci, co, lv, _ in denoiser_layers(widths))


def lean_flops(h, w, width):
    """Algorithmic FLOPs of the lean network on an h x w image: the 3x3
    convolution of the 7 used features into `width` channels and the 1x1
    layer into the 3 used outputs."""
    return 2.0 * (9 * 7 * width + width * 3) * h * w


def denoiser_bytes(height, width, h, w, widths):
    """Algorithmic HBM bytes of one denoiser pass (input features, forward,
    output) for an h x w image in a height x width slot, tensors as the runner
    stores them (bf16, channels padded to 16 at input and output): the
    feature kernel reads Linear, Guide, Normal and Hit (64 B a pixel) and
    writes the 16-channel input; each convolution reads its input once (the
    concatenating ones the coarse tensor at its own size) and writes its
    output; pooling reads a skip and writes a quarter; the output kernel reads
    3 channels and writes float4 radiance."""
    px = lambda lv: (height >> lv) * (width >> lv)
    total = 64 * h * w + 32 * px(0) + 6 * h * w + 16 * h * w
    layers = denoiser_layers(widths)
    for i, (ci, co, lv, cat) in enumerate(layers):
        cin = 16 if i == 0 else ci
        cout = 16 if i == len(layers) - 1 else co
        total += 2 * (cin - cat) * px(lv) + 2 * cat * px(lv + 1) + 2 * cout * px(lv)
    for k in range(len(widths) - 1):  # 2x2 pooling of every skip
        total += 2 * widths[k] * px(k) + 2 * widths[k] * px(k + 1)
    return float(total)


def bench_shadows4k(args):
    """3840x2160 at 4 samples per frame with sun visibility (one shadow ray,
    and up to 3 more through leaves, per hit pixel), sun at 35 deg, breeze on
    (wind time +1/60 s per frame), through the engine's multi-sample frame."""
    import numpy as np
    import torch

    import pyarbor as px

    W, H, S, K, Wm = 3840, 2160, 4, args.steps, args.warmup
    eng = px.Engine(W, H, args.seed)
    C = px._wrap_device(eng.buffer_ptr("C"), 256).view(torch.float32)
    C[13] = 0.610865  # 35 deg sun elevation
    torch.cuda.synchronize()
    first = torch.zeros(32, dtype=torch.float32, device="cuda")
    first[7] = 1.0  # breeze on (updateCamera I[7])
    still = torch.zeros(32, dtype=torch.float32, device="cuda")
    eng.frame_samples(S, 1.0 / 60.0, first)
    for _ in range(Wm):
        eng.frame_samples(S, 1.0 / 60.0, still)
    eng.sync()
    c = eng.download("C", 64)
    assert c[22] == 1.0 and abs(c[11] - 0.6) < 1e-6, "sun visibility and breeze must be on"
    # The SAH-walk autotuner times every trace launch shape on this view (the
    # engine's own buffers) and keeps the fastest, before the timed frames.
    import ctypes as ct

    tj = ct.create_string_buffer(4096)
    bp = [ct.c_void_p(eng.buffer_ptr(n)) for n in ("P", "Nodes", "Order", "Accel", "C", "Hit")]
    px.check(px.lib().arbor_autotune_trace_accel(*bp, W, H, 1, px._stream(), tj, 4096), "arbor_autotune_trace_accel")
    tuned = json.loads(tj.value.decode())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(K):
        eng.frame_samples(S, 1.0 / 60.0, still)
    eng.sync()
    elapsed = time.perf_counter() - t0
    eng.close()
    rays = W * H * S
    result = {
        "metric": "Mrays/s, ms/frame @3840x2160 4 spp sun visibility on",
        "value": round(rays * K / elapsed / 1e6, 3), "unit": "Mrays/s (primary samples)",
        "n_gpus": 1, "steps": K, "warmup": Wm, "ms_per_step": round(elapsed / K * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (procedural forest from seed, static initial camera)",
        "config": {"workload": "324 trees, 3840x2160, 4 jittered samples per frame, tree-to-tree sun visibility on, "
                               "sun elevation 35 deg, breeze on (wind time +1/60 s per frame)",
                   "width": W, "height": H, "samples": S, "seed": args.seed, "frames": K},
        "autotune_accel_variant": tuned,
    }
    print(json.dumps(result), flush=True)


def profile_denoiser(args):
    import torch

    import pyarbor as px

    W, H = 1920, 1080
    path = args.denoiser.split(",")[0] if args.denoiser else DENOISER_WEIGHTS
    inf = px.Denoiser.from_checkpoint(path, 1, 1088, 1920)
    g = torch.Generator(device="cuda").manual_seed(0)
    bufs = [torch.rand(W * H * 4, device="cuda", generator=g) for _ in range(4)]
    out = torch.zeros(W * H * 4, device="cuda")
    levels, widths = inf.shape()
    feat = torch.rand(W * H * 8, device="cuda", generator=g).to(torch.bfloat16)  # the lean network's input
    for _ in range(args.profile_denoiser):
        if levels == 0:
            inf.denoise_features(feat, W, H, out)
        else:
            inf.denoise(bufs[0], bufs[1], bufs[2], bufs[3], W, H, out)
    torch.cuda.synchronize()
    inf.close()
    print("profile driver: %d passes of the %d-level %s denoiser (%s) at %dx%d"
          % (args.profile_denoiser, levels, widths, os.path.basename(path), W, H))


def bench_denoiser(args):
    """Learned denoisers vs filterLighting on held-out frames: quality against
    the 64-sample image (PSNR, SSIM after resolvePixels' display mapping) and
    time per 1920x1080 frame (HIP events) of each network; each network's
    FLOP rate against the BF16 MFMA peak and its byte rate against HBM."""
    import numpy as np
    import torch

    import dn_train
    import pyarbor as px

    W, H, spp = 1920, 1080, 64
    paths = [p for p in args.denoiser.split(",") if p]
    scene = px.DeviceScene().generate(args.seed)
    nets = [px.Denoiser.from_checkpoint(p, 1, 1088, 1920) for p in paths]
    eng = px.Engine(W, H, args.seed)
    eng.set_input(px.walk_input(args.seed, 0))
    eng.frame(1.0 / 60.0)
    cams = [("walk frame 0", eng.download("C", 64))]
    eng.close()
    vc, _ = px.batch_walk_views(11 + 777777, args.seed, 2)
    cams += [("held-out view %d" % i, vc[i]) for i in range(2)]
    rows = []
    for name, cam in cams:
        cam = cam.copy()
        cam[22] = 1.0
        fb, normal, Cd = dn_train.render_inputs(px, torch, scene, cam, W, H)
        target = dn_train.render_targets_batched(px, torch, scene, cam[None, :], W, H, spp)[0]
        ref = dn_train.display(px, torch, target, Cd, W, H)
        out = torch.zeros(W * H * 4, dtype=torch.float32, device="cuda")
        row = {"camera": name}
        for k, buf in (("1 spp", fb.Linear), ("filterLighting", fb.Filtered)):
            img = dn_train.display(px, torch, buf, Cd, W, H)
            row[k] = {"psnr": round(dn_train.psnr(img, ref), 3), "ssim": round(dn_train.ssim(img, ref), 4)}
        row["filter_ms"] = round(dn_train.event_ms(torch, lambda: px.filter_lighting(Cd, fb)), 4)
        # the lean network's input: the compact features the shading pass
        # writes in place of Linear and Guide; both shading forms are timed
        # with sun visibility off (no shadow rays), so that the difference is
        # what they write
        feat = torch.zeros(W * H * 8, dtype=torch.bfloat16, device="cuda")
        px.shade_features(scene, Cd, fb, feat)
        lin_keep, gui_keep = fb.Linear.clone(), fb.Guide.clone()
        dark = Cd.clone()
        dark[22] = 0.0
        scratch = torch.zeros_like(feat)
        row["shade_ms"] = round(dn_train.event_ms(torch, lambda: px.shade(scene, dark, fb), reps=10), 4)
        row["shade_features_ms"] = round(dn_train.event_ms(torch, lambda: px.shade_features(scene, dark, fb, scratch),
                                                           reps=10), 4)
        fb.Linear.copy_(lin_keep)
        fb.Guide.copy_(gui_keep)
        row["networks"] = []
        for path, inf in zip(paths, nets):
            def net():
                if inf.shape()[0] == 0:
                    inf.denoise_features(feat, W, H, out)
                else:
                    inf.denoise(fb.Linear, fb.Guide, normal, fb.Hit, W, H, out)

            def layered():
                inf.set_input(0, fb.Linear, fb.Guide, normal, fb.Hit, W, H)
                inf.forward(1)
                inf.output(0, out, W, H)

            net()
            img = dn_train.display(px, torch, out, Cd, W, H)
            row["networks"].append({
                "checkpoint": os.path.basename(path),
                "psnr": round(dn_train.psnr(img, ref), 3), "ssim": round(dn_train.ssim(img, ref), 4),
                "denoiser_ms": round(dn_train.event_ms(torch, net, reps=10), 4),
                "layered_ms": round(dn_train.event_ms(torch, layered, reps=10), 4),
                "forward_ms": round(dn_train.event_ms(torch, lambda: inf.forward(1), reps=10), 4)})
        rows.append(row)
    flt = float(np.mean([r["filter_ms"] for r in rows]))
    summary = []
    for i, (path, inf) in enumerate(zip(paths, nets)):
        levels, widths = inf.shape()
        den = float(np.mean([r["networks"][i]["denoiser_ms"] for r in rows]))
        lay = float(np.mean([r["networks"][i]["layered_ms"] for r in rows]))
        fwd = float(np.mean([r["networks"][i]["forward_ms"] for r in rows]))
        lean = levels == 0
        one_pass = lean or (levels == 1 and widths == [16])
        flops = lean_flops(H, W, widths[0]) if lean else denoiser_flops(1088, 1920, widths)
        # one pass: the lean kernel reads the 16-byte compact features and
        # writes Filtered; the 1-level kernel reads the four frame buffers
        nbytes = (32.0 if lean else 80.0) * W * H if one_pass else denoiser_bytes(1088, 1920, H, W, widths)
        tflops = flops / ((den if one_pass else fwd) * 1e-3) / 1e12
        gbs = nbytes / (den * 1e-3) / 1e9
        beats = all(r["networks"][i]["psnr"] >= r["filterLighting"]["psnr"] for r in rows)
        summary.append({
            "checkpoint": os.path.basename(path), "levels": levels, "widths": widths,
            "parameters": int(inf.parameters()), "one_pass": one_pass, "denoiser_ms": round(den, 4),
            "layered_ms": round(lay, 4), "forward_ms": round(fwd, 4),
            "vs_filter_time": round(den / flt, 2),
            "mean_psnr": round(float(np.mean([r["networks"][i]["psnr"] for r in rows])), 3),
            "psnr_at_or_above_filter_on_every_frame": beats,
            "flop_per_frame": flops, "forward_tflops": round(tflops, 2),
            "mfma_frac": round(tflops / BF16_PEAK_TFLOPS, 4),
            "algorithmic_bytes_per_frame": nbytes, "pass_gbs": round(gbs, 1),
            "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
            # operational intensity against the ridge point decides the bound
            "bound": "mfma" if flops / nbytes > BF16_PEAK_TFLOPS * 1e12 / (HBM_PEAK_GBS * 1e9) else "hbm"})
        inf.close()
    first = summary[0]
    if first["bound"] == "hbm":
        kernel = ("lean denoiser pass (compact features, 3x3 and 1x1 convolutions, output: one kernel), %d channels"
                  % first["widths"][0] if first["levels"] == 0 else
                  "denoiser pass (features, %d convolutions, output%s), %d-level %s"
                  % (len(denoiser_layers(first["widths"])), ": one kernel" if first["one_pass"] else "",
                     first["levels"], first["widths"]))
        roof = {"bound": "hbm", "kernel": kernel,
                "achieved": first["pass_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": first["hbm_frac"],
                "algorithmic_bytes": first["algorithmic_bytes_per_frame"], "traffic": None}
    else:
        roof = {"bound": "mfma", "kernel": "denoiser forward (%d convolutions + resampling), %d-level %s"
                % (len(denoiser_layers(first["widths"])), first["levels"], first["widths"]),
                "achieved": first["forward_tflops"], "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": first["mfma_frac"], "flop_per_frame": first["flop_per_frame"], "traffic": None}
    result = {
        "metric": "learned vs filter denoise ms at equal PSNR",
        "value": first["denoiser_ms"],
        "unit": "ms/frame (learned denoiser, 1920x1080)",
        "n_gpus": 1, "steps": len(rows), "warmup": 1, "ms_per_step": first["denoiser_ms"],
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
        "data": "synthetic (tracer-rendered held-out frames of the seed-7 forest, sun visibility on)",
        "config": {"workload": "learned denoisers (BF16 MFMA networks) vs filterLighting on 1-sample 1920x1080 "
                               "frames, against 64-sample targets",
                   "checkpoints": [os.path.basename(p) for p in paths]},
        "filter_ms": round(flt, 4),
        "learned_vs_filter_time": first["vs_filter_time"],
        "shading_sun_off_ms": {"Linear_Guide": round(float(np.mean([r["shade_ms"] for r in rows])), 4),
                               "compact_features": round(float(np.mean([r["shade_features_ms"] for r in rows])), 4)},
        "networks": summary,
        "quality": rows,
        "roofline": roof,
    }
    print(json.dumps(result), flush=True)


def bench_views(args):
    """Batched views: N cameras at seeded positions inside the 324-tree forest,
    resolutions drawn from {256x256, 640x360, 1280x720, 1920x1080, 2560x1440};
    a step is one replay of the captured batch (fused trace+shade over every
    view's tiles, then filterLighting and resolvePixels per view). The same
    views rendered one after another through the single-view entry points are
    timed beside it."""
    import torch

    import pyarbor as px

    n, K, Wm = args.views, args.steps, args.warmup
    scene = px.DeviceScene().generate(args.seed)
    cams, sizes = px.batch_walk_views(args.views_seed, args.seed, n)
    rays = sum(w * h for w, h in sizes)
    accel = px.TraceAccel(scene)
    b = px.Batch(n, rays)
    b.set_views(cams, sizes)
    b.set_accel(accel)
    tuned = b.autotune(scene, reps=1)
    stream = torch.cuda.Stream()
    b.capture(scene, stream)
    for _ in range(Wm):
        b.launch(stream)
    stream.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(K):
        b.launch(stream)
    ev1.record(stream)
    stream.synchronize()
    elapsed = time.perf_counter() - t0
    graph_ms = ev0.elapsed_time(ev1) / K

    # the same views, one at a time through the single-view entry points
    # (each view with a tile schedule of its own, as the engine traces)
    fbs, cds, scheds = {}, [], []
    for i, (w, h) in enumerate(sizes):
        fbs.setdefault((w, h), px.FrameBuffers(w, h))
        cds.append(torch.from_numpy(cams[i].copy()).cuda())
        scheds.append(px.TileSchedule(w, h))
    for rep in range(2):  # the first pass warms the single-view kernels and seeds the schedules
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for i, (w, h) in enumerate(sizes):
            fb = fbs[(w, h)]
            px.trace_accel(scene, accel, cds[i], fb, scheds[i])
            px.shade(scene, cds[i], fb)
            px.filter_lighting(cds[i], fb)
            px.resolve(cds[i], fb)
        torch.cuda.synchronize()
        seq_s = time.perf_counter() - t1
    b.close()
    result = {
        "metric": "Mrays/s (batched views)",
        "value": round(rays * K / elapsed / 1e6, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": Wm,
        "ms_per_step": round(elapsed / K * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (procedural forest from seed; seeded view positions and resolutions)",
        "config": {"workload": "batched views: %d cameras inside the 324-tree forest, resolutions from "
                               "{256x256, 640x360, 1280x720, 1920x1080, 2560x1440}, 1 spp primary+shade+filter+"
                               "resolve, shadows off, one captured hipGraph per step, trace over the SAH "
                               "hierarchy" % n,
                   "views": n, "pixels": rays, "seed": args.seed, "views_seed": args.views_seed},
        "graph_ms_per_step": round(graph_ms, 4),
        "autotune": tuned,
        "sequential_single_view_mrays_per_s": round(rays / seq_s / 1e6, 3),
        "batch_vs_sequential": round(seq_s / (elapsed / K), 3),
    }
    print(json.dumps(result), flush=True)


def profile_trace(args):
    import torch

    import pyarbor as px

    W, H = args.width, args.height
    eng = px.Engine(W, H, args.seed)
    inputs = torch.from_numpy(walk_inputs(px, args.seed, 1)).cuda()
    eng.frame(1.0 / 60.0, inputs[0])  # camera of walk frame 0 (also one full frame)
    eng.sync()
    lib = px.lib()
    s = lib.arbor_engine_stream(eng.h)
    P, Nodes, Order, C, Hit = [eng.buffer_ptr(n) for n in ("P", "Nodes", "Order", "C", "Hit")]
    accel, sched = eng.buffer_ptr("Accel"), eng.buffer_ptr("Schedule")
    for _ in range(args.profile_trace):  # the engine's own trace call: SAH walk, scheduled, same camera
        px.check(lib.arbor_tracePixelsAccel(P, Nodes, Order, accel, C, Hit, W, H, sched, s), "arbor_tracePixelsAccel")
    eng.sync()
    a = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    eng.close()
    print("profile driver: %d tracePixels launches at %dx%d, then a 256 MiB copy" % (args.profile_trace, W, H))


def walk_inputs(px, seed, n):
    import numpy as np

    rows = [px.walk_input(seed, f) for f in range(n)]
    rows[0][10] = 1.0  # frame 0 toggles tree-to-tree sun visibility off (updateCamera I[10])
    return np.stack(rows).astype(np.float32)


# Walk frames the CPU baseline renders (the GPU's time on the same frames is
# reported beside it): the first frame and the middle of the 120-frame walk.
CPU_FRAMES = (0, 60)


def _cpu_camera(cpu, pyarbor, seed, frame, width, height):
    """Camera of walk frame `frame`: initCamera, then updateCamera over the
    walk's inputs 0..frame (as the engine runs them)."""
    C = cpu.init_camera(seed)
    for f, I in enumerate(walk_inputs(pyarbor, seed, frame + 1)):
        cpu.update_camera(C, I, 1.0 / 60.0, width, height)
    return C


def _cpu_frame(cpu, scene, C, width, height, threads):
    """One frame on a CPU implementation: (seconds for trace+shade+filter+
    resolve, seconds for the trace)."""
    import numpy as np

    t0 = time.perf_counter()
    H = cpu.trace(scene, C, width, height, threads)
    t1 = time.perf_counter()
    L, G = cpu.shade(scene, C, H, width, height, threads)
    F = cpu.filter(L, G, H, C, width, height, threads)
    hist = np.zeros_like(F)
    cpu.resolve(F, hist, C, width, height, threads)
    t2 = time.perf_counter()
    return t2 - t0, t1 - t0


def _cpu_timed(cpu, seed, width, height, threads):
    """Warm-up (walk frame 0 once), then the best of 3 runs of each frame in
    CPU_FRAMES: per frame (frame s, trace s)."""
    import pyarbor

    scene = cpu.generate(seed)
    cams = {f: _cpu_camera(cpu, pyarbor, seed, f, width, height) for f in CPU_FRAMES}
    _cpu_frame(cpu, scene, cams[CPU_FRAMES[0]], width, height, threads)
    out = {}
    for f in CPU_FRAMES:
        runs = [_cpu_frame(cpu, scene, cams[f], width, height, threads) for _ in range(3)]
        out[f] = min(runs)
    return out


def _host_cpu():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model


def cpu_baseline(seed, width, height, threads):
    """The reference CPU harness (compiled from its sources into
    oracle/_ref/libarborref.so) on this host's cores: walk frames 0 and 60 of
    the same workload (trace + shade + filter + resolve), warm-up then best of
    3 per frame. The oracle's float restatement (same outputs bit for bit) is
    timed beside it as "port". Threads: the CPU share this process is given
    (OMP_NUM_THREADS where set, else every core in its affinity mask)."""
    sys.path.insert(0, os.path.join(HERE, "tests"))
    import oracle_lib

    affinity = len(os.sched_getaffinity(0))
    if threads <= 0:
        share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        threads = min(affinity, share) if share > 0 else affinity
    rays = width * height
    port = _cpu_timed(oracle_lib.Oracle(), seed, width, height, threads)
    kind = "reference" if oracle_lib.ref_available() else "port"
    ref = _cpu_timed(oracle_lib.RefHarness(), seed, width, height, threads) if kind == "reference" else port
    frame_s = sum(v[0] for v in ref.values()) / len(ref)
    trace_s = sum(v[1] for v in ref.values()) / len(ref)
    port_s = sum(v[0] for v in port.values()) / len(port)
    port_trace_s = sum(v[1] for v in port.values()) / len(port)
    return {"value": round(rays / frame_s / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": kind,
            "sample": "walk frames %s of the same workload (%dx%d, seed %d, shadows off), trace+shade+filter+"
                      "resolve, warm-up then best of 3 per frame, mean over the frames; OpenMP "
                      "schedule(dynamic,2) rows" % (list(CPU_FRAMES), width, height, seed),
            "nproc": os.cpu_count(), "affinity_cores": affinity,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "cpu_model": _host_cpu(),
            "ms_per_frame": round(frame_s * 1e3, 1), "trace_ms": round(trace_s * 1e3, 1),
            "ms_by_frame": {str(f): round(v[0] * 1e3, 1) for f, v in ref.items()},
            "port": {"value": round(rays / port_s / 1e6, 4), "ms_per_frame": round(port_s * 1e3, 1),
                     "trace_ms": round(port_trace_s * 1e3, 1)}}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.accel_variant >= 0:
        import pyarbor

        pyarbor.tune_set("accel_variant", args.accel_variant)
    if args.accel_vis_variant >= 0:
        import pyarbor

        pyarbor.tune_set("accel_vis_variant", args.accel_vis_variant)
    if args.batch_variant >= 0:
        import pyarbor

        pyarbor.tune_set("batch_variant", args.batch_variant)
    if args.profile_trace:
        profile_trace(args)
        return
    if args.profile_denoiser:
        profile_denoiser(args)
        return
    if args.shadows4k:
        bench_shadows4k(args)
        return
    if args.views:
        bench_views(args)
        return
    if args.denoiser:
        bench_denoiser(args)
        return
    import numpy as np
    import torch

    import pyarbor as px

    W, H, K, Wm = args.width, args.height, args.steps, args.warmup
    eng = px.Engine(W, H, args.seed)
    inputs = torch.from_numpy(walk_inputs(px, args.seed, max(K, Wm, WALK_FRAMES, 1))).cuda()
    dt = 1.0 / 60.0

    # warmup: the first Wm walk frames, then back to the initial camera
    for f in range(Wm):
        eng.frame(dt, inputs[f])
    eng.sync()
    eng.reset_camera()
    eng.sync()

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for f in range(K):
        eng.frame(dt, inputs[f])
    eng.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / K * 1e3

    # per-stage kernel times: the whole 120-frame walk replayed with HIP
    # events around each stage on the engine's stream (it synchronises per
    # frame, so it runs outside the timed loop), over the SAH hierarchy (the
    # engine's trace), then again over the heap BVH (arbor_tracePixels' walk)
    walk = torch.from_numpy(walk_inputs(px, args.seed, WALK_FRAMES)).cuda()

    def stage_pass():
        eng.reset_camera()
        eng.set_timing(True)
        rows = []
        for f in range(WALK_FRAMES):
            eng.frame(dt, walk[f])
            rows.append(eng.stage_ms())
        eng.set_timing(False)
        return np.array(rows)

    st = stage_pass()
    eng.set_accel(False)
    st_heap = stage_pass()
    eng.set_accel(True)
    mean = st.mean(axis=0)
    trace_ms = float(mean[1])
    heap_trace_ms = float(st_heap[:, 1].mean())

    roofline = {"bound": "fp32_valu", "unit": "TFLOP/s", "peak": FP32_PEAK_TFLOPS,
                "kernel": "tracePixels (SAH walk, arbor_tracePixelsAccel)",
                "compute": "FP32 vector ALU (ray traversal has no MFMA shape); peak is MI355X's FP32 157.3 TFLOP/s",
                "duration_ms": round(trace_ms, 4), "duration_over": "mean of the %d-frame walk" % WALK_FRAMES,
                "traffic": None}
    std_res = (W, H) == (1920, 1080) and args.seed == 7
    if std_res:
        achieved = TRACE_FLOP_PER_RAY * W * H / (trace_ms * 1e-3) / 1e12
        roofline.update({"achieved": round(achieved, 3), "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                         "flop_per_ray": TRACE_FLOP_PER_RAY,
                         "flop_per_ray_heap_walk": TRACE_FLOP_PER_RAY_HEAP,
                         "frac_of_single_issue_rate": round(achieved / FP32_SINGLE_ISSUE_TOPS, 4),
                         "traffic": int(TRACE_HBM_BYTES),
                         "traffic_algorithmic": int(16 * W * H + 5373952)})
    else:
        roofline.update({"achieved": None, "frac": None})

    rays = W * H
    walk_ms = float(mean[5])
    result = {
        "metric": "Mrays/s, ms/frame @1080p 324 trees",
        "value": round(rays * K / elapsed / 1e6, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": Wm,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (procedural forest from seed; seeded walk-path inputs)",
        "config": {"workload": "full forest plain method: 324 trees, 1920x1080, 1 spp primary+shade+filter+resolve, "
                               "shadows off, seeded 120-frame walk", "width": W, "height": H, "trees": 324,
                   "seed": args.seed, "spp": 1, "frames": K, "timed_walk_frames": [0, K - 1]},
        "ms_per_frame": round(ms_per_step, 4),
        "walk120": {"frames": WALK_FRAMES, "gpu_ms_per_frame": round(walk_ms, 4),
                    "mrays_per_s": round(rays / (walk_ms * 1e-3) / 1e6, 2),
                    "note": "HIP-event time of every frame of the walk (the timed steps cover frames 0..steps-1)"},
        "trace_mrays_per_s": round(rays / (trace_ms * 1e-3) / 1e6, 2),
        "heap_walk_trace": {"entry_point": "arbor_tracePixels' heap BVH walk, tile-scheduled",
                            "ms": round(heap_trace_ms, 4),
                            "mrays_per_s": round(rays / (heap_trace_ms * 1e-3) / 1e6, 2)},
        "launches_per_frame": 5,
        "stage_ms": {"prologue": round(float(mean[0]), 4), "trace": round(trace_ms, 4),
                     "shade": round(float(mean[2]), 4), "filter_resolve": round(float(mean[3]), 4),
                     "frame": round(walk_ms, 4)},
        "roofline": roofline,
    }
    # HBM figure of the bandwidth-bound fused filter + resolve: 100 algorithmic
    # B/px (Linear, Guide, Hit and History read; Filtered, History and Pixels written)
    fr_ms = float(mean[3])
    result["hbm_kernels"] = {"filterLighting+resolvePixels": {
        "GB/s": round(100 * rays / (fr_ms * 1e-3) / 1e9, 1),
        "frac": round(100 * rays / (fr_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}
    eng.close()
    if not args.no_cpu:
        cb = cpu_baseline(args.seed, W, H, args.cpu_threads)
        result["cpu_baseline"] = cb
        result["gpu_vs_cpu"] = round(result["value"] / cb["value"], 1)
        result["gpu_vs_cpu_port"] = round(result["value"] / cb["port"]["value"], 1)
        same = float(np.mean([st[f, 5] for f in CPU_FRAMES if f < WALK_FRAMES]))
        result["gpu_vs_cpu_same_frames"] = {"frames": list(CPU_FRAMES), "gpu_ms_per_frame": round(same, 4),
                                           "ratio": round(cb["ms_per_frame"] / same, 1)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
