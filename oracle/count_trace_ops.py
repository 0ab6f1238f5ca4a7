"""Algorithmic FP32 operations per primary ray of the benchmark workload.

Runs the oracle's counting instantiation (oracle/flops.cpp) on the
full-forest workload bench.py times: seed 7, 324 trees, 1920x1080, shadows
off, the seeded 120-frame walk. Counts tracePixels with the GPU's traversal
(the engine's walk over the SAH hierarchy, including its tie-order box tests;
with --heap the lean heap walk), bounding-sphere rejections, early exits past
the caller's best
the leaf Newton solve's fixed-point/2-cycle exits, the skips of
band-weighted terms with a zero band weight and the stored leaf basis with
the breeze off, sampling every `--frame-step`-th
walk frame and every `--stride`-th pixel in x and y, and prints the mean.

Test / measurement infrastructure: it calls the oracle, never the product.
Usage: python oracle/count_trace_ops.py [--frame-step 8] [--stride 8]
"""
import argparse
import ctypes as ct
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(REPO, "tests"))
from oracle_lib import Oracle  # noqa: E402


def walk_input(seed, frame):
    """The engine's seeded walk (arbor_walk_input), restated for the oracle."""
    def h(x):
        x &= 0xFFFFFFFF
        x ^= x >> 16
        x = (x * 2146121005) & 0xFFFFFFFF
        x ^= x >> 15
        x = (x * 2221713035) & 0xFFFFFFFF
        x ^= x >> 16
        return x

    I = np.zeros(32, np.float32)
    v = h(((frame * 2654435761) & 0xFFFFFFFF) ^ h(seed + 40503))
    I[15] = (1.0, 1.0, 0.0, -1.0)[v & 3]
    I[13] = (0.0, 1.0, 0.0, -1.0)[(v >> 2) & 3]
    I[0] = np.float32(np.float32((v >> 8) & 0xFFFF) / np.float32(65536.0) * np.float32(40.0) - np.float32(20.0))
    return I


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--frame-step", type=int, default=8)
    ap.add_argument("--stride", type=int, default=8)
    ap.add_argument("--frames", type=int, default=120)
    ap.add_argument("--heap", action="store_true",
                    help="count the heap BVH walk (arbor_tracePixels) instead of the SAH walk the engine runs")
    args = ap.parse_args()
    o = Oracle()
    lib = o.lib
    lib.orc_set_cull.argtypes = [ct.c_int]
    lib.orc_set_skip.argtypes = [ct.c_int]
    lib.orc_set_newton_exits.argtypes = [ct.c_int]
    lib.orc_set_band_skips.argtypes = [ct.c_int]
    lib.orc_set_wind_shortcut.argtypes = [ct.c_int]
    W, H, seed = 1920, 1080, 7
    s = o.generate(seed)
    C = o.init_camera(seed)
    lib.orc_set_sah.argtypes = [ct.c_void_p]
    accel = None if args.heap else o.build_sah(s)
    lib.orc_set_sah(None if accel is None else accel.ctypes.data)
    print("walk:", "heap BVH" if accel is None else "SAH hierarchy", flush=True)
    per_frame = []
    kinds = None
    for f in range(args.frames):
        I = walk_input(seed, f)
        if f == 0:
            I[10] = 1.0
        o.update_camera(C, I, 1.0 / 60.0, W, H)
        if f % args.frame_step:
            continue
        lib.orc_set_cull(1)
        lib.orc_set_skip(1)
        lib.orc_set_newton_exits(1)
        lib.orc_set_band_skips(1)
        lib.orc_set_wind_shortcut(1)
        r = o.count_trace_ops(s, C, W, H, args.stride)
        lib.orc_set_cull(0)
        lib.orc_set_skip(0)
        lib.orc_set_newton_exits(0)
        lib.orc_set_band_skips(0)
        lib.orc_set_wind_shortcut(0)
        per_frame.append(r["ops_per_ray"])
        kinds = r if kinds is None else {k: kinds[k] + r[k] for k in r}
        print("frame %3d: %.1f ops/ray" % (f, r["ops_per_ray"]), flush=True)
    n = len(per_frame)
    print("mean over %d frames: %.1f ops/ray" % (n, float(np.mean(per_frame))))
    print("mix per ray:", {k: round(v / n, 1) for k, v in kinds.items() if k not in ("rays", "ops_per_ray")})


if __name__ == "__main__":
    main()
