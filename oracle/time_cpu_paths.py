"""CPU cross-check of the two CPU implementations of the frame pipeline.

Times the oracle's float restatement (oracle/liboracle.so) and the reference
harness compiled from its own sources (oracle/_ref/libarborref.so) on the
same frame and the same threads: the full-forest workload (seed 7, 324 trees,
1920x1080, shadows off), walk frame 0's camera, one tracePixels,
shadePixels, filterLighting and resolvePixels pass each, best of `--reps`
after one warm-up. Both emulate the harness's launch (OpenMP rows,
schedule(dynamic, 2)). A ratio near 1 shows the restatement is a fair
stand-in for the reference's CPU cost; the outputs are also compared bit for
bit.

Test / measurement infrastructure: it calls the oracle and the reference
build, never the product.
Usage: python oracle/time_cpu_paths.py [--threads 8] [--reps 3] [--width 1920 --height 1080]
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "tests"))
sys.path.insert(0, HERE)
from count_trace_ops import walk_input  # noqa: E402
from oracle_lib import Oracle, RefHarness, ref_available  # noqa: E402


def best_of(fn, reps):
    fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - t0)
    return min(ts), out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count())
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    args = ap.parse_args()
    if not ref_available():
        print("reference harness build (oracle/_ref/libarborref.so) not present")
        return 1
    W, H, T = args.width, args.height, args.threads
    o, r = Oracle(), RefHarness()
    s = o.generate(7)
    C = o.init_camera(7)
    I = walk_input(7, 0)
    I[10] = 1.0  # sun visibility off, as the workload runs
    o.update_camera(C, I, 1.0 / 60.0, W, H)
    results = {}
    for name, impl in (("oracle", o), ("reference", r)):
        t_trace, Hit = best_of(lambda: impl.trace(s, C, W, H, T), args.reps)
        t_shade, (L, G) = best_of(lambda: impl.shade(s, C, Hit, W, H, T), args.reps)
        t_filter, F = best_of(lambda: impl.filter(L, G, Hit, C, W, H, T), args.reps)
        hist = np.zeros(W * H * 4, np.float32)
        t_resolve, px = best_of(lambda: impl.resolve(F, hist.copy(), C, W, H, T), args.reps)
        results[name] = dict(trace=t_trace, shade=t_shade, filter=t_filter, resolve=t_resolve,
                             frame=t_trace + t_shade + t_filter + t_resolve, out=(Hit, L, F, px))
        print("%-9s trace %.3f s  shade %.3f s  filter %.3f s  resolve %.3f s  frame %.3f s" % (
            name, t_trace, t_shade, t_filter, t_resolve, results[name]["frame"]), flush=True)
    same = all(np.array_equal(a.view(np.uint32), b.view(np.uint32))
               for a, b in zip(results["oracle"]["out"], results["reference"]["out"]))
    ratio = results["oracle"]["frame"] / results["reference"]["frame"]
    print("threads %d, %dx%d: oracle/reference frame time ratio %.3f; outputs bit-identical: %s" % (
        T, W, H, ratio, same))
    return 0


if __name__ == "__main__":
    sys.exit(main())
