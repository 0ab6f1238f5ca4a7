"""Writes the golden fixtures under tests/golden/ from the REFERENCE harness.

Runs oracle/_ref/libarborref.so (tests/native_harness.cpp + Arbor.cu compiled
from the reference checkout by oracle/Makefile) and records:
  * sha256 digests of B, P, Keys, Order, Nodes for seeds 1, 7, 42, 1788, with
    the root box and root count (native_harness.cpp:57-58,68);
  * the 120-stage bitonic network result digest (native_harness.cpp:76-80);
  * camera state C after initCamera and a scripted input sequence (Arbor.cu:181-182);
  * the harness's 100 random BVH rays (native_harness.cpp:81-88): t and id;
  * the bark footprint sweep maximum increment (native_harness.cpp:89-91);
  * the 50 leaf near/far intersections (native_harness.cpp:92-93);
  * small frames (trace, shade with and without sun visibility, filter,
    resolve) as digests plus raw 16x16 crops, and a single-tree 64x64 frame;
  * frames of the shadowed, windy forest setting: breeze on (C[11]=0.6) at two
    wind times, a progressive 4-sample sequence (C[9]=0..3, reset only on the
    first) whose jittered rays and sun-disk samples change every sample and
    whose History accumulates, the sun at 35 degrees with visibility on and
    off, one resolve over a seeded non-zero History, sample numbers past 3,
    and the three debug views (Arbor.cu:50-51,171-172,256-304).

Usage: make -C oracle && python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from oracle_lib import RefHarness, digest, seeded_history, c3_camera, C3_CASES, OBJECT_COUNT, PS  # noqa: E402

SEEDS = (1, 7, 42, 1788)
FRAME_W, FRAME_H = 96, 64
CROP = (slice(24, 40), slice(40, 56))  # rows, cols


def input_script():
    """Scripted updateCamera inputs: presets, toggles, mouse and walking."""
    seq = []
    for step in range(24):
        I = np.zeros(32, np.float32)
        if step == 0:
            I[5] = 2
        if step == 3:
            I[7] = 1  # breeze on
        if step == 5:
            I[9] = 1  # debug mode
        if step == 6:
            I[10] = 1  # sun visibility toggle
        if step == 8:
            I[8] = 0.5  # sun azimuth
        if step == 9:
            I[12] = -0.7  # exposure
        if 10 <= step < 16:
            I[15] = 1.0
            I[0] = 13.0 - step
        if 16 <= step < 20:
            I[13] = -1.0
            I[16] = 1.0
            I[1] = 4.0
        if step == 21:
            I[5] = 4
        seq.append(I)
    return np.array(seq, np.float32)


def main():
    r = RefHarness()
    out = {"source": "reference harness tests/native_harness.cpp + Arbor.cu, compiled by oracle/Makefile",
           "scenes": {}}
    crops = {}
    for seed in SEEDS:
        s = r.generate(seed)
        out["scenes"][str(seed)] = {
            "B": digest(s.B), "P": digest(s.P), "Keys": digest(s.Keys), "Order": digest(s.Order),
            "Nodes": digest(s.Nodes),
            "root_lo": [float(x) for x in s.Nodes[8:11]], "root_hi": [float(x) for x in s.Nodes[12:15]],
            "root_count": float(s.Nodes[11]),
            "counts": [int((s.P[3::PS] == t).sum()) for t in (0, 1, 2)],
        }
        if seed == 1788:
            sb = r.generate(seed, bitonic=True)
            out["bitonic_1788"] = {"Keys": digest(sb.Keys), "Order": digest(sb.Order)}
            crops["P_1788_head"] = s.P[: 64 * PS].copy()
            crops["Keys_1788_head"] = s.Keys[:256].copy()
            crops["Order_1788_head"] = s.Order[:256].copy()
            crops["Nodes_1788_top"] = s.Nodes[: 64 * 8].copy()

    # camera state machine
    C = r.init_camera(7)
    out["camera_init_7"] = [float(x) for x in C]
    script = input_script()
    states = []
    for I in script:
        r.update_camera(C, I, 1.0 / 60.0, 320, 240)
        states.append(C.copy())
    crops["camera_inputs"] = script
    crops["camera_states"] = np.array(states, np.float32)

    # harness rays / sweeps on seed 1788 (the harness Renderer default)
    s = r.generate(1788)
    ro, rd = r.harness_rays()
    tr = r.trace_tree_rays(s, ro, rd, cone=0.0005, best=100000.0)
    crops["rays_ro"], crops["rays_rd"], crops["rays_hit"] = ro, rd, tr
    out["bark_sweep_max_increment"] = float(r.bark_sweep(s.P))
    crops["leaf_near_far"] = r.leaf_probes(s.P)

    # small frames, seed 7 forest, preset views
    s7 = r.generate(7)
    frames = {}
    for view, light in ((1, 0), (3, 1), (4, 0)):
        C = r.init_camera(7)
        I = np.zeros(32, np.float32)
        I[5] = view
        r.update_camera(C, I, 1.0 / 60.0, FRAME_W, FRAME_H)
        C[22] = light
        C[9] = 0
        C[10] = 1
        H = r.trace(s7, C, FRAME_W, FRAME_H)
        L, G = r.shade(s7, C, H, FRAME_W, FRAME_H)
        F = r.filter(L, G, H, C, FRAME_W, FRAME_H)
        Hist = np.zeros_like(F)
        px = r.resolve(F, Hist, C, FRAME_W, FRAME_H)
        key = "v%d_l%d" % (view, light)
        frames[key] = {"C": [float(x) for x in C], "Hit": digest(H), "Linear": digest(L), "Guide": digest(G),
                       "Filtered": digest(F), "History": digest(Hist), "Pixels": digest(px)}
        for name, arr in (("Hit", H), ("Linear", L), ("Filtered", F)):
            crops["%s_%s" % (key, name)] = arr.reshape(FRAME_H, FRAME_W, 4)[CROP].copy()
        crops["%s_Pixels" % key] = px.reshape(FRAME_H, FRAME_W)[CROP].copy()
    out["frames"] = {"seed": 7, "width": FRAME_W, "height": FRAME_H, "views": frames,
                     "crop_rows": [CROP[0].start, CROP[0].stop], "crop_cols": [CROP[1].start, CROP[1].stop]}

    # shadowed, windy forest frames; a case with the same sequence name carries
    # its History from the previous case (the progressive samples of one frame)
    c3 = {}
    hist = {}
    for case in C3_CASES:
        C = c3_camera(r, case, FRAME_W, FRAME_H)
        H = r.trace(s7, C, FRAME_W, FRAME_H)
        L, G = r.shade(s7, C, H, FRAME_W, FRAME_H)
        F = r.filter(L, G, H, C, FRAME_W, FRAME_H)
        seq = case["sequence"]
        if seq not in hist:
            hist[seq] = seeded_history(FRAME_W, FRAME_H, case["history_seed"])
        Hist = hist[seq]
        px = r.resolve(F, Hist, C, FRAME_W, FRAME_H)
        name = case["name"]
        c3[name] = {"C": [float(x) for x in C], "Hit": digest(H), "Linear": digest(L), "Guide": digest(G),
                    "Filtered": digest(F), "History": digest(Hist), "Pixels": digest(px)}
        for buf, arr in (("Hit", H), ("Linear", L), ("Filtered", F), ("History", Hist)):
            crops["c3_%s_%s" % (name, buf)] = arr.reshape(FRAME_H, FRAME_W, 4)[CROP].copy()
        crops["c3_%s_Pixels" % name] = px.reshape(FRAME_H, FRAME_W)[CROP].copy()
    out["frames_c3"] = c3

    # single-tree frame (seed 1, camera 8 m from the trunk at eye height 1.72)
    s1 = r.generate(1)
    C = r.init_camera(1)
    C[0:3] = [0.0, 1.72, 8.0]
    C[4:7] = [0.0, 1.72, 7.0]
    C[22] = 0
    Ht = r.trace_tree(s1, C, 64, 64)
    out["single_tree_64"] = {"C": [float(x) for x in C], "Hit": digest(Ht)}
    crops["single_tree_64_Hit"] = Ht.copy()

    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    np.savez_compressed(os.path.join(HERE, "crops.npz"), **crops)
    print("wrote golden.json and crops.npz")


if __name__ == "__main__":
    main()
