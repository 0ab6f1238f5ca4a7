"""CPU checks of the C-ABI boundary: the library loads, exports every symbol
the public headers declare, and rejects bad arguments with status codes
before touching a device (the reference's host-forcible failures:
engine.js:28 seed range, engine.js:36-37 dimensions, runtime.js:352 limits)."""
import ctypes as ct
import os
import re

import numpy as np
import pytest

import pyarbor

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NULL = None
DUMMY = ct.c_void_p(4096)  # never dereferenced: every case below fails validation first


def declared_symbols():
    names = set()
    for h in ("arbor.h", "arbor_engine.h", "arbor_batch.h", "arbor_denoise.h"):
        with open(os.path.join(REPO, "include", h)) as f:
            text = f.read()
        names |= set(re.findall(r"\b(arbor_\w+)\s*\(", text))
    return sorted(names)


@pytest.fixture(scope="module")
def L():
    return pyarbor.lib()


def test_every_declared_symbol_is_exported(L):
    names = declared_symbols()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
        assert n in pyarbor.SIGNATURES, n


def test_twelve_reference_entry_points_present(L):
    for n in ("growTree", "makePrimitives", "makeKeys", "sortPass", "makeBounds", "reduceBounds", "initCamera",
              "updateCamera", "tracePixels", "shadePixels", "filterLighting", "resolvePixels"):
        assert hasattr(L, "arbor_" + n)


def test_null_pointers_rejected(L):
    E = -1
    assert L.arbor_growTree(NULL, 0, 1, NULL) == E
    assert L.arbor_makePrimitives(DUMMY, NULL, 1, NULL) == E
    assert L.arbor_makeKeys(DUMMY, DUMMY, NULL, NULL) == E
    assert L.arbor_sortPass(NULL, DUMMY, 1, 2, NULL) == E
    assert L.arbor_makeBounds(DUMMY, NULL, DUMMY, NULL) == E
    assert L.arbor_reduceBounds(NULL, 1, NULL) == E
    assert L.arbor_initCamera(NULL, NULL, 1, NULL) == E
    assert L.arbor_updateCamera(DUMMY, NULL, NULL, 0.01, 64, 64, NULL) == E
    assert L.arbor_tracePixels(DUMMY, DUMMY, DUMMY, DUMMY, NULL, 64, 64, NULL) == E
    assert L.arbor_shadePixels(DUMMY, DUMMY, DUMMY, DUMMY, DUMMY, DUMMY, NULL, 64, 64, NULL) == E
    assert L.arbor_filterLighting(DUMMY, DUMMY, DUMMY, NULL, DUMMY, 64, 64, NULL) == E
    assert L.arbor_resolvePixels(DUMMY, DUMMY, NULL, DUMMY, 64, 64, NULL) == E
    assert L.arbor_sortKeys(NULL, DUMMY, NULL) == E
    assert L.arbor_buildBVH(DUMMY, DUMMY, DUMMY, NULL, NULL) == E
    assert L.arbor_traceTreePixels(NULL, DUMMY, DUMMY, DUMMY, DUMMY, 8, 8, NULL) == E


def test_scalar_ranges_rejected(L):
    A = -2
    assert L.arbor_growTree(DUMMY, 5, 1, NULL) == A
    assert L.arbor_growTree(DUMMY, -1, 1, NULL) == A
    assert L.arbor_growTree(DUMMY, 0, -1, NULL) == A  # engine.js:28 seed range
    assert L.arbor_growTree(DUMMY, 0, 1000001, NULL) == A
    assert L.arbor_makePrimitives(DUMMY, DUMMY, -5, NULL) == A
    for d, s in ((3, 8), (2, 2), (4, 2), (1, 65536), (0, 2)):
        assert L.arbor_sortPass(DUMMY, DUMMY, d, s, NULL) == A, (d, s)
    for first in (0, 3, 32768, -4):
        assert L.arbor_reduceBounds(DUMMY, first, NULL) == A, first
    for w, h in ((0, 10), (10, 0), (8193, 10), (-1, -1)):
        assert L.arbor_tracePixels(DUMMY, DUMMY, DUMMY, DUMMY, DUMMY, w, h, NULL) == A
        assert L.arbor_resolvePixels(DUMMY, DUMMY, DUMMY, DUMMY, w, h, NULL) == A
        assert L.arbor_updateCamera(DUMMY, DUMMY, NULL, 0.0, w, h, NULL) == A
    assert L.arbor_initCamera(DUMMY, NULL, 1000001, NULL) == A
    assert L.arbor_generate(DUMMY, DUMMY, DUMMY, DUMMY, DUMMY, DUMMY, -1, NULL, 0, 0, NULL) == A


def test_engine_argument_checks(L):
    h = ct.c_void_p()
    assert L.arbor_engine_create(None, 64, 64, 1, 0, NULL) == -1
    assert L.arbor_engine_create(ct.byref(h), 64, 64, -1, 0, NULL) == -2  # engine.js:28
    assert not h.value
    assert L.arbor_engine_frame(NULL, 0.016, NULL) == -1
    assert L.arbor_engine_destroy(NULL) == -1
    assert L.arbor_engine_buffer(NULL, b"P") is None


def test_engine_resize_rounding(L):
    # engine.js:36: width -> multiple of 64 (min 320), height -> multiple of 8 (min 192)
    for (w, h), want in (((1920, 1080), (1920, 1080)), ((1000, 701), (1024, 704)), ((100, 100), (320, 192)),
                         ((3840, 2160), (3840, 2160)), ((256, 256), (320, 256))):
        cw, ch = ct.c_int(w), ct.c_int(h)
        L.arbor_engine_round_size(ct.byref(cw), ct.byref(ch))
        assert (cw.value, ch.value) == want


def test_engine_record_replay_argument_checks(L, tmp_path):
    assert L.arbor_engine_record(None, b"x.rec") == -1
    n = ct.c_int(5)
    assert L.arbor_engine_replay(None, b"x.rec", 0, 0, ct.byref(n)) == -1
    assert L.arbor_engine_replay(DUMMY, None, 0, 0, ct.byref(n)) == -1


def test_tuning_keys_validated(L):
    assert L.arbor_tune_set(None, 1) == -1
    assert L.arbor_tune_set(b"no_such_key", 1) == -2
    assert L.arbor_tune_set(b"trace_variant", 99) == -2
    assert L.arbor_tune_set(b"trace_variant", -1) == -2
    assert L.arbor_tune_set(b"grid_cap", 0) == -2
    old = L.arbor_tune_get(b"grid_cap")
    assert L.arbor_tune_set(b"grid_cap", 777) == 0 and L.arbor_tune_get(b"grid_cap") == 777
    assert L.arbor_tune_set(b"grid_cap", old) == 0
    assert L.arbor_tune_get(b"bogus") == -2


def test_batch_argument_checks(L):
    h = ct.c_void_p()
    assert L.arbor_batch_create(None, 4, 1000) == -1
    assert L.arbor_batch_create(ct.byref(h), 0, 1000) == -2
    assert L.arbor_batch_create(ct.byref(h), 257, 1000) == -2
    assert L.arbor_batch_create(ct.byref(h), 4, 0) == -2
    assert not h.value
    assert L.arbor_batch_destroy(None) == -1
    assert L.arbor_batch_render(None, DUMMY, DUMMY, DUMMY, 7, None) == -1
    assert L.arbor_batch_launch(None, None) == -1
    assert L.arbor_batch_buffer(None, 0, b"Hit") is None
    assert L.arbor_autotune_trace(DUMMY, DUMMY, DUMMY, DUMMY, NULL, 8, 8, 2, NULL, None, 0) == -1
    assert L.arbor_autotune_trace(DUMMY, DUMMY, DUMMY, DUMMY, DUMMY, 0, 8, 2, NULL, None, 0) == -2


def test_tune_cache_roundtrip(L, tmp_path):
    p = str(tmp_path / "tune.txt").encode()
    old = L.arbor_tune_get(b"sched_tiles")
    assert L.arbor_tune_set(b"sched_tiles", 33) == 0
    assert L.arbor_tune_save(p) == 0
    assert L.arbor_tune_set(b"sched_tiles", 5) == 0
    assert L.arbor_tune_load(p) == 0 and L.arbor_tune_get(b"sched_tiles") == 33
    bad = tmp_path / "bad.txt"
    bad.write_text("sched_tiles 7\nno_such_key 3\n")
    assert L.arbor_tune_load(str(bad).encode()) == -2
    assert L.arbor_tune_get(b"sched_tiles") == 33  # rejected whole
    assert L.arbor_tune_load(str(tmp_path / "missing").encode()) == -2
    assert L.arbor_tune_set(b"sched_tiles", old) == 0


def test_trace_accel_size_and_checks(L):
    # header + 4 float4 per interior node of a binary tree over CAPACITY
    # leaves, then a u32 heap slot per primitive id; the build's workspace
    # (two arrays of CAPACITY primitives of 40 bytes, the node queues)
    # follows those words
    words = (4 + 4 * (32768 - 1)) * 16 + 4 * 32768
    assert words + 2 * 32768 * 40 < L.arbor_trace_accel_bytes() < words + 4 * 1024 * 1024
    E = -1
    assert L.arbor_buildTraceAccel(NULL, DUMMY, DUMMY, NULL) == E
    assert L.arbor_buildTraceAccel(DUMMY, DUMMY, NULL, NULL) == E
    assert L.arbor_buildTraceAccelLimit(NULL, DUMMY, DUMMY, 32, NULL) == E
    for bad in (14, 33):
        assert L.arbor_buildTraceAccelLimit(DUMMY, DUMMY, DUMMY, bad, NULL) == -2
    assert L.arbor_tracePixelsAccel(DUMMY, DUMMY, DUMMY, NULL, DUMMY, DUMMY, 64, 64, NULL, NULL) == E
    assert L.arbor_tracePixelsAccel(DUMMY, DUMMY, DUMMY, DUMMY, DUMMY, DUMMY, 0, 64, NULL, NULL) == -2
    assert L.arbor_sunVisibilityAccel(DUMMY, DUMMY, DUMMY, NULL, DUMMY, DUMMY, DUMMY, 64, 64, NULL, NULL) == E
    assert L.arbor_sunVisibilityAccel(DUMMY, DUMMY, DUMMY, DUMMY, DUMMY, DUMMY, DUMMY, 64, 9000, NULL, NULL) == -2
    assert L.arbor_engine_set_accel(NULL, 1) == E
    assert L.arbor_autotune_trace_accel(DUMMY, DUMMY, DUMMY, NULL, DUMMY, DUMMY, 64, 64, 1, NULL, NULL, 0) == E
    assert L.arbor_autotune_trace_accel(DUMMY, DUMMY, DUMMY, DUMMY, DUMMY, DUMMY, 64, 64, 0, NULL, NULL, 0) == -2
    for key, n in ((b"accel_variant", 35), (b"accel_vis_variant", 28)):
        old = L.arbor_tune_get(key)
        assert L.arbor_tune_set(key, n) == -2 and L.arbor_tune_set(key, -1) == -2
        assert L.arbor_tune_set(key, n - 1) == 0 and L.arbor_tune_get(key) == n - 1
        assert L.arbor_tune_set(key, old) == 0


def test_tile_schedule_size_and_checks(L):
    # 18,112-byte header (previous camera, valid flag, accumulator parity, two sets of
    # 128 bucket accumulators: tile count and scatter cursor; 324 tree records of 48 B)
    # + cost, order and key per 8x8 tile
    header = 2560 + 324 * 48
    assert L.arbor_tile_schedule_bytes(1920, 1080) == header + 12 * 240 * 135
    assert L.arbor_tile_schedule_bytes(33, 9) == header + 12 * 5 * 2
    assert L.arbor_tile_schedule_bytes(0, 10) == -2
    assert L.arbor_tracePixelsScheduled(DUMMY, DUMMY, DUMMY, DUMMY, DUMMY, 64, 64, NULL, NULL) == -1
    assert L.arbor_tracePixelsScheduled(DUMMY, DUMMY, DUMMY, DUMMY, DUMMY, 0, 64, DUMMY, NULL) == -2


def test_ppm_writer_format(L, tmp_path):
    # native_harness.cpp:53: "P6\n<w> <h>\n255\n" then R, G, B = the low three bytes
    px = np.array([0xFF030201, 0x00FFFEFD, 0x12345678, 0x0], np.uint32)
    path = tmp_path / "a.ppm"
    pyarbor.write_ppm(path, px, 2, 2)
    data = path.read_bytes()
    assert data == b"P6\n2 2\n255\n" + bytes([1, 2, 3, 0xFD, 0xFE, 0xFF, 0x78, 0x56, 0x34, 0, 0, 0])
    assert L.arbor_write_ppm(None, None, 2, 2) == -1
    assert L.arbor_write_ppm(str(path).encode(), px.ctypes.data_as(ct.POINTER(ct.c_uint)), 0, 2) == -2
    assert L.arbor_engine_render(None, 1, 4, 1) == -1
    assert L.arbor_engine_save_ppm(None, b"x.ppm") == -1


def test_camera_pose_matches_init_camera(L, oracle):
    # arbor_camera_pose at initCamera's pose gives initCamera's state with the
    # accumulation reset flag C[10] set.
    C = pyarbor.camera_pose(7, (2.0, 1.72, 18.0), np.float32(3.14159265359), 0.0)
    want = oracle.init_camera(7)
    want[10] = 1.0
    assert np.array_equal(C.view(np.uint32), want.view(np.uint32))


def test_batch_walk_views(L):
    cams, sizes = pyarbor.batch_walk_views(3, 7, 64)
    assert cams.shape == (64, 64) and len(sizes) == 64
    assert set(sizes) <= {(256, 256), (640, 360), (1280, 720), (1920, 1080), (2560, 1440)}
    assert len(set(sizes)) >= 4
    assert (np.abs(cams[:, 0]) < 240).all() and (np.abs(cams[:, 2]) < 240).all() and (cams[:, 1] == 1.72).all()
    assert (cams[:, 31] == 7).all() and (cams[:, 22] == 0).all()
    c2, s2 = pyarbor.batch_walk_views(3, 7, 64)
    assert np.array_equal(cams, c2) and sizes == s2


def test_walk_input_deterministic(L):
    a = pyarbor.walk_input(7, 3)
    b = pyarbor.walk_input(7, 3)
    assert np.array_equal(a, b)
    seen = set()
    for f in range(200):
        I = pyarbor.walk_input(7, f)
        assert I[15] in (-1.0, 0.0, 1.0) and I[13] in (-1.0, 0.0, 1.0)
        assert -20.0 <= I[0] < 20.0
        assert (np.delete(I, [0, 13, 15]) == 0).all()
        seen.add((I[15], I[13]))
    assert len(seen) >= 6
    assert not np.array_equal(pyarbor.walk_input(7, 3), pyarbor.walk_input(8, 3))
    assert L.arbor_walk_input(7, -1, (ct.c_float * 32)()) == -2


def test_committed_denoiser_checkpoints_and_costs(L):
    """The committed checkpoints' shapes as read from their files (host code
    only), the null checks of the slot query and the engine's denoiser
    option, and the bench's FLOP count of the 5-level network (379.0 GFLOP
    at 1088x1920), of the 1-level one and of the lean one."""
    import sys

    sys.path.insert(0, REPO)
    import bench

    weights = os.path.join(REPO, "arbor-cuda_amd", "weights")
    assert pyarbor.checkpoint_shape(os.path.join(weights, "denoiser_forest.bin")) == (5, [32, 48, 64, 96, 128])
    assert pyarbor.checkpoint_shape(os.path.join(weights, "denoiser_forest_small.bin")) == (1, [16])
    assert pyarbor.checkpoint_shape(os.path.join(weights, "denoiser_forest_lean.bin")) == (0, [16])
    n = ct.c_int()
    assert L.arbor_denoiser_slot(NULL, ct.byref(n), ct.byref(n), ct.byref(n)) == -1
    assert L.arbor_engine_set_denoiser(NULL, NULL) == -1
    assert round(bench.denoiser_flops(1088, 1920) / 1e9, 4) == 379.0209
    # 1 level: 10 -> 16, 16 -> 16, 16 -> 3 used channels, 18 FLOP per tap pair
    assert bench.denoiser_flops(1088, 1920, [16]) == 18.0 * (10 * 16 + 16 * 16 + 16 * 3) * 1088 * 1920
    # lean: 7 used features -> 16 (3x3), 16 -> 3 used outputs (1x1)
    assert bench.lean_flops(1080, 1920, 16) == 2.0 * (9 * 7 * 16 + 16 * 3) * 1080 * 1920
    assert [l[:3] for l in bench.denoiser_layers([16, 32])] == [(10, 16, 0), (16, 16, 0), (16, 32, 1), (32, 32, 1),
                                                                 (48, 16, 0), (16, 16, 0), (16, 3, 0)]
