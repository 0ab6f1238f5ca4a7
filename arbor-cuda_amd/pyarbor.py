"""Python binding of libarbor.so (the HIP render path) over ctypes.

Device memory and streams come from PyTorch (ROCm build); every compute call
goes through the C-ABI declared in include/arbor.h. There is no fallback: if
the library or a HIP device is missing, calls raise.
"""
import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libarbor.so")

CAPACITY = 32768
NODE_COUNT = 65536
PS = 24
CS = 20
CURVES = 2109
OBJECT_COUNT = 29412

STATUS = {0: "ARBOR_OK", -1: "ARBOR_ERR_NULL", -2: "ARBOR_ERR_ARG", -3: "ARBOR_ERR_LAUNCH",
          -4: "ARBOR_ERR_DEVICE", -5: "ARBOR_ERR_STATE"}

_vp = ct.c_void_p
_i = ct.c_int
_f = ct.c_float

# name -> argtypes (all return int status unless listed in _RESTYPES)
SIGNATURES = {
    "arbor_growTree": [_vp, _i, _i, _vp],
    "arbor_makePrimitives": [_vp, _vp, _i, _vp],
    "arbor_makeKeys": [_vp, _vp, _vp, _vp],
    "arbor_sortPass": [_vp, _vp, _i, _i, _vp],
    "arbor_makeBounds": [_vp, _vp, _vp, _vp],
    "arbor_reduceBounds": [_vp, _i, _vp],
    "arbor_initCamera": [_vp, _vp, _i, _vp],
    "arbor_updateCamera": [_vp, _vp, _vp, _f, _i, _i, _vp],
    "arbor_tracePixels": [_vp, _vp, _vp, _vp, _vp, _i, _i, _vp],
    "arbor_dn_conv3x3": [_vp, _i, _i, _i, _i, _vp, _vp, _i, _vp, _i, _vp],
    "arbor_dn_wgrad_partial_floats": [_i, _i, _i, _i, _i],
    "arbor_dn_conv1x1": [_vp, ct.c_longlong, _i, _vp, _vp, _i, _vp, _i, _vp],
    "arbor_dn_conv3x3_cat": [_vp, _i, _vp, _i, _i, _i, _i, _vp, _vp, _i, _vp, _i, _vp],
    "arbor_dn_conv3x3_wgrad": [_vp, _vp, _i, _i, _i, _i, _i, _vp, ct.c_longlong, _vp, _vp, _vp],
    "arbor_dn_flip_weights": [_vp, _i, _i, _vp, _vp],
    "arbor_dn_pool2": [_vp, _i, _i, _i, _i, _vp, _vp],
    "arbor_dn_pool2_bwd": [_vp, _i, _i, _i, _i, _vp, _vp],
    "arbor_dn_up2": [_vp, _i, _i, _i, _i, _vp, _vp],
    "arbor_dn_up2_bwd": [_vp, _i, _i, _i, _i, _vp, _vp],
    "arbor_dn_concat": [_vp, _i, _vp, _i, ct.c_longlong, _vp, _vp],
    "arbor_dn_split": [_vp, _i, _i, ct.c_longlong, _vp, _vp, _vp],
    "arbor_dn_relu_bwd": [_vp, _vp, ct.c_longlong, _vp],
    "arbor_dn_add": [_vp, _vp, ct.c_longlong, _vp],
    "arbor_shadePixelsAOV": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp],
    "arbor_denoiser_create": [ct.POINTER(_vp), _i, _i, _i, _i, _i],
    "arbor_denoiser_create_net": [ct.POINTER(_vp), _i, _i, _i, _i, _i, _i, ct.POINTER(_i)],
    "arbor_denoiser_shape": [_vp, ct.POINTER(_i), ct.POINTER(_i)],
    "arbor_denoiser_checkpoint_shape": [ct.c_char_p, ct.POINTER(_i), ct.POINTER(_i)],
    "arbor_denoiser_slot": [_vp, ct.POINTER(_i), ct.POINTER(_i), ct.POINTER(_i)],
    "arbor_denoiser_denoise": [_vp, _vp, _vp, _vp, _vp, _i, _i, _vp, _vp],
    "arbor_denoiser_denoise_features": [_vp, _vp, _i, _i, _vp, _vp],
    "arbor_shadePixelsFeatures": [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp],
    "arbor_denoiser_destroy": [_vp],
    "arbor_denoiser_parameters": [_vp],
    "arbor_denoiser_set_input": [_vp, _i, _vp, _vp, _vp, _vp, _i, _i, _vp],
    "arbor_denoiser_set_target": [_vp, _i, _vp, _i, _i, _vp],
    "arbor_denoiser_forward": [_vp, _i, _vp],
    "arbor_denoiser_get_output": [_vp, _i, _vp, _i, _i, _vp],
    "arbor_denoiser_train_step": [_vp, _i, ct.c_float, _vp],
    "arbor_denoiser_loss": [_vp, ct.POINTER(ct.c_float)],
    "arbor_denoiser_save": [_vp, ct.c_char_p],
    "arbor_denoiser_load": [_vp, ct.c_char_p],
    "arbor_denoiser_tensor": [_vp, ct.c_char_p],
    "arbor_engine_render": [_vp, _i, _i, _i],
    "arbor_engine_frame_samples": [_vp, ct.c_float, _vp, _i],
    "arbor_engine_save_ppm": [_vp, ct.c_char_p],
    "arbor_engine_record": [_vp, ct.c_char_p],
    "arbor_engine_replay": [_vp, ct.c_char_p, _i, _i, ct.POINTER(ct.c_int)],
    "arbor_write_ppm": [ct.c_char_p, ct.POINTER(ct.c_uint), _i, _i],
    "arbor_tile_schedule_bytes": [_i, _i],
    "arbor_tracePixelsScheduled": [_vp, _vp, _vp, _vp, _vp, _i, _i, _vp, _vp],
    "arbor_sunVisibility": [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp],
    "arbor_sunVisibilityScheduled": [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp, _vp],
    "arbor_trace_accel_bytes": [],
    "arbor_buildTraceAccel": [_vp, _vp, _vp, _vp],
    "arbor_buildTraceAccelLimit": [_vp, _vp, _vp, ct.c_int, _vp],
    "arbor_tracePixelsAccel": [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp, _vp],
    "arbor_no_alloc_scratch_bytes": [],
    "arbor_tracePixelsAccelNoAlloc": [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp],
    "arbor_sunVisibilityAccel": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp, _vp],
    "arbor_autotune_trace_accel": [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _vp, ct.c_char_p, _i],
    "arbor_shadePixelsVis": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp],
    "arbor_shadePixels": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp],
    "arbor_filterLighting": [_vp, _vp, _vp, _vp, _vp, _i, _i, _vp],
    "arbor_resolvePixels": [_vp, _vp, _vp, _vp, _i, _i, _vp],
    "arbor_sortKeys": [_vp, _vp, _vp],
    "arbor_buildBVH": [_vp, _vp, _vp, _vp, _vp],
    "arbor_generate": [_vp, _vp, _vp, _vp, _vp, _vp, _i, _vp, _i, _i, _vp],
    "arbor_traceTreePixels": [_vp, _vp, _vp, _vp, _vp, _i, _i, _vp],
    "arbor_version": [],
    "arbor_tune_set": [ct.c_char_p, _i],
    "arbor_tune_get": [ct.c_char_p],
    "arbor_tune_save": [ct.c_char_p],
    "arbor_tune_load": [ct.c_char_p],
    "arbor_autotune_trace": [_vp, _vp, _vp, _vp, _vp, _i, _i, _i, _vp, ct.c_char_p, _i],
    "arbor_batch_create": [ct.POINTER(_vp), _i, ct.c_longlong],
    "arbor_batch_destroy": [_vp],
    "arbor_batch_set_views": [_vp, _i, ct.POINTER(_f), ct.POINTER(_i), ct.POINTER(_i)],
    "arbor_batch_render": [_vp, _vp, _vp, _vp, _i, _vp],
    "arbor_batch_capture": [_vp, _vp, _vp, _vp, _i, _vp],
    "arbor_batch_launch": [_vp, _vp],
    "arbor_batch_set_accel": [_vp, _vp],
    "arbor_batch_buffer": [_vp, _i, ct.c_char_p],
    "arbor_batch_pixels": [_vp],
    "arbor_batch_tiles": [_vp],
    "arbor_autotune_batch": [_vp, _vp, _vp, _vp, _i, _vp, ct.c_char_p, _i],
    "arbor_camera_pose": [ct.POINTER(_f), _i, _f, _f, _f, _f, _f],
    "arbor_batch_walk_views": [_i, _i, _i, ct.POINTER(_f), ct.POINTER(_i), ct.POINTER(_i)],
    "arbor_engine_round_size": [ct.POINTER(_i), ct.POINTER(_i)],
    "arbor_engine_create": [ct.POINTER(_vp), _i, _i, _i, _i, _vp],
    "arbor_engine_destroy": [_vp],
    "arbor_engine_generate": [_vp, _i],
    "arbor_engine_resize": [_vp, _i, _i, _i],
    "arbor_engine_reset_camera": [_vp],
    "arbor_engine_set_input": [_vp, ct.POINTER(_f)],
    "arbor_engine_set_timing": [_vp, _i],
    "arbor_engine_set_accel": [_vp, _i],
    "arbor_engine_set_denoiser": [_vp, _vp],
    "arbor_engine_frame": [_vp, _f, _vp],
    "arbor_engine_stage_ms": [_vp, ct.POINTER(_f)],
    "arbor_engine_sync": [_vp],
    "arbor_engine_buffer": [_vp, ct.c_char_p],
    "arbor_engine_stream": [_vp],
    "arbor_engine_size": [_vp, ct.POINTER(_i), ct.POINTER(_i)],
    "arbor_engine_inspect": [_vp, ct.c_char_p, _i],
    "arbor_walk_input": [_i, _i, ct.POINTER(_f)],
}
_RESTYPES = {"arbor_version": ct.c_char_p, "arbor_engine_round_size": None, "arbor_engine_buffer": _vp,
             "arbor_engine_stream": _vp, "arbor_batch_buffer": _vp, "arbor_batch_pixels": ct.c_longlong,
             "arbor_tile_schedule_bytes": ct.c_longlong, "arbor_dn_wgrad_partial_floats": ct.c_longlong,
             "arbor_trace_accel_bytes": ct.c_longlong, "arbor_no_alloc_scratch_bytes": ct.c_longlong,
             "arbor_denoiser_parameters": ct.c_longlong, "arbor_denoiser_tensor": _vp}
BATCH_TRACE_SHADE, BATCH_FILTER, BATCH_RESOLVE, BATCH_ALL = 1, 2, 4, 7


class ArborError(RuntimeError):
    pass


def load_library(path=LIB_PATH):
    """Loads libarbor.so and binds every C-ABI symbol. Importing torch first
    lets the library share torch's HIP runtime (same soname)."""
    if not os.path.exists(path):
        raise ArborError("libarbor.so is not built (%s): run __graft_entry__.build()" % path)
    lib = ct.CDLL(path)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, _i)
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        import torch  # noqa: F401  (shares the HIP runtime with the library)

        _LIB = load_library()
    return _LIB


def check(rc, what=""):
    if rc != 0:
        raise ArborError("%s failed: %s" % (what, STATUS.get(rc, rc)))


def _ptr(t):
    return None if t is None else ct.c_void_p(t.data_ptr())


def _stream(stream=None):
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return ct.c_void_p(s.cuda_stream)


def call(name, *args, stream=None):
    """Calls C-ABI entry `name` with torch tensors / scalars; raises on a
    non-zero status. The HIP stream is appended (torch's current stream)."""
    conv = []
    for a in args:
        if hasattr(a, "data_ptr"):
            conv.append(_ptr(a))
        else:
            conv.append(a)
    rc = getattr(lib(), name)(*conv, _stream(stream))
    check(rc, name)


class DeviceScene:
    """Device buffers B, P, Keys, Order, Nodes, C of one scene."""

    def __init__(self, device="cuda"):
        import torch

        z = dict(device=device)
        self.B = torch.zeros(CURVES * CS, dtype=torch.float32, **z)
        self.P = torch.zeros(CAPACITY * PS, dtype=torch.float32, **z)
        self.Keys = torch.zeros(CAPACITY, dtype=torch.int32, **z)
        self.Order = torch.zeros(CAPACITY, dtype=torch.int32, **z)
        self.Nodes = torch.zeros(NODE_COUNT * 8, dtype=torch.float32, **z)
        self.C = torch.zeros(64, dtype=torch.float32, **z)

    def generate(self, seed, history=None, width=0, height=0):
        call("arbor_generate", self.B, self.P, self.Keys, self.Order, self.Nodes, self.C, seed, history, width, height)
        return self

    def host(self, name):
        import torch

        t = getattr(self, name)
        torch.cuda.synchronize()
        a = t.cpu().numpy()
        return a.view(np.uint32) if t.dtype == torch.int32 else a


class FrameBuffers:
    """Per-view device buffers Hit, Linear, Guide, Filtered, History, Pixels."""

    def __init__(self, width, height, device="cuda"):
        import torch

        n = width * height
        self.width, self.height = width, height
        self.Hit = torch.zeros(n * 4, dtype=torch.float32, device=device)
        self.Linear = torch.zeros(n * 4, dtype=torch.float32, device=device)
        self.Guide = torch.zeros(n * 4, dtype=torch.float32, device=device)
        self.Filtered = torch.zeros(n * 4, dtype=torch.float32, device=device)
        self.History = torch.zeros(n * 4, dtype=torch.float32, device=device)
        self.Pixels = torch.zeros(n, dtype=torch.int32, device=device)


def trace(scene, C, fb):
    call("arbor_tracePixels", scene.P, scene.Nodes, scene.Order, C, fb.Hit, fb.width, fb.height)


class TileSchedule:
    """Device workspace of arbor_tracePixelsScheduled for one view size:
    a header, then per 8x8 tile its recorded cost, its place in the order and
    its bucket key (u32 each). Zeroed at creation, as the library requires."""

    def __init__(self, width, height, device="cuda"):
        import torch

        n = lib().arbor_tile_schedule_bytes(width, height)
        check(n if n < 0 else 0, "arbor_tile_schedule_bytes")
        self.HEADER = lib().arbor_tile_schedule_bytes(8, 8) - 12  # one tile: 12 bytes after the header
        self.buf = torch.zeros(n, dtype=torch.uint8, device=device)
        self.width, self.height = width, height

    def order(self):
        """Host copy of the tile order the last call traced in."""
        a = self.buf.cpu().numpy()
        tiles = ((self.width + 7) // 8) * ((self.height + 7) // 8)
        return a[self.HEADER + 4 * tiles: self.HEADER + 8 * tiles].view(np.uint32).copy()

    def cost(self):
        a = self.buf.cpu().numpy()
        tiles = ((self.width + 7) // 8) * ((self.height + 7) // 8)
        return a[self.HEADER: self.HEADER + 4 * tiles].view(np.uint32).copy()


def trace_scheduled(scene, C, fb, schedule):
    call("arbor_tracePixelsScheduled", scene.P, scene.Nodes, scene.Order, C, fb.Hit, fb.width, fb.height,
         schedule.buf)


class TraceAccel:
    """Device SAH hierarchy of one scene (arbor_buildTraceAccel; with
    depth_limit, arbor_buildTraceAccelLimit)."""

    # u32 words of the hierarchy at the start of the buffer (the build's
    # workspace follows them)
    WORDS = 16 + 16 * (CAPACITY - 1) + CAPACITY

    def __init__(self, scene, device="cuda", depth_limit=None):
        import torch

        n = lib().arbor_trace_accel_bytes()
        self.buf = torch.zeros(n, dtype=torch.uint8, device=device)
        if depth_limit is None:
            call("arbor_buildTraceAccel", scene.Nodes, scene.Order, self.buf)
        else:
            call("arbor_buildTraceAccelLimit", scene.Nodes, scene.Order, self.buf, int(depth_limit))

    def words(self):
        """The hierarchy's u32 words, on the host."""
        import numpy as np

        return self.buf[: 4 * self.WORDS].cpu().numpy().view(np.uint32).copy()


def trace_accel(scene, accel, C, fb, schedule=None):
    call("arbor_tracePixelsAccel", scene.P, scene.Nodes, scene.Order, accel.buf, C, fb.Hit, fb.width, fb.height,
         None if schedule is None else schedule.buf)


def autotune_trace_accel(scene, accel, C, fb, reps=2):
    """Times every accel_variant on this view and keeps the fastest."""
    import json

    buf = ct.create_string_buffer(4096)
    check(lib().arbor_autotune_trace_accel(_ptr(scene.P), _ptr(scene.Nodes), _ptr(scene.Order), _ptr(accel.buf),
                                           _ptr(C), _ptr(fb.Hit), fb.width, fb.height, reps, _stream(), buf, 4096),
          "arbor_autotune_trace_accel")
    return json.loads(buf.value.decode())


def sun_visibility_accel(scene, accel, C, fb, vis, schedule=None):
    call("arbor_sunVisibilityAccel", scene.P, scene.Nodes, scene.Order, accel.buf, C, fb.Hit, vis, fb.width,
         fb.height, None if schedule is None else schedule.buf)


def trace_tree(scene, C, fb):
    call("arbor_traceTreePixels", scene.P, scene.Nodes, scene.Order, C, fb.Hit, fb.width, fb.height)


def shade(scene, C, fb):
    call("arbor_shadePixels", scene.P, scene.Nodes, scene.Order, C, fb.Hit, fb.Linear, fb.Guide, fb.width, fb.height)


def filter_lighting(C, fb):
    call("arbor_filterLighting", fb.Linear, fb.Guide, fb.Hit, fb.Filtered, C, fb.width, fb.height)


def resolve(C, fb):
    call("arbor_resolvePixels", fb.Filtered, fb.History, fb.Pixels, C, fb.width, fb.height)


def tune_set(key, value):
    check(lib().arbor_tune_set(key.encode(), int(value)), "arbor_tune_set(%s)" % key)


def tune_get(key):
    v = lib().arbor_tune_get(key.encode())
    if v < 0:
        raise ArborError("unknown tuning key " + key)
    return v


def autotune_trace(scene, C, fb, reps=3):
    """Runs the single-view autotuner on these buffers; returns its JSON."""
    import json

    buf = ct.create_string_buffer(4096)
    s = _stream()
    rc = lib().arbor_autotune_trace(_ptr(scene.P), _ptr(scene.Nodes), _ptr(scene.Order), _ptr(C), _ptr(fb.Hit),
                                    fb.width, fb.height, reps, s, buf, 4096)
    check(rc, "arbor_autotune_trace")
    return json.loads(buf.value.decode())


def camera_pose(seed, eye, yaw, pitch):
    C = (ct.c_float * 64)()
    check(lib().arbor_camera_pose(C, seed, float(eye[0]), float(eye[1]), float(eye[2]), float(yaw), float(pitch)),
          "arbor_camera_pose")
    return np.frombuffer(C, dtype=np.float32).copy()


def batch_walk_views(seed, world_seed, n):
    """The batched-views workload: (cameras [n,64], [(w, h)] * n)."""
    cams = (ct.c_float * (64 * n))()
    w = (ct.c_int * n)()
    h = (ct.c_int * n)()
    check(lib().arbor_batch_walk_views(seed, world_seed, n, cams, w, h), "arbor_batch_walk_views")
    return np.frombuffer(cams, dtype=np.float32).reshape(n, 64).copy(), list(zip(list(w), list(h)))


class Batch:
    """Batched multi-view rendering (include/arbor_batch.h)."""

    def __init__(self, max_views, max_pixels):
        h = ct.c_void_p()
        check(lib().arbor_batch_create(ct.byref(h), max_views, max_pixels), "arbor_batch_create")
        self.h = h
        self.sizes = []

    def close(self):
        if self.h:
            check(lib().arbor_batch_destroy(self.h), "arbor_batch_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_views(self, cameras, sizes):
        cams = np.ascontiguousarray(cameras, np.float32).reshape(-1)
        n = len(sizes)
        w = (ct.c_int * n)(*[s[0] for s in sizes])
        hh = (ct.c_int * n)(*[s[1] for s in sizes])
        check(lib().arbor_batch_set_views(self.h, n, cams.ctypes.data_as(ct.POINTER(ct.c_float)), w, hh),
              "arbor_batch_set_views")
        self.sizes = list(sizes)

    def render(self, scene, stages=BATCH_ALL, stream=None):
        check(lib().arbor_batch_render(self.h, _ptr(scene.P), _ptr(scene.Nodes), _ptr(scene.Order), stages,
                                       _stream(stream)), "arbor_batch_render")

    def capture(self, scene, stream, stages=BATCH_ALL):
        check(lib().arbor_batch_capture(self.h, _ptr(scene.P), _ptr(scene.Nodes), _ptr(scene.Order), stages,
                                        _stream(stream)), "arbor_batch_capture")

    def launch(self, stream=None):
        check(lib().arbor_batch_launch(self.h, _stream(stream)), "arbor_batch_launch")

    def set_accel(self, accel):
        """Trace over a TraceAccel's SAH hierarchy (None: the heap BVH)."""
        self.accel = accel  # keeps the device buffer alive while the batch uses it
        check(lib().arbor_batch_set_accel(self.h, None if accel is None else _ptr(accel.buf)),
              "arbor_batch_set_accel")

    def autotune(self, scene, reps=2):
        import json

        buf = ct.create_string_buffer(4096)
        check(lib().arbor_autotune_batch(self.h, _ptr(scene.P), _ptr(scene.Nodes), _ptr(scene.Order), reps,
                                         _stream(), buf, 4096), "arbor_autotune_batch")
        return json.loads(buf.value.decode())

    def pixels(self):
        return lib().arbor_batch_pixels(self.h)

    def download(self, view, name):
        """Host copy of view `view`'s slice of buffer `name`."""
        import torch

        torch.cuda.synchronize()
        w, h = self.sizes[view]
        per = {"Pixels": 4, "C": 256}.get(name, 16)
        n = 64 * 4 if name == "C" else w * h * per
        ptr = lib().arbor_batch_buffer(self.h, view, name.encode())
        if not ptr:
            raise ArborError("no buffer %s for view %d" % (name, view))
        host = torch.empty(n, dtype=torch.uint8)
        host.copy_(_wrap_device(ptr, n))
        a = host.numpy()
        return a.view(np.uint32).copy() if name == "Pixels" else a.view(np.float32).copy()


def write_ppm(path, pixels, width, height):
    """Binary PPM (P6) of packed RGBA8 host pixels, as native_harness.cpp:53."""
    px = np.ascontiguousarray(pixels, dtype=np.uint32)
    check(lib().arbor_write_ppm(str(path).encode(), px.ctypes.data_as(ct.POINTER(ct.c_uint)), width, height),
          "arbor_write_ppm")


def sun_visibility(scene, C, fb, vis, schedule=None):
    """Sun-visibility pass into vis (float4 per pixel); schedule: a TileSchedule."""
    if schedule is None:
        call("arbor_sunVisibility", scene.P, scene.Nodes, scene.Order, C, fb.Hit, vis, fb.width, fb.height)
    else:
        call("arbor_sunVisibilityScheduled", scene.P, scene.Nodes, scene.Order, C, fb.Hit, vis, fb.width,
             fb.height, schedule.buf)


def shade_vis(scene, C, fb, vis):
    call("arbor_shadePixelsVis", scene.P, scene.Nodes, scene.Order, C, fb.Hit, vis, fb.Linear, fb.Guide,
         fb.width, fb.height)


def shade_aov(scene, C, fb, normal):
    call("arbor_shadePixelsAOV", scene.P, scene.Nodes, scene.Order, C, fb.Hit, fb.Linear, fb.Guide, normal,
         fb.width, fb.height)


def shade_features(scene, C, fb, feat):
    """shadePixels writing the lean denoiser's compact features (8 bf16 per
    pixel) into feat (a device tensor of at least 16 bytes per pixel)."""
    call("arbor_shadePixelsFeatures", scene.P, scene.Nodes, scene.Order, C, fb.Hit, feat, fb.width, fb.height)


def checkpoint_shape(path):
    """(levels, widths) of the network a denoiser checkpoint was saved from."""
    lv = ct.c_int()
    w = (ct.c_int * 5)()
    check(lib().arbor_denoiser_checkpoint_shape(str(path).encode(), ct.byref(lv), w), "arbor_denoiser_checkpoint_shape")
    return lv.value, list(w)[:max(lv.value, 1)]


class Denoiser:
    """The learned denoiser's network runner (include/arbor_denoise.h). levels
    and widths choose the U-Net (default: 5 levels of 32/48/64/96/128), or
    with levels 0 the lean network of widths[0] (16 or 32) channels."""

    def __init__(self, max_batch, height, width, train=True, seed=1, levels=None, widths=None):
        h = ct.c_void_p()
        if levels is None:
            check(lib().arbor_denoiser_create(ct.byref(h), max_batch, height, width, int(train), seed),
                  "arbor_denoiser_create")
        else:
            w = (ct.c_int * 5)(*(list(widths) + [0] * (5 - len(widths))))
            check(lib().arbor_denoiser_create_net(ct.byref(h), max_batch, height, width, int(train), seed, levels, w),
                  "arbor_denoiser_create_net")
        self.h, self.max_batch, self.height, self.width, self.train = h, max_batch, height, width, train

    @classmethod
    def from_checkpoint(cls, path, max_batch, height, width, train=False, seed=0):
        levels, widths = checkpoint_shape(path)
        d = cls(max_batch, height, width, train=train, seed=seed, levels=levels, widths=widths)
        d.load(path)
        return d

    def shape(self):
        lv = ct.c_int()
        w = (ct.c_int * 5)()
        check(lib().arbor_denoiser_shape(self.h, ct.byref(lv), w), "arbor_denoiser_shape")
        return lv.value, list(w)[:max(lv.value, 1)]

    def slot(self):
        """(max_batch, height, width) of the runner's slots."""
        n, h, w = ct.c_int(), ct.c_int(), ct.c_int()
        check(lib().arbor_denoiser_slot(self.h, ct.byref(n), ct.byref(h), ct.byref(w)), "arbor_denoiser_slot")
        return n.value, h.value, w.value

    def close(self):
        if self.h:
            check(lib().arbor_denoiser_destroy(self.h), "arbor_denoiser_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def parameters(self):
        return lib().arbor_denoiser_parameters(self.h)

    def set_input(self, index, linear, guide, normal, hit, w, h):
        call("arbor_denoiser_set_input", self.h, index, linear, guide, normal, hit, w, h)

    def set_target(self, index, radiance, w, h):
        call("arbor_denoiser_set_target", self.h, index, radiance, w, h)

    def forward(self, n):
        call("arbor_denoiser_forward", self.h, n)

    def output(self, index, out, w, h):
        call("arbor_denoiser_get_output", self.h, index, out, w, h)

    def denoise(self, linear, guide, normal, hit, w, h, out):
        """One frame: features, forward and output (slot 0) into out."""
        call("arbor_denoiser_denoise", self.h, linear, guide, normal, hit, w, h, out)

    def denoise_features(self, feat, w, h, out):
        """The lean network's one pass from compact features into out."""
        call("arbor_denoiser_denoise_features", self.h, feat, w, h, out)

    def train_step(self, n, lr):
        call("arbor_denoiser_train_step", self.h, n, ct.c_float(lr))

    def loss(self):
        v = ct.c_float()
        check(lib().arbor_denoiser_loss(self.h, ct.byref(v)), "arbor_denoiser_loss")
        return v.value

    def save(self, path):
        check(lib().arbor_denoiser_save(self.h, str(path).encode()), "arbor_denoiser_save")

    def load(self, path):
        check(lib().arbor_denoiser_load(self.h, str(path).encode()), "arbor_denoiser_load")

    def tensor_ptr(self, name):
        """Device address of a named tensor (arbor_denoiser_tensor)."""
        ptr = lib().arbor_denoiser_tensor(self.h, name.encode())
        if not ptr:
            raise ArborError("no denoiser tensor " + name)
        return ptr

    def tensor(self, name, count, dtype):
        """Host copy of `count` elements of a named device tensor."""
        import torch

        ptr = lib().arbor_denoiser_tensor(self.h, name.encode())
        if not ptr:
            raise ArborError("no denoiser tensor " + name)
        nbytes = count * np.dtype(dtype).itemsize
        torch.cuda.synchronize()
        host = torch.empty(nbytes, dtype=torch.uint8)
        host.copy_(_wrap_device(ptr, nbytes))
        return host.numpy().view(dtype).copy()


def walk_input(seed, frame):
    out = (ct.c_float * 32)()
    check(lib().arbor_walk_input(seed, frame, out), "arbor_walk_input")
    return np.frombuffer(out, dtype=np.float32).copy()


class Engine:
    """The C++ engine (include/arbor_engine.h): generate / resize / frame."""

    def __init__(self, width, height, seed, round_size=False, stream=None):
        self._lib = lib()
        h = ct.c_void_p()
        s = _stream(stream) if stream is not None else _stream()
        check(self._lib.arbor_engine_create(ct.byref(h), width, height, seed, int(round_size), s),
              "arbor_engine_create")
        self.h = h
        self._dn = None

    def close(self):
        if self.h:
            check(self._lib.arbor_engine_destroy(self.h), "arbor_engine_destroy")
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def generate(self, seed):
        check(self._lib.arbor_engine_generate(self.h, seed), "arbor_engine_generate")

    def resize(self, w, h, round_size=False):
        check(self._lib.arbor_engine_resize(self.h, w, h, int(round_size)), "arbor_engine_resize")

    def size(self):
        w, h = ct.c_int(), ct.c_int()
        check(self._lib.arbor_engine_size(self.h, ct.byref(w), ct.byref(h)), "arbor_engine_size")
        return w.value, h.value

    def reset_camera(self):
        check(self._lib.arbor_engine_reset_camera(self.h), "arbor_engine_reset_camera")

    def set_input(self, inp):
        arr = (ct.c_float * 32)(*[float(v) for v in inp])
        check(self._lib.arbor_engine_set_input(self.h, arr), "arbor_engine_set_input")

    def set_timing(self, on):
        check(self._lib.arbor_engine_set_timing(self.h, int(on)), "arbor_engine_set_timing")

    def set_accel(self, on):
        """Trace over the SAH hierarchy (default) or over the heap BVH."""
        check(self._lib.arbor_engine_set_accel(self.h, int(on)), "arbor_engine_set_accel")

    def set_denoiser(self, dn):
        """Denoise frames that fit dn's slot with the network (a Denoiser, or
        None for filterLighting); the engine keeps a reference while attached."""
        check(self._lib.arbor_engine_set_denoiser(self.h, None if dn is None else dn.h), "arbor_engine_set_denoiser")
        self._dn = dn

    def frame(self, dt=1.0 / 60.0, device_input=None):
        dev = None if device_input is None else ct.c_void_p(device_input.data_ptr())
        check(self._lib.arbor_engine_frame(self.h, dt, dev), "arbor_engine_frame")

    def stage_ms(self):
        out = (ct.c_float * 6)()
        check(self._lib.arbor_engine_stage_ms(self.h, out), "arbor_engine_stage_ms")
        return list(out)

    def sync(self):
        check(self._lib.arbor_engine_sync(self.h), "arbor_engine_sync")

    def buffer_ptr(self, name):
        return self._lib.arbor_engine_buffer(self.h, name.encode())

    def frame_samples(self, samples, dt=1.0 / 60.0, device_input=None):
        ptr = None if device_input is None else ct.c_void_p(device_input.data_ptr())
        check(self._lib.arbor_engine_frame_samples(self.h, ct.c_float(dt), ptr, samples), "arbor_engine_frame_samples")

    def render(self, view, samples, light=True):
        """The reference harness's render mode (view preset, `samples`
        jittered samples accumulated by resolvePixels)."""
        check(self._lib.arbor_engine_render(self.h, view, samples, int(bool(light))), "arbor_engine_render")

    def save_ppm(self, path):
        check(self._lib.arbor_engine_save_ppm(self.h, str(path).encode()), "arbor_engine_save_ppm")

    def record(self, path):
        """Record every later frame's inputs to path (None stops)."""
        check(self._lib.arbor_engine_record(self.h, None if path is None else str(path).encode()),
              "arbor_engine_record")

    def replay(self, path, first=0, count=0):
        """Run recorded frames; returns how many ran."""
        n = ct.c_int(0)
        check(self._lib.arbor_engine_replay(self.h, str(path).encode(), first, count, ct.byref(n)),
              "arbor_engine_replay")
        return n.value

    def inspect(self):
        import json

        buf = ct.create_string_buffer(4096)
        check(self._lib.arbor_engine_inspect(self.h, buf, 4096), "arbor_engine_inspect")
        return json.loads(buf.value.decode())

    def download(self, name, count, dtype=np.float32):
        """Copies `count` elements of a named engine buffer to the host."""
        import torch

        self.sync()
        ptr = self.buffer_ptr(name)
        if not ptr:
            raise ArborError("unknown buffer " + name)
        nbytes = count * np.dtype(dtype).itemsize
        host = torch.empty(nbytes, dtype=torch.uint8)
        # device -> host through a torch tensor view of the engine allocation
        dev = _wrap_device(ptr, nbytes)
        host.copy_(dev)
        return host.numpy().view(dtype).copy()


def _wrap_device(ptr, nbytes):
    """A uint8 torch tensor aliasing `nbytes` of device memory at `ptr`."""
    import torch

    class _Iface:
        def __init__(self):
            self.__cuda_array_interface__ = {"shape": (nbytes,), "typestr": "|u1", "data": (ptr, False),
                                             "version": 3, "strides": None}

    return torch.as_tensor(_Iface(), device="cuda")
