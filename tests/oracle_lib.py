"""ctypes bindings for the oracle (test infrastructure only).

`Oracle` wraps oracle/liboracle.so, the C++ restatement of the reference
kernels (float and double instantiations). `RefHarness` wraps
oracle/_ref/libarborref.so, the reference CPU harness compiled from its own
sources; it exists only where it was built (it is never required on a GPU box).
Both follow the launch emulation of tests/native_harness.cpp:35-39.
"""
import ctypes as ct
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "liboracle.so")
REF_SO = os.path.join(REPO, "oracle", "_ref", "libarborref.so")

CAPACITY = 32768
NODE_COUNT = 65536
PS = 24
CS = 20
CURVES = 2109
OBJECT_COUNT = 29412
FAR = np.float32(100000.0)

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i = ct.c_int
_f = ct.c_float


class Scene:
    """Host copies of the build outputs B, P, Keys, Order, Nodes."""

    def __init__(self):
        self.B = np.zeros(CURVES * CS, np.float32)
        self.P = np.zeros(CAPACITY * PS, np.float32)
        self.Keys = np.zeros(CAPACITY, np.uint32)
        self.Order = np.zeros(CAPACITY, np.uint32)
        self.Nodes = np.zeros(NODE_COUNT * 8, np.float32)


def _bind(lib, name, res, args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = args
    return fn


class _Common:
    """Entry points both libraries share (prefix orc_ or ref_)."""

    def _setup_common(self, p):
        L = self.lib
        self._generate = _bind(L, p + "generate", _i, [_i, _f32p, _f32p, _u32p, _u32p, _f32p, _i])
        self._init_camera = _bind(L, p + "init_camera", _i, [_f32p, _i])
        self._update_camera = _bind(L, p + "update_camera", _i, [_f32p, _f32p, _f, _i, _i])
        self._trace = _bind(L, p + "trace_pixels" + self._suf, _i, [_f32p, _f32p, _u32p, _f32p, _f32p, _i, _i, _i])
        self._trace_tree = _bind(L, p + "trace_tree_pixels" + self._suf, _i,
                                 [_f32p, _f32p, _u32p, _f32p, _f32p, _i, _i, _i])
        self._shade = _bind(L, p + "shade_pixels" + self._suf, _i,
                            [_f32p, _f32p, _u32p, _f32p, _f32p, _f32p, _f32p, _i, _i, _i])
        self._filter = _bind(L, p + "filter_lighting" + self._suf, _i,
                             [_f32p, _f32p, _f32p, _f32p, _f32p, _i, _i, _i])
        self._resolve = _bind(L, p + "resolve_pixels" + self._suf, _i, [_f32p, _f32p, _u32p, _f32p, _i, _i, _i])
        self._tree_rays = _bind(L, p + "trace_tree_rays", _i,
                                [_f32p, _f32p, _u32p, _f32p, _f32p, _i, _f, _f, _f, _f, _i, _f32p])
        self._bark_cut = _bind(L, p + "bark_cut", _f, [_f32p, _i, _f, _f, _f, _f])
        self._leaf_t = _bind(L, p + "leaf_t", _f, [_f32p, _i, _f32p, _f32p, _f, _f, _f])
        self._grow = _bind(L, p + "grow_tree", _i, [_f32p, _i, _i])
        self._hash_u = _bind(L, p + "hash_u", ct.c_uint, [ct.c_uint])
        self._bark_sweep = _bind(L, p + "bark_sweep", _f, [_f32p])
        self._leaf_probes = _bind(L, p + "leaf_probes", _i, [_f32p, _f32p])

    def generate(self, seed, bitonic=False):
        s = Scene()
        rc = self._generate(seed, s.B, s.P, s.Keys, s.Order, s.Nodes, int(bitonic))
        assert rc == 0
        return s

    def init_camera(self, seed):
        C = np.zeros(64, np.float32)
        self._init_camera(C, seed)
        return C

    def update_camera(self, C, I, delta, w, h):
        self._update_camera(C, np.ascontiguousarray(I, np.float32), delta, w, h)

    def trace(self, s, C, w, h, threads=0):
        Hit = np.zeros(w * h * 4, np.float32)
        self._trace(s.P, s.Nodes, s.Order, C, Hit, w, h, threads)
        return Hit

    def trace_tree(self, s, C, w, h, threads=0):
        Hit = np.zeros(w * h * 4, np.float32)
        self._trace_tree(s.P, s.Nodes, s.Order, C, Hit, w, h, threads)
        return Hit

    def shade(self, s, C, Hit, w, h, threads=0):
        Linear = np.zeros(w * h * 4, np.float32)
        Guide = np.zeros(w * h * 4, np.float32)
        self._shade(s.P, s.Nodes, s.Order, C, Hit, Linear, Guide, w, h, threads)
        return Linear, Guide

    def filter(self, Linear, Guide, Hit, C, w, h, threads=0):
        out = np.zeros(w * h * 4, np.float32)
        self._filter(Linear, Guide, Hit, out, C, w, h, threads)
        return out

    def resolve(self, Filtered, History, C, w, h, threads=0):
        Pixels = np.zeros(w * h, np.uint32)
        self._resolve(Filtered, History, Pixels, C, w, h, threads)
        return Pixels

    def trace_tree_rays(self, s, ro, rd, cone=0.0005, time=0.0, wind=0.0, best=float(FAR), shadows=False):
        ro = np.ascontiguousarray(ro, np.float32).reshape(-1)
        rd = np.ascontiguousarray(rd, np.float32).reshape(-1)
        n = ro.size // 3
        out = np.zeros(2 * n, np.float32)
        self._tree_rays(s.P, s.Nodes, s.Order, ro, rd, n, cone, time, wind, best, int(shadows), out)
        return out.reshape(n, 2)

    def bark_cut(self, P, oid, p, footprint):
        return self._bark_cut(P, oid, float(p[0]), float(p[1]), float(p[2]), footprint)

    def leaf_t(self, P, oid, ro, rd, footprint, time=0.0, wind=0.0):
        return self._leaf_t(P, oid, np.asarray(ro, np.float32), np.asarray(rd, np.float32), footprint, time, wind)

    def bark_sweep(self, P):
        return self._bark_sweep(P)

    def leaf_probes(self, P):
        out = np.zeros(100, np.float32)
        self._leaf_probes(P, out)
        return out.reshape(50, 2)

    def grow_tree(self, B, level, seed):
        return self._grow(B, level, seed)

    def hash_u(self, x):
        return self._hash_u(x)


class Oracle(_Common):
    """The C++ restatement (oracle/liboracle.so)."""

    _suf = "_f32"

    def __init__(self, path=ORACLE_SO):
        self.lib = ct.CDLL(path)
        self._setup_common("orc_")
        L = self.lib
        self._trace64 = _bind(L, "orc_trace_pixels_f64", _i, [_f64p, _f64p, _u32p, _f64p, _f64p, _i, _i, _i])
        self._trace_list = _bind(L, "orc_trace_pixel_list_f32", _i,
                                 [_f32p, _f32p, _u32p, _f32p, _i32p, _i, _i, _i, _f32p, _i])
        self._trace_list64 = _bind(L, "orc_trace_pixel_list_f64", _i,
                                   [_f64p, _f64p, _u32p, _f64p, _i32p, _i, _i, _i, _f64p, _i])
        self._trace_off = _bind(L, "orc_trace_pixel_offsets_f32", _i,
                                [_f32p, _f32p, _u32p, _f32p, _i32p, _f32p, _i, _i, _i, _f32p, _i, _i])
        self._trace_off64 = _bind(L, "orc_trace_pixel_offsets_f64", _i,
                                  [_f64p, _f64p, _u32p, _f64p, _i32p, _f32p, _i, _i, _i, _f64p, _i, _i])
        self._trace_tree64 = _bind(L, "orc_trace_tree_pixels_f64", _i, [_f64p, _f64p, _u32p, _f64p, _f64p, _i, _i, _i])
        self._shade64 = _bind(L, "orc_shade_pixels_f64", _i,
                              [_f64p, _f64p, _u32p, _f64p, _f64p, _f64p, _f64p, _i, _i, _i])
        self._filter64 = _bind(L, "orc_filter_lighting_f64", _i, [_f64p, _f64p, _f64p, _f64p, _f64p, _i, _i, _i])
        self._resolve64 = _bind(L, "orc_resolve_pixels_f64", _i, [_f64p, _f64p, _u32p, _f64p, _i, _i, _i])
        self._make_keys = _bind(L, "orc_make_keys", _i, [_f32p, _u32p, _u32p])
        self._sort_pass = _bind(L, "orc_sort_pass", _i, [_u32p, _u32p, _i, _i])
        self._sort_pairs = _bind(L, "orc_sort_pairs", _i, [_u32p, _u32p, _i])
        self._make_bounds = _bind(L, "orc_make_bounds", _i, [_f32p, _u32p, _f32p])
        self._reduce_bounds = _bind(L, "orc_reduce_bounds", _i, [_f32p, _i])
        self._make_prims = _bind(L, "orc_make_primitives", _i, [_f32p, _f32p, _i])
        self._exhaustive = _bind(L, "orc_exhaustive_rays", _i, [_f32p, _f32p, _f32p, _i, _f, _f32p])
        self._spread = _bind(L, "orc_spread", ct.c_uint, [ct.c_uint])
        self._band = _bind(L, "orc_band_weight", _f, [_f, _f])
        self._obj_t = _bind(L, "orc_object_t", _f, [_f32p, _i, _f32p, _f32p, _f, _f, _f, _i])
        self._sphere_t = _bind(L, "orc_sphere_t", _f, [_f32p, _f32p, _f32p, _f])
        self._tube_t = _bind(L, "orc_tube_t", _f, [_f32p, _f32p, _f32p, _f32p, _f, _f])
        self._aces = _bind(L, "orc_aces", _f, [_f])
        self._forest = _bind(L, "orc_forest_table", _i, [_i, _f32p])
        self._count_ops = _bind(L, "orc_count_trace_ops", _i,
                                [_f32p, _f32p, _u32p, _f32p, _i, _i, _i, _i,
                                 np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")])
        self._lean_mismatch = _bind(L, "orc_lean_walk_mismatches", _i, [_f32p, _f32p, _u32p, _f32p, _i, _i, _i])
        _i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
        self._build_sah = _bind(L, "orc_build_sah", _i, [_f32p, _u32p, _u32p, ct.c_longlong])
        self._build_sah_limit = _bind(L, "orc_build_sah_limit", _i, [_f32p, _u32p, _u32p, ct.c_longlong, _i])
        self._sah_mismatch = _bind(L, "orc_sah_walk_mismatches", _i,
                                   [_f32p, _f32p, _u32p, _f32p, _i, _i, _u32p, _i, _i64p])
        self._set_sah_tie_reversed = _bind(L, "orc_set_sah_tie_reversed", None, [_i])
        self._cull_mismatch = _bind(L, "orc_cull_mismatches", _i, [_f32p, _f32p, _u32p, _f32p, _i, _i])
        self._skip_mismatch = _bind(L, "orc_skip_mismatches", _i, [_f32p, _f32p, _u32p, _f32p, _i, _i])
        self._newton_mismatch = _bind(L, "orc_newton_exit_mismatches", _i, [_f32p, _f32p, _u32p, _f32p, _i, _i])
        self._band_mismatch = _bind(L, "orc_band_skip_mismatches", _i, [_f32p, _f32p, _u32p, _f32p, _i, _i])
        self._band_object_mismatch = _bind(L, "orc_band_skip_object_mismatches", _i, [_f32p, _i])
        self._wind_mismatch = _bind(L, "orc_wind_shortcut_mismatches", _i, [_f32p, _f32p, _u32p, _f32p, _i, _i])
        self._generate_cr = _bind(L, "orc_generate_cr", _i, [_i, _f32p, _f32p, _u32p, _u32p, _f32p])
        self._cull_stress = _bind(L, "orc_cull_stress", _i,
                                  [_f32p, _i, _i, _f, _f, np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")])
        self._cull_pad_scale = _bind(L, "orc_set_cull_pad_scale", None, [_f])
        self._exit_stress = _bind(L, "orc_exit_stress", _i,
                                  [_f32p, _i, _i, _f, _f, np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")])

    def cull_stress(self, P, rays_per_object, seed, time, wind):
        """Grazing-ray stress of the rejection spheres (flops.cpp orc_cull_stress):
        (rays, rays with a hit, rays whose results differ, nearest distance of a
        removed hit)."""
        out = np.zeros(4, np.int64)
        assert self._cull_stress(P, rays_per_object, seed, time, wind, out) == 0
        return tuple(int(v) for v in out)

    def exit_stress(self, P, rays_per_object, seed, time, wind):
        """The early exits past the caller's best under the grazing rays, at
        five bests per test (flops.cpp orc_exit_stress): (evaluations, those
        whose full result is below the best, violations, answers an exit
        changed at or above the best, nearest distance of a violation)."""
        out = np.zeros(5, np.int64)
        assert self._exit_stress(P, rays_per_object, seed, time, wind, out) == 0
        return tuple(int(v) for v in out)

    def set_cull_pad_scale(self, scale):
        self._cull_pad_scale(scale)

    def generate_cr(self, seed):
        """The build with correctly rounded sin/cos/pow instead of glibc's."""
        s = Scene()
        assert self._generate_cr(seed, s.B, s.P, s.Keys, s.Order, s.Nodes) == 0
        return s

    def count_trace_ops(self, s, C, w, h, stride, single_tree=False):
        """Operations per primary ray (lean walk), sampled every `stride` pixels.
        Returns dict with rays, ops_per_ray and the per-kind breakdown per ray."""
        out = np.zeros(9, np.float64)
        rc = self._count_ops(s.P, s.Nodes, s.Order, np.ascontiguousarray(C, np.float32), w, h, stride,
                             int(single_tree), out)
        assert rc == 0
        rays = out[0]
        kinds = ("add", "mul", "div", "sqrt", "minmax", "absfloor", "libm")
        return {"rays": int(rays), "ops_per_ray": out[1] / rays,
                **{k: out[2 + i] / rays for i, k in enumerate(kinds)}}

    def cull_mismatches(self, s, C, w, h):
        """Pixels whose Hit or Linear changes when woodT/leafT first reject
        rays missing their padded bounding spheres (the GPU's culls)."""
        return self._cull_mismatch(s.P, s.Nodes, s.Order, np.ascontiguousarray(C, np.float32), w, h)

    def newton_exit_mismatches(self, s, C, w, h):
        """Pixels whose Hit or Linear changes when the leaf Newton solve stops
        at a bit-exact fixed point or 2-cycle (0 expected)."""
        return self._newton_mismatch(s.P, s.Nodes, s.Order, np.ascontiguousarray(C, np.float32), w, h)

    def band_skip_mismatches(self, s, C, w, h):
        """Pixels whose Hit or Linear changes when band-weighted leaf terms
        with a zero band weight are skipped (0 expected)."""
        return self._band_mismatch(s.P, s.Nodes, s.Order, np.ascontiguousarray(C, np.float32), w, h)

    def band_skip_object_mismatches(self, s, n):
        """Single-object form of band_skip_mismatches: n seeded leaf tests
        (with leafHeight/leafWidth) and n wood cases (barkCut, woodField,
        woodT) across the bands' transitions."""
        return self._band_object_mismatch(s.P, n)

    def wind_shortcut_mismatches(self, s, C, w, h):
        """Pixels whose Hit or Linear changes when the leaf basis is taken as
        stored with the breeze off (0 expected)."""
        return self._wind_mismatch(s.P, s.Nodes, s.Order, np.ascontiguousarray(C, np.float32), w, h)

    def skip_mismatches(self, s, C, w, h):
        """Pixels whose Hit or Linear changes when woodT/leafT stop once their
        answer is known to be >= the caller's best (the GPU's early exits)."""
        return self._skip_mismatch(s.P, s.Nodes, s.Order, np.ascontiguousarray(C, np.float32), w, h)

    SAH_WORDS = 16 + 16 * (32768 - 1) + 32768

    def build_sah(self, s, depth_limit=None):
        """The GPU accel walk's SAH hierarchy of this scene (u32 words); with
        depth_limit, the builder's bound set lower than the walk's 32."""
        out = np.zeros(self.SAH_WORDS, np.uint32)
        if depth_limit is None:
            assert self._build_sah(s.Nodes, s.Order, out, out.size) == 0
        else:
            assert self._build_sah_limit(s.Nodes, s.Order, out, out.size, depth_limit) == 0
        return out

    def sah_walk_mismatches(self, s, C, w, h, accel, mode=0, tie_reversed=False):
        """(pixels differing, of them walk-order-dependent object hits, first
        differing pixel) between the SAH walk and the heap walk: mode 0 Hit,
        mode 1 Linear with sun visibility as C sets it."""
        out = np.zeros(3, np.int64)
        self._set_sah_tie_reversed(int(tie_reversed))
        try:
            self._sah_mismatch(s.P, s.Nodes, s.Order, np.ascontiguousarray(C, np.float32), w, h, accel, mode, out)
        finally:
            self._set_sah_tie_reversed(0)
        return tuple(int(v) for v in out)

    def lean_walk_mismatches(self, s, C, w, h, single_tree=False):
        return self._lean_mismatch(s.P, s.Nodes, s.Order, np.ascontiguousarray(C, np.float32), w, h,
                                   int(single_tree))

    # -- FP64 run: float inputs promoted to double, arithmetic in double --
    def trace_pixels(self, s, C, xy, w, h, threads=0, fp64=False):
        """Hit records of the listed pixels (xy: n x 2 ints) of a w x h view,
        float32 or (fp64) float64."""
        xy = np.ascontiguousarray(xy, np.int32).reshape(-1)
        n = xy.size // 2
        if fp64:
            out = np.zeros(4 * n, np.float64)
            self._trace_list64(s.P.astype(np.float64), s.Nodes.astype(np.float64), s.Order, C.astype(np.float64), xy,
                               n, w, h, out, threads)
        else:
            out = np.zeros(4 * n, np.float32)
            self._trace_list(s.P, s.Nodes, s.Order, np.ascontiguousarray(C, np.float32), xy, n, w, h, out, threads)
        return out

    def trace_pixel_offsets(self, s, C, xy, dxy, w, h, threads=0, fp64=False, single=False):
        """Hit records of primary rays through points offset by dxy (n x 2, in
        pixels) from the listed pixels' sample positions; dxy = 0 gives the
        pixels' own Hit records. single: the one-tree trace (traceTreePixels)."""
        xy = np.ascontiguousarray(xy, np.int32).reshape(-1)
        dxy = np.ascontiguousarray(dxy, np.float32).reshape(-1)
        n = xy.size // 2
        assert dxy.size == 2 * n
        if fp64:
            out = np.zeros(4 * n, np.float64)
            self._trace_off64(s.P.astype(np.float64), s.Nodes.astype(np.float64), s.Order, C.astype(np.float64), xy,
                              dxy, n, w, h, out, threads, int(single))
        else:
            out = np.zeros(4 * n, np.float32)
            self._trace_off(s.P, s.Nodes, s.Order, np.ascontiguousarray(C, np.float32), xy, dxy, n, w, h, out,
                            threads, int(single))
        return out.reshape(n, 4)

    def tie_answers(self, s, C, xy, w, h, eps, single=False):
        """For each listed pixel, the set of (id, tree) answers that the FP32
        and FP64 oracle give for the pixel's ray and for the 8 rays through
        points eps pixels away from its sample position (axes and diagonals).
        A ray is within eps of a tie when this set has more than one answer."""
        xy = np.asarray(xy, np.int32).reshape(-1, 2)
        offs = [(0.0, 0.0)] + [(eps * a, eps * b) for a in (-1, 0, 1) for b in (-1, 0, 1) if (a, b) != (0, 0)]
        allxy = np.repeat(xy, len(offs), axis=0)
        alld = np.tile(np.array(offs, np.float32), (len(xy), 1))
        answers = [set() for _ in range(len(xy))]
        for fp64 in (False, True):
            hits = self.trace_pixel_offsets(s, C, allxy, alld, w, h, fp64=fp64, single=single)
            hits = hits.reshape(len(xy), len(offs), 4)
            for i in range(len(xy)):
                for hrec in hits[i]:
                    answers[i].add((int(hrec[1]), int(hrec[2])))
        return answers

    def trace64(self, s, C, w, h, threads=0):
        Hit = np.zeros(w * h * 4, np.float64)
        self._trace64(s.P.astype(np.float64), s.Nodes.astype(np.float64), s.Order, C.astype(np.float64), Hit, w, h,
                      threads)
        return Hit

    def trace_tree64(self, s, C, w, h, threads=0):
        Hit = np.zeros(w * h * 4, np.float64)
        self._trace_tree64(s.P.astype(np.float64), s.Nodes.astype(np.float64), s.Order, C.astype(np.float64), Hit,
                           w, h, threads)
        return Hit

    def shade64(self, s, C, Hit, w, h, threads=0):
        Linear = np.zeros(w * h * 4, np.float64)
        Guide = np.zeros(w * h * 4, np.float64)
        self._shade64(s.P.astype(np.float64), s.Nodes.astype(np.float64), s.Order, C.astype(np.float64),
                      np.ascontiguousarray(Hit, np.float64), Linear, Guide, w, h, threads)
        return Linear, Guide

    def filter64(self, Linear, Guide, Hit, C, w, h, threads=0):
        out = np.zeros(w * h * 4, np.float64)
        self._filter64(np.ascontiguousarray(Linear, np.float64), np.ascontiguousarray(Guide, np.float64),
                       np.ascontiguousarray(Hit, np.float64), out, C.astype(np.float64), w, h, threads)
        return out

    def make_keys(self, P):
        k = np.zeros(CAPACITY, np.uint32)
        o = np.zeros(CAPACITY, np.uint32)
        self._make_keys(P, k, o)
        return k, o

    def sort_pass(self, k, o, distance, span):
        self._sort_pass(k, o, distance, span)

    def sort_pairs(self, k, o):
        self._sort_pairs(k, o, k.size)

    def make_bounds(self, P, Order, Nodes):
        self._make_bounds(P, Order, Nodes)

    def reduce_bounds(self, Nodes, first):
        self._reduce_bounds(Nodes, first)

    def make_primitives(self, B, seed):
        P = np.zeros(CAPACITY * PS, np.float32)
        self._make_prims(B, P, seed)
        return P

    def exhaustive_rays(self, P, ro, rd, cone=0.0005):
        ro = np.ascontiguousarray(ro, np.float32).reshape(-1)
        rd = np.ascontiguousarray(rd, np.float32).reshape(-1)
        n = ro.size // 3
        out = np.zeros(2 * n, np.float32)
        self._exhaustive(P, ro, rd, n, cone, out)
        return out.reshape(n, 2)

    def spread(self, x):
        return self._spread(x)

    def band_weight(self, period, fp):
        return self._band(period, fp)

    def object_t(self, P, oid, ro, rd, fp, time=0.0, wind=0.0, refine=True):
        return self._obj_t(P, oid, np.asarray(ro, np.float32), np.asarray(rd, np.float32), fp, time, wind, int(refine))

    def sphere_t(self, ro, rd, c, r):
        return self._sphere_t(np.asarray(ro, np.float32), np.asarray(rd, np.float32), np.asarray(c, np.float32), r)

    def tube_t(self, ro, rd, a, b, ra, rb):
        return self._tube_t(*(np.asarray(v, np.float32) for v in (ro, rd, a, b)), ra, rb)

    def aces(self, x):
        return self._aces(x)

    def forest_table(self, world_seed):
        out = np.zeros(324 * 7, np.float32)
        self._forest(world_seed, out)
        return out.reshape(324, 7)


class RefHarness(_Common):
    """The reference harness compiled from its sources (oracle/_ref/libarborref.so)."""

    _suf = ""

    def __init__(self, path=REF_SO):
        self.lib = ct.CDLL(path)
        self._setup_common("ref_")
        self._rnd = _bind(self.lib, "ref_rnd", _f, [_i, _i])
        self._rays = _bind(self.lib, "ref_harness_rays", _i, [_f32p, _f32p])

    def rnd(self, n, seed):
        return self._rnd(n, seed)

    def harness_rays(self):
        ro = np.zeros(300, np.float32)
        rd = np.zeros(300, np.float32)
        self._rays(ro, rd)
        return ro.reshape(100, 3), rd.reshape(100, 3)


def ref_available():
    return os.path.exists(REF_SO)


def fnv1a64(arr):
    """FNV-1a 64-bit over the raw little-endian bytes of an array."""
    data = np.ascontiguousarray(arr).view(np.uint8)
    h = np.uint64(0xCBF29CE484222325)
    prime = np.uint64(0x100000001B3)
    # vectorised in chunks is not possible for FNV; do it in Python over bytes
    hv = int(h)
    for b in data.tobytes():
        hv ^= b
        hv = (hv * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return "%016x" % hv


def digest(arr):
    """Fast content digest (sha256 of raw bytes), used by the golden fixtures."""
    import hashlib

    return hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest()


# Frames of the shadowed, windy forest setting (seed 7, 96x64), written by the
# reference harness into the golden fixtures and replayed by the oracle tests.
# Cases sharing a "sequence" run in order and carry History from one to the
# next: "prog" is one pixel's four progressive samples (C[9]=0..3, reset only
# on the first), the form the 4-sample shadowed workload renders per frame.
SUN_35_DEG = 0.610865
C3_CASES = (
    dict(name="prog0", sequence="prog", view=3, time=0.75, sample=0, reset=1, light=1, wind=0.6,
         sun=SUN_35_DEG, debug=0, history_seed=11),
    dict(name="prog1", sequence="prog", view=3, time=0.75, sample=1, reset=0, light=1, wind=0.6,
         sun=SUN_35_DEG, debug=0, history_seed=11),
    dict(name="prog2", sequence="prog", view=3, time=0.75, sample=2, reset=0, light=1, wind=0.6,
         sun=SUN_35_DEG, debug=0, history_seed=11),
    dict(name="prog3", sequence="prog", view=3, time=0.75, sample=3, reset=0, light=1, wind=0.6,
         sun=SUN_35_DEG, debug=0, history_seed=11),
    dict(name="hist2", sequence="hist2", view=1, time=1.3, sample=2, reset=0, light=1, wind=0.6,
         sun=SUN_35_DEG, debug=0, history_seed=23),
    dict(name="late37", sequence="late37", view=4, time=1.3, sample=37, reset=0, light=0, wind=0.6,
         sun=SUN_35_DEG, debug=0, history_seed=5),
    dict(name="calm1", sequence="calm1", view=1, time=0.0, sample=1, reset=0, light=1, wind=0.0,
         sun=None, debug=0, history_seed=7),
    dict(name="late200", sequence="late200", view=3, time=2.9, sample=200, reset=0, light=1, wind=0.6,
         sun=SUN_35_DEG, debug=0, history_seed=3),
    dict(name="debug1", sequence="debug1", view=3, time=0.75, sample=1, reset=0, light=0, wind=0.6,
         sun=None, debug=1, history_seed=13),
    dict(name="debug2", sequence="debug2", view=3, time=0.75, sample=1, reset=0, light=0, wind=0.6,
         sun=None, debug=2, history_seed=13),
    dict(name="debug3", sequence="debug3", view=3, time=0.75, sample=1, reset=0, light=1, wind=0.6,
         sun=None, debug=3, history_seed=13),
)


def c3_camera(lib, case, w, h, seed=7):
    """Camera state of one C3_CASES entry: initCamera, one updateCamera with the
    preset view, then the wind time, sample, reset, sun and debug slots set
    directly (Arbor.cu:181-182 slot meanings)."""
    C = lib.init_camera(seed)
    I = np.zeros(32, np.float32)
    I[5] = case["view"]
    lib.update_camera(C, I, 1.0 / 60.0, w, h)
    C[8] = case["time"]
    C[11] = case["wind"]
    C[9] = case["sample"]
    C[10] = case["reset"]
    if case["sun"] is not None:
        C[13] = case["sun"]
    C[22] = case["light"]
    C[15] = case["debug"]
    return C


def seeded_history(w, h, seed):
    """A deterministic non-zero History buffer: rgb in [0, 2), w = a sample count."""
    n = np.arange(w * h * 4, dtype=np.uint64)
    v = ((n * np.uint64(2654435761) + np.uint64(seed) * np.uint64(40503)) % np.uint64(4093)).astype(np.float32)
    out = v / np.float32(2048.0)
    out[3::4] = np.float32(seed % 5)
    return out


def replay_c3(lib, scene, w, h):
    """Runs every C3_CASES entry through trace, shade, filter and resolve of
    `lib` (oracle or reference harness), carrying History along each sequence
    as the fixture generator does. Returns per case C, Hit, Linear, Guide,
    Filtered, the History before and after resolve, and Pixels."""
    hist, out = {}, {}
    for case in C3_CASES:
        C = c3_camera(lib, case, w, h)
        H = lib.trace(scene, C, w, h)
        L, G = lib.shade(scene, C, H, w, h)
        F = lib.filter(L, G, H, C, w, h)
        seq = case["sequence"]
        if seq not in hist:
            hist[seq] = seeded_history(w, h, case["history_seed"])
        Hist = hist[seq]
        before = Hist.copy()
        px = lib.resolve(F, Hist, C, w, h)
        out[case["name"]] = dict(C=C, Hit=H, Linear=L, Guide=G, Filtered=F, before=before, History=Hist.copy(),
                                 Pixels=px)
    return out
