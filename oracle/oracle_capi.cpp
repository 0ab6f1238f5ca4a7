// C entry points of the ORACLE (test infrastructure only; see arbor_oracle.hpp).
//
// Launch emulation follows tests/native_harness.cpp:35-39: 1-D kernels run
// serially in thread order, 2-D kernels run rows under OpenMP
// schedule(dynamic,2). Whole-build order follows native_harness.cpp:46-48
// (growTree 0..4, makePrimitives, makeKeys, std::sort of (key,index) pairs,
// makeBounds, reduceBounds 16384..1, initCamera).
#include "arbor_oracle.hpp"

#include <omp.h>

using namespace orc;
using A32 = Arbor<float>;
using A64 = Arbor<double>;

namespace {
const int kGrowCounts[5] = {9, 12, 72, 288, 1728};

template <class F>
void rows(int w, int h, int threads, F f) {
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 2)
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) f(x, y);
}

// tracePixels for a list of pixels (x, y) of a w x h view: out[4i..] is pixel
// i's Hit record (the full view's Hit at that pixel), for parity checks at
// full view sizes on a sample of the pixels.
template <class R>
int trace_pixel_list(const R* P, const R* Nodes, const unsigned* Order, const R* C, const int* xy, int n, int w,
                     int h, R* out, int threads) {
  using AR = Arbor<R>;
  std::vector<R> hit((size_t)w * h * 4);
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < n; i++) {
    const int x = xy[2 * i], y = xy[2 * i + 1];
    if (x < 0 || y < 0 || x >= w || y >= h) continue;
    AR::tracePixel(P, Nodes, Order, C, hit.data(), w, h, x, y);
    for (int k = 0; k < 4; k++) out[4 * i + k] = hit[((size_t)y * w + x) * 4 + k];
  }
  return 0;
}
// Primary rays through points offset from the pixels' sample positions: ray i
// is pixel (xy[2i], xy[2i+1])'s ray moved by (dxy[2i], dxy[2i+1]) pixels;
// single != 0 traces the single specimen in tree-local space (traceTreePixels).
template <class R>
int trace_pixel_offsets(const R* P, const R* Nodes, const unsigned* Order, const R* C, const int* xy,
                        const float* dxy, int n, int w, int h, R* out, int threads, int single) {
  using AR = Arbor<R>;
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < n; i++) {
    const int x = xy[2 * i], y = xy[2 * i + 1];
    if (x < 0 || y < 0 || x >= w || y >= h) continue;
    AR::tracePixelOffset(P, Nodes, Order, C, w, h, x, y, R(dxy[2 * i]), R(dxy[2 * i + 1]), out + 4 * i, single != 0);
  }
  return 0;
}
template <class R>
void generate_impl(int seed, R* B, R* P, unsigned* Keys, unsigned* Order, R* Nodes, int bitonic) {
  using Ar = Arbor<R>;
  for (int level = 0; level < 5; level++)
    for (int i = 0; i < kGrowCounts[level]; i++) Ar::growTree(B, level, seed, i);
  for (int i = 0; i < CAPACITY; i++) Ar::makePrimitive(B, P, seed, i);
  for (int i = 0; i < CAPACITY; i++) Ar::makeKey(P, Keys, Order, i);
  if (bitonic) {
    for (int span = 2; span <= CAPACITY; span *= 2)
      for (int distance = span / 2; distance > 0; distance /= 2)
        for (int i = 0; i < CAPACITY; i++) Ar::sortPass(Keys, Order, distance, span, i);
  } else {
    std::vector<std::pair<unsigned, unsigned>> a(CAPACITY);
    for (int i = 0; i < CAPACITY; i++) a[i] = {Keys[i], Order[i]};
    std::sort(a.begin(), a.end());
    for (int i = 0; i < CAPACITY; i++) { Keys[i] = a[i].first; Order[i] = a[i].second; }
  }
  for (int i = 0; i < CAPACITY; i++) Ar::makeBound(P, Order, Nodes, i);
  for (int first = CAPACITY / 2; first >= 1; first /= 2)
    for (int i = 0; i < first; i++) Ar::reduceBound(Nodes, first, i);
}
}  // namespace

extern "C" {

int orc_version(void) { return 1; }

// ---- build stages (float) ----
int orc_grow_tree(float* B, int level, int seed) {
  if (level < 0 || level > 4) return -1;
  for (int i = 0; i < kGrowCounts[level]; i++) A32::growTree(B, level, seed, i);
  return 0;
}
int orc_make_primitives(const float* B, float* P, int seed) {
  for (int i = 0; i < CAPACITY; i++) A32::makePrimitive(B, P, seed, i);
  return 0;
}
int orc_make_keys(const float* P, unsigned* Keys, unsigned* Order) {
  for (int i = 0; i < CAPACITY; i++) A32::makeKey(P, Keys, Order, i);
  return 0;
}
int orc_sort_pass(unsigned* Keys, unsigned* Order, int distance, int span) {
  for (int i = 0; i < CAPACITY; i++) A32::sortPass(Keys, Order, distance, span, i);
  return 0;
}
// std::sort of (key, index) pairs, as native_harness.cpp:47 does.
int orc_sort_pairs(unsigned* Keys, unsigned* Order, int n) {
  std::vector<std::pair<unsigned, unsigned>> a(n);
  for (int i = 0; i < n; i++) a[i] = {Keys[i], Order[i]};
  std::sort(a.begin(), a.end());
  for (int i = 0; i < n; i++) { Keys[i] = a[i].first; Order[i] = a[i].second; }
  return 0;
}
int orc_make_bounds(const float* P, const unsigned* Order, float* Nodes) {
  for (int i = 0; i < CAPACITY; i++) A32::makeBound(P, Order, Nodes, i);
  return 0;
}
int orc_reduce_bounds(float* Nodes, int first) {
  for (int i = 0; i < first; i++) A32::reduceBound(Nodes, first, i);
  return 0;
}
int orc_generate(int seed, float* B, float* P, unsigned* Keys, unsigned* Order, float* Nodes, int bitonic) {
  generate_impl<float>(seed, B, P, Keys, Order, Nodes, bitonic);
  return 0;
}
// The same build with correctly rounded sin/cos/pow (see g_crMath).
int orc_generate_cr(int seed, float* B, float* P, unsigned* Keys, unsigned* Order, float* Nodes) {
  g_crMath = true;
  generate_impl<float>(seed, B, P, Keys, Order, Nodes, 0);
  g_crMath = false;
  return 0;
}
int orc_init_camera_cr(float* C, int seed) {
  g_crMath = true;
  A32::initCamera(C, seed);
  g_crMath = false;
  return 0;
}
unsigned orc_spread(unsigned x) { return A32::spread(x); }

// ---- camera ----
int orc_init_camera(float* C, int seed) { A32::initCamera(C, seed); return 0; }
int orc_update_camera(float* C, const float* I, float delta, int width, int height) {
  A32::updateCamera(C, I, delta, width, height);
  return 0;
}

// ---- per-pixel stages, float and double ----
#define ORC_PIXEL_STAGES(SUF, R, AR)                                                                          \
  int orc_trace_pixels_##SUF(const R* P, const R* Nodes, const unsigned* Order, const R* C, R* Hit, int w,    \
                             int h, int threads) {                                                             \
    rows(w, h, threads, [&](int x, int y) { AR::tracePixel(P, Nodes, Order, C, Hit, w, h, x, y); });          \
    return 0;                                                                                                  \
  }                                                                                                            \
  int orc_trace_tree_pixels_##SUF(const R* P, const R* Nodes, const unsigned* Order, const R* C, R* Hit,       \
                                  int w, int h, int threads) {                                                 \
    rows(w, h, threads, [&](int x, int y) { AR::traceTreePixel(P, Nodes, Order, C, Hit, w, h, x, y); });      \
    return 0;                                                                                                  \
  }                                                                                                            \
  int orc_shade_pixels_##SUF(const R* P, const R* Nodes, const unsigned* Order, const R* C, const R* Hit,      \
                             R* Linear, R* Guide, int w, int h, int threads) {                                 \
    rows(w, h, threads,                                                                                        \
         [&](int x, int y) { AR::shadePixel(P, Nodes, Order, C, Hit, Linear, Guide, w, h, x, y); });          \
    return 0;                                                                                                  \
  }                                                                                                            \
  int orc_filter_lighting_##SUF(const R* Linear, const R* Guide, const R* Hit, R* Filtered, const R* C, int w, \
                                int h, int threads) {                                                          \
    rows(w, h, threads, [&](int x, int y) { AR::filterPixel(Linear, Guide, Hit, Filtered, C, w, h, x, y); }); \
    return 0;                                                                                                  \
  }                                                                                                            \
  int orc_resolve_pixels_##SUF(const R* Filtered, R* History, unsigned* Pixels, const R* C, int w, int h,      \
                               int threads) {                                                                  \
    rows(w, h, threads, [&](int x, int y) { AR::resolvePixel(Filtered, History, Pixels, C, w, h, x, y); });   \
    return 0;                                                                                                  \
  }

ORC_PIXEL_STAGES(f32, float, A32)
ORC_PIXEL_STAGES(f64, double, A64)

// tracePixels for a list of pixels of a w x h view (trace_pixel_list).
int orc_trace_pixel_list_f32(const float* P, const float* Nodes, const unsigned* Order, const float* C, const int* xy,
                             int n, int w, int h, float* out, int threads) {
  return trace_pixel_list(P, Nodes, Order, C, xy, n, w, h, out, threads);
}
int orc_trace_pixel_offsets_f32(const float* P, const float* Nodes, const unsigned* Order, const float* C,
                                const int* xy, const float* dxy, int n, int w, int h, float* out, int threads,
                                int single) {
  return trace_pixel_offsets(P, Nodes, Order, C, xy, dxy, n, w, h, out, threads, single);
}
int orc_trace_pixel_offsets_f64(const double* P, const double* Nodes, const unsigned* Order, const double* C,
                                const int* xy, const float* dxy, int n, int w, int h, double* out, int threads,
                                int single) {
  return trace_pixel_offsets(P, Nodes, Order, C, xy, dxy, n, w, h, out, threads, single);
}
int orc_trace_pixel_list_f64(const double* P, const double* Nodes, const unsigned* Order, const double* C,
                             const int* xy, int n, int w, int h, double* out, int threads) {
  return trace_pixel_list(P, Nodes, Order, C, xy, n, w, h, out, threads);
}

// ---- ray-level probes (float) used by the harness-mirroring checks ----
// traceTree for n rays in tree-local space; out[2i] = t, out[2i+1] = id.
int orc_trace_tree_rays(const float* P, const float* Nodes, const unsigned* Order, const float* ro,
                        const float* rd, int n, float cone, float time, float wind, float best, int shadows,
                        float* out) {
  for (int i = 0; i < n; i++) {
    auto h = A32::traceTree(P, Nodes, Order, {ro[3 * i], ro[3 * i + 1], ro[3 * i + 2]},
                            {rd[3 * i], rd[3 * i + 1], rd[3 * i + 2]}, cone, time, wind, best, -1, shadows != 0);
    out[2 * i] = h.x;
    out[2 * i + 1] = h.y;
  }
  return 0;
}
// Exhaustive nearest hit over every object (native_harness.cpp:85).
int orc_exhaustive_rays(const float* P, const float* ro, const float* rd, int n, float cone, float* out) {
  for (int r = 0; r < n; r++) {
    V3<float> o{ro[3 * r], ro[3 * r + 1], ro[3 * r + 2]}, d{rd[3 * r], rd[3 * r + 1], rd[3 * r + 2]};
    float nearest = A32::FAR();
    int winner = -1;
    for (int id = 0; id < OBJECT_COUNT; id++) {
      auto a = A32::boxSpan(o, d, A32::boundLow(P, id), A32::boundHigh(P, id));
      if (a.y < std::max(.0001f, a.x) || a.x > nearest) continue;
      float fp = std::max(.000005f, cone * std::max(.05f, a.x));
      float t = A32::objectT(P, id, o, d, fp, 0, 0, true);
      if (t < nearest) { nearest = t; winner = id; }
    }
    out[2 * r] = nearest;
    out[2 * r + 1] = (float)winner;
  }
  return 0;
}
float orc_bark_cut(const float* P, int id, float px, float py, float pz, float footprint) {
  return A32::barkCut(P, id, {px, py, pz}, footprint);
}
float orc_band_weight(float period, float footprint) { return A32::bandWeight(period, footprint); }
float orc_leaf_t(const float* P, int id, const float* ro, const float* rd, float footprint, float time, float wind) {
  return A32::leafT(P, id, {ro[0], ro[1], ro[2]}, {rd[0], rd[1], rd[2]}, footprint, time, wind);
}
float orc_object_t(const float* P, int id, const float* ro, const float* rd, float footprint, float time, float wind,
                   int refine) {
  return A32::objectT(P, id, {ro[0], ro[1], ro[2]}, {rd[0], rd[1], rd[2]}, footprint, time, wind, refine != 0);
}
float orc_sphere_t(const float* ro, const float* rd, const float* c, float radius) {
  return A32::sphereT({ro[0], ro[1], ro[2]}, {rd[0], rd[1], rd[2]}, {c[0], c[1], c[2]}, radius);
}
float orc_tube_t(const float* ro, const float* rd, const float* a, const float* b, float ra, float rb) {
  return A32::tubeT({ro[0], ro[1], ro[2]}, {rd[0], rd[1], rd[2]}, {a[0], a[1], a[2]}, {b[0], b[1], b[2]}, ra, rb);
}
float orc_aces(float x) { return A32::aces(x); }
// Largest barkCut increment over the 10,000-footprint sweep on object 4
// (the check at native_harness.cpp:89-91).
float orc_bark_sweep(const float* P) {
  const int id = 4;
  V3<float> point = A32::ld3(P, id * PS) + V3<float>{0.0f, 0.0f, P[id * PS + 7]};
  float prev = A32::barkCut(P, id, point, .000001f), jump = 0;
  for (int i = 0; i < 10000; i++) {
    float f = 0.000001f * powf(100000.0f, float(i) / 9999);
    float cut = A32::barkCut(P, id, point, f);
    jump = std::max(jump, std::abs(cut - prev));
    prev = cut;
  }
  return jump;
}
// The 50 leaf near/far probes (the check at native_harness.cpp:92-93).
int orc_leaf_probes(const float* P, float* out) {
  for (int k = 0; k < 50; k++) {
    int id = WOOD_COUNT + k * 431, b = id * PS;
    V3<float> u = A32::ld3(P, b + 4), n = A32::ld3(P, b + 12);
    V3<float> p = A32::ld3(P, b) + u * (P[b + 16] + P[b + 7] * .5f) + n * .10f;
    V3<float> rd = n * -1.0f;
    out[2 * k] = A32::leafT(P, id, p, rd, .00001f, 0, 0);
    out[2 * k + 1] = A32::leafT(P, id, p, rd, .05f, 0, 0);
  }
  return 0;
}
unsigned orc_hash_u(unsigned x) { return A32::hashU(x); }
// Per-tree instancing values {pos.xyz, scale, yaw, phase, seed} for 324 trees.
int orc_forest_table(int worldSeed, float* out) {
  for (int t = 0; t < FOREST_COUNT; t++) {
    auto p = A32::forestPos(t, worldSeed);
    out[7 * t + 0] = p.x; out[7 * t + 1] = p.y; out[7 * t + 2] = p.z;
    out[7 * t + 3] = A32::forestScale(t, worldSeed);
    out[7 * t + 4] = A32::forestYaw(t, worldSeed);
    out[7 * t + 5] = (float)(A32::forestSeed(t, worldSeed) & 8191) * 0.0017f;
    out[7 * t + 6] = (float)A32::forestSeed(t, worldSeed);
  }
  return 0;
}
}  // extern "C"
