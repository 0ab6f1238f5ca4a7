// Builds the REFERENCE native CPU harness (tests/native_harness.cpp, which
// #includes Arbor.cu) from where it lies under the reference checkout, as a
// shared library with C entry points. Test infrastructure only: it produces the
// golden fixtures that pin the oracle restatement, and serves as the timed CPU
// baseline ("reference" kind). No reference source is copied; the Makefile
// passes the harness path as ARBOR_REF_HARNESS.
//
// The harness's own main() (native_harness.cpp:55) is renamed so that the
// kernels, CUDA shims and dispatch emulation (native_harness.cpp:15-39) can be
// called directly.
#define main arbor_reference_harness_main
#include ARBOR_REF_HARNESS
#undef main

#include <omp.h>

namespace {
const int kGrow[5] = {9, 12, 72, 288, 1728};
template <class F>
void rows2(int w, int h, int threads, F f) {
  if (threads > 0) omp_set_num_threads(threads);
  dispatch2(w, h, f);
}
}  // namespace

extern "C" {
int ref_available(void) { return 1; }

// Whole build exactly as Renderer::Renderer (native_harness.cpp:46-48); with
// bitonic != 0 the 120 sortPass stages replace std::sort (native_harness.cpp:77-79).
int ref_generate(int seed, float* B, float* P, unsigned* Keys, unsigned* Order, float* Nodes, int bitonic) {
  for (int level = 0; level < 5; level++) dispatch1(kGrow[level], [&] { growTree(B, level, seed); });
  dispatch1(CAPACITY, [&] { makePrimitives(B, P, seed); });
  dispatch1(CAPACITY, [&] { makeKeys(P, Keys, Order); });
  if (bitonic) {
    for (int span = 2; span <= CAPACITY; span *= 2)
      for (int distance = span / 2; distance > 0; distance /= 2)
        dispatch1(CAPACITY, [&] { sortPass(Keys, Order, distance, span); });
  } else {
    std::vector<std::pair<unsigned, unsigned>> a;
    for (int i = 0; i < CAPACITY; i++) a.push_back({Keys[i], Order[i]});
    std::sort(a.begin(), a.end());
    for (int i = 0; i < CAPACITY; i++) { Keys[i] = a[i].first; Order[i] = a[i].second; }
  }
  dispatch1(CAPACITY, [&] { makeBounds(P, Order, Nodes); });
  for (int first = CAPACITY / 2; first >= 1; first /= 2) dispatch1(first, [&] { reduceBounds(Nodes, first); });
  return 0;
}
int ref_grow_tree(float* B, int level, int seed) {
  dispatch1(kGrow[level], [&] { growTree(B, level, seed); });
  return 0;
}
int ref_init_camera(float* C, int seed) {
  scalar();
  initCamera(C, nullptr, seed);
  return 0;
}
int ref_update_camera(float* C, const float* I, float delta, int width, int height) {
  scalar();
  updateCamera(C, I, nullptr, delta, width, height);
  return 0;
}
int ref_trace_pixels(const float* P, const float* Nodes, const unsigned* Order, const float* C, float* Hit, int w,
                     int h, int threads) {
  rows2(w, h, threads, [&] { tracePixels(P, Nodes, Order, C, Hit, w, h); });
  return 0;
}
int ref_shade_pixels(const float* P, const float* Nodes, const unsigned* Order, const float* C, const float* Hit,
                     float* Linear, float* Guide, int w, int h, int threads) {
  rows2(w, h, threads, [&] { shadePixels(P, Nodes, Order, C, Hit, Linear, Guide, w, h); });
  return 0;
}
int ref_filter_lighting(const float* Linear, const float* Guide, const float* Hit, float* Filtered, const float* C,
                        int w, int h, int threads) {
  rows2(w, h, threads, [&] { filterLighting(Linear, Guide, Hit, Filtered, C, w, h); });
  return 0;
}
int ref_resolve_pixels(const float* Filtered, float* History, unsigned* Pixels, const float* C, int w, int h,
                       int threads) {
  rows2(w, h, threads, [&] { resolvePixels(Filtered, History, Pixels, C, w, h); });
  return 0;
}
// Single-tree primary rays: reference cameraRay + traceTree in tree-local space
// with the ground plane as initial best (the call at native_harness.cpp:84, per pixel).
int ref_trace_tree_pixels(const float* P, const float* Nodes, const unsigned* Order, const float* C, float* Hit,
                          int w, int h, int threads) {
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 2)
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      int b = (y * w + x) * 4;
      float3 ro = cameraPos(C), rd = cameraRay(C, x, y, w, h);
      float ground = rd.y < -0.00001f ? -ro.y / rd.y : FAR;
      float2 r = traceTree(P, Nodes, Order, ro, rd, C[14] / (float)h, C[8], C[11], ground > 0.0f ? ground : FAR, -1,
                           false);
      Hit[b] = r.x; Hit[b + 1] = r.y; Hit[b + 2] = -1.0f; Hit[b + 3] = C[14] * r.x / (float)h;
    }
  return 0;
}
int ref_trace_tree_rays(const float* P, const float* Nodes, const unsigned* Order, const float* ro, const float* rd,
                        int n, float cone, float time, float wind, float best, int shadows, float* out) {
  for (int i = 0; i < n; i++) {
    float2 r = traceTree(P, Nodes, Order, {ro[3 * i], ro[3 * i + 1], ro[3 * i + 2]},
                         {rd[3 * i], rd[3 * i + 1], rd[3 * i + 2]}, cone, time, wind, best, -1, shadows != 0);
    out[2 * i] = r.x;
    out[2 * i + 1] = r.y;
  }
  return 0;
}
float ref_bark_cut(const float* P, int id, float px, float py, float pz, float footprint) {
  return barkCut(P, id, {px, py, pz}, footprint);
}
float ref_leaf_t(const float* P, int id, const float* ro, const float* rd, float footprint, float time, float wind) {
  return leafT(P, id, {ro[0], ro[1], ro[2]}, {rd[0], rd[1], rd[2]}, footprint, time, wind);
}
unsigned ref_hash_u(unsigned x) { return hashU(x); }
// The 100 random rays of native_harness.cpp:82-83 (origins and unit directions).
int ref_harness_rays(float* ro, float* rd) {
  for (int ray = 0; ray < 100; ray++) {
    float3 o = {rnd(ray, 198) * 24 - 12, 1 + rnd(ray, 31) * 14, rnd(ray, 78) * 24 - 12};
    float3 t = {rnd(ray, 211) * 10 - 5, 2 + rnd(ray, 433) * 8, rnd(ray, 123) * 10 - 5};
    float3 d = norm3(t - o);
    ro[3 * ray] = o.x; ro[3 * ray + 1] = o.y; ro[3 * ray + 2] = o.z;
    rd[3 * ray] = d.x; rd[3 * ray + 1] = d.y; rd[3 * ray + 2] = d.z;
  }
  return 0;
}
// Footprint sweep of native_harness.cpp:89-90 on object 4: returns the largest
// barkCut increment over 10,000 log-spaced footprints.
float ref_bark_sweep(const float* P) {
  int id = 4;
  float3 point = read3(P, id * PS) + make_float3(0.0f, 0.0f, P[id * PS + 7]);
  float prev = barkCut(P, id, point, .000001f), jump = 0;
  for (int i = 0; i < 10000; i++) {
    float f = 0.000001f * powf(100000.0f, float(i) / 9999);
    float cut = barkCut(P, id, point, f);
    jump = std::max(jump, std::abs(cut - prev));
    prev = cut;
  }
  return jump;
}
// The 50 leaf near/far probes of native_harness.cpp:92: out[2k], out[2k+1].
int ref_leaf_probes(const float* P, float* out) {
  for (int k = 0; k < 50; k++) {
    int id = WOOD_COUNT + k * 431, b = id * PS;
    float3 u = read3(P, b + 4), n = read3(P, b + 12);
    float3 p = read3(P, b) + u * (P[b + 16] + P[b + 7] * .5f) + n * .10f;
    float3 rd = n * -1.0f;
    out[2 * k] = leafT(P, id, p, rd, .00001f, 0, 0);
    out[2 * k + 1] = leafT(P, id, p, rd, .05f, 0, 0);
  }
  return 0;
}
float ref_rnd(int n, int seed) { return rnd(n, seed); }
}  // extern "C"
