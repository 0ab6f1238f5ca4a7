// Operation counter for the oracle (test / measurement infrastructure only).
//
// Instantiates the restatement with a float wrapper that counts every
// arithmetic operation it performs, so the algorithmic work per primary ray
// of a real workload is measured, not estimated. Counted as one operation
// each: + - * / (and their compound forms), sqrt, min, max, abs, floor, and
// every libm call (sin, cos, exp, pow, atan2); comparisons, negation and
// conversions are free. The traversal counted is the lean walk the GPU runs
// (same visiting order and results as the reference's stack loop, without
// its re-tests of directly taken nodes).
#include <cstdint>

#include "arbor_oracle.hpp"
#include "sah_build.hpp"

namespace orc {

struct OpCount {
  uint64_t add = 0, mul = 0, div = 0, sqrt = 0, minmax = 0, misc = 0, trans = 0;
};
static thread_local OpCount g_ops;  // per thread: counting runs may use OpenMP

struct CF {
  float v;
  CF() = default;
  constexpr CF(float x) : v(x) {}
  constexpr CF(double x) : v((float)x) {}
  constexpr CF(int x) : v((float)x) {}
  constexpr CF(unsigned x) : v((float)x) {}
  explicit operator int() const { return (int)v; }
  explicit operator unsigned() const { return (unsigned)v; }
  CF& operator+=(CF o) { g_ops.add++; v += o.v; return *this; }
  CF& operator-=(CF o) { g_ops.add++; v -= o.v; return *this; }
  CF& operator*=(CF o) { g_ops.mul++; v *= o.v; return *this; }
  CF& operator/=(CF o) { g_ops.div++; v /= o.v; return *this; }
};
inline CF operator+(CF a, CF b) { g_ops.add++; return CF(a.v + b.v); }
inline CF operator-(CF a, CF b) { g_ops.add++; return CF(a.v - b.v); }
inline CF operator*(CF a, CF b) { g_ops.mul++; return CF(a.v * b.v); }
inline CF operator/(CF a, CF b) { g_ops.div++; return CF(a.v / b.v); }
inline CF operator-(CF a) { return CF(-a.v); }
inline bool operator<(CF a, CF b) { return a.v < b.v; }
inline bool operator>(CF a, CF b) { return a.v > b.v; }
inline bool operator<=(CF a, CF b) { return a.v <= b.v; }
inline bool operator>=(CF a, CF b) { return a.v >= b.v; }
inline bool operator==(CF a, CF b) { return a.v == b.v; }
inline bool operator!=(CF a, CF b) { return a.v != b.v; }
inline CF m_sin(CF x) { g_ops.trans++; return CF(::sinf(x.v)); }
inline CF m_cos(CF x) { g_ops.trans++; return CF(::cosf(x.v)); }
inline CF m_exp(CF x) { g_ops.trans++; return CF(::expf(x.v)); }
inline CF m_pow(CF x, CF y) { g_ops.trans++; return CF(::powf(x.v, y.v)); }
inline CF m_atan2(CF y, CF x) { g_ops.trans++; return CF(::atan2f(y.v, x.v)); }
inline CF m_sqrt(CF x) { g_ops.sqrt++; return CF(::sqrtf(x.v)); }
inline CF m_floor(CF x) { g_ops.misc++; return CF(::floorf(x.v)); }
inline CF m_abs(CF x) { g_ops.misc++; return CF(::fabsf(x.v)); }
inline CF m_min(CF a, CF b) { g_ops.minmax++; return CF(::fminf(a.v, b.v)); }
inline CF m_max(CF a, CF b) { g_ops.minmax++; return CF(::fmaxf(a.v, b.v)); }

}  // namespace orc

using namespace orc;
using ACF = Arbor<CF>;

namespace {
std::vector<CF> wrap(const float* a, size_t n) {
  std::vector<CF> v(n);
  for (size_t i = 0; i < n; i++) v[i] = CF(a[i]);
  return v;
}
}  // namespace

extern "C" {
// Bounding-sphere rejections on/off for the counting instantiation.
void orc_set_cull(int on) { ACF::cullSpheres = on != 0; }
// Early exits of woodT/leafT past the caller's best on/off for the counting instantiation.
void orc_set_skip(int on) { ACF::skipBeyondBest = on != 0; }
// Early exits of the leaf Newton solve (fixed point, 2-cycle) on/off for the counting instantiation.
void orc_set_newton_exits(int on) { ACF::newtonExits = on != 0; }
// Skips of band-weighted terms whose band weight is 0 on/off for the counting instantiation.
void orc_set_band_skips(int on) { ACF::bandSkips = on != 0; }
// The breeze-off shortcut of the leaf basis on/off for the counting instantiation.
void orc_set_wind_shortcut(int on) { ACF::windShortcut = on != 0; }

// Counts the operations of tracePixels (forest, lean walk) over every
// stride-th pixel in x and y of a width x height view. out[0] rays, out[1]
// total operations, out[2..8] add, mul, div, sqrt, min/max, abs/floor, libm.
int orc_count_trace_ops(const float* P, const float* Nodes, const unsigned* Order, const float* C, int width,
                        int height, int stride, int single_tree, double* out) {
  if (stride < 1) return -1;
  auto p = wrap(P, (size_t)CAPACITY * PS), n = wrap(Nodes, (size_t)NODE_COUNT * 8), c = wrap(C, 64);
  std::vector<CF> hit((size_t)width * height * 4);
  ACF::leanWalk = true;
  g_ops = OpCount();
  double rays = 0;
  for (int y = stride / 2; y < height; y += stride)
    for (int x = stride / 2; x < width; x += stride) {
      CF* h = hit.data();
      if (single_tree) ACF::traceTreePixel(p.data(), n.data(), Order, c.data(), h, width, height, x, y);
      else ACF::tracePixel(p.data(), n.data(), Order, c.data(), h, width, height, x, y);
      rays += 1;
    }
  ACF::leanWalk = false;
  const OpCount& o = g_ops;
  out[0] = rays;
  out[1] = (double)(o.add + o.mul + o.div + o.sqrt + o.minmax + o.misc + o.trans);
  out[2] = (double)o.add; out[3] = (double)o.mul; out[4] = (double)o.div; out[5] = (double)o.sqrt;
  out[6] = (double)o.minmax; out[7] = (double)o.misc; out[8] = (double)o.trans;
  return 0;
}

// Operations of every pixel's primary ray (forest, lean walk) into out[w*h]:
// the per-ray work distribution, e.g. how uneven the rays of one tile are.
int orc_trace_ops_per_pixel(const float* P, const float* Nodes, const unsigned* Order, const float* C, int width,
                            int height, float* out) {
  auto p = wrap(P, (size_t)CAPACITY * PS), n = wrap(Nodes, (size_t)NODE_COUNT * 8), c = wrap(C, 64);
  std::vector<CF> hit((size_t)width * height * 4);
  ACF::leanWalk = true;
#pragma omp parallel for schedule(dynamic, 2)
  for (int y = 0; y < height; y++)
    for (int x = 0; x < width; x++) {
      g_ops = OpCount();
      ACF::tracePixel(p.data(), n.data(), Order, c.data(), hit.data(), width, height, x, y);
      const OpCount& o = g_ops;
      out[(size_t)y * width + x] = (float)(o.add + o.mul + o.div + o.sqrt + o.minmax + o.misc + o.trans);
    }
  ACF::leanWalk = false;
  return 0;
}

// Full intersection tests vs tests that stop once their answer is known to be
// >= the caller's best (the GPU's woodT/leafT), over whole frames (trace, and
// shade with sun visibility on): the number of pixels whose Hit or Linear
// differs in any bit (0 expected).
// Three runs: the full tests; with the exits; with the exits and the
// bounding-sphere rejections (whose entry-at-or-beyond-best exits need both).
int orc_skip_mismatches(const float* P, const float* Nodes, const unsigned* Order, const float* C, int width,
                        int height) {
  using A = Arbor<float>;
  size_t n = (size_t)width * height * 4;
  std::vector<float> h[3] = {std::vector<float>(n), std::vector<float>(n), std::vector<float>(n)};
  std::vector<float> l[3] = {std::vector<float>(n), std::vector<float>(n), std::vector<float>(n)}, g(n);
  for (int run = 0; run < 3; run++) {
    A::skipBeyondBest = run > 0;
    A::cullSpheres = run > 1;
#pragma omp parallel for schedule(dynamic, 2)
    for (int y = 0; y < height; y++)
      for (int x = 0; x < width; x++) A::tracePixel(P, Nodes, Order, C, h[run].data(), width, height, x, y);
#pragma omp parallel for schedule(dynamic, 2)
    for (int y = 0; y < height; y++)
      for (int x = 0; x < width; x++)
        A::shadePixel(P, Nodes, Order, C, h[run].data(), l[run].data(), g.data(), width, height, x, y);
  }
  A::skipBeyondBest = A::cullSpheres = false;
  int bad = 0;
  for (size_t i = 0; i < n; i += 4)
    for (int run = 1; run < 3; run++)
      if (std::memcmp(&h[0][i], &h[run][i], 16) != 0 || std::memcmp(&l[0][i], &l[run][i], 16) != 0) {
        bad++;
        break;
      }
  return bad;
}

// Full 7-iteration leaf Newton solves vs solves that stop at a bit-exact
// fixed point or 2-cycle (the GPU's leafNewton), over whole frames (trace,
// and shade with sun visibility on): pixels whose Hit or Linear differs in
// any bit (0 expected).
int orc_newton_exit_mismatches(const float* P, const float* Nodes, const unsigned* Order, const float* C, int width,
                               int height) {
  using A = Arbor<float>;
  size_t n = (size_t)width * height * 4;
  std::vector<float> h[2] = {std::vector<float>(n), std::vector<float>(n)};
  std::vector<float> l[2] = {std::vector<float>(n), std::vector<float>(n)}, g(n);
  for (int on = 0; on < 2; on++) {
    A::newtonExits = on != 0;
#pragma omp parallel for schedule(dynamic, 2)
    for (int y = 0; y < height; y++)
      for (int x = 0; x < width; x++) A::tracePixel(P, Nodes, Order, C, h[on].data(), width, height, x, y);
#pragma omp parallel for schedule(dynamic, 2)
    for (int y = 0; y < height; y++)
      for (int x = 0; x < width; x++)
        A::shadePixel(P, Nodes, Order, C, h[on].data(), l[on].data(), g.data(), width, height, x, y);
  }
  A::newtonExits = false;
  int bad = 0;
  for (size_t i = 0; i < n; i += 4)
    if (std::memcmp(&h[0][i], &h[1][i], 16) != 0 || std::memcmp(&l[0][i], &l[1][i], 16) != 0) bad++;
  return bad;
}

// Band-weighted leaf terms always evaluated vs skipped when their band weight
// is 0 (the GPU's leafNewton and leafWidth), over whole frames (trace, and
// shade with sun visibility on): pixels whose Hit or Linear differs in any
// bit (0 expected).
int orc_band_skip_mismatches(const float* P, const float* Nodes, const unsigned* Order, const float* C, int width,
                               int height) {
  using A = Arbor<float>;
  size_t n = (size_t)width * height * 4;
  std::vector<float> h[2] = {std::vector<float>(n), std::vector<float>(n)};
  std::vector<float> l[2] = {std::vector<float>(n), std::vector<float>(n)}, g(n);
  for (int on = 0; on < 2; on++) {
    A::bandSkips = on != 0;
#pragma omp parallel for schedule(dynamic, 2)
    for (int y = 0; y < height; y++)
      for (int x = 0; x < width; x++) A::tracePixel(P, Nodes, Order, C, h[on].data(), width, height, x, y);
#pragma omp parallel for schedule(dynamic, 2)
    for (int y = 0; y < height; y++)
      for (int x = 0; x < width; x++)
        A::shadePixel(P, Nodes, Order, C, h[on].data(), l[on].data(), g.data(), width, height, x, y);
  }
  A::bandSkips = false;
  int bad = 0;
  for (size_t i = 0; i < n; i += 4)
    if (std::memcmp(&h[0][i], &h[1][i], 16) != 0 || std::memcmp(&l[0][i], &l[1][i], 16) != 0) bad++;
  return bad;
}

// The leaf basis rotated by a zero wind angle vs taken as stored (the GPU's
// leafFrame with the breeze off), over whole frames (trace, and shade with
// sun visibility on): pixels whose Hit or Linear differs in any bit (0
// expected).
int orc_wind_shortcut_mismatches(const float* P, const float* Nodes, const unsigned* Order, const float* C, int width,
                               int height) {
  using A = Arbor<float>;
  size_t n = (size_t)width * height * 4;
  std::vector<float> h[2] = {std::vector<float>(n), std::vector<float>(n)};
  std::vector<float> l[2] = {std::vector<float>(n), std::vector<float>(n)}, g(n);
  for (int on = 0; on < 2; on++) {
    A::windShortcut = on != 0;
#pragma omp parallel for schedule(dynamic, 2)
    for (int y = 0; y < height; y++)
      for (int x = 0; x < width; x++) A::tracePixel(P, Nodes, Order, C, h[on].data(), width, height, x, y);
#pragma omp parallel for schedule(dynamic, 2)
    for (int y = 0; y < height; y++)
      for (int x = 0; x < width; x++)
        A::shadePixel(P, Nodes, Order, C, h[on].data(), l[on].data(), g.data(), width, height, x, y);
  }
  A::windShortcut = false;
  int bad = 0;
  for (size_t i = 0; i < n; i += 4)
    if (std::memcmp(&h[0][i], &h[1][i], 16) != 0 || std::memcmp(&l[0][i], &l[1][i], 16) != 0) bad++;
  return bad;
}

// The same check on single objects: n leaf tests with seeded leaves, rays
// aimed at the leaf surface from 0.02-0.5 m off it, footprints log-uniform in
// [1e-5, 1e-2] (across both bands' transitions) and wind on, plus
// leafHeight/leafWidth at seeded (u, v); and n wood cases: barkCut and
// woodField at seeded points around a seeded wood segment, and its refined
// woodT along a ray through them, footprints log-uniform in [1e-5, 1e-1]
// (across the plate, splinter and pore bands). Returns the number of cases
// whose results differ in any bit between skips off and on (0 expected).
int orc_band_skip_object_mismatches(const float* P, int n) {
  using A = Arbor<float>;
  int bad = 0;
  for (int i = 0; i < n; i++) {
    auto r = [&](int k) { return A::rnd(i * 7 + k, 4242); };
    const int id = WOOD_COUNT + (int)(r(0) * (float)LEAF_COUNT) % LEAF_COUNT, b = id * PS;
    const float time = r(1) * 10.0f, wind = 0.6f, fp = std::pow(10.0f, -5.0f + 3.0f * r(2));
    V3<float> u = A::leafU(P, id, time, wind), n3 = A::leafN(P, id, time, wind), v = A::ld3(P, b + 8);
    V3<float> target = A::ld3(P, b) + u * (P[b + 16] + P[b + 7] * r(3)) + v * (P[b + 11] * (r(4) - 0.5f));
    V3<float> off = n3 * (0.02f + 0.48f * r(5)) + u * (r(6) - 0.5f) * 0.1f + v * (r(7) - 0.5f) * 0.1f;
    V3<float> ro = target + off, rd = A::unit(target - ro);
    const float uu = P[b + 7] * (1.4f * r(8) - 0.2f), vv = P[b + 11] * (2.0f * r(9) - 1.0f);
    float res[2][3];
    for (int on = 0; on < 2; on++) {
      A::bandSkips = on != 0;
      res[on][0] = A::leafT(P, id, ro, rd, fp, time, wind);
      res[on][1] = A::leafHeight(uu, vv, P[b + 7], P[b + 11], P[b + 17], P[b + 18], fp);
      res[on][2] = A::leafWidth(uu, P[b + 7], P[b + 11], fp);
    }
    A::bandSkips = false;
    if (std::memcmp(res[0], res[1], sizeof(res[0])) != 0) bad++;
  }
  for (int i = 0; i < n; i++) {
    auto r = [&](int k) { return A::rnd(i * 11 + k, 9191); };
    const int id = (int)(r(0) * (float)WOOD_COUNT) % WOOD_COUNT, b = id * PS;
    const float fp = std::pow(10.0f, -5.0f + 4.0f * r(1));
    V3<float> a = A::ld3(P, b), e = A::ld3(P, b + 4);
    V3<float> axis = e - a;
    V3<float> side = A::unit(A::c3(r(2) - 0.5f, r(3) - 0.5f, r(4) - 0.5f));
    const float rad = std::fmax(P[b + 7], P[b + 11]);
    V3<float> p = a + axis * r(5) + side * (rad * (0.7f + 0.6f * r(6)));
    V3<float> ro = p + side * (0.05f + 0.3f * r(7)), rd = A::unit(p - ro);
    float res[2][3];
    for (int on = 0; on < 2; on++) {
      A::bandSkips = on != 0;
      res[on][0] = A::barkCut(P, id, p, fp);
      res[on][1] = A::woodField(P, id, p, fp);
      res[on][2] = A::woodT(P, id, ro, rd, fp, true);
    }
    A::bandSkips = false;
    if (std::memcmp(res[0], res[1], sizeof(res[0])) != 0) bad++;
  }
  return bad;
}

// Reference intersection tests vs the same tests behind the GPU's
// bounding-sphere rejections, over whole frames (trace, and shade with sun
// visibility on so shadow rays use the non-refined wood path): the number of
// pixels whose Hit or Linear differs in any bit (0 expected).
int orc_cull_mismatches(const float* P, const float* Nodes, const unsigned* Order, const float* C, int width,
                        int height) {
  using A = Arbor<float>;
  size_t n = (size_t)width * height * 4;
  std::vector<float> h[2] = {std::vector<float>(n), std::vector<float>(n)};
  std::vector<float> l[2] = {std::vector<float>(n), std::vector<float>(n)}, g(n);
  for (int cull = 0; cull < 2; cull++) {
    A::cullSpheres = cull != 0;
#pragma omp parallel for schedule(dynamic, 2)
    for (int y = 0; y < height; y++)
      for (int x = 0; x < width; x++) A::tracePixel(P, Nodes, Order, C, h[cull].data(), width, height, x, y);
#pragma omp parallel for schedule(dynamic, 2)
    for (int y = 0; y < height; y++)
      for (int x = 0; x < width; x++)
        A::shadePixel(P, Nodes, Order, C, h[cull].data(), l[cull].data(), g.data(), width, height, x, y);
  }
  A::cullSpheres = false;
  int bad = 0;
  for (size_t i = 0; i < n; i += 4)
    if (std::memcmp(&h[0][i], &h[1][i], 16) != 0 || std::memcmp(&l[0][i], &l[1][i], 16) != 0) bad++;
  return bad;
}

// Grazing-ray stress of the bounding-sphere rejections in woodT / leafT:
// raysPerObject rays per wood and leaf object, from a distance log-uniform in
// [1, 500] m along a seeded direction, aimed either at a seeded point near
// the object's surface (a capsule point of the wood segment; a point of the
// leaf's Newton slab or petiole, swayed by the wind) jittered by up to
// 1 mm + 2e-3 * distance, or at a point 0.9-1.1 times the padded radius of
// one of its rejection spheres from the centre. Each ray's objectT, refined
// (primary rays) and not (shadow rays), with the rejections off and on, at a
// footprint of 8.33e-4 * distance (1080p pixels). out[0] = rays, out[1] =
// rays with a hit without the rejections, out[2] = rays whose results differ
// in any bit (0 expected), out[3] = the nearest distance (m) of a ray whose
// hit a rejection removed (0 if none).
void orc_set_cull_pad_scale(float s) { Arbor<float>::cullPadScale = s; }
namespace {
struct StressRay { V3<float> ro, rd; float fp, dist; int id; };
// The grazing rays of the stress tests below.
std::vector<StressRay> stressRays(const float* P, int raysPerObject, int seed, float time, float wind) {
  using A = Arbor<float>;
  using Ray = StressRay;
  const int objects = WOOD_COUNT + LEAF_COUNT;
  const size_t n = (size_t)objects * raysPerObject;
  std::vector<Ray> rays(n);
#pragma omp parallel for schedule(static)
  for (int id = 0; id < objects; id++) {
    const int b = id * PS;
    for (int k = 0; k < raysPerObject; k++) {
      auto r = [&](int j) { return A::rnd((id * raysPerObject + k) * 16 + j, 7001 + seed); };
      const float dist = std::exp(std::log(500.0f) * r(0));
      V3<float> target, centre;
      float radius;
      if (id < WOOD_COUNT) {
        V3<float> a = A::ld3(P, b), e = A::ld3(P, b + 4), half = (e - a) * 0.5f;
        centre = a + half;
        radius = A::len(half) + std::fmax(P[b + 7], P[b + 11]);
        const float s = r(1);
        V3<float> axis = A::unit(e - a);
        V3<float> side = A::unit(A::cross(axis, A::unit(A::c3(r(2) - 0.5f, r(3) - 0.5f, r(4) - 0.5f))));
        target = a + (e - a) * s + side * (P[b + 7] + (P[b + 11] - P[b + 7]) * s);
      } else {
        V3<float> u = A::leafU(P, id, time, wind), nn = A::leafN(P, id, time, wind), v = A::ld3(P, b + 8);
        const float stem = P[b + 16], ln = P[b + 7], w = P[b + 11];
        if (r(1) < 0.25f) {  // petiole
          centre = A::ld3(P, b) + u * (stem * 0.5f);
          radius = stem * 0.5f + 0.0011f;
          target = A::ld3(P, b) + u * (stem * r(2));
        } else {  // Newton slab
          const float hx = ln * 0.5f + 0.001f;
          centre = A::ld3(P, b) + u * (stem + ln * 0.5f) + nn * 0.015f;
          radius = std::sqrt(hx * hx + w * w + 0.05f * 0.05f);
          target = A::ld3(P, b) + u * (stem + ln * r(2)) + v * (w * (2.0f * r(3) - 1.0f)) + nn * (0.1f * r(4) - 0.035f);
        }
      }
      if (r(5) < 0.5f) {
        const float jit = 0.001f + 0.002f * dist;
        target = target + A::c3(r(6) - 0.5f, r(7) - 0.5f, r(8) - 0.5f) * (2.0f * jit);
      } else {
        const float padded = radius + 0.001f + 0.004f * dist;
        target = centre + A::unit(A::c3(r(6) - 0.5f, r(7) - 0.5f, r(8) - 0.5f)) * (padded * (0.9f + 0.2f * r(9)));
      }
      Ray& ray = rays[(size_t)id * raysPerObject + k];
      ray.rd = A::unit(A::c3(r(10) - 0.5f, r(11) - 0.5f, r(12) - 0.5f));
      ray.ro = target - ray.rd * dist;
      ray.fp = 8.33e-4f * dist;
      ray.dist = dist;
      ray.id = id;
    }
  }
  return rays;
}
}  // namespace
int orc_cull_stress(const float* P, int raysPerObject, int seed, float time, float wind, long long* out) {
  using A = Arbor<float>;
  const std::vector<StressRay> rays = stressRays(P, raysPerObject, seed, time, wind);
  const size_t n = rays.size();
  std::vector<float> res[2] = {std::vector<float>(2 * n), std::vector<float>(2 * n)};
  for (int on = 0; on < 2; on++) {
    A::cullSpheres = on != 0;
#pragma omp parallel for schedule(dynamic, 256)
    for (long long i = 0; i < (long long)n; i++) {
      const StressRay& ray = rays[i];
      res[on][2 * i] = A::objectT(P, ray.id, ray.ro, ray.rd, ray.fp, time, wind, true);
      res[on][2 * i + 1] = A::objectT(P, ray.id, ray.ro, ray.rd, ray.fp, time, wind, false);
    }
  }
  A::cullSpheres = false;
  long long hits = 0, bad = 0;
  float worst = 0.0f;
  for (size_t i = 0; i < n; i++) {
    if (res[0][2 * i] < A::FAR() || res[0][2 * i + 1] < A::FAR()) hits++;
    if (std::memcmp(&res[0][2 * i], &res[1][2 * i], 2 * sizeof(float)) != 0) {
      bad++;
      worst = worst == 0.0f ? rays[i].dist : std::fmin(worst, rays[i].dist);
    }
  }
  out[0] = (long long)n;
  out[1] = hits;
  out[2] = bad;
  out[3] = (long long)worst;
  return 0;
}

// The early exits past the caller's best under stress: the grazing rays of
// orc_cull_stress, each object test (refined and not) run in full, then with
// the rejections and the exits against five bests: one ulp above the full
// result (the SAH walk's tie form), 1e-4 relative below it, and 0.6, 1.0
// and 1.4 times the ray's distance to its aim point. Contract: a full result
// below the best comes back bit for bit, any other result stays at or above
// the best. out[0] = evaluations, out[1] = those with the full result below
// the best, out[2] = violations (0 expected), out[3] = evaluations whose
// answer an exit or rejection changed (results at or above the best), out[4]
// = the nearest distance (m) of a violation (0 if none).
int orc_exit_stress(const float* P, int raysPerObject, int seed, float time, float wind, long long* out) {
  using A = Arbor<float>;
  const std::vector<StressRay> rays = stressRays(P, raysPerObject, seed, time, wind);
  const long long n = (long long)rays.size();
  constexpr int kBests = 5;
  std::vector<float> full(2 * n), bests(2 * n * kBests), got(2 * n * kBests);
#pragma omp parallel for schedule(dynamic, 256)
  for (long long i = 0; i < n; i++) {
    const StressRay& ray = rays[i];
    for (int refine = 0; refine < 2; refine++) {
      const float f = A::objectT(P, ray.id, ray.ro, ray.rd, ray.fp, time, wind, refine != 0);
      full[2 * i + refine] = f;
      float* b = &bests[(2 * i + refine) * kBests];
      b[0] = f < A::FAR() ? std::nextafter(f, 2.0f * A::FAR()) : A::FAR();
      b[1] = f < A::FAR() ? f * (1.0f - 1e-4f) : A::FAR();
      b[2] = 0.6f * ray.dist;
      b[3] = 1.0f * ray.dist;
      b[4] = 1.4f * ray.dist;
    }
  }
  A::cullSpheres = A::skipBeyondBest = true;
#pragma omp parallel for schedule(dynamic, 256)
  for (long long i = 0; i < n; i++) {
    const StressRay& ray = rays[i];
    for (int refine = 0; refine < 2; refine++)
      for (int k = 0; k < kBests; k++) {
        const size_t e = (2 * i + refine) * kBests + k;
        got[e] = A::objectT(P, ray.id, ray.ro, ray.rd, ray.fp, time, wind, refine != 0, bests[e]);
      }
  }
  A::cullSpheres = A::skipBeyondBest = false;
  long long kept = 0, bad = 0, changed = 0;
  float worst = 0.0f;
  for (long long i = 0; i < n; i++)
    for (int refine = 0; refine < 2; refine++)
      for (int k = 0; k < kBests; k++) {
        const size_t e = (2 * i + refine) * kBests + k;
        const float f = full[2 * i + refine], b = bests[e], g = got[e];
        bool wrong;
        if (f < b) {
          kept++;
          wrong = std::memcmp(&g, &f, sizeof(float)) != 0;
        } else {
          wrong = g < b;
          if (!wrong && std::memcmp(&g, &f, sizeof(float)) != 0) changed++;
        }
        if (wrong) {
          bad++;
          worst = worst == 0.0f ? rays[i].dist : std::fmin(worst, rays[i].dist);
        }
      }
  out[0] = 2 * n * kBests;
  out[1] = kept;
  out[2] = bad;
  out[3] = changed;
  out[4] = (long long)worst;
  return 0;
}

// Reference walk vs lean walk over whole frames: returns the number of
// pixels whose Hit differs in any bit (0 expected).
// The SAH hierarchy of the GPU's accel walk over this heap BVH (the
// host restatement of the device builder, sah_build.hpp) into out[words]: 16 + 16 * (CAPACITY -
// 1) + CAPACITY words. Returns 0, or -1 if out is too small or the depth
// limit (32 as the walk's stack; lower in the limit form, to test the
// builder's bound) failed.
int orc_build_sah_limit(const float* Nodes, const unsigned* Order, uint32_t* out, long long words, int depthLimit) {
  const long long need = 16 + 16LL * (CAPACITY - 1) + CAPACITY;
  if (words < need || depthLimit < 15) return -1;
  std::vector<uint32_t> w =
      arbor::sah::buildAccel(Nodes + (size_t)CAPACITY * 8, Order, CAPACITY, depthLimit, 100000.0f);
  if (w.empty() || (long long)w.size() != need) return -1;
  std::memcpy(out, w.data(), w.size() * 4);
  return 0;
}
int orc_build_sah(const float* Nodes, const unsigned* Order, uint32_t* out, long long words) {
  return orc_build_sah_limit(Nodes, Order, out, words, 32);
}
// Operation counts walk the SAH hierarchy `accel` (NULL: the heap).
void orc_set_sah(const uint32_t* accel) { ACF::sahAccel = accel; }
// Test hook: invert the SAH walk's tie order (a planted mistake).
void orc_set_sah_tie_reversed(int on) { Arbor<float>::sahTieReversed = on != 0; }

// The SAH walk against the heap walk over whole frames, with the GPU's
// intersection shortcuts: mode 0 compares tracePixels' Hit, mode 1
// shadePixels' Linear with the same Hit (its sun visibility walks the
// hierarchy too). out[0] pixels differing, out[1] of them where both record
// object hits within 1e-3 relative in t (mode 0 only; the walk-order-dependent
// hits of swaying leaves), out[2] index of the first differing pixel or -1.
int orc_sah_walk_mismatches(const float* P, const float* Nodes, const unsigned* Order, const float* C, int width,
                            int height, const uint32_t* accel, int mode, long long* out) {
  using A = Arbor<float>;
  A::cullSpheres = A::skipBeyondBest = A::newtonExits = A::bandSkips = A::windShortcut = true;
  A::leanWalk = true;
  const size_t n = (size_t)width * height * 4;
  std::vector<float> hit(n), a(n), b(n), guide(n);
#pragma omp parallel for schedule(dynamic, 2)
  for (int y = 0; y < height; y++)
    for (int x = 0; x < width; x++) A::tracePixel(P, Nodes, Order, C, hit.data(), width, height, x, y);
  for (int sah = 0; sah < 2; sah++) {
    A::sahAccel = sah ? accel : nullptr;
    float* H = sah ? b.data() : a.data();
#pragma omp parallel for schedule(dynamic, 2)
    for (int y = 0; y < height; y++)
      for (int x = 0; x < width; x++) {
        if (mode == 0) A::tracePixel(P, Nodes, Order, C, H, width, height, x, y);
        else A::shadePixel(P, Nodes, Order, C, hit.data(), H, guide.data(), width, height, x, y);
      }
  }
  A::sahAccel = nullptr;
  A::leanWalk = false;
  A::cullSpheres = A::skipBeyondBest = A::newtonExits = A::bandSkips = A::windShortcut = false;
  out[0] = out[1] = 0;
  out[2] = -1;
  for (size_t i = 0; i < n; i += 4) {
    if (std::memcmp(&a[i], &b[i], 16) == 0) continue;
    if (out[2] < 0) out[2] = (long long)(i / 4);
    out[0]++;
    if (mode == 0 && a[i + 1] >= 0 && b[i + 1] >= 0 && std::fabs(a[i] - b[i]) <= 1e-3f * a[i]) out[1]++;
  }
  return 0;
}

int orc_lean_walk_mismatches(const float* P, const float* Nodes, const unsigned* Order, const float* C, int width,
                             int height, int single_tree) {
  using A = Arbor<float>;
  std::vector<float> a((size_t)width * height * 4), b(a.size());
  for (int lean = 0; lean < 2; lean++) {
    A::leanWalk = lean != 0;
    float* H = lean ? b.data() : a.data();
#pragma omp parallel for schedule(dynamic, 2)
    for (int y = 0; y < height; y++)
      for (int x = 0; x < width; x++)
        if (single_tree) A::traceTreePixel(P, Nodes, Order, C, H, width, height, x, y);
        else A::tracePixel(P, Nodes, Order, C, H, width, height, x, y);
  }
  A::leanWalk = false;
  int bad = 0;
  for (size_t i = 0; i < a.size(); i += 4)
    if (std::memcmp(&a[i], &b[i], 16) != 0) bad++;
  return bad;
}
}  // extern "C"
