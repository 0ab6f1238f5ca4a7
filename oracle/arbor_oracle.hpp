// ARBOR render-path ORACLE: a plain CPU restatement of the reference kernels.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product (the HIP library, its host
// engine, bench.py's GPU leg) may include, link or call this file; only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the checker.
//
// It restates Arbor.cu (the reference's file of record) function by function,
// templated on the arithmetic type R:
//   R = float  : bit-exact restatement of the reference CPU harness when built
//                with g++ -O2 -ffp-contract=off against glibc libm (pinned by the
//                golden fixtures produced from the compiled reference harness);
//   R = double : the FP64 run the GPU trace/shade tolerances are measured
//                against (same algorithm, double arithmetic and double libm).
// Every literal goes through K(), so R=float uses exactly the reference's float
// literal and R=double its decimal value. The order of every floating-point
// operation follows the reference expression order; that is what makes the
// float instantiation bit-exact, so it is kept alike on purpose.
//
// Citations are Arbor.cu line numbers (kernels/*.cu bodies concatenated).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>

namespace orc {

// ---- ABI constants (Arbor.cu:5-15, 54-56) --------------------------------
constexpr int CAPACITY = 32768;
constexpr int NODE_COUNT = 65536;
constexpr int PS = 24;
constexpr int CS = 20;
constexpr int CURVES = 2109;
constexpr int WOOD_COUNT = 5124;
constexpr int LEAF_COUNT = 24192;
constexpr int ROCK_COUNT = 96;
constexpr int OBJECT_COUNT = 29412;
constexpr int FOREST_SIDE = 18;
constexpr int FOREST_COUNT = 324;

#define K(x) (sizeof(R) == 4 ? R(x##f) : R(x))

// ---- libm dispatch: glibc float entry points for float, double for double --
// With g_crMath set, float sin/cos/pow are the correctly rounded values
// (evaluated in double, rounded once) instead of glibc's, which differ from
// correct rounding on about 1.4% of sinf/cosf inputs. The algorithm is
// unchanged; only the libm is swapped. This is the libm the GPU build kernels
// use, so the GPU's B and P can be checked bit-for-bit against it.
inline bool g_crMath = false;
inline float m_sin(float x) { return g_crMath ? (float)std::sin((double)x) : ::sinf(x); }
inline double m_sin(double x) { return std::sin(x); }
inline float m_cos(float x) { return g_crMath ? (float)std::cos((double)x) : ::cosf(x); }
inline double m_cos(double x) { return std::cos(x); }
inline float m_exp(float x) { return ::expf(x); }
inline double m_exp(double x) { return std::exp(x); }
inline float m_pow(float x, float y) { return g_crMath ? (float)std::pow((double)x, (double)y) : ::powf(x, y); }
inline double m_pow(double x, double y) { return std::pow(x, y); }
inline float m_atan2(float y, float x) { return ::atan2f(y, x); }
inline double m_atan2(double y, double x) { return std::atan2(y, x); }
inline float m_sqrt(float x) { return ::sqrtf(x); }
inline double m_sqrt(double x) { return std::sqrt(x); }
inline float m_floor(float x) { return ::floorf(x); }
inline double m_floor(double x) { return std::floor(x); }
inline float m_abs(float x) { return ::fabsf(x); }
inline double m_abs(double x) { return std::fabs(x); }
inline float m_min(float a, float b) { return ::fminf(a, b); }
inline double m_min(double a, double b) { return std::fmin(a, b); }
inline float m_max(float a, float b) { return ::fmaxf(a, b); }
inline double m_max(double a, double b) { return std::fmax(a, b); }

template <class R> struct V3 { R x, y, z; };
template <class R> struct V2 { R x, y; };
template <class R> inline V3<R> v3(R x, R y, R z) { return {x, y, z}; }
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }

// Signed-int arithmetic of the reference wraps in practice; do it in unsigned.
inline int wmul(int a, int b) { return (int)((unsigned)a * (unsigned)b); }
inline int wadd(int a, int b) { return (int)((unsigned)a + (unsigned)b); }

template <class R> struct Arbor {
  using F3 = V3<R>;
  using F2 = V2<R>;
  static constexpr R FAR() { return R(100000); }
  static R PI() { return K(3.14159265359); }

  // -- scalar helpers (Arbor.cu:17-31) --
  static R clampf(R x, R a, R b) { return m_min(b, m_max(a, x)); }
  static R sat(R x) { return clampf(x, K(0.0), K(1.0)); }
  static R lerp(R a, R b, R t) { return a + (b - a) * t; }
  static R frac(R a) { return a - m_floor(a); }
  static R smooth(R a, R b, R x) {
    R t = sat((x - a) / (b - a));
    return t * t * (K(3.0) - K(2.0) * t);
  }
  static R dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
  static F3 cross(F3 a, F3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
  }
  static R len(F3 a) { return m_sqrt(dot(a, a)); }
  static F3 unit(F3 a) { return a / m_max(K(0.0000001), len(a)); }
  static F3 vmin(F3 a, F3 b) { return {m_min(a.x, b.x), m_min(a.y, b.y), m_min(a.z, b.z)}; }
  static F3 vmax(F3 a, F3 b) { return {m_max(a.x, b.x), m_max(a.y, b.y), m_max(a.z, b.z)}; }
  static F3 mix(F3 a, F3 b, R t) { return a + (b - a) * t; }
  static F3 ld3(const R* a, int i) { return {a[i], a[i + 1], a[i + 2]}; }
  static void st3(R* a, int i, F3 v) { a[i] = v.x; a[i + 1] = v.y; a[i + 2] = v.z; }
  static F3 c3(R x, R y, R z) { return {x, y, z}; }

  // -- integer hash and value noise (Arbor.cu:32-36) --
  static unsigned hashU(unsigned x) {
    x ^= x >> 16; x *= 2146121005u; x ^= x >> 15; x *= 2221713035u; x ^= x >> 16;
    return x;
  }
  static R rnd(int n, int seed) {
    return R(hashU((unsigned)n ^ hashU((unsigned)seed)) & 16777215u) / K(16777216.0);
  }
  static R hash2(int x, int y, int seed) { return rnd(wadd(wmul(x, 1973), wmul(y, 9277)), seed); }
  static R noise(R x, R y, int seed) {
    int ix = (int)m_floor(x), iy = (int)m_floor(y);
    R fx = frac(x), fy = frac(y);
    fx = fx * fx * (K(3.0) - K(2.0) * fx);
    fy = fy * fy * (K(3.0) - K(2.0) * fy);
    return lerp(lerp(hash2(ix, iy, seed), hash2(ix + 1, iy, seed), fx),
                lerp(hash2(ix, iy + 1, seed), hash2(ix + 1, iy + 1, seed), fx), fy);
  }
  static R noise3(F3 p, int seed) {
    int iz = (int)m_floor(p.z);
    R f = frac(p.z);
    f = f * f * (K(3.0) - K(2.0) * f);
    return lerp(noise(p.x, p.y, wadd(seed, wmul(iz, 131))), noise(p.x, p.y, wadd(seed, wmul(iz + 1, 131))), f);
  }

  // -- geometry helpers (Arbor.cu:37-44) --
  static R invSafe(R x) {
    return m_abs(x) < K(0.00000001) ? (x < K(0.0) ? K(-100000000.0) : K(100000000.0)) : K(1.0) / x;
  }
  static F2 boxSpan(F3 ro, F3 rd, F3 lo, F3 hi) {
    F3 inv = c3(invSafe(rd.x), invSafe(rd.y), invSafe(rd.z));
    F3 a = (lo - ro) * inv;
    F3 b = (hi - ro) * inv;
    F3 mn = vmin(a, b), mx = vmax(a, b);
    return {m_max(mn.x, m_max(mn.y, mn.z)), m_min(mx.x, m_min(mx.y, mx.z))};
  }
  static F3 basisU(F3 n) {
    return unit(cross(m_abs(n.y) < K(0.91) ? c3(K(0.0), K(1.0), K(0.0)) : c3(K(0.0), K(0.0), K(1.0)), n));
  }
  static F3 curveAt(const R* B, int id, R t) {
    int b = id * CS;
    R q = K(1.0) - t;
    return ld3(B, b) * (q * q * q) + ld3(B, b + 3) * (K(3.0) * q * q * t) +
           ld3(B, b + 6) * (K(3.0) * q * t * t) + ld3(B, b + 9) * (t * t * t);
  }
  static F3 curveTangent(const R* B, int id, R t) {
    int b = id * CS;
    R q = K(1.0) - t;
    return unit((ld3(B, b + 3) - ld3(B, b)) * (q * q) + (ld3(B, b + 6) - ld3(B, b + 3)) * (K(2.0) * q * t) +
                (ld3(B, b + 9) - ld3(B, b + 6)) * (t * t));
  }
  static R curveRadius(const R* B, int id, R t) { return lerp(B[id * CS + 12], B[id * CS + 13], t); }
  static R bandWeight(R period, R footprint) {
    return K(1.0) - smooth(period * K(0.12), period * K(0.65), footprint);
  }

  // -- camera helpers (Arbor.cu:45-51) --
  static F3 cameraPos(const R* C) { return ld3(C, 0); }
  static F3 forwardC(const R* C) { return unit(ld3(C, 4) - ld3(C, 0)); }
  static F3 rightC(const R* C) { return unit(cross(forwardC(C), c3(K(0.0), K(1.0), K(0.0)))); }
  static F3 sunC(const R* C) {
    return unit(c3(m_cos(C[12]) * m_cos(C[13]), m_sin(C[13]), m_sin(C[12]) * m_cos(C[13])));
  }
  static unsigned packRGB(F3 c) {
    unsigned r = (unsigned)(sat(c.x) * K(255.0) + K(0.5));
    unsigned g = (unsigned)(sat(c.y) * K(255.0) + K(0.5));
    unsigned b = (unsigned)(sat(c.z) * K(255.0) + K(0.5));
    return r | (g << 8) | (b << 16) | 4278190080u;
  }
  static R halton(int n, int base) {
    R f = K(1.0), r = K(0.0);
    for (int j = 0; j < 12; j++) {
      if (n <= 0) break;
      f /= R(base);
      r += f * R(n % base);
      n /= base;
    }
    return r;
  }
  static F3 cameraRay(const R* C, int x, int y, int width, int height) {
    int sample = (int)C[9];
    R jx = halton(sample + 1, 2) - K(0.5);
    R jy = halton(sample + 1, 3) - K(0.5);
    R u = (R(x) + K(0.5) + jx - R(width) * K(0.5)) / R(height);
    R v = (R(height) * K(0.5) - R(y) - K(0.5) - jy) / R(height);
    F3 f = forwardC(C), r = rightC(C), up = cross(r, f);
    return unit(f + r * (u * C[14]) + up * (v * C[14]));
  }
  // cameraRay through a point (dx, dy) pixels away from the sample position.
  // Test infrastructure: defines "a ray within eps of the pixel's ray". With
  // dx = dy = 0 it is cameraRay bit for bit (adding +0 changes no value).
  static F3 cameraRayOffset(const R* C, int x, int y, int width, int height, R dx, R dy) {
    int sample = (int)C[9];
    R jx = halton(sample + 1, 2) - K(0.5) + dx;
    R jy = halton(sample + 1, 3) - K(0.5) + dy;
    R u = (R(x) + K(0.5) + jx - R(width) * K(0.5)) / R(height);
    R v = (R(height) * K(0.5) - R(y) - K(0.5) - jy) / R(height);
    F3 f = forwardC(C), r = rightC(C), up = cross(r, f);
    return unit(f + r * (u * C[14]) + up * (v * C[14]));
  }

  // -- forest instancing (Arbor.cu:57-69) --
  static int forestSeed(int tree, int ws) {
    return (int)(hashU((unsigned)wadd(wadd(wmul(tree, 92821), wmul(ws, 68917)), 17)) & 1048575u);
  }
  static F3 forestPos(int tree, int ws) {
    int x = tree % FOREST_SIDE - FOREST_SIDE / 2;
    int z = tree / FOREST_SIDE - FOREST_SIDE / 2;
    int s = forestSeed(tree, ws);
    R jx = (rnd(1, s) - K(0.5)) * K(10.0);
    R jz = (rnd(2, s) - K(0.5)) * K(10.0);
    R warpX = (noise(R(x) * K(0.19), R(z) * K(0.19), ws + 311) - K(0.5)) * K(3.0);
    R warpZ = (noise(R(x) * K(0.19) + K(17.0), R(z) * K(0.19) - K(9.0), ws + 733) - K(0.5)) * K(3.0);
    return c3(R(x) * K(28.0) + jx + warpX, K(0.0), R(z) * K(28.0) + jz + warpZ);
  }
  static R forestScale(int tree, int ws) { return K(0.72) + rnd(3, forestSeed(tree, ws)) * K(0.62); }
  static R forestYaw(int tree, int ws) { return rnd(4, forestSeed(tree, ws)) * PI() * K(2.0); }
  static F3 toLocalPoint(F3 p, int tree, int ws) {
    F3 q = p - forestPos(tree, ws);
    R a = -forestYaw(tree, ws);
    R c = m_cos(a), s = m_sin(a);
    R sc = forestScale(tree, ws);
    return c3(q.x * c - q.z * s, q.y, q.x * s + q.z * c) / sc;
  }
  static F3 toLocalDir(F3 d, int tree, int ws) {
    R a = -forestYaw(tree, ws);
    R c = m_cos(a), s = m_sin(a);
    return unit(c3(d.x * c - d.z * s, d.y, d.x * s + d.z * c));
  }
  static F3 toWorldNormal(F3 n, int tree, int ws) {
    R a = forestYaw(tree, ws);
    R c = m_cos(a), s = m_sin(a);
    return unit(c3(n.x * c - n.z * s, n.y, n.x * s + n.z * c));
  }

  // ====================== growTree (Arbor.cu:75-102) ======================
  static void storeCurve(R* B, int id, F3 a, F3 b, F3 c, F3 d, R ra, R rb, R arc, int seed) {
    int o = id * CS;
    st3(B, o, a); st3(B, o + 3, b); st3(B, o + 6, c); st3(B, o + 9, d);
    B[o + 12] = ra; B[o + 13] = rb; B[o + 14] = arc; B[o + 15] = R(seed); B[o + 16] = len(d - a);
  }
  // One thread i of growTree(B, level, seed).
  static void growTree(R* B, int level, int seed, int i) {
    const R pi = PI();
    if (level == 0) {
      if (i == 0)
        storeCurve(B, 0, c3(K(0.0), K(-0.15), K(0.0)), c3(K(-0.22), K(1.1), K(0.16)), c3(K(0.23), K(2.9), K(-0.19)),
                   c3(K(0.04), K(4.2), K(0.05)), K(0.68), K(0.23), K(0.0), seed);
      if (i < 8) {
        R a = R(i) * pi * K(0.25) + (rnd(i, seed) - K(0.5)) * K(0.5);
        F3 d = c3(m_cos(a), K(0.0), m_sin(a));
        R ln = K(2.0) + rnd(i + 10, seed) * K(1.4);
        F3 p = d * K(0.32) + c3(K(0.0), K(0.27), K(0.0));
        storeCurve(B, i + 1, p, d * K(0.8) + c3(K(0.0), K(0.10), K(0.0)), d * (ln * K(0.68)) + c3(K(0.0), K(-0.06), K(0.0)),
                   d * ln + c3(K(0.0), K(-0.16), K(0.0)), K(0.235), K(0.012), K(0.0), seed + i * 37);
      }
      return;
    }
    if (level == 1 && i < 12) {
      R t = K(0.31) + K(0.65) * R(i) / K(11.0);
      R a = R(i) * K(2.399963) + K(0.31);
      R r = rnd(i + 37, seed);
      R ln = K(3.2) + K(2.1) * r;
      F3 base = curveAt(B, 0, t);
      F3 d = c3(m_cos(a), K(0.0), m_sin(a));
      F3 side = c3(-d.z, K(0.0), d.x);
      F3 end = base + d * (ln * (K(0.87) - K(0.30) * t)) + c3(K(0.0), ln * (K(0.60) + K(0.40) * t), K(0.0)) +
               side * ((r - K(0.5)) * K(1.2));
      R ra = curveRadius(B, 0, t) * (K(0.46) + K(0.13) * r);
      storeCurve(B, 9 + i, base, base + d * (ln * K(0.29)) + c3(K(0.0), ln * K(0.12), K(0.0)),
                 end - d * (ln * K(0.17)) - c3(K(0.0), ln * K(0.26), K(0.0)), end, ra, ra * K(0.14),
                 B[14] + t * K(4.2), seed + i * 19);
      return;
    }
    int count = level == 2 ? 72 : (level == 3 ? 288 : 1728);
    if (i >= count) return;
    int children = level == 2 ? 6 : (level == 3 ? 4 : 6);
    int offset = level == 2 ? 21 : (level == 3 ? 93 : 381);
    int parentOffset = level == 2 ? 9 : (level == 3 ? 21 : 93);
    int parent = parentOffset + i / children, child = i % children;
    R h = rnd(i + level * 730, seed);
    R t = K(0.23) + K(0.72) * (R(child) + K(0.35)) / R(children);
    F3 p = curveAt(B, parent, t);
    F3 tangent = curveTangent(B, parent, t);
    F3 u = basisU(tangent);
    F3 v = cross(tangent, u);
    R angle = R(child) * K(2.399963) + h * K(0.8);
    F3 lateral = u * m_cos(angle) + v * m_sin(angle);
    F3 d = unit(tangent * K(0.52) + lateral * (level == 4 ? K(0.64) : K(0.80)) + c3(K(0.0), K(0.26), K(0.0)));
    R ln = level == 2 ? K(1.65) + K(1.10) * h : (level == 3 ? K(0.85) + K(0.65) * h : K(0.43) + K(0.37) * h);
    ln *= K(0.84) + K(0.25) * (K(1.0) - t);
    F3 end = p + d * ln + c3(K(0.0), ln * K(0.16), K(0.0));
    R radius = curveRadius(B, parent, t) * (level == 4 ? K(0.30) : K(0.40));
    radius = m_max(level == 4 ? K(0.0035) : K(0.008), radius);
    storeCurve(B, offset + i, p, p + tangent * (ln * K(0.25)) + lateral * (ln * K(0.11)),
               end - d * (ln * K(0.28)) + c3(K(0.0), ln * K(0.05), K(0.0)), end, radius,
               radius * (level == 4 ? K(0.30) : K(0.22)), B[parent * CS + 14] + t * B[parent * CS + 16],
               seed + (offset + i) * 31);
  }

  // ==================== makePrimitives (Arbor.cu:103-131) ====================
  static void makePrimitive(const R* B, R* P, int seed, int i) {
    if (i >= CAPACITY) return;
    int b = i * PS;
    for (int j = 0; j < PS; j++) P[b + j] = K(0.0);
    P[b + 3] = K(-1.0);
    if (i >= OBJECT_COUNT) return;
    if (i < WOOD_COUNT) {
      int c, segment, segments;
      if (i < 12) { c = 0; segment = i; segments = 12; }
      else if (i < 60) { c = 1 + (i - 12) / 6; segment = (i - 12) % 6; segments = 6; }
      else if (i < 156) { c = 9 + (i - 60) / 8; segment = (i - 60) % 8; segments = 8; }
      else if (i < 516) { c = 21 + (i - 156) / 5; segment = (i - 156) % 5; segments = 5; }
      else if (i < 1668) { c = 93 + (i - 516) / 4; segment = (i - 516) % 4; segments = 4; }
      else { c = 381 + (i - 1668) / 2; segment = (i - 1668) % 2; segments = 2; }
      R t0 = R(segment) / R(segments), t1 = R(segment + 1) / R(segments);
      st3(P, b, curveAt(B, c, t0));
      P[b + 3] = K(0.0);
      st3(P, b + 4, curveAt(B, c, t1));
      P[b + 7] = curveRadius(B, c, t0);
      P[b + 11] = curveRadius(B, c, t1);
      P[b + 15] = B[c * CS + 15];
      R arclength = B[c * CS + 14];
      for (int j = 0; j < 12; j++)
        if (j < segment)
          arclength += len(curveAt(B, c, R(j + 1) / R(segments)) - curveAt(B, c, R(j) / R(segments)));
      P[b + 16] = arclength;
      P[b + 17] = R(c);
      P[b + 18] = B[c * CS + 12];
      st3(P, b + 8, basisU(curveTangent(B, c, K(0.0))));
      return;
    }
    if (i < WOOD_COUNT + LEAF_COUNT) {
      int leaf = i - WOOD_COUNT, twig = leaf / 14, k = leaf % 14, c = 381 + twig;
      R h = rnd(leaf + 13, seed);
      R t = K(0.16) + K(0.81) * (R(k) + K(0.4)) / K(14.0);
      F3 anchor = curveAt(B, c, t);
      F3 axis = curveTangent(B, c, t);
      F3 bu = basisU(axis);
      F3 bv = cross(axis, bu);
      R angle = R(k) * K(2.399963) + rnd(twig, seed) * K(6.28);
      F3 out = unit(bu * m_cos(angle) + bv * m_sin(angle));
      F3 u = unit(out * K(0.82) + axis * K(0.35) + c3(K(0.0), K(0.20), K(0.0)));
      F3 normal = unit(c3(K(0.0), K(1.0), K(0.0)) + out * (K(0.20) + h * K(0.38)));
      F3 v = unit(cross(normal, u));
      normal = unit(cross(u, v));
      st3(P, b, anchor);
      P[b + 3] = K(1.0);
      st3(P, b + 4, u);
      P[b + 7] = K(0.145) + K(0.075) * h;
      st3(P, b + 8, v);
      P[b + 11] = P[b + 7] * (K(0.34) + K(0.07) * rnd(leaf + 37, seed));
      st3(P, b + 12, normal);
      P[b + 15] = R(leaf + seed * 11);
      P[b + 16] = K(0.023) + K(0.024) * rnd(leaf + 149, seed);
      P[b + 17] = K(0.013) + K(0.016) * h;
      P[b + 18] = (rnd(leaf + 777, seed) - K(0.5)) * K(0.10);
      P[b + 19] = R(twig);
      return;
    }
    int rock = i - WOOD_COUNT - LEAF_COUNT;
    R a = rnd(rock * 3 + 1, seed) * K(6.2831853);
    R r = K(0.8) + K(4.5) * m_sqrt(rnd(rock * 3 + 2, seed));
    R size = K(0.045) + K(0.19) * m_pow(rnd(rock + 371, seed), K(3.0));
    st3(P, b, c3(m_cos(a) * r, size * K(0.30) - K(0.015), m_sin(a) * r));
    P[b + 3] = K(2.0);
    st3(P, b + 4, c3(size, size * K(0.55), size * (K(0.7) + rnd(rock + 16, seed) * K(0.5))));
    P[b + 15] = R(rock + seed);
  }
  static F3 boundLow(const R* P, int id) {
    int b = id * PS, type = (int)P[b + 3];
    if (type < 0) return c3(FAR(), FAR(), FAR());
    F3 a = ld3(P, b);
    if (type == 0) {
      R r = m_max(P[b + 7], P[b + 11]) + K(0.002);
      return vmin(a, ld3(P, b + 4)) - c3(r, r, r);
    }
    if (type == 2) return a - ld3(P, b + 4);
    R radius = P[b + 7] + P[b + 16] + K(0.025);
    return a - c3(radius, radius, radius);
  }
  static F3 boundHigh(const R* P, int id) {
    int b = id * PS, type = (int)P[b + 3];
    if (type < 0) return c3(-FAR(), -FAR(), -FAR());
    F3 a = ld3(P, b);
    if (type == 0) {
      R r = m_max(P[b + 7], P[b + 11]) + K(0.002);
      return vmax(a, ld3(P, b + 4)) + c3(r, r, r);
    }
    if (type == 2) return a + ld3(P, b + 4);
    R radius = P[b + 7] + P[b + 16] + K(0.025);
    return a + c3(radius, radius, radius);
  }

  // ===================== BVH build (Arbor.cu:138-142) =====================
  static unsigned spread(unsigned x) {
    x &= 1023u;
    x = (x | (x << 16)) & 50331903u;
    x = (x | (x << 8)) & 50393103u;
    x = (x | (x << 4)) & 51130563u;
    x = (x | (x << 2)) & 153391689u;
    return x;
  }
  static void makeKey(const R* P, unsigned* Keys, unsigned* Order, int i) {
    Order[i] = (unsigned)i;
    if (P[i * PS + 3] < K(0.0)) { Keys[i] = 4294967295u; return; }
    F3 p = (boundLow(P, i) + boundHigh(P, i)) * K(0.5);
    unsigned x = (unsigned)(sat((p.x + K(12.0)) / K(24.0)) * K(1023.0));
    unsigned y = (unsigned)(sat((p.y + K(1.0)) / K(18.0)) * K(1023.0));
    unsigned z = (unsigned)(sat((p.z + K(12.0)) / K(24.0)) * K(1023.0));
    Keys[i] = spread(x) | (spread(y) << 1) | (spread(z) << 2);
  }
  static void sortPass(unsigned* Keys, unsigned* Order, int distance, int span, int i) {
    int j = i ^ distance;
    if (j <= i) return;
    unsigned a = Keys[i], b = Keys[j], ia = Order[i], ib = Order[j];
    bool greater = a > b || (a == b && ia > ib);
    bool ascending = (i & span) == 0;
    if (greater == ascending) { Keys[i] = b; Keys[j] = a; Order[i] = ib; Order[j] = ia; }
  }
  static void makeBound(const R* P, const unsigned* Order, R* Nodes, int i) {
    int id = (int)Order[i], b = (CAPACITY + i) * 8;
    st3(Nodes, b, boundLow(P, id));
    st3(Nodes, b + 4, boundHigh(P, id));
    Nodes[b + 3] = P[id * PS + 3] >= K(0.0) ? K(1.0) : K(0.0);
    Nodes[b + 7] = R(id);
  }
  static void reduceBound(R* Nodes, int first, int i) {
    int b = (first + i) * 8, a = b * 2, c = a + 8;
    st3(Nodes, b, vmin(ld3(Nodes, a), ld3(Nodes, c)));
    st3(Nodes, b + 4, vmax(ld3(Nodes, a + 4), ld3(Nodes, c + 4)));
    Nodes[b + 3] = Nodes[a + 3] + Nodes[c + 3];
    Nodes[b + 7] = K(-1.0);
  }

  // ===================== surface fields (Arbor.cu:148-174) =====================
  static F2 woodUV(const R* P, int id, F3 p) {
    int b = id * PS;
    F3 a = ld3(P, b);
    F3 d = unit(ld3(P, b + 4) - a);
    F3 initial = ld3(P, b + 8);
    F3 u = unit(initial - d * dot(initial, d));
    F3 v = cross(d, u);
    F3 q = p - a;
    R z = dot(q, d);
    R theta = m_atan2(dot(q, v), dot(q, u));
    return {theta, P[b + 16] + z};
  }
  static R barkCut(const R* P, int id, F3 p, R footprint) {
    const R pi = PI();
    int b = id * PS;
    F3 a = ld3(P, b);
    F3 d = ld3(P, b + 4) - a;
    R along = sat(dot(p - a, d) / m_max(K(0.000001), dot(d, d)));
    R radius = lerp(P[b + 7], P[b + 11], along);
    R amp = m_min(K(0.021), radius * K(0.055));
    int columns = (int)m_max(K(7.0), m_floor(P[b + 18] * K(2.0) * pi / K(0.065) + K(0.5)));
    R period = radius * K(2.0) * pi / R(columns);
    R coarse = bandWeight(period, footprint);
    R fine = bandWeight(K(0.010), footprint);
    R grain = bandWeight(K(0.0025), footprint);
    if (coarse < K(0.000001) && fine < K(0.000001)) return amp * K(0.32);
    F2 uv = woodUV(P, id, p);
    R theta = uv.x, arc = uv.y;
    int seed = (int)P[b + 15];
    // With bandSkips, a term whose band weight is 0 is a signed zero, and the
    // running sum (from 0.32, never -0) is unchanged by adding it: the term's
    // inputs (the plates' cell coordinates and Voronoi search, or a noise)
    // are then not evaluated.
    R sum = K(0.32);
    if (!(bandSkips && coarse == K(0.0))) {
      R x = (theta / (K(2.0) * pi) + K(0.5)) * R(columns) + K(0.32) * m_sin(arc * K(6.7) + m_cos(theta) * K(3.0)) +
            K(0.15) * m_sin(arc * K(19.0) + m_sin(theta) * K(2.0));
      R y = arc / K(0.19) + K(0.25) * m_sin(theta * K(7.0)) + K(0.1) * m_sin(theta * K(19.0));
      int ix = (int)m_floor(x), iy = (int)m_floor(y);
      R first = K(100.0), second = K(100.0);
      for (int j = -1; j <= 1; j++)
        for (int k = -1; k <= 1; k++) {
          int cell = ix + k;
          int wrapped = ((cell % columns) + columns) % columns;
          R xx = R(cell) + K(0.18) + K(0.64) * hash2(wrapped, iy + j, seed) - x;
          R yy = R(iy + j) + K(0.18) + K(0.64) * hash2(wrapped + columns, iy + j, seed + 31) - y;
          R dd = xx * xx + yy * yy;
          if (dd < first) { second = first; first = dd; }
          else if (dd < second) second = dd;
        }
      R gap = m_sqrt(second) - m_sqrt(first);
      R fissure = K(1.0) - smooth(K(0.015), K(0.17), gap);
      R plates = fissure * fissure;
      sum = sum + coarse * (plates - K(0.37)) * K(0.69);
    }
    F3 coord = c3(m_cos(theta) * K(9.0), m_sin(theta) * K(9.0), arc * K(23.0));
    if (!(bandSkips && fine == K(0.0))) sum = sum + fine * (noise3(coord, seed) - K(0.5)) * K(0.24);
    if (!(bandSkips && grain == K(0.0))) sum = sum + grain * (noise3(coord * K(4.0), seed + 731) - K(0.5)) * K(0.08);
    return amp * clampf(sum, K(0.005), K(1.0));
  }
  static R woodField(const R* P, int id, F3 p, R footprint) {
    int b = id * PS;
    F3 a = ld3(P, b);
    F3 ab = ld3(P, b + 4) - a;
    R t = sat(dot(p - a, ab) / m_max(K(0.0000001), dot(ab, ab)));
    R radius = lerp(P[b + 7], P[b + 11], t);
    return len(p - a - ab * t) - radius + barkCut(P, id, p, footprint);
  }
  static R windAngle(const R* P, int id, R time, R wind) {
    int b = id * PS;
    R phase = R(((int)P[b + 15]) % 8192) * K(0.618);
    return wind * (K(0.11) * m_sin(time * K(2.1) + phase) + K(0.045) * m_sin(time * K(5.8) + phase * K(1.7)));
  }
  // When set, leafU/leafN return the stored basis when the breeze is off
  // (the GPU's leafFrame): the angle wind * (...) is then +-0, cos = 1 and
  // sin = +-0, so the rotated vectors equal the stored ones in value.
  // Results must be bit-identical with the flag off (checked on whole frames).
  static inline bool windShortcut = false;
  static F3 leafU(const R* P, int id, R time, R wind) {
    int b = id * PS;
    F3 u = ld3(P, b + 4), n = ld3(P, b + 12);
    if (windShortcut && wind == K(0.0)) return u;
    R a = windAngle(P, id, time, wind);
    return u * m_cos(a) + n * m_sin(a);
  }
  static F3 leafN(const R* P, int id, R time, R wind) {
    int b = id * PS;
    F3 u = ld3(P, b + 4), n = ld3(P, b + 12);
    if (windShortcut && wind == K(0.0)) return n;
    R a = windAngle(P, id, time, wind);
    return n * m_cos(a) - u * m_sin(a);
  }
  // When set, a band-weighted detail term whose band weight is exactly 0 is
  // not evaluated: leafWidth's fine lobes (bw * 0.015 * sin(...) is a signed
  // zero, and lobes + (+-0) = lobes as lobes >= 0.54), leafHeight's midrib,
  // whose value (m * s) * 0 with m > 0 is s * 0 for any positive m, and
  // barkCut's plates, splinters and pores (below).
  // Results must be bit-identical with the flag off.
  static inline bool bandSkips = false;
  static R leafWidth(R u, R length, R halfWidth, R footprint) {
    const R pi = PI();
    R q = sat(u / length);
    R s = m_sin(pi * q);
    R lobes = K(0.77) + K(0.23) * m_cos(q * pi * K(10.0) + K(0.45));
    R bw = bandWeight(K(0.002), footprint);
    R fine = bandSkips && bw == K(0.0) ? K(0.0) : bw * K(0.015) * m_sin(q * pi * K(160.0));
    return halfWidth * m_pow(m_max(K(0.0), s), K(0.72)) * (lobes + fine);
  }
  static R leafHeight(R u, R v, R ln, R halfWidth, R curl, R twist, R footprint) {
    R q = clampf(u / ln, K(-0.2), K(1.2));
    R s = m_sin(PI() * q);
    R coarse = curl * s + twist * (q - K(0.4)) * v + K(0.009) * (v * v / (halfWidth * halfWidth)) * s;
    R bw = bandWeight(K(0.007), footprint);
    R m = bandSkips && bw == K(0.0) ? K(1.0) : K(0.0012) * m_exp(-m_abs(v) / K(0.005));
    R mid = m * s * bw;
    return coarse + mid;
  }

  // ===================== camera (Arbor.cu:179-182) =====================
  static void fpsPose(R* C, F3 eye, R yaw, R pitch) {
    C[16] = yaw; C[17] = pitch; C[27] = yaw; C[28] = pitch;
    st3(C, 0, eye);
    R cp = m_cos(pitch);
    F3 f = unit(c3(m_sin(yaw) * cp, m_sin(pitch), m_cos(yaw) * cp));
    st3(C, 4, eye + f);
    C[3] = K(1.0); C[18] = K(1.0);
  }
  static void viewPreset(R* C, int preset) {
    const R pi = PI();
    if (preset == 1) fpsPose(C, c3(K(2.0), K(1.72), K(18.0)), pi, K(0.0));
    if (preset == 2) fpsPose(C, c3(K(-1.4), K(1.72), K(7.0)), pi * K(0.94), K(0.02));
    if (preset == 3) fpsPose(C, c3(K(8.0), K(1.72), K(-3.0)), pi * K(0.55), K(-0.02));
    if (preset == 4) fpsPose(C, c3(K(-11.0), K(1.72), K(-7.0)), K(0.65), K(0.03));
    if (preset == 5) fpsPose(C, c3(K(0.0), K(12.0), K(72.0)), pi, K(-0.12));
    C[9] = K(0.0); C[10] = K(1.0);
  }
  static void initCamera(R* C, int seed) {
    for (int i = 0; i < 64; i++) C[i] = K(0.0);
    C[12] = K(-0.45); C[13] = K(0.63); C[14] = K(0.90); C[20] = K(1.18); C[22] = K(1.0); C[23] = K(1.0);
    C[31] = R(seed); C[44] = K(1.0);
    fpsPose(C, c3(K(2.0), K(1.72), K(18.0)), PI(), K(0.0));
  }
  static void updateCamera(R* C, const R* I, R delta, int width, int height) {
    C[10] = K(0.0);
    C[24] += K(1.0);
    R dt = clampf(delta, K(0.0), K(0.1));
    F3 old = ld3(C, 0);
    if (I[5] > K(0.0)) viewPreset(C, (int)I[5]);
    if (I[7] > K(0.0)) { C[11] = C[11] > K(0.0) ? K(0.0) : K(0.60); C[10] = K(1.0); }
    if (I[9] > K(0.0)) { C[15] = R((((int)C[15]) + 1) % 4); C[10] = K(1.0); }
    if (I[10] > K(0.0)) { C[22] = K(1.0) - C[22]; C[10] = K(1.0); }
    if (I[8] != K(0.0)) { C[12] += I[8] * dt * K(0.6); C[10] = K(1.0); }
    if (I[12] != K(0.0)) { C[20] = clampf(C[20] + I[12] * dt, K(0.25), K(3.0)); C[10] = K(1.0); }
    if (C[11] > K(0.0)) { C[8] += dt; C[10] = K(1.0); }
    C[16] -= I[0] * K(0.00225);
    C[17] = clampf(C[17] - I[1] * K(0.0020), K(-1.45), K(1.45));
    R yaw = C[16], cp = m_cos(C[17]);
    F3 look = unit(c3(m_sin(yaw) * cp, m_sin(C[17]), m_cos(yaw) * cp));
    F3 walkF = unit(c3(m_sin(yaw), K(0.0), m_cos(yaw)));
    F3 right = unit(c3(walkF.z, K(0.0), -walkF.x));
    R speed = I[16] > K(0.5) ? K(10.5) : K(4.2);
    F3 move = walkF * I[15] + right * I[13];
    if (dot(move, move) > K(0.0001)) move = unit(move) * speed * dt;
    F3 eye = ld3(C, 0) + move;
    eye.y = K(1.72);
    st3(C, 0, eye);
    st3(C, 4, eye + look);
    C[27] = C[16]; C[28] = C[17]; C[3] = K(1.0); C[18] = K(1.0);
    if (len(eye - old) > K(0.000001) || I[0] != K(0.0) || I[1] != K(0.0)) C[10] = K(1.0);
    if (C[32] != R(width) || C[33] != R(height)) { C[10] = K(1.0); C[32] = R(width); C[33] = R(height); }
    C[9] = C[10] > K(0.5) ? K(0.0) : m_min(K(255.0), C[9] + K(1.0));
  }

  // ===================== intersection (Arbor.cu:186-219) =====================
  static R sphereT(F3 ro, F3 rd, F3 p, R radius) {
    F3 q = ro - p;
    R b = dot(q, rd);
    R c = dot(q, q) - radius * radius;
    R d = b * b - c;
    if (d < K(0.0)) return FAR();
    R t = -b - m_sqrt(d);
    if (t <= K(0.0001)) t = -b + m_sqrt(d);
    return t > K(0.0001) ? t : FAR();
  }
  static R tubeT(F3 ro, F3 rd, F3 a, F3 b, R ra, R rb) {
    F3 axis = b - a;
    R ln = len(axis);
    axis = axis / m_max(ln, K(0.000001));
    F3 q = ro - a;
    R z = dot(q, axis), dz = dot(rd, axis);
    R k = (rb - ra) / m_max(K(0.000001), ln);
    F3 v = q - axis * z, d = rd - axis * dz;
    R radius = ra + k * z;
    R aa = dot(d, d) - k * k * dz * dz;
    R bb = dot(v, d) - radius * k * dz;
    R cc = dot(v, v) - radius * radius;
    R best = m_min(sphereT(ro, rd, a, ra), sphereT(ro, rd, b, rb));
    R disc = bb * bb - aa * cc;
    if (disc >= K(0.0) && m_abs(aa) > K(0.0000001)) {
      for (int j = 0; j < 2; j++) {
        R t = (-bb + (j == 0 ? K(-1.0) : K(1.0)) * m_sqrt(disc)) / aa;
        R h = z + t * dz;
        if (t > K(0.0001) && h >= K(0.0) && h <= ln) best = m_min(best, t);
      }
    }
    return best;
  }
  // When set, woodT and leafT first reject rays that miss a padded bounding
  // sphere of everything they can hit (the GPU's early-out); results must be
  // bit-identical with the flag off, which the tests check on whole frames.
  static inline bool cullSpheres = false;
  // Padding of the rejection spheres. sphereT/tubeT decide hits from
  // discriminants that cancel catastrophically far from the object, so the
  // reference reports hits up to ~8e-4*|q| beyond a small sphere or tube at
  // distance |q|; tests that end in such a discriminant are padded by
  // 1 mm + 4e-3*|q|. The leaf slab (boxSpan) and the wood SDF march (which
  // must reach within 2.5e-4 m of the surface) are well conditioned: 1 mm.
  // Test infrastructure: scales both pads (1 = the GPU's; a test lowers it to
  // show that the grazing-ray stress test catches a too-tight rejection).
  static inline R cullPadScale = R(1);
  static bool sphereMiss(F3 ro, F3 rd, F3 centre, R radius, R padPerMetre) {
    F3 q = ro - centre;
    R b = dot(q, rd);
    F3 p = q - rd * b;
    R r = radius + (K(0.001) + padPerMetre * len(q)) * cullPadScale;
    return dot(p, p) > r * r || b > r;
  }

  // When set, woodT and leafT stop once their answer is known to be >= the
  // caller's current best (a result the caller discards: it keeps strictly
  // nearer hits): the wood march never moves before the tube hit and never
  // back from any step, the leaf Newton solution never before the slab entry. Results must be
  // bit-identical with the flag off (checked on whole frames).
  static inline bool skipBeyondBest = false;
  // When set, the leaf Newton solve stops as soon as its remaining iterations
  // are known: an update that leaves t unchanged bit for bit repeats forever
  // (each iteration is a pure function of t), and one that returns t to its
  // value two updates earlier alternates between the last two values until
  // the 7th, whose value the parity of the remaining count selects. The
  // derivative break cannot fire later in either case: every value involved
  // already passed it. Results must be bit-identical with the flag off.
  static inline bool newtonExits = false;
  static bool sameBits(R a, R b) { return std::memcmp(&a, &b, sizeof(R)) == 0; }
  static R woodT(const R* P, int id, F3 ro, F3 rd, R footprint, bool refine, R bestSoFar = FAR()) {
    int b = id * PS;
    F3 a = ld3(P, b), end = ld3(P, b + 4);
    if (cullSpheres) {
      F3 half = (end - a) * K(0.5);
      if (sphereMiss(ro, rd, a + half, len(half) + m_max(P[b + 7], P[b + 11]), refine ? K(0.0) : K(0.004)))
        return FAR();
    }
    if (!refine) {
      R ra = P[b + 7] - m_min(K(0.021), P[b + 7] * K(0.055)) * K(0.32);
      R rb = P[b + 11] - m_min(K(0.021), P[b + 11] * K(0.055)) * K(0.32);
      return tubeT(ro, rd, a, end, ra, rb);
    }
    R t = tubeT(ro, rd, a, end, P[b + 7], P[b + 11]);
    if (t >= FAR()) return FAR();
    if (skipBeyondBest && t >= bestSoFar) return FAR();
    F2 span = boxSpan(ro, rd, boundLow(P, id), boundHigh(P, id));
    R eps = m_max(K(0.000025), m_min(K(0.00025), footprint * K(0.13)));
    for (int k = 0; k < 64; k++) {
      R f = woodField(P, id, ro + rd * t, footprint);
      if (f < eps) return t;
      t += m_max(eps * K(0.35), f * K(0.27));
      if (t > span.y) return FAR();
      if (skipBeyondBest && t >= bestSoFar) return FAR();  // t only grows: any hit is >= best
    }
    return FAR();
  }
  static R leafT(const R* P, int id, F3 ro, F3 rd, R footprint, R time, R wind, R bestSoFar = FAR()) {
    int b = id * PS;
    F3 origin = ld3(P, b);
    F3 u = leafU(P, id, time, wind);
    F3 v = ld3(P, b + 8);
    F3 n = leafN(P, id, time, wind);
    R stem = P[b + 16], ln = P[b + 7], width = P[b + 11];
    bool petiole = true, slab = true;
    if (cullSpheres) {
      // petiole capsule: radius <= 1.1 mm along u over [0, stem]
      petiole = !sphereMiss(ro, rd, origin + u * (stem * K(0.5)), stem * K(0.5) + K(0.0011), K(0.004));
      // Newton slab: u in [stem-0.001, stem+len+0.001], |v| <= width, n in [-0.035, 0.065]
      R hx = ln * K(0.5) + K(0.001);
      slab = !sphereMiss(ro, rd, origin + u * (stem + ln * K(0.5)) + n * K(0.015),
                         m_sqrt(hx * hx + width * width + K(0.05) * K(0.05)), K(0.00001));
      if (!petiole && !slab) return FAR();
    }
    F3 q = ro - origin;
    R x = dot(q, u) - stem, y = dot(q, v), z = dot(q, n);
    R dx = dot(rd, u), dy = dot(rd, v), dz = dot(rd, n);
    R best = petiole ? tubeT(ro, rd, origin, origin + u * stem, K(0.0011), K(0.00065)) : FAR();
    if (!slab) return best;
    F2 span = boxSpan(c3(x, y, z), c3(dx, dy, dz), c3(K(-0.001), -width, K(-0.035)), c3(ln + K(0.001), width, K(0.065)));
    if (span.y < m_max(K(0.0001), span.x)) return best;
    if (skipBeyondBest && span.x >= bestSoFar) return best;
    R t = m_abs(dz) > K(0.000001) ? -z / dz : (span.x + span.y) * K(0.5);
    t = clampf(t, m_max(K(0.0001), span.x), span.y);
    const R curl = P[b + 17], twist = P[b + 18];
    R tPrev = t;  // t before the previous update
    for (int k = 0; k < 7; k++) {
      R xx = x + dx * t, yy = y + dy * t;
      R height = leafHeight(xx, yy, ln, width, curl, twist, footprint);
      R f = z + dz * t - height;
      R e = K(0.00015);
      R hx = (leafHeight(xx + e, yy, ln, width, curl, twist, footprint) -
              leafHeight(xx - e, yy, ln, width, curl, twist, footprint)) / (K(2.0) * e);
      R hy = (leafHeight(xx, yy + e, ln, width, curl, twist, footprint) -
              leafHeight(xx, yy - e, ln, width, curl, twist, footprint)) / (K(2.0) * e);
      R deriv = dz - hx * dx - hy * dy;
      if (m_abs(deriv) < K(0.00001)) break;
      R tn = t - clampf(f / deriv, K(-0.13), K(0.13));
      tn = clampf(tn, span.x, span.y);
      if (newtonExits) {
        if (sameBits(tn, t)) break;                          // fixed point
        if (sameBits(tn, tPrev)) { if (((6 - k) & 1) == 0) t = tn; break; }  // 2-cycle
      }
      tPrev = t;
      t = tn;
    }
    R xx = x + dx * t, yy = y + dy * t;
    R err = m_abs(z + dz * t - leafHeight(xx, yy, ln, width, curl, twist, footprint));
    if (t > K(0.0001) && xx > K(0.0) && xx < ln && m_abs(yy) < leafWidth(xx, ln, width, footprint) && err < K(0.0003))
      best = m_min(best, t);
    return best;
  }
  static R objectT(const R* P, int id, F3 ro, F3 rd, R footprint, R time, R wind, bool refine, R best = FAR()) {
    int b = id * PS, type = (int)P[b + 3];
    if (type < 0) return FAR();
    if (type == 0) return woodT(P, id, ro, rd, footprint, refine, best);
    if (type == 1) return leafT(P, id, ro, rd, footprint, time, wind, best);
    F3 h = ld3(P, b + 4);
    F3 q = ro - ld3(P, b);
    F3 p = c3(q.x / h.x, q.y / h.y, q.z / h.z);
    F3 d = c3(rd.x / h.x, rd.y / h.y, rd.z / h.z);
    R aa = dot(d, d), bb = dot(p, d), cc = dot(p, p) - K(1.0);
    R disc = bb * bb - aa * cc;
    if (disc < K(0.0)) return FAR();
    R t = (-bb - m_sqrt(disc)) / aa;
    return t > K(0.0001) ? t : FAR();
  }

  // When set, traceTree runs the lean equivalent walk below instead of the
  // reference's stack loop (used to cross-check the GPU's traversal order and
  // to count its operations; results are bit-identical).
  static inline bool leanWalk = false;

  // Heap BVH walk with the traversal stack as a bit per heap depth: the
  // reference pushes only the far sibling of the child it descends into, so
  // the stack always holds far siblings along the current root path, deepest
  // on top. Popping takes the deepest set bit d and visits the sibling of the
  // path's depth-d node; a popped node is re-tested with the current best as
  // the reference does. A child reached directly keeps the span its parent
  // computed, and the ray reciprocal is computed once (identical values).
  static F2 traceTreeLean(const R* P, const R* Nodes, const unsigned* Order, F3 ro, F3 rd, R cone, R time, R wind,
                          R best, int ignore, bool shadows) {
    const F3 inv = c3(invSafe(rd.x), invSafe(rd.y), invSafe(rd.z));
    auto span_of = [&](unsigned n) {
      F3 a = (ld3(Nodes, n * 8) - ro) * inv, b = (ld3(Nodes, n * 8 + 4) - ro) * inv;
      F3 mn = vmin(a, b), mx = vmax(a, b);
      return F2{m_max(mn.x, m_max(mn.y, mn.z)), m_min(mx.x, m_min(mx.y, mx.z))};
    };
    auto live = [&](F2 s, unsigned n) {
      return Nodes[n * 8 + 3] > K(0.0) && s.y >= m_max(K(0.0001), s.x) && s.x < best;
    };
    auto depth = [](unsigned n) { return 31 - __builtin_clz(n); };
    int found = -1;
    unsigned node = 1, pending = 0;
    F2 span = span_of(1);
    bool hit = live(span, 1);
    while (true) {
      if (hit) {
        if (node < (unsigned)CAPACITY) {
          unsigned l = node * 2, r = l + 1;
          F2 ls = span_of(l), rs = span_of(r);
          bool lh = live(ls, l), rh = live(rs, r);
          if (lh || rh) {
            bool takeLeft = lh && (!rh || ls.x < rs.x);
            if (lh && rh) pending |= 1u << (depth(node) + 1);
            node = takeLeft ? l : r;
            span = takeLeft ? ls : rs;
            continue;
          }
        } else {
          int id = (int)Order[node - CAPACITY];
          if (id != ignore) {
            R footprint = m_max(K(0.000005), cone * m_max(K(0.05), span.x));
            R t = objectT(P, id, ro, rd, footprint, time, wind, !shadows, best);
            if (t < best) { best = t; found = id; }
          }
        }
      }
      if (pending == 0) break;
      int d = 31 - __builtin_clz(pending);
      pending &= ~(1u << d);
      node = (node >> (depth(node) - d)) ^ 1u;
      span = span_of(node);
      hit = live(span, node);
    }
    return {best, R(found)};
  }

  // ---- walk over a SAH hierarchy (no counterpart in the reference) ----
  // The GPU's accel walk, restated in this file's arithmetic: a binary
  // hierarchy over the heap's leaf boxes in the device layout of
  // arbor-cuda_amd/csrc/sah_build.hpp (sahAccel, set by the caller), walked
  // near-first with a stack of the far siblings and their entry t. The object
  // tests run with their early exits made strict (best raised by one ulp), so
  // an object with t equal to the best reports it; of two such, the one the
  // heap walk tests first is kept: their order is decided at their lowest
  // common heap ancestor, by the entry t of its two children (nearer first,
  // the right one on equal t). Used to check that walk against the heap walk
  // on whole frames and to count its operations.
  static inline const uint32_t* sahAccel = nullptr;
  static inline bool sahTieReversed = false;  // planted mistake for the tests: the tie order inverted
  static R nextUpR(R x) {  // x > 0 and finite: the next representable value
    if constexpr (sizeof(R) == 8) {
      uint64_t u;
      std::memcpy(&u, &x, 8);
      u++;
      std::memcpy(&x, &u, 8);
    } else {
      uint32_t u;
      std::memcpy(&u, &x, 4);
      u++;
      std::memcpy(&x, &u, 4);
    }
    return x;
  }
  static F3 sahV(const uint32_t* w) {
    float f[3];
    std::memcpy(f, w, sizeof f);
    return c3(R(f[0]), R(f[1]), R(f[2]));
  }
  static F2 spanInv(F3 ro, F3 inv, F3 lo, F3 hi) {
    F3 a = (lo - ro) * inv, b = (hi - ro) * inv;
    F3 mn = vmin(a, b), mx = vmax(a, b);
    return F2{m_max(mn.x, m_max(mn.y, mn.z)), m_min(mx.x, m_min(mx.y, mx.z))};
  }
  static bool heapLeafFirst(const R* Nodes, F3 ro, F3 inv, unsigned sa, unsigned sb) {
    const unsigned na = CAPACITY + sa, nb = CAPACITY + sb;
    const int k = 32 - __builtin_clz(na ^ nb);
    const unsigned ca = na >> (k - 1), cb = nb >> (k - 1);
    const unsigned l = std::min(ca, cb), r = l ^ 1u;
    F2 ls = spanInv(ro, inv, ld3(Nodes, l * 8), ld3(Nodes, l * 8 + 4));
    F2 rs = spanInv(ro, inv, ld3(Nodes, r * 8), ld3(Nodes, r * 8 + 4));
    return ((ls.x < rs.x ? l : r) == ca) != sahTieReversed;
  }
  static F2 traceTreeSah(const R* P, const R* Nodes, F3 ro, F3 rd, R cone, R time, R wind, R best, int ignore,
                         bool shadows) {
    const uint32_t* A = sahAccel;
    const uint32_t* slotOf = A + 16 + 16 * (CAPACITY - 1);
    const F3 inv = c3(invSafe(rd.x), invSafe(rd.y), invSafe(rd.z));
    auto live = [&](F2 sp) { return sp.y >= m_max(K(0.0001), sp.x) && sp.x < best; };
    int found = -1;
    F2 span = spanInv(ro, inv, sahV(A), sahV(A + 4));
    int ref = (int)A[3];
    bool hit = live(span);
    std::vector<std::pair<int, R>> stack;
    while (true) {
      if (hit) {
        if (ref >= 0) {
          const uint32_t* nd = A + 16 + 16 * (size_t)ref;
          F2 ls = spanInv(ro, inv, sahV(nd), sahV(nd + 4)), rs = spanInv(ro, inv, sahV(nd + 8), sahV(nd + 12));
          bool lh = live(ls), rh = live(rs);
          if (lh || rh) {
            bool takeLeft = lh && (!rh || ls.x < rs.x);
            int lref = (int)nd[3], rref = (int)nd[11];
            if (lh && rh) stack.push_back({takeLeft ? rref : lref, takeLeft ? rs.x : ls.x});
            ref = takeLeft ? lref : rref;
            span = takeLeft ? ls : rs;
            continue;
          }
        } else {
          const int id = ~ref;
          if (id != ignore) {
            R footprint = m_max(K(0.000005), cone * m_max(K(0.05), span.x));
            R t = objectT(P, id, ro, rd, footprint, time, wind, !shadows, nextUpR(best));
            if (t < best) {
              best = t;
              found = id;
            } else if (t == best && found >= 0 && heapLeafFirst(Nodes, ro, inv, slotOf[id], slotOf[found])) {
              found = id;
            }
          }
        }
      }
      hit = false;
      while (!stack.empty()) {
        auto e = stack.back();
        stack.pop_back();
        if (e.second < best) {
          ref = e.first;
          span.x = e.second;
          hit = true;
          break;
        }
      }
      if (!hit) break;
    }
    return {best, R(found)};
  }

  // Heap BVH walk, explicit stack (Arbor.cu:220-230).
  static F2 traceTree(const R* P, const R* Nodes, const unsigned* Order, F3 ro, F3 rd, R cone, R time, R wind,
                      R best, int ignore, bool shadows) {
    if (sahAccel) return traceTreeSah(P, Nodes, ro, rd, cone, time, wind, best, ignore, shadows);
    if (leanWalk) return traceTreeLean(P, Nodes, Order, ro, rd, cone, time, wind, best, ignore, shadows);
    int stack[32], top = 0, node = 1, found = -1;
    for (int visit = 0; visit < NODE_COUNT; visit++) {
      if (node <= 0) break;
      int b = node * 8;
      F2 span = boxSpan(ro, rd, ld3(Nodes, b), ld3(Nodes, b + 4));
      bool hit = Nodes[b + 3] > K(0.0) && span.y >= m_max(K(0.0001), span.x) && span.x < best;
      if (hit && node < CAPACITY) {
        int left = node * 2, right = left + 1, lb = left * 8, rb = right * 8;
        F2 ls = boxSpan(ro, rd, ld3(Nodes, lb), ld3(Nodes, lb + 4));
        F2 rs = boxSpan(ro, rd, ld3(Nodes, rb), ld3(Nodes, rb + 4));
        bool lh = Nodes[lb + 3] > K(0.0) && ls.y >= m_max(K(0.0001), ls.x) && ls.x < best;
        bool rh = Nodes[rb + 3] > K(0.0) && rs.y >= m_max(K(0.0001), rs.x) && rs.x < best;
        if (lh && rh) {
          bool l = ls.x < rs.x;
          stack[top++] = l ? right : left;
          node = l ? left : right;
          continue;
        }
        if (lh) { node = left; continue; }
        if (rh) { node = right; continue; }
      } else if (hit) {
        int id = (int)Order[node - CAPACITY];
        if (id != ignore) {
          R footprint = m_max(K(0.000005), cone * m_max(K(0.05), span.x));
          R t = objectT(P, id, ro, rd, footprint, time, wind, !shadows, best);
          if (t < best) { best = t; found = id; }
        }
      }
      node = 0;
      if (top > 0) node = stack[--top];
    }
    return {best, R(found)};
  }
  static F3 objectNormal(const R* P, int id, F3 p, R footprint, R time, R wind) {
    int b = id * PS, type = (int)P[b + 3];
    if (type == 0) {
      R e = m_max(K(0.000035), m_min(K(0.0005), footprint * K(0.18)));
      F3 dx = c3(e, K(0.0), K(0.0)), dy = c3(K(0.0), e, K(0.0)), dz = c3(K(0.0), K(0.0), e);
      return unit(c3(woodField(P, id, p + dx, footprint) - woodField(P, id, p - dx, footprint),
                     woodField(P, id, p + dy, footprint) - woodField(P, id, p - dy, footprint),
                     woodField(P, id, p + dz, footprint) - woodField(P, id, p - dz, footprint)));
    }
    if (type == 1) {
      F3 u = leafU(P, id, time, wind);
      F3 v = ld3(P, b + 8);
      F3 n = leafN(P, id, time, wind);
      F3 q = p - ld3(P, b);
      R x = dot(q, u) - P[b + 16], y = dot(q, v);
      if (x < K(0.0)) {
        F3 centre = ld3(P, b) + u * (x + P[b + 16]);
        return unit(p - centre);
      }
      R e = K(0.0002);
      R a = (leafHeight(x + e, y, P[b + 7], P[b + 11], P[b + 17], P[b + 18], footprint) -
             leafHeight(x - e, y, P[b + 7], P[b + 11], P[b + 17], P[b + 18], footprint)) / (K(2.0) * e);
      R c = (leafHeight(x, y + e, P[b + 7], P[b + 11], P[b + 17], P[b + 18], footprint) -
             leafHeight(x, y - e, P[b + 7], P[b + 11], P[b + 17], P[b + 18], footprint)) / (K(2.0) * e);
      return unit(n - u * a - v * c);
    }
    F3 q = p - ld3(P, b), h = ld3(P, b + 4);
    return unit(c3(q.x / (h.x * h.x), q.y / (h.y * h.y), q.z / (h.z * h.z)));
  }
  // Forest-grid 2D DDA over instanced trees (Arbor.cu:235-251).
  static F2 traceForest(const R* P, const R* Nodes, const unsigned* Order, F3 ro, F3 rd, R cone, R time, R wind,
                        R best, int ws, int* treeOut, bool shadows) {
    const R spacing = K(28.0);
    int foundTree = -1;
    R found = K(-1.0);
    R half = R(FOREST_SIDE) * spacing * K(0.5);
    F2 worldSpan = boxSpan(ro, rd, c3(-half, K(-1.2), -half), c3(half, K(13.5), half));
    R enter = m_max(K(0.0), worldSpan.x), exit = m_min(best, worldSpan.y);
    if (exit < enter) { *treeOut = -1; return {best, K(-1.0)}; }
    F3 pos = ro + rd * (enter + K(0.0002));
    int ix = (int)m_floor(pos.x / spacing + R(FOREST_SIDE) * K(0.5) + K(0.5));
    int iz = (int)m_floor(pos.z / spacing + R(FOREST_SIDE) * K(0.5) + K(0.5));
    int sx = rd.x >= K(0.0) ? 1 : -1, sz = rd.z >= K(0.0) ? 1 : -1;
    R invx = invSafe(rd.x), invz = invSafe(rd.z);
    R bx = (R(ix - FOREST_SIDE / 2) + (sx > 0 ? K(0.5) : K(-0.5))) * spacing;
    R bz = (R(iz - FOREST_SIDE / 2) + (sz > 0 ? K(0.5) : K(-0.5))) * spacing;
    R tx = (bx - ro.x) * invx, tz = (bz - ro.z) * invz;
    R dx = spacing * m_abs(invx), dz = spacing * m_abs(invz);
    for (int step = 0; step < 48; step++) {
      if (ix < 0 || iz < 0 || ix >= FOREST_SIDE || iz >= FOREST_SIDE) break;
      int tree = iz * FOREST_SIDE + ix;
      R sc = forestScale(tree, ws);
      F3 centre = forestPos(tree, ws) + c3(K(0.0), K(4.4) * sc, K(0.0));
      R radius = K(6.6) * sc;
      R broad = sphereT(ro, rd, centre, radius);
      if (broad < best) {
        F3 lr = toLocalPoint(ro, tree, ws);
        F3 ld = toLocalDir(rd, tree, ws);
        R phase = R(forestSeed(tree, ws) & 8191) * K(0.0017);
        F2 h = traceTree(P, Nodes, Order, lr, ld, cone / sc, time + phase, wind, best / sc, -1, shadows);
        if (h.y >= K(0.0) && h.x * sc < best) { best = h.x * sc; found = h.y; foundTree = tree; }
      }
      R next = tx < tz ? tx : tz;
      if (next > exit || next > best) break;
      if (tx < tz) { ix += sx; tx += dx; } else { iz += sz; tz += dz; }
    }
    *treeOut = foundTree;
    return {best, found};
  }
  // One pixel of tracePixels (Arbor.cu:252).
  static void tracePixel(const R* P, const R* Nodes, const unsigned* Order, const R* C, R* Hit, int width, int height,
                         int x, int y) {
    int b = (y * width + x) * 4;
    F3 ro = cameraPos(C), rd = cameraRay(C, x, y, width, height);
    R ground = rd.y < K(-0.00001) ? -ro.y / rd.y : FAR();
    int tree = -1;
    F2 h = traceForest(P, Nodes, Order, ro, rd, C[14] / R(height), C[8], C[11], ground > K(0.0) ? ground : FAR(),
                       (int)C[31], &tree, false);
    Hit[b] = h.x; Hit[b + 1] = h.y; Hit[b + 2] = R(tree); Hit[b + 3] = C[14] * h.x / R(height);
  }
  // tracePixel along cameraRayOffset: the primary ray through a point (dx, dy)
  // pixels from the pixel's sample position; out = the Hit record.
  static void tracePixelOffset(const R* P, const R* Nodes, const unsigned* Order, const R* C, int width, int height,
                               int x, int y, R dx, R dy, R* out, bool single = false) {
    F3 ro = cameraPos(C), rd = cameraRayOffset(C, x, y, width, height, dx, dy);
    R ground = rd.y < K(-0.00001) ? -ro.y / rd.y : FAR();
    int tree = -1;
    F2 h = single ? traceTree(P, Nodes, Order, ro, rd, C[14] / R(height), C[8], C[11],
                              ground > K(0.0) ? ground : FAR(), -1, false)
                  : traceForest(P, Nodes, Order, ro, rd, C[14] / R(height), C[8], C[11],
                                ground > K(0.0) ? ground : FAR(), (int)C[31], &tree, false);
    out[0] = h.x; out[1] = h.y; out[2] = R(tree); out[3] = C[14] * h.x / R(height);
  }
  // Single-tree primary ray in tree-local space: traceTree with the ground plane
  // as the initial best and no forest transform (the single-specimen workload
  // that native_harness.cpp:84 traces). Hit layout as tracePixels, tree = -1.
  static void traceTreePixel(const R* P, const R* Nodes, const unsigned* Order, const R* C, R* Hit, int width,
                             int height, int x, int y) {
    int b = (y * width + x) * 4;
    F3 ro = cameraPos(C), rd = cameraRay(C, x, y, width, height);
    R ground = rd.y < K(-0.00001) ? -ro.y / rd.y : FAR();
    F2 h = traceTree(P, Nodes, Order, ro, rd, C[14] / R(height), C[8], C[11], ground > K(0.0) ? ground : FAR(), -1,
                     false);
    Hit[b] = h.x; Hit[b + 1] = h.y; Hit[b + 2] = K(-1.0); Hit[b + 3] = C[14] * h.x / R(height);
  }

  // ===================== shading (Arbor.cu:256-289) =====================
  static F3 sky(F3 rd, F3 sun) {
    R h = sat(rd.y);
    F3 c = mix(c3(K(0.75), K(0.79), K(0.77)), c3(K(0.19), K(0.35), K(0.54)), m_pow(h, K(0.48)));
    R warm = m_pow(sat(dot(rd, sun)), K(32.0));
    c = c + c3(K(0.48), K(0.26), K(0.075)) * warm;
    R disc = smooth(K(0.99996), K(0.999984), dot(rd, sun));
    c = c + c3(K(8.0), K(6.7), K(4.9)) * disc;
    R clouds = noise(rd.x * K(5.0) / (K(0.2) + h), rd.z * K(5.0) / (K(0.2) + h), 231);
    R veil = smooth(K(0.62), K(0.81), clouds) * smooth(K(0.02), K(0.18), h) * K(0.38);
    return mix(c, c3(K(1.1), K(1.10), K(1.05)), veil);
  }
  static F3 barkColor(const R* P, int id, F3 p, R footprint) {
    int b = id * PS, seed = (int)P[b + 15];
    F2 uv = woodUV(P, id, p);
    F3 coord = c3(m_cos(uv.x) * K(9.0), m_sin(uv.x) * K(9.0), uv.y * K(7.0));
    R mottle = noise3(p * K(3.0), 714);
    R rough = noise3(coord, seed);
    F3 d = ld3(P, b + 4) - ld3(P, b);
    R t = sat(dot(p - ld3(P, b), d) / m_max(K(0.000001), dot(d, d)));
    R amp = m_min(K(0.021), lerp(P[b + 7], P[b + 11], t) * K(0.055));
    R cut = barkCut(P, id, p, footprint) / m_max(K(0.00001), amp);
    F3 c = mix(c3(K(0.10), K(0.075), K(0.050)), c3(K(0.25), K(0.21), K(0.16)), K(0.25) + rough * K(0.62));
    c = c * (K(0.75) + K(0.35) * mottle);
    c = c * (K(1.0) - K(0.62) * smooth(K(0.28), K(0.85), cut));
    R moss = smooth(K(0.52), K(0.72), noise3(p * K(5.0), 189)) * (K(1.0) - smooth(K(0.15), K(1.4), p.y));
    c = mix(c, c3(K(0.045), K(0.064), K(0.018)), moss * K(0.62));
    R lichen = smooth(K(0.65), K(0.77), noise3(p * K(23.0), 27)) * smooth(K(0.25), K(0.65), noise3(p * K(90.0), 712)) *
               bandWeight(K(0.018), footprint);
    c = mix(c, c3(K(0.31), K(0.33), K(0.23)), lichen * K(0.5));
    R grain = (noise3(coord * K(15.0), seed) - K(0.5)) * bandWeight(K(0.0025), footprint);
    return c * (K(1.0) + grain * K(0.38));
  }
  static R leafVein(const R* P, int id, F3 p, R footprint, R time, R wind) {
    int b = id * PS;
    F3 q = p - ld3(P, b);
    R x = dot(q, leafU(P, id, time, wind)) - P[b + 16];
    R y = dot(q, ld3(P, b + 8));
    R ln = P[b + 7];
    R mid = m_exp(-m_abs(y) / m_max(K(0.00065), footprint * K(0.55)));
    R fy = frac((x - m_abs(y) * K(0.70)) / ln * K(7.0));
    R dd = m_min(fy, K(1.0) - fy) * ln / K(7.0);
    R lateral = m_exp(-dd / m_max(K(0.00033), footprint * K(0.5)));
    return sat(mid * K(0.82) + lateral * K(0.48)) * bandWeight(K(0.003), footprint);
  }
  static F3 leafColor(const R* P, int id, F3 p, R fp, R time, R wind) {
    int b = id * PS;
    R h = rnd((int)P[b + 15], 887);
    F3 q = p - ld3(P, b);
    R x = dot(q, leafU(P, id, time, wind)) - P[b + 16];
    if (x < K(0.0)) return c3(K(0.17), K(0.15), K(0.048));
    R y = dot(q, ld3(P, b + 8));
    R v = leafVein(P, id, p, fp, time, wind);
    F3 c = mix(c3(K(0.035), K(0.11), K(0.008)), c3(K(0.15), K(0.255), K(0.025)), h);
    c = mix(c, c3(K(0.22), K(0.26), K(0.055)), v * K(0.57));
    R speck = (noise(x * K(700.0), y * K(700.0), (int)P[b + 15]) - K(0.5)) * bandWeight(K(0.0017), fp);
    R edge = smooth(K(0.78), K(1.0), m_abs(y) / m_max(K(0.001), leafWidth(x, P[b + 7], P[b + 11], fp)));
    c = c * (K(1.0) + speck * K(0.27));
    c = mix(c, c3(K(0.23), K(0.20), K(0.05)), edge * K(0.22));
    return c;
  }
  static F3 groundColor(F3 p, R fp) {
    R a = noise(p.x * K(3.0), p.z * K(3.0), 819);
    R b = noise(p.x * K(15.0), p.z * K(15.0), 28);
    R r = m_sqrt(p.x * p.x + p.z * p.z);
    R soil = K(1.0) - smooth(K(4.1), K(6.8), r + noise(p.x, p.z, 13));
    F3 grass = mix(c3(K(0.085), K(0.115), K(0.035)), c3(K(0.23), K(0.24), K(0.092)), a);
    F3 dirt = mix(c3(K(0.075), K(0.055), K(0.035)), c3(K(0.16), K(0.13), K(0.080)), b);
    R moss = smooth(K(0.53), K(0.73), a);
    dirt = mix(dirt, c3(K(0.049), K(0.079), K(0.022)), moss * K(0.72));
    R grit = (noise(p.x * K(200.0), p.z * K(200.0), 127) - K(0.5)) * bandWeight(K(0.007), fp);
    return mix(grass, dirt, soil) * (K(1.0) + grit * K(0.40));
  }
  static F3 surfaceColor(const R* P, int id, F3 p, R fp, R time, R wind) {
    if (id < 0) return groundColor(p, fp);
    int b = id * PS, type = (int)P[b + 3];
    if (type == 0) return barkColor(P, id, p, fp);
    if (type == 1) return leafColor(P, id, p, fp, time, wind);
    R f = noise3(p * K(42.0), (int)P[b + 15]);
    return mix(c3(K(0.09), K(0.10), K(0.086)), c3(K(0.31), K(0.30), K(0.24)), f);
  }
  static F3 lightVisibility(const R* P, const R* Nodes, const unsigned* Order, F3 p, F3 direction, R time, R wind,
                            int ws) {
    F3 transmission = c3(K(1.0), K(1.0), K(1.0));
    F3 ro = p;
    for (int j = 0; j < 4; j++) {
      int tree = -1;
      F2 h = traceForest(P, Nodes, Order, ro, direction, K(0.0001), time, wind, K(36.0), ws, &tree, true);
      int id = (int)h.y;
      if (id < 0) break;
      if (P[id * PS + 3] != K(1.0)) return c3(K(0.0), K(0.0), K(0.0));
      transmission = transmission * c3(K(0.18), K(0.29), K(0.065));
      ro = ro + direction * (h.x + K(0.0018));
    }
    return transmission;
  }
  static R specGGX(F3 n, F3 v, F3 l, R rough) {
    F3 h = unit(v + l);
    R nh = sat(dot(n, h));
    R nv = m_max(K(0.002), sat(dot(n, v)));
    R nl = sat(dot(n, l));
    R vh = sat(dot(v, h));
    R a = rough * rough, aa = a * a;
    R den = nh * nh * (aa - K(1.0)) + K(1.0);
    R D = aa / (PI() * den * den);
    R k = (rough + K(1.0)) * (rough + K(1.0)) * K(0.125);
    R G = nv / (nv * (K(1.0) - k) + k) * nl / (nl * (K(1.0) - k) + k);
    R F = K(0.035) + K(0.965) * m_pow(K(1.0) - vh, K(5.0));
    return D * G * F / (K(4.0) * nv + K(0.0001));
  }
  // One pixel of shadePixels (Arbor.cu:277-289).
  static void shadePixel(const R* P, const R* Nodes, const unsigned* Order, const R* C, const R* Hit, R* Linear,
                         R* Guide, int width, int height, int x, int y) {
    const R pi = PI();
    int pixel = y * width + x, b = pixel * 4;
    R t = Hit[b];
    int id = (int)Hit[b + 1], tree = (int)Hit[b + 2];
    F3 ro = cameraPos(C), rd = cameraRay(C, x, y, width, height);
    F3 sun = sunC(C);
    F3 col = sky(rd, sun);
    F3 guide = c3(K(1.0), K(1.0), K(1.0));
    int ws = (int)C[31];
    if (t < FAR()) {
      F3 p = ro + rd * t;
      R fp = m_max(K(0.000005), Hit[b + 3]);
      F3 localP = p, n = c3(K(0.0), K(1.0), K(0.0));
      R localFp = fp, phase = K(0.0);
      if (id >= 0 && tree >= 0) {
        R sc = forestScale(tree, ws);
        localP = toLocalPoint(p, tree, ws);
        localFp = fp / sc;
        phase = R(forestSeed(tree, ws) & 8191) * K(0.0017);
        n = toWorldNormal(objectNormal(P, id, localP, localFp, C[8] + phase, C[11]), tree, ws);
      }
      bool leaf = id >= 0 && P[id * PS + 3] == K(1.0);
      if (dot(n, rd) > K(0.0)) n = n * K(-1.0);
      if (id < 0) {
        R e = m_max(K(0.001), fp);
        R nx = (noise((p.x + e) * K(18.0), p.z * K(18.0), 133) - noise((p.x - e) * K(18.0), p.z * K(18.0), 133)) * K(0.18);
        R nz = (noise(p.x * K(18.0), (p.z + e) * K(18.0), 133) - noise(p.x * K(18.0), (p.z - e) * K(18.0), 133)) * K(0.18);
        n = unit(c3(nx, K(1.0), nz));
      }
      F3 albedo = surfaceColor(P, id, id >= 0 ? localP : p, localFp, C[8] + phase, C[11]);
      if (tree >= 0) {
        int ts = forestSeed(tree, ws);
        R hue = rnd(8, ts), age = rnd(9, ts), season = rnd(10, ts);
        R dark = K(0.76) + age * K(0.38);
        F3 tint = c3(K(0.88) + K(0.22) * hue, K(0.84) + K(0.28) * (K(1.0) - m_abs(hue - K(0.55))),
                     K(0.72) + K(0.25) * (K(1.0) - hue));
        if (leaf) tint = mix(tint, c3(K(1.12), K(0.88), K(0.48)), smooth(K(0.78), K(0.98), season) * K(0.38));
        else tint = mix(tint, c3(K(0.78), K(0.66), K(0.53)), season * K(0.18));
        albedo = albedo * tint * dark;
      }
      guide = albedo;
      int sample = (int)C[9];
      R r1 = rnd(pixel + sample * 1823, 991), r2 = rnd(pixel + sample * 971, 731);
      F3 su = basisU(sun), sv = cross(sun, su);
      R angle = r1 * K(6.28318), disk = m_sqrt(r2) * K(0.012);
      F3 light = unit(sun + su * (m_cos(angle) * disk) + sv * (m_sin(angle) * disk));
      R offset = id >= 0 ? K(0.0014) : K(0.004);
      F3 visibility = c3(K(1.0), K(1.0), K(1.0));
      if (C[22] > K(0.5)) visibility = lightVisibility(P, Nodes, Order, p + light * offset, light, C[8], C[11], ws);
      R nl = sat(dot(n, light)), back = sat(-dot(n, light));
      F3 direct = albedo * (nl / pi);
      if (leaf) {
        R vein = leafVein(P, id, localP, localFp, C[8] + phase, C[11]);
        direct = direct + albedo * (back * (K(0.75) - K(0.25) * vein) + m_pow(sat(dot(rd, light)), K(9.0)) * K(0.12));
      }
      R spec = specGGX(n, rd * K(-1.0), light, leaf ? K(0.39) : K(0.82));
      direct = direct + c3(spec, spec, spec);
      R ao = K(1.0);
      F3 ambient = c3(K(0.30), K(0.37), K(0.45)) * (K(0.45) + K(0.55) * sat(n.y)) * ao +
                   c3(K(0.08), K(0.068), K(0.035)) * sat(-n.y);
      if (leaf) ambient = ambient + c3(K(0.04), K(0.055), K(0.022)) * ao;
      col = direct * visibility * c3(K(4.8), K(4.2), K(3.1)) + albedo * ambient;
      R haze = K(1.0) - m_exp(-t * K(0.0018));
      col = mix(col, sky(rd, sun), haze);
      if (C[15] == K(1.0)) {
        R w0 = bandWeight(K(0.065), localFp), w1 = bandWeight(K(0.005), localFp), w2 = bandWeight(K(0.001), localFp);
        col = c3(w0 * K(0.80), w1 * K(0.85), w2 * K(0.9));
      }
      if (C[15] == K(2.0)) {
        R h = rnd(id + 177 + (tree + 1) * 71, 32);
        col = id < 0 ? c3(K(0.10), K(0.10), K(0.10))
                     : (leaf ? c3(K(0.1), K(0.3) + h * K(0.5), K(0.17)) : c3(K(0.47) + h * K(0.3), K(0.25), K(0.10)));
      }
      if (C[15] == K(3.0)) col = n * K(0.5) + c3(K(0.5), K(0.5), K(0.5));
    }
    st3(Linear, b, col); Linear[b + 3] = K(1.0);
    st3(Guide, b, guide); Guide[b + 3] = K(1.0);
  }
  // One pixel of filterLighting (Arbor.cu:292-302).
  static void filterPixel(const R* Linear, const R* Guide, const R* Hit, R* Filtered, const R* C, int width, int height,
                          int x, int y) {
    int b = (y * width + x) * 4;
    F3 col = ld3(Linear, b);
    F3 base = ld3(Guide, b) + c3(K(0.02), K(0.02), K(0.02));
    F3 sum = c3(K(0.0), K(0.0), K(0.0));
    R weight = K(0.0);
    int id = (int)Hit[b + 1];
    R depth = Hit[b];
    if (depth < FAR() && C[15] < K(0.5)) {
      for (int j = -2; j <= 2; j++)
        for (int k = -2; k <= 2; k++) {
          int xx = x + k, yy = y + j;
          if (xx < 0 || yy < 0 || xx >= width || yy >= height) continue;
          int q = (yy * width + xx) * 4;
          if ((int)Hit[q + 1] != id || (int)Hit[q + 2] != (int)Hit[b + 2] ||
              m_abs(Hit[q] - depth) > m_max(K(0.005), Hit[b + 3] * K(6.0)))
            continue;
          R w = m_exp(-R(j * j + k * k) * K(0.30));
          F3 c = ld3(Linear, q);
          F3 g = ld3(Guide, q) + c3(K(0.02), K(0.02), K(0.02));
          sum = sum + c3(c.x / g.x, c.y / g.y, c.z / g.z) * w;
          weight += w;
        }
      if (weight > K(0.0)) col = mix(col, (sum / weight) * base, K(0.86));
    }
    st3(Filtered, b, col); Filtered[b + 3] = K(1.0);
  }
  static R aces(R x) { return sat((x * (K(2.51) * x + K(0.03))) / (x * (K(2.43) * x + K(0.59)) + K(0.14))); }
  // One pixel of resolvePixels (Arbor.cu:304).
  static void resolvePixel(const R* Filtered, R* History, unsigned* Pixels, const R* C, int width, int height, int x,
                           int y) {
    int i = y * width + x, b = i * 4;
    F3 current = ld3(Filtered, b);
    R count = C[10] > K(0.5) ? K(0.0) : C[9];
    F3 mean = count < K(0.5) ? current : mix(ld3(History, b), current, K(1.0) / (count + K(1.0)));
    st3(History, b, mean);
    History[b + 3] = count + K(1.0);
    F3 col = mean * C[20];
    col = c3(aces(col.x), aces(col.y), aces(col.z));
    R g = K(1.0) / K(2.2);
    col = c3(m_pow(col.x, g), m_pow(col.y, g), m_pow(col.z, g));
    R u = (R(x) / R(width) - K(0.5)) * K(2.0), v = (R(y) / R(height) - K(0.5)) * K(2.0);
    col = col * (K(1.0) - K(0.09) * (u * u + v * v));
    Pixels[i] = packRGB(col);
  }
};

#undef K
}  // namespace orc
