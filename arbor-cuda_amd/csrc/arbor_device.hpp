// Device-side arithmetic of the ARBOR render path for gfx950.
//
// Every function here evaluates the reference formula (Arbor.cu, cited per
// function) in the same floating-point operation order, because bit-exact keys,
// permutation and node boxes require identical rounding; the library is built
// with -ffp-contract=off so no multiply-add is fused. What differs from the
// reference is the execution structure around it (stackless heap traversal,
// per-block forest tables, SoA/float4 access), documented where it occurs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "arbor.h"
#include "arbor_math.hpp"

namespace arbor {

#define AD __device__ __forceinline__

constexpr float kFar = 100000.0f;
constexpr float kPi = 3.14159265359f;
constexpr int kCapacity = ARBOR_CAPACITY;
constexpr int kPS = ARBOR_PS;
constexpr int kCS = ARBOR_CS;
constexpr int kObjects = ARBOR_OBJECT_COUNT;
constexpr int kWood = ARBOR_WOOD_COUNT;
constexpr int kLeaves = ARBOR_LEAF_COUNT;
constexpr int kSide = ARBOR_FOREST_SIDE;
constexpr int kTrees = ARBOR_FOREST_COUNT;
constexpr float kSpacing = 28.0f;

struct v3 { float x, y, z; };
struct v2 { float x, y; };
AD v3 mk(float x, float y, float z) { return {x, y, z}; }
AD v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
AD v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
AD v3 operator*(v3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
AD v3 operator*(v3 a, v3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
AD v3 operator/(v3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }

// ---- scalar helpers (Arbor.cu:17-31) ----
AD float clampf(float x, float a, float b) { return fminf(b, fmaxf(a, x)); }
AD float sat(float x) { return clampf(x, 0.0f, 1.0f); }
AD float lerp(float a, float b, float t) { return a + (b - a) * t; }
AD float frac(float a) { return a - floorf(a); }
AD float smooth(float a, float b, float x) {
  float t = sat((x - a) / (b - a));
  return t * t * (3.0f - 2.0f * t);
}
AD float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
AD v3 cross(v3 a, v3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
AD float length(v3 a) { return sqrtf(dot(a, a)); }
AD v3 unit(v3 a) { return a / fmaxf(0.0000001f, length(a)); }
AD v3 vmin(v3 a, v3 b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
AD v3 vmax(v3 a, v3 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
AD v3 mix(v3 a, v3 b, float t) { return a + (b - a) * t; }
AD v3 ld3(const float* a, int i) { return {a[i], a[i + 1], a[i + 2]}; }
AD void st3(float* a, int i, v3 v) { a[i] = v.x; a[i + 1] = v.y; a[i + 2] = v.z; }

// Correctly rounded single-precision transcendentals for the one-time build and
// per-block tables: evaluated in double and rounded once, which agrees with
// glibc's float routines wherever those are correctly rounded.
AD float sin_cr(float x) { return (float)sin((double)x); }
AD float cos_cr(float x) { return (float)cos((double)x); }
AD float pow_cr(float x, float y) { return (float)pow((double)x, (double)y); }

// ---- integer hash and value noise (Arbor.cu:32-36) ----
AD uint32_t hashU(uint32_t x) {
  x ^= x >> 16; x *= 2146121005u; x ^= x >> 15; x *= 2221713035u; x ^= x >> 16;
  return x;
}
AD float rnd(int n, int seed) {
  return (float)(hashU((uint32_t)n ^ hashU((uint32_t)seed)) & 16777215u) / 16777216.0f;
}
AD int wmul(int a, int b) { return (int)((uint32_t)a * (uint32_t)b); }
AD int wadd(int a, int b) { return (int)((uint32_t)a + (uint32_t)b); }
AD float hash2(int x, int y, int seed) { return rnd(wadd(wmul(x, 1973), wmul(y, 9277)), seed); }
AD float noise(float x, float y, int seed) {
  int ix = (int)floorf(x), iy = (int)floorf(y);
  float fx = frac(x), fy = frac(y);
  fx = fx * fx * (3.0f - 2.0f * fx);
  fy = fy * fy * (3.0f - 2.0f * fy);
  return lerp(lerp(hash2(ix, iy, seed), hash2(ix + 1, iy, seed), fx),
              lerp(hash2(ix, iy + 1, seed), hash2(ix + 1, iy + 1, seed), fx), fy);
}
AD float noise3(v3 p, int seed) {
  int iz = (int)floorf(p.z);
  float f = frac(p.z);
  f = f * f * (3.0f - 2.0f * f);
  return lerp(noise(p.x, p.y, wadd(seed, wmul(iz, 131))), noise(p.x, p.y, wadd(seed, wmul(iz + 1, 131))), f);
}

// ---- geometry helpers (Arbor.cu:37-44) ----
AD float invSafe(float x) { return fabsf(x) < 0.00000001f ? (x < 0.0f ? -100000000.0f : 100000000.0f) : 1.0f / x; }
AD v3 invDir(v3 d) { return {invSafe(d.x), invSafe(d.y), invSafe(d.z)}; }
// boxSpan with the ray's reciprocal hoisted (same values as Arbor.cu:38).
AD v2 boxSpanInv(v3 ro, v3 inv, v3 lo, v3 hi) {
  v3 a = (lo - ro) * inv, b = (hi - ro) * inv;
  v3 mn = vmin(a, b), mx = vmax(a, b);
  return {fmaxf(mn.x, fmaxf(mn.y, mn.z)), fminf(mx.x, fminf(mx.y, mx.z))};
}
AD v2 boxSpan(v3 ro, v3 rd, v3 lo, v3 hi) { return boxSpanInv(ro, invDir(rd), lo, hi); }
// The BVH walk's slab test: each plane distance (p - o) / d as one fused
// multiply-add p * (1/d) - o * (1/d), with roi = o * (1/d) computed once per
// tree entry. Equal to boxSpanInv's value up to rounding (a node box face
// within ~1e-5 relative of the ray can be entered or missed differently);
// every walk of the library uses it, so they all test the same nodes.
AD v2 slabSpan(v3 roi, v3 inv, v3 lo, v3 hi) {
  v3 a = mk(fmaf(lo.x, inv.x, -roi.x), fmaf(lo.y, inv.y, -roi.y), fmaf(lo.z, inv.z, -roi.z));
  v3 b = mk(fmaf(hi.x, inv.x, -roi.x), fmaf(hi.y, inv.y, -roi.y), fmaf(hi.z, inv.z, -roi.z));
  v3 mn = vmin(a, b), mx = vmax(a, b);
  return {fmaxf(mn.x, fmaxf(mn.y, mn.z)), fminf(mx.x, fminf(mx.y, mx.z))};
}
AD v3 basisU(v3 n) { return unit(cross(fabsf(n.y) < 0.91f ? mk(0.0f, 1.0f, 0.0f) : mk(0.0f, 0.0f, 1.0f), n)); }
AD v3 curveAt(const float* B, int id, float t) {
  int b = id * kCS;
  float q = 1.0f - t;
  return ld3(B, b) * (q * q * q) + ld3(B, b + 3) * (3.0f * q * q * t) + ld3(B, b + 6) * (3.0f * q * t * t) +
         ld3(B, b + 9) * (t * t * t);
}
AD v3 curveTangent(const float* B, int id, float t) {
  int b = id * kCS;
  float q = 1.0f - t;
  return unit((ld3(B, b + 3) - ld3(B, b)) * (q * q) + (ld3(B, b + 6) - ld3(B, b + 3)) * (2.0f * q * t) +
              (ld3(B, b + 9) - ld3(B, b + 6)) * (t * t));
}
AD float curveRadius(const float* B, int id, float t) { return lerp(B[id * kCS + 12], B[id * kCS + 13], t); }
AD float bandWeight(float period, float footprint) {
  return 1.0f - smooth(period * 0.12f, period * 0.65f, footprint);
}

// ---- camera helpers (Arbor.cu:45-51) ----
AD v3 forwardC(const float* C) { return unit(ld3(C, 4) - ld3(C, 0)); }
AD v3 rightC(const float* C) { return unit(cross(forwardC(C), mk(0.0f, 1.0f, 0.0f))); }
AD uint32_t packRGB(v3 c) {
  uint32_t r = (uint32_t)(sat(c.x) * 255.0f + 0.5f);
  uint32_t g = (uint32_t)(sat(c.y) * 255.0f + 0.5f);
  uint32_t b = (uint32_t)(sat(c.z) * 255.0f + 0.5f);
  return r | (g << 8) | (b << 16) | 4278190080u;
}
AD float halton(int n, int base) {
  float f = 1.0f, r = 0.0f;
  for (int j = 0; j < 12; j++) {
    if (n <= 0) break;
    f /= (float)base;
    r += f * (float)(n % base);
    n /= base;
  }
  return r;
}

// Per-view camera frame, computed once per block instead of per pixel (the
// values are the same as cameraRay's, Arbor.cu:51).
struct View {
  v3 eye, f, r, up;
  float jx, jy, scale;
};
AD View makeView(const float* C) {
  View v;
  int sample = (int)C[9];
  v.jx = halton(sample + 1, 2) - 0.5f;
  v.jy = halton(sample + 1, 3) - 0.5f;
  v.eye = ld3(C, 0);
  v.f = forwardC(C);
  v.r = rightC(C);
  v.up = cross(v.r, v.f);
  v.scale = C[14];
  return v;
}
AD v3 viewRay(const View& v, int x, int y, int width, int height) {
  float u = ((float)x + 0.5f + v.jx - (float)width * 0.5f) / (float)height;
  float w = ((float)height * 0.5f - (float)y - 0.5f - v.jy) / (float)height;
  return unit(v.f + v.r * (u * v.scale) + v.up * (w * v.scale));
}

// ---- forest instancing (Arbor.cu:57-69) ----
AD int forestSeed(int tree, int ws) {
  return (int)(hashU((uint32_t)wadd(wadd(wmul(tree, 92821), wmul(ws, 68917)), 17)) & 1048575u);
}
AD v3 forestPos(int tree, int ws) {
  int x = tree % kSide - kSide / 2, z = tree / kSide - kSide / 2;
  int s = forestSeed(tree, ws);
  float jx = (rnd(1, s) - 0.5f) * 10.0f;
  float jz = (rnd(2, s) - 0.5f) * 10.0f;
  float warpX = (noise((float)x * 0.19f, (float)z * 0.19f, ws + 311) - 0.5f) * 3.0f;
  float warpZ = (noise((float)x * 0.19f + 17.0f, (float)z * 0.19f - 9.0f, ws + 733) - 0.5f) * 3.0f;
  return mk((float)x * kSpacing + jx + warpX, 0.0f, (float)z * kSpacing + jz + warpZ);
}
AD float forestScale(int tree, int ws) { return 0.72f + rnd(3, forestSeed(tree, ws)) * 0.62f; }
AD float forestYaw(int tree, int ws) { return rnd(4, forestSeed(tree, ws)) * kPi * 2.0f; }

// One tree's instancing values. Computed once per block into LDS; every field
// is the value the reference recomputes per DDA cell / per shaded pixel.
struct TreeInst {
  float px, pz;    // forestPos (y = 0)
  float sc;        // forestScale
  float cl, sl;    // cos/sin(-yaw): world -> local
  float cw, sw;    // cos/sin(+yaw): local normal -> world
  float phase;     // (forestSeed & 8191) * 0.0017
  int seed;        // forestSeed
};
AD TreeInst makeTree(int tree, int ws) {
  TreeInst t;
  v3 p = forestPos(tree, ws);
  t.px = p.x;
  t.pz = p.z;
  t.sc = forestScale(tree, ws);
  float yaw = forestYaw(tree, ws);
  float a = -yaw;
  t.cl = cos_cr(a);
  t.sl = sin_cr(a);
  t.cw = cos_cr(yaw);
  t.sw = sin_cr(yaw);
  t.seed = forestSeed(tree, ws);
  t.phase = (float)(t.seed & 8191) * 0.0017f;
  return t;
}
// The same records in device memory for the kernels that keep LDS for their
// traversal stacks: three float4 per tree, {px, pz, sc, cl} {sl, cw, sw,
// phase} {seed (int bits), 0, 0, 0}, written by the frame prologue.
constexpr int kTreeRecF4 = 3;
struct TreeRecs {
  const float4* rec;
};
// What a DDA cell's broad test reads ({px, pz, sc} and cl).
struct TreeHead {
  float px, pz, sc, cl;
};
AD TreeHead treeHead(const TreeInst* t, int i) { return {t[i].px, t[i].pz, t[i].sc, t[i].cl}; }
AD TreeInst treeRecord(const TreeInst* t, int i, const TreeHead&) { return t[i]; }
AD TreeHead treeHead(TreeRecs t, int i) {
  const float4 a = t.rec[kTreeRecF4 * i];
  return {a.x, a.y, a.z, a.w};
}
AD TreeInst treeRecord(TreeRecs t, int i, const TreeHead& h) {
  const float4 b = t.rec[kTreeRecF4 * i + 1];
  TreeInst r;
  r.px = h.px;
  r.pz = h.pz;
  r.sc = h.sc;
  r.cl = h.cl;
  r.sl = b.x;
  r.cw = b.y;
  r.sw = b.z;
  r.phase = b.w;
  r.seed = __float_as_int(t.rec[kTreeRecF4 * i + 2].x);
  return r;
}
AD void storeTreeRec(float4* rec, int i, const TreeInst& t) {
  rec[kTreeRecF4 * i] = make_float4(t.px, t.pz, t.sc, t.cl);
  rec[kTreeRecF4 * i + 1] = make_float4(t.sl, t.cw, t.sw, t.phase);
  rec[kTreeRecF4 * i + 2] = make_float4(__int_as_float(t.seed), 0.0f, 0.0f, 0.0f);
}
AD v3 toLocalPoint(v3 p, const TreeInst& t) {
  v3 q = p - mk(t.px, 0.0f, t.pz);
  return mk(q.x * t.cl - q.z * t.sl, q.y, q.x * t.sl + q.z * t.cl) / t.sc;
}
AD v3 toLocalDir(v3 d, const TreeInst& t) { return unit(mk(d.x * t.cl - d.z * t.sl, d.y, d.x * t.sl + d.z * t.cl)); }
AD v3 toWorldNormal(v3 n, const TreeInst& t) { return unit(mk(n.x * t.cw - n.z * t.sw, n.y, n.x * t.sw + n.z * t.cw)); }

// ---- primitive bounds (Arbor.cu:132-133) ----
AD v3 boundLow(const float* P, int id) {
  int b = id * kPS, type = (int)P[b + 3];
  if (type < 0) return mk(kFar, kFar, kFar);
  v3 a = ld3(P, b);
  if (type == 0) {
    float r = fmaxf(P[b + 7], P[b + 11]) + 0.002f;
    return vmin(a, ld3(P, b + 4)) - mk(r, r, r);
  }
  if (type == 2) return a - ld3(P, b + 4);
  float radius = P[b + 7] + P[b + 16] + 0.025f;
  return a - mk(radius, radius, radius);
}
AD v3 boundHigh(const float* P, int id) {
  int b = id * kPS, type = (int)P[b + 3];
  if (type < 0) return mk(-kFar, -kFar, -kFar);
  v3 a = ld3(P, b);
  if (type == 0) {
    float r = fmaxf(P[b + 7], P[b + 11]) + 0.002f;
    return vmax(a, ld3(P, b + 4)) + mk(r, r, r);
  }
  if (type == 2) return a + ld3(P, b + 4);
  float radius = P[b + 7] + P[b + 16] + 0.025f;
  return a + mk(radius, radius, radius);
}

// ---- surface fields (Arbor.cu:148-174) ----
AD v2 woodUV(const float* P, int id, v3 p) {
  int b = id * kPS;
  v3 a = ld3(P, b);
  v3 d = unit(ld3(P, b + 4) - a);
  v3 initial = ld3(P, b + 8);
  v3 u = unit(initial - d * dot(initial, d));
  v3 v = cross(d, u);
  v3 q = p - a;
  float z = dot(q, d);
  float theta = atan2f(dot(q, v), dot(q, u));
  return {theta, P[b + 16] + z};
}
AD float windAngle(float seedField, float time, float wind) {
  float phase = (float)(((int)seedField) % 8192) * 0.618f;
  return wind * (0.11f * sinBounded(time * 2.1f + phase) + 0.045f * sinBounded(time * 5.8f + phase * 1.7f));
}
AD v3 leafU(const float* P, int id, float time, float wind) {
  int b = id * kPS;
  SinCos a = sincosBounded(windAngle(P[b + 15], time, wind));
  return ld3(P, b + 4) * a.c + ld3(P, b + 12) * a.s;
}
AD v3 leafN(const float* P, int id, float time, float wind) {
  int b = id * kPS;
  SinCos a = sincosBounded(windAngle(P[b + 15], time, wind));
  return ld3(P, b + 12) * a.c - ld3(P, b + 4) * a.s;
}
AD float leafWidth(float u, float len, float halfWidth, float footprint) {
  float q = sat(u / len);
  float s = sinBounded(kPi * q);
  float lobes = 0.77f + 0.23f * cosBounded(q * kPi * 10.0f + 0.45f);
  // with a zero band weight the fine lobes are a signed zero and
  // lobes + (+-0) = lobes (lobes >= 0.54): the sine is not evaluated
  const float bw = bandWeight(0.002f, footprint);
  float fine = bw == 0.0f ? 0.0f : bw * 0.015f * sinBounded(q * kPi * 160.0f);
  return halfWidth * powf(fmaxf(0.0f, s), 0.72f) * (lobes + fine);
}
AD float leafHeight(float u, float v, float len, float halfWidth, float curl, float twist, float footprint) {
  float q = clampf(u / len, -0.2f, 1.2f);
  float s = sinBounded(kPi * q);
  float coarse = curl * s + twist * (q - 0.4f) * v + 0.009f * (v * v / (halfWidth * halfWidth)) * s;
  const float bw = bandWeight(0.007f, footprint);  // 0: (m * s) * 0 is s * 0 for any m > 0
  float m = bw == 0.0f ? 1.0f : 0.0012f * expf(-fabsf(v) / 0.005f);
  float mid = m * s * bw;
  return coarse + mid;
}
// leafHeight split into its u-only and v-only factors (same expressions; the
// divisions by the leaf's length and squared half-width use their reciprocals
// rlen = 1/len and rhw2 = 1/hw2 and give the same correctly rounded quotients).
struct LeafU { float s, tw; };  // sin(PI*q), twist*(q-0.4)
struct LeafV { float v, w2, m; };  // v, 0.009*(v*v/hw^2), 0.0012*exp(-|v|/0.005)
AD LeafU leafUTerms(float u, float len, float rlen, float twist) {
  float q = clampf(divKnown(u, len, rlen), -0.2f, 1.2f);
  return {sinBounded(kPi * q), twist * (q - 0.4f)};
}
// mid = false when the midrib's band weight is 0: the midrib term is then
// (m * s) * 0, a zero with the sign of s for any positive m, so m = 1 stands
// in for the exponential.
AD LeafV leafVTerms(float v, float hw2, float rhw2, bool mid = true) {
  return {v, 0.009f * divKnown(v * v, hw2, rhw2),
          mid ? 0.0012f * expf(-divKnown(fabsf(v), 0.005f, 1.0f / 0.005f)) : 1.0f};
}
AD float leafHeightOf(LeafU a, LeafV b, float curl, float band) {
  float coarse = curl * a.s + a.tw * b.v + b.w2 * a.s;
  float mid = b.m * a.s * band;
  return coarse + mid;
}

// ---- intersection (Arbor.cu:186-219) ----
AD float sphereT(v3 ro, v3 rd, v3 p, float radius) {
  v3 q = ro - p;
  float b = dot(q, rd);
  float c = dot(q, q) - radius * radius;
  float d = b * b - c;
  if (d < 0.0f) return kFar;
  float t = -b - sqrtf(d);
  if (t <= 0.0001f) t = -b + sqrtf(d);
  return t > 0.0001f ? t : kFar;
}
}  // namespace arbor
