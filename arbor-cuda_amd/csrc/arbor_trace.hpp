// Ray queries of the ARBOR render path on gfx950: primitive records, surface
// fields, exact intersections, the stackless heap-BVH walk and the forest DDA.
//
// Arithmetic follows Arbor.cu expression by expression (cited per function).
// Structure differs: a primitive's 20 used floats are loaded once as five
// float4 (the reference re-reads P inside every helper), the ray reciprocal is
// hoisted out of the box tests, and the traversal stack is a 16-bit mask over
// heap depths instead of stack[32] (see traceTree below).
#pragma once
#include "arbor_device.hpp"

namespace arbor {

// One primitive record, floats 0..19 of its 24 (ARCHITECTURE.md:13).
struct Prim {
  float4 q0, q1, q2, q3, q4;
  AD v3 a() const { return mk(q0.x, q0.y, q0.z); }   // wood start / leaf anchor / stone centre
  AD float type() const { return q0.w; }
  AD v3 b() const { return mk(q1.x, q1.y, q1.z); }   // wood end / leaf u / stone half-axes
  AD float ra() const { return q1.w; }                // wood radius at a / leaf length
  AD v3 c() const { return mk(q2.x, q2.y, q2.z); }   // wood transported axis / leaf v
  AD float rb() const { return q2.w; }                // wood radius at b / leaf half-width
  AD v3 n() const { return mk(q3.x, q3.y, q3.z); }   // leaf normal
  AD float seed() const { return q3.w; }
  AD float f16() const { return q4.x; }               // wood arc offset / leaf petiole
  AD float f17() const { return q4.y; }               // wood curve id / leaf camber
  AD float f18() const { return q4.z; }               // wood curve radius / leaf twist
};
AD Prim loadPrim(const float* P, int id) {
  const float4* r = reinterpret_cast<const float4*>(P + id * kPS);
  Prim p;
  p.q0 = r[0]; p.q1 = r[1]; p.q2 = r[2]; p.q3 = r[3]; p.q4 = r[4];
  return p;
}
// boundLow/boundHigh (Arbor.cu:132-133) from a loaded record.
AD void primBounds(const Prim& p, v3& lo, v3& hi) {
  int type = (int)p.type();
  if (type < 0) { lo = mk(kFar, kFar, kFar); hi = mk(-kFar, -kFar, -kFar); return; }
  v3 a = p.a();
  if (type == 0) {
    float r = fmaxf(p.ra(), p.rb()) + 0.002f;
    lo = vmin(a, p.b()) - mk(r, r, r);
    hi = vmax(a, p.b()) + mk(r, r, r);
    return;
  }
  if (type == 2) { lo = a - p.b(); hi = a + p.b(); return; }
  float radius = p.ra() + p.f16() + 0.025f;
  lo = a - mk(radius, radius, radius);
  hi = a + mk(radius, radius, radius);
}

// ---- wood surface (Arbor.cu:148-170) ----
// The terms of woodField, barkCut and woodUV that depend on the primitive and
// the footprint but not on the sample point, each computed by the same
// operations as in those functions. A march or a gradient evaluates many
// points of one primitive with one frame.
struct WoodFrame {
  v3 a, ab;          // a, b - a
  float fieldLen2, rField;  // fmaxf(1e-7, dot(ab, ab)) and its reciprocal: woodField's segment parameter
  float cutLen2, rCut;      // fmaxf(1e-6, dot(ab, ab)): barkCut's
  float ra, rb, arc0;
  v3 d, u, v;        // woodUV's axis frame
  int columns;
  float rColumns;  // 1 / columns
  uint32_t seedHash, seedHash31;  // hashU(seed), hashU(seed + 31) of the Voronoi hashes
  int seed;
  float fine, grain, footprint;
};
AD WoodFrame woodFrame(const Prim& pr, float footprint) {
  WoodFrame f;
  f.a = pr.a();
  f.ab = pr.b() - f.a;
  float len2 = dot(f.ab, f.ab);
  f.fieldLen2 = fmaxf(0.0000001f, len2);
  f.cutLen2 = fmaxf(0.000001f, len2);
  f.rField = 1.0f / f.fieldLen2;
  f.rCut = 1.0f / f.cutLen2;
  f.ra = pr.ra();
  f.rb = pr.rb();
  f.arc0 = pr.f16();
  f.d = unit(f.ab);
  v3 initial = pr.c();
  f.u = unit(initial - f.d * dot(initial, f.d));
  f.v = cross(f.d, f.u);
  f.columns = (int)fmaxf(7.0f, floorf(pr.f18() * 2.0f * kPi / 0.065f + 0.5f));
  f.rColumns = 1.0f / (float)f.columns;
  f.seed = (int)pr.seed();
  f.seedHash = hashU((uint32_t)f.seed);
  f.seedHash31 = hashU((uint32_t)(f.seed + 31));
  f.fine = bandWeight(0.010f, footprint);
  f.grain = bandWeight(0.0025f, footprint);
  f.footprint = footprint;
  return f;
}
// hash2(x, y, seed) given hashU(seed)
AD float hash2h(int x, int y, uint32_t seedHash) {
  return (float)(hashU((uint32_t)wadd(wmul(x, 1973), wmul(y, 9277)) ^ seedHash) & 16777215u) / 16777216.0f;
}
AD v2 woodUV(const WoodFrame& f, v3 p) {
  v3 q = p - f.a;
  float z = dot(q, f.d);
  float theta = atan2f(dot(q, f.v), dot(q, f.u));
  return {theta, f.arc0 + z};
}
__device__ inline float barkCut(const WoodFrame& f, v3 p) {
  float along = sat(divKnown(dot(p - f.a, f.ab), f.cutLen2, f.rCut));
  float radius = lerp(f.ra, f.rb, along);
  float amp = fminf(0.021f, radius * 0.055f);
  const int columns = f.columns;
  float period = divKnown(radius * 2.0f * kPi, (float)columns, f.rColumns);
  float coarse = bandWeight(period, f.footprint);
  if (coarse < 0.000001f && f.fine < 0.000001f) return amp * 0.32f;
  v2 uv = woodUV(f, p);
  float theta = uv.x, arc = uv.y;
  const SinCos tc = sincosBounded(theta);
  float st = tc.s, ct = tc.c;
  // A term whose band weight is 0 is a signed zero, and the running sum
  // (from 0.32, never -0) is unchanged by adding it: the plates (the cell
  // coordinates and the 3x3 Voronoi search) and the two noises are only
  // evaluated under a nonzero band weight.
  float sum = 0.32f;
  if (coarse != 0.0f) {
    float x = (divKnown(theta, 2.0f * kPi, 1.0f / (2.0f * kPi)) + 0.5f) * (float)columns +
              0.32f * sinBounded(arc * 6.7f + ct * 3.0f) + 0.15f * sinBounded(arc * 19.0f + st * 2.0f);
    float y = divKnown(arc, 0.19f, 1.0f / 0.19f) + 0.25f * sinBounded(theta * 7.0f) + 0.1f * sinBounded(theta * 19.0f);
    int ix = (int)floorf(x), iy = (int)floorf(y);
    // the three columns ix-1, ix, ix+1 wrapped into [0, columns)
    int wrap[3];
    wrap[0] = (((ix - 1) % columns) + columns) % columns;
    wrap[1] = wrap[0] + 1 == columns ? 0 : wrap[0] + 1;
    wrap[2] = wrap[1] + 1 == columns ? 0 : wrap[1] + 1;
    float first = 100.0f, second = 100.0f;
#pragma unroll
    for (int j = -1; j <= 1; j++)
#pragma unroll
      for (int k = -1; k <= 1; k++) {
        int cell = ix + k;
        int wrapped = wrap[k + 1];
        float xx = (float)cell + 0.18f + 0.64f * hash2h(wrapped, iy + j, f.seedHash) - x;
        float yy = (float)(iy + j) + 0.18f + 0.64f * hash2h(wrapped + columns, iy + j, f.seedHash31) - y;
        float dd = xx * xx + yy * yy;
        if (dd < first) { second = first; first = dd; }
        else if (dd < second) second = dd;
      }
    float gap = sqrtf(second) - sqrtf(first);
    float fissure = 1.0f - smooth(0.015f, 0.17f, gap);
    float plates = fissure * fissure;
    sum = sum + coarse * (plates - 0.37f) * 0.69f;
  }
  v3 coord = mk(ct * 9.0f, st * 9.0f, arc * 23.0f);
  if (f.fine != 0.0f) sum = sum + f.fine * (noise3(coord, f.seed) - 0.5f) * 0.24f;
  if (f.grain != 0.0f) sum = sum + f.grain * (noise3(coord * 4.0f, f.seed + 731) - 0.5f) * 0.08f;
  return amp * clampf(sum, 0.005f, 1.0f);
}
__device__ inline float woodField(const WoodFrame& f, v3 p) {
  float t = sat(divKnown(dot(p - f.a, f.ab), f.fieldLen2, f.rField));
  float radius = lerp(f.ra, f.rb, t);
  return length(p - f.a - f.ab * t) - radius + barkCut(f, p);
}
AD v2 woodUV(const Prim& pr, v3 p) { return woodUV(woodFrame(pr, 1.0f), p); }
__device__ inline float barkCut(const Prim& pr, v3 p, float footprint) { return barkCut(woodFrame(pr, footprint), p); }

// ---- leaf frame (Arbor.cu:171-172) ----
struct LeafFrame { v3 u, n; };
AD LeafFrame leafFrame(const Prim& pr, float time, float wind) {
  LeafFrame f;
  // Breeze off: the rotation angle is wind*(...) = 0, cos = 1 and sin = 0, so
  // the rotated basis equals the stored one.
  if (wind == 0.0f) { f.u = pr.b(); f.n = pr.n(); return f; }
  const SinCos w = sincosBounded(windAngle(pr.seed(), time, wind));
  float ca = w.c, sa = w.s;
  f.u = pr.b() * ca + pr.n() * sa;
  f.n = pr.n() * ca - pr.b() * sa;
  return f;
}

// Conservative rejection: true when the ray's line passes farther than the
// padded radius from `centre`, or the whole padded sphere lies behind the
// origin. The perpendicular offset is formed as q - rd*(q.rd) (no
// cancellation of two large squares). The pad is 1 mm plus padPerMetre*|q|:
// sphereT and tubeT decide hits from discriminants that cancel
// catastrophically far from a small object, so the reference reports hits up
// to ~8e-4*|q| outside a thin tube or sphere at distance |q|; tests ending in
// such a discriminant use 4e-3 per metre. The leaf slab (boxSpan) and the wood
// SDF march (which must get within 2.5e-4 m of the surface) are well
// conditioned and use 1 mm. A rejected object is one the exact tests miss.
AD bool sphereMiss(v3 ro, v3 rd, v3 centre, float radius, float padPerMetre) {
  v3 q = ro - centre;
  float b = dot(q, rd);
  v3 p = q - rd * b;
  float r = radius + 0.001f + padPerMetre * length(q);
  return dot(p, p) > r * r || b > r;
}
constexpr float kDiscriminantPad = 0.004f;
constexpr float kSlabPad = 0.00001f;

// ---- exact intersections (Arbor.cu:188-219) ----
__device__ inline float tubeT(v3 ro, v3 rd, v3 a, v3 b, float ra, float rb) {
  v3 axis = b - a;
  float len = length(axis);
  axis = axis / fmaxf(len, 0.000001f);
  v3 q = ro - a;
  float z = dot(q, axis), dz = dot(rd, axis);
  float k = (rb - ra) / fmaxf(0.000001f, len);
  v3 v = q - axis * z, d = rd - axis * dz;
  float radius = ra + k * z;
  float aa = dot(d, d) - k * k * dz * dz;
  float bb = dot(v, d) - radius * k * dz;
  float cc = dot(v, v) - radius * radius;
  float best = fminf(sphereT(ro, rd, a, ra), sphereT(ro, rd, b, rb));
  float disc = bb * bb - aa * cc;
  if (disc >= 0.0f && fabsf(aa) > 0.0000001f) {
    float sq = sqrtf(disc);
    for (int j = 0; j < 2; j++) {
      float t = (-bb + (j == 0 ? -1.0f : 1.0f) * sq) / aa;
      float h = z + t * dz;
      if (t > 0.0001f && h >= 0.0f && h <= len) best = fminf(best, t);
    }
  }
  return best;
}
__device__ inline float woodT(const Prim& pr, v3 ro, v3 rd, float footprint, bool refine, bool cull, float best) {
  v3 a = pr.a(), end = pr.b();
  // Every surface woodT can return lies inside the radius-max(ra,rb) capsule
  // (bark only cuts inward); a ray missing its bounding sphere returns FAR.
  if (cull) {
    v3 half = (end - a) * 0.5f;
    float r = length(half) + fmaxf(pr.ra(), pr.rb());
    if (sphereMiss(ro, rd, a + half, r, refine ? 0.0f : kDiscriminantPad)) return kFar;
  }
  if (!refine) {
    float ra = pr.ra() - fminf(0.021f, pr.ra() * 0.055f) * 0.32f;
    float rb = pr.rb() - fminf(0.021f, pr.rb() * 0.055f) * 0.32f;
    return tubeT(ro, rd, a, end, ra, rb);
  }
  float t = tubeT(ro, rd, a, end, pr.ra(), pr.rb());
  if (t >= kFar) return kFar;
  // The march only moves t forward from the tube hit: from t >= best (at the
  // tube hit or at any later step) it can only return a value the caller
  // discards (it keeps strictly nearer hits).
  if (t >= best) return kFar;
  v3 lo, hi;
  primBounds(pr, lo, hi);
  v2 span = boxSpan(ro, rd, lo, hi);
  float eps = fmaxf(0.000025f, fminf(0.00025f, footprint * 0.13f));
  const WoodFrame wf = woodFrame(pr, footprint);
  for (int k = 0; k < 64; k++) {
    float f = woodField(wf, ro + rd * t);
    if (f < eps) return t;
    t += fmaxf(eps * 0.35f, f * 0.27f);
    if (t > span.y || t >= best) return kFar;
  }
  return kFar;
}
// The leaf's Newton solve and acceptance test (Arbor.cu:205-214) in its local
// frame (x along u from the petiole end, y along v, z along n; d* the ray
// direction there), from the clamped start t inside the slab span: seven
// steps of central-difference Newton on z + dz*t = leafHeight(x + dx*t,
// y + dy*t); returns t when the end point lies on the blade within 3e-4, else
// FAR. leafHeight(u, v) = curl*s(u) + tw(u)*v + w2(v)*s(u) + m(v)*s(u)*band
// with s, tw functions of u only and w2, m of v only, so the five evaluations
// of a step share 3 sines and 3 exponentials; each term is computed as
// leafHeight computes it, and the divisions by len, width^2, 0.005 and 2e by
// their reciprocals (same correctly rounded quotients).
struct LeafLocal {
  float x, y, z, dx, dy, dz, len, width, curl, twist, footprint;
};
__device__ inline float leafNewton(const LeafLocal& L, float t, float spanLo, float spanHi) {
  const float band = bandWeight(0.007f, L.footprint);
  const bool mid = band != 0.0f;
  const float hw2 = L.width * L.width;
  const float rlen = 1.0f / L.len, rhw2 = 1.0f / hw2;
  const float e = 0.00015f;
  const float twoE = 2.0f * e, rTwoE = 1.0f / (2.0f * e);
  // Each iteration is a pure function of t, so the solve stops once the rest
  // is known: an update that leaves t's bits unchanged repeats forever, and
  // one that brings t back to its value two updates earlier alternates
  // between the last two values through the 7th iteration, whose value the
  // parity of the remaining count picks. The derivative break cannot fire
  // later: every value the rest would visit has already passed it.
  float tPrev = t;
  for (int k = 0; k < 7; k++) {
    float xx = L.x + L.dx * t, yy = L.y + L.dy * t;
    LeafU u0 = leafUTerms(xx, L.len, rlen, L.twist), up = leafUTerms(xx + e, L.len, rlen, L.twist),
          um = leafUTerms(xx - e, L.len, rlen, L.twist);
    LeafV v0 = leafVTerms(yy, hw2, rhw2, mid), vp = leafVTerms(yy + e, hw2, rhw2, mid),
          vm = leafVTerms(yy - e, hw2, rhw2, mid);
    float height = leafHeightOf(u0, v0, L.curl, band);
    float f = L.z + L.dz * t - height;
    float hx = divKnown(leafHeightOf(up, v0, L.curl, band) - leafHeightOf(um, v0, L.curl, band), twoE, rTwoE);
    float hy = divKnown(leafHeightOf(u0, vp, L.curl, band) - leafHeightOf(u0, vm, L.curl, band), twoE, rTwoE);
    float deriv = L.dz - hx * L.dx - hy * L.dy;
    if (fabsf(deriv) < 0.00001f) break;
    const float tn = clampf(t - clampf(f / deriv, -0.13f, 0.13f), spanLo, spanHi);
    const uint32_t bn = __float_as_uint(tn);
    if (bn == __float_as_uint(t)) break;
    if (bn == __float_as_uint(tPrev)) {
      if (((6 - k) & 1) == 0) t = tn;
      break;
    }
    tPrev = t;
    t = tn;
  }
  float xx = L.x + L.dx * t, yy = L.y + L.dy * t;
  float err = fabsf(L.z + L.dz * t -
                    leafHeightOf(leafUTerms(xx, L.len, rlen, L.twist), leafVTerms(yy, hw2, rhw2, mid), L.curl, band));
  // same predicate as the reference, cheapest terms first
  if (t > 0.0001f && xx > 0.0f && xx < L.len && err < 0.0003f && fabsf(yy) < leafWidth(xx, L.len, L.width, L.footprint))
    return t;
  return kFar;
}
__device__ inline float leafT(const Prim& pr, v3 ro, v3 rd, float footprint, float time, float wind, bool cull,
                             float bestSoFar) {
  v3 origin = pr.a();
  LeafFrame fr = leafFrame(pr, time, wind);
  v3 u = fr.u, v = pr.c(), n = fr.n;
  float stem = pr.f16(), len = pr.ra(), width = pr.rb();
  float curl = pr.f17(), twist = pr.f18();
  // Separate rejections for the petiole capsule (radius <= 1.1 mm along u
  // over [0, stem]) and the Newton slab (u in [stem-0.001, stem+len+0.001],
  // |v| <= width, n in [-0.035, 0.065]); a ray that misses both returns FAR.
  bool petiole = true, slab = true;
  if (cull) {
    petiole = !sphereMiss(ro, rd, origin + u * (stem * 0.5f), stem * 0.5f + 0.0011f, kDiscriminantPad);
    float hx = len * 0.5f + 0.001f;
    slab = !sphereMiss(ro, rd, origin + u * (stem + len * 0.5f) + n * 0.015f,
                       sqrtf(hx * hx + width * width + 0.05f * 0.05f), kSlabPad);
    if (!petiole && !slab) return kFar;
  }
  v3 q = ro - origin;
  float x = dot(q, u) - stem, y = dot(q, v), z = dot(q, n);
  float dx = dot(rd, u), dy = dot(rd, v), dz = dot(rd, n);
  float best = petiole ? tubeT(ro, rd, origin, origin + u * stem, 0.0011f, 0.00065f) : kFar;
  if (!slab) return best;
  v2 span = boxSpan(mk(x, y, z), mk(dx, dy, dz), mk(-0.001f, -width, -0.035f), mk(len + 0.001f, width, 0.065f));
  if (span.y < fmaxf(0.0001f, span.x)) return best;
  // The Newton solution is clamped to [span.x, span.y]: from span.x >= the
  // caller's best it cannot be kept, and min(petiole, it) is decided by the
  // petiole alone.
  if (span.x >= bestSoFar) return best;
  float t = fabsf(dz) > 0.000001f ? -z / dz : (span.x + span.y) * 0.5f;
  t = clampf(t, fmaxf(0.0001f, span.x), span.y);
  LeafLocal L{x, y, z, dx, dy, dz, len, width, curl, twist, footprint};
  return fminf(best, leafNewton(L, t, span.x, span.y));
}
// `best` is the caller's current nearest hit: a result >= best is discarded
// by every caller, so a test may give up (returning FAR) once its answer is
// known to be >= best; any result < best is exact.
__device__ inline float objectT(const Prim& pr, v3 ro, v3 rd, float footprint, float time, float wind, bool refine,
                               bool cull, float best = kFar) {
  int type = (int)pr.type();
  if (type < 0) return kFar;
  if (type == 0) return woodT(pr, ro, rd, footprint, refine, cull, best);
  if (type == 1) return leafT(pr, ro, rd, footprint, time, wind, cull, best);
  v3 h = pr.b();
  v3 q = ro - pr.a();
  v3 p = mk(q.x / h.x, q.y / h.y, q.z / h.z);
  v3 d = mk(rd.x / h.x, rd.y / h.y, rd.z / h.z);
  float aa = dot(d, d), bb = dot(p, d), cc = dot(p, p) - 1.0f;
  float disc = bb * bb - aa * cc;
  if (disc < 0.0f) return kFar;
  float t = (-bb - sqrtf(disc)) / aa;
  return t > 0.0001f ? t : kFar;
}

// ---- heap-BVH walk (Arbor.cu:220-230) ----
// Node n's record is two float4 {lo, count} {hi, id}; children 2n, 2n+1 are
// adjacent, so an interior visit reads one 64-byte pair.
struct NodeBox { float4 lo, hi; };
// Evaluated without short-circuit branches (same truth value): the three
// tests are cheap, and a branch between them splits the slab test into two
// blocks, where the compiler re-canonicalises its operands for the max.
AD bool spanHit(v2 s, float count, float best) {
  return (count > 0.0f) & (s.y >= fmaxf(0.0001f, s.x)) & (s.x < best);
}
AD int heapDepth(uint32_t node) { return 31 - __builtin_clz(node); }

// Compile-time configuration of the walk. Every configuration visits the same
// nodes in the same order and returns the same (best, id); they differ in how
// the data reaches the lanes.
//   kCull:     bounding-sphere rejections inside woodT / leafT;
//   kPrefetch: while testing a node's children, load its four grandchildren
//              (the next step's children whichever child is taken), so each
//              dependent fetch overlaps one step of box tests;
//   kLdsTop:   heap levels 0-7 (nodes 0..255, 8 KiB) are copied into LDS once
//              per block and read from there.
//   kSah:      the tree is walked over the SAH hierarchy of traceTreeSah
//              below instead of the heap; the walk's `top` pointer is then
//              that hierarchy (device memory).
//   kLdsStack: with kSah, the first kLdsStack entries of each lane's stack
//              live in the block's dynamic LDS (the launch provides
//              waves x kLdsStack x 64 x 8 bytes), the rest in scratch.
//   kGather:   with kSah and kGather = g > 0, lanes keep stepping through
//              interior nodes until at least g/8 of the wave's live lanes
//              wait at a leaf (or none has an interior step left); the object
//              tests of the waiting lanes then run together. Each lane
//              performs the same steps in the same order as with g = 0; only
//              their interleaving within the wave changes.
//   kSteps:    with kGather, the interior steps a stepping lane takes per
//              round of the gathered loop (each round ends in two ballots).
template <bool kCull, bool kPrefetch, bool kLdsTop, bool kSah = false, int kLdsStack = 0, int kGather = 0,
          int kSteps = 1>
struct Walk {
  static constexpr bool cull = kCull, prefetch = kPrefetch, ldsTop = kLdsTop, sah = kSah;
  static constexpr int ldsStack = kLdsStack, gather = kGather, steps = kSteps;
};
// Dynamic LDS of a kernel launch (the SAH walk's stack entries, kLdsStack).
extern __shared__ uint64_t sahLdsStack[];
constexpr int kTopNodes = 256;

template <class W>
AD float4 nodeQ(const float4* Nodes4, const float4* top, uint32_t q) {
  if (W::ldsTop && q < (uint32_t)(2 * kTopNodes)) return top[q];
  return Nodes4[q];
}

// The reference's stack[32] loop (Arbor.cu:220-230) with its visiting rules.
// The node slab test is the fused form slabSpan (rounding at box faces may
// differ from the reference's; object results do not depend on it beyond
// which boxes a grazing ray enters). The reference pushes the far child only when both children are hit
// and always takes the nearer one (ties -> right); the stack is therefore
// always "the far siblings along the current root path", deepest on top. A bit
// per heap depth records them: popping takes the deepest set bit d and goes
// to the sibling of the current path's depth-d ancestor. A popped node is
// re-tested against the current best exactly as the reference does; a node
// reached directly reuses the span its parent computed (identical values).
// Always inlined: as a called function it is register-allocated on its own
// under the calling convention and costs the kernels a third of their
// occupancy (183 vs 154 VGPRs: 2 vs 3 waves per SIMD).
template <class W>
__device__ __forceinline__ v2 traceTreeSah(const float* P, const float4* Nodes4, const float4* accel,
                                           const uint32_t* Order, v3 ro, v3 rd, float cone, float time, float wind,
                                           float best, int ignore, bool shadows);
template <class W>
__device__ __forceinline__ v2 traceTree(const float* P, const float4* Nodes4, const float4* top, const uint32_t* Order, v3 ro,
                               v3 rd, float cone, float time, float wind, float best, int ignore, bool shadows) {
  if constexpr (W::sah) return traceTreeSah<W>(P, Nodes4, top, Order, ro, rd, cone, time, wind, best, ignore, shadows);
  const v3 inv = invDir(rd), roi = ro * inv;
  int found = -1;
  uint32_t node = 1, pending = 0;
  float4 rl = nodeQ<W>(Nodes4, top, 2), rh = nodeQ<W>(Nodes4, top, 3);
  v2 span = slabSpan(roi, inv, mk(rl.x, rl.y, rl.z), mk(rh.x, rh.y, rh.z));
  bool live = spanHit(span, rl.w, best);
  float4 k0, k1, k2, k3;  // children boxes of `node` when haveKids
  bool haveKids = false;
  while (true) {
    if (live) {
      if (node < (uint32_t)kCapacity) {
        if (!haveKids) {
          // node loads at a raised wave priority, the object test at the
          // highest (as in traceTreeSah; issue order only)
          __builtin_amdgcn_s_setprio(2);
          k0 = nodeQ<W>(Nodes4, top, node * 4);
          k1 = nodeQ<W>(Nodes4, top, node * 4 + 1);
          k2 = nodeQ<W>(Nodes4, top, node * 4 + 2);
          k3 = nodeQ<W>(Nodes4, top, node * 4 + 3);
          __builtin_amdgcn_s_setprio(0);
        }
        // grandchildren exist below depth 14 (node < CAPACITY/2)
        const bool pf = W::prefetch && node < (uint32_t)(kCapacity / 2);
        float4 g[8];
        if (pf) {
#pragma unroll
          for (int j = 0; j < 8; j++) g[j] = nodeQ<W>(Nodes4, top, node * 8 + j);
        }
        v2 ls = slabSpan(roi, inv, mk(k0.x, k0.y, k0.z), mk(k1.x, k1.y, k1.z));
        v2 rs = slabSpan(roi, inv, mk(k2.x, k2.y, k2.z), mk(k3.x, k3.y, k3.z));
        bool lh = spanHit(ls, k0.w, best), rhit = spanHit(rs, k2.w, best);
        if (lh || rhit) {
          bool takeLeft = lh && (!rhit || ls.x < rs.x);
          if (lh && rhit) pending |= 1u << (heapDepth(node) + 1);
          node = node * 2 + (takeLeft ? 0u : 1u);
          span = takeLeft ? ls : rs;
          if (pf) {
            k0 = takeLeft ? g[0] : g[4];
            k1 = takeLeft ? g[1] : g[5];
            k2 = takeLeft ? g[2] : g[6];
            k3 = takeLeft ? g[3] : g[7];
          }
          haveKids = pf;
          continue;
        }
      } else {
        int id = (int)Order[node - kCapacity];
        if (id != ignore) {
          float footprint = fmaxf(0.000005f, cone * fmaxf(0.05f, span.x));
          __builtin_amdgcn_s_setprio(3);
          Prim pr = loadPrim(P, id);
          float t = objectT(pr, ro, rd, footprint, time, wind, !shadows, W::cull, best);
          __builtin_amdgcn_s_setprio(0);
          if (t < best) { best = t; found = id; }
        }
      }
    }
    haveKids = false;
    if (pending == 0) break;
    int d = 31 - __builtin_clz(pending);
    pending &= ~(1u << d);
    node = (node >> (heapDepth(node) - d)) ^ 1u;
    __builtin_amdgcn_s_setprio(2);
    float4 b0 = nodeQ<W>(Nodes4, top, node * 2), b1 = nodeQ<W>(Nodes4, top, node * 2 + 1);
    __builtin_amdgcn_s_setprio(0);
    span = slabSpan(roi, inv, mk(b0.x, b0.y, b0.z), mk(b1.x, b1.y, b1.z));
    live = spanHit(span, b0.w, best);
  }
  return {best, (float)found};
}

// ---- the same tree over a SAH hierarchy (no counterpart in the reference) ----
// The heap BVH has the topology of the Morton order; a binned-SAH hierarchy
// over the same leaf boxes (built once per scene on the host, accel.hip)
// needs about a third of its box tests and, entering near children first
// with tighter boxes, finds the nearest object sooner and tests about 30%
// fewer objects. Layout (float4 units):
//   [0] {root lo, root ref}  [1] {root hi, depth}  [2], [3] header words
//   node j at 4 + 4j: {left lo, left ref} {left hi, -} {right lo, right ref}
//   {right hi, -}, for j < CAPACITY - 1
//   then CAPACITY u32: the heap leaf slot of each primitive id
// A ref >= 0 is an interior node, ref < 0 the primitive ~ref, whose box is
// the heap's leaf box of that primitive: every object is tested with the
// footprint the heap walk gives it. The depth is at most kSahDepth, so the
// far siblings pending along the current path fit a stack of that size.
//
// Which objects are tested, and in which order, differs from the heap walk;
// the result does not, except as follows.
// * Exact ties. Of objects with equal t the heap walk keeps the first it
//   tests, which is the first in its near-first depth-first order (nearer
//   child first, the right one when both are entered at the same t). The
//   object tests here run with their early exits made strict (the best so
//   far raised by one ulp), so an object tested with t equal to the best
//   reports it, and of the two the one that order puts first is kept: the
//   order of two heap leaves is decided at their lowest common ancestor, by
//   the entry t of its two children (two box tests). Ties are common: where
//   a tree is scaled below 1, the tree-local best of a ray that has hit
//   nothing (1e5 / scale) exceeds the no-hit value 1e5 the object tests
//   return, so the reference reports the first object it tests as a hit at
//   t = 1e5 and every later miss ties with it.
// * Hits before the object's own box entry. A swaying leaf can be hit up to
//   a few centimetres outside its rest-pose box; which of two walks tests it
//   then depends on the best each has at that moment (about 1 ray in 2
//   million at 1920x1080). Both results are valid hits of the reference's
//   tests; the heap walk's is the reference's.
constexpr int kAccelHeader = 4;
constexpr int kAccelSlots = kAccelHeader + 4 * (kCapacity - 1);  // float4 index of the slot table
constexpr int kSahDepth = 32;
AD float nextUp(float x) { return __int_as_float(__float_as_int(x) + 1); }  // x > 0, finite

// Whether heap leaf slot sa precedes slot sb in the heap walk's order for
// this ray (the slots differ).
AD bool heapLeafFirst(const float4* Nodes4, v3 roi, v3 inv, uint32_t sa, uint32_t sb) {
  const uint32_t na = (uint32_t)kCapacity + sa, nb = (uint32_t)kCapacity + sb;
  const int k = 32 - __builtin_clz(na ^ nb);
  const uint32_t ca = na >> (k - 1), cb = nb >> (k - 1);  // the ancestors that are children of the LCA
  const uint32_t l = min(ca, cb), r = l ^ 1u;
  const float4 l0 = Nodes4[2 * l], l1 = Nodes4[2 * l + 1], r0 = Nodes4[2 * r], r1 = Nodes4[2 * r + 1];
  const v2 ls = slabSpan(roi, inv, mk(l0.x, l0.y, l0.z), mk(l1.x, l1.y, l1.z));
  const v2 rs = slabSpan(roi, inv, mk(r0.x, r0.y, r0.z), mk(r1.x, r1.y, r1.z));
  return (ls.x < rs.x ? l : r) == ca;
}

template <class W>
__device__ __forceinline__ v2 traceTreeSah(const float* P, const float4* Nodes4, const float4* accel,
                                           const uint32_t* Order, v3 ro, v3 rd, float cone, float time, float wind,
                                           float best, int ignore, bool shadows) {
  (void)Order;
  const uint32_t* slotOf = reinterpret_cast<const uint32_t*>(accel + kAccelSlots);
  const v3 inv = invDir(rd), roi = ro * inv;
  int found = -1;
  const float4 h0 = accel[0], h1 = accel[1];
  v2 span = slabSpan(roi, inv, mk(h0.x, h0.y, h0.z), mk(h1.x, h1.y, h1.z));
  int ref = __float_as_int(h0.w);
  bool live = spanHit(span, 1.0f, best);
  // Pending siblings {ref, entry t} in per-lane scratch memory. The array's
  // address escapes into an empty asm statement, so it is not promoted to
  // registers (64 of them, indexed per lane); its accesses stay ordinary
  // cached scratch accesses (a volatile array would make them system-coherent,
  // bypassing the caches).
  uint64_t stack[kSahDepth];
  asm volatile("" : : "v"(stack) : "memory");
  int top = 0;
  // entries [0, kLdsStack) in LDS: lane-interleaved per wave
  constexpr int kL = W::ldsStack;
  uint64_t* lds = kL > 0 ? sahLdsStack + (size_t)(threadIdx.x >> 6) * kL * 64 + (threadIdx.x & 63) : nullptr;
  // the most recent push stays in registers until the next push or its pop
  uint64_t cached = 0;
  bool haveCached = false;
  // pop the deepest pending sibling still entered before the best (the
  // other two liveness conditions held when it was pushed)
  auto popNext = [&]() -> bool {
    while (haveCached || top > 0) {
      uint64_t e = cached;
      if (!haveCached) {
        --top;
        e = (kL > 0 && top < kL) ? lds[top * 64] : stack[top];
      }
      haveCached = false;
      const float t = __uint_as_float((uint32_t)e);
      if (t < best) {
        ref = (int)(uint32_t)(e >> 32);
        span.x = t;
        return true;
      }
    }
    return false;
  };
  // one interior step: false when neither child is entered. The node's
  // loads are issued at a raised wave priority (s_setprio), so a wave about
  // to wait on memory gets its loads out ahead of the SIMD's other waves'
  // arithmetic; priorities change only the issue order, never a result.
  auto stepNode = [&]() -> bool {
    __builtin_amdgcn_s_setprio(2);
    const float4* nd = accel + kAccelHeader + 4 * ref;
    const float4 k0 = nd[0], k1 = nd[1], k2 = nd[2], k3 = nd[3];
    __builtin_amdgcn_s_setprio(0);
    const v2 ls = slabSpan(roi, inv, mk(k0.x, k0.y, k0.z), mk(k1.x, k1.y, k1.z));
    const v2 rs = slabSpan(roi, inv, mk(k2.x, k2.y, k2.z), mk(k3.x, k3.y, k3.z));
    const bool lh = spanHit(ls, 1.0f, best), rhit = spanHit(rs, 1.0f, best);
    if (!(lh || rhit)) return false;
    const bool takeLeft = lh && (!rhit || ls.x < rs.x);
    const int lref = __float_as_int(k0.w), rref = __float_as_int(k2.w);
    if (lh && rhit) {
      const uint32_t farRef = (uint32_t)(takeLeft ? rref : lref);
      const uint32_t farT = __float_as_uint(takeLeft ? rs.x : ls.x);
      if (haveCached) {
        if (kL > 0 && top < kL) lds[top * 64] = cached;
        else stack[top] = cached;
        top++;
      }
      cached = ((uint64_t)farRef << 32) | farT;
      haveCached = true;
    }
    ref = takeLeft ? lref : rref;
    span = takeLeft ? ls : rs;
    return true;
  };
  auto testLeaf = [&]() {
    const int id = ~ref;
    if (id != ignore) {
      const float footprint = fmaxf(0.000005f, cone * fmaxf(0.05f, span.x));
      const Prim pr = loadPrim(P, id);
      const float t = objectT(pr, ro, rd, footprint, time, wind, !shadows, W::cull, nextUp(best));
      if (t < best) {
        best = t;
        found = id;
      } else if (t == best && found >= 0 && heapLeafFirst(Nodes4, roi, inv, slotOf[id], slotOf[found])) {
        found = id;
      }
    }
  };
  if constexpr (W::gather > 0) {
    while (true) {
      // interior steps (and pops after a missed node) until enough of the
      // live lanes wait at a leaf; the exit test is uniform over the wave
      while (true) {
        const bool step = live && ref >= 0;
        const uint64_t stepping = __ballot(step);
        if (stepping == 0) break;
        const uint64_t waiting = __ballot(live && ref < 0);
        if (8 * __popcll(waiting) >= W::gather * __popcll(waiting | stepping)) break;
        if (step) {
#pragma unroll
          for (int k = 0; k < W::steps; k++) {
            if (k > 0 && !(live && ref >= 0)) break;
            if (!stepNode()) live = popNext();
          }
        }
      }
      // the gathered object tests at the highest priority: a wave holding
      // most of its lanes there finishes them before the stepping waves
      if (live && ref < 0) {
        __builtin_amdgcn_s_setprio(3);
        testLeaf();
        __builtin_amdgcn_s_setprio(0);
        live = popNext();
      }
      if (!live) break;
    }
    return {best, (float)found};
  }
  while (true) {
    if (live && ref >= 0) {
      if (stepNode()) continue;
      live = false;  // fall through to the pop
    }
    if (live) testLeaf();  // at a leaf
    live = popNext();
    if (!live) break;
  }
  return {best, (float)found};
}

// ---- forest DDA (Arbor.cu:235-251) ----
// `trees` holds the 324 instancing records (a TreeInst array in LDS, or
// TreeRecs in device memory); every value is what the reference recomputes
// inline per cell.
template <class W, class Trees>
__device__ inline v2 traceForest(const float* P, const float4* Nodes4, const float4* top, const uint32_t* Order,
                                 Trees trees, v3 ro, v3 rd, float cone, float time, float wind, float best,
                                 int* treeOut, bool shadows) {
  int foundTree = -1;
  float found = -1.0f;
  const float half = (float)kSide * kSpacing * 0.5f;
  v2 worldSpan = boxSpan(ro, rd, mk(-half, -1.2f, -half), mk(half, 13.5f, half));
  float enter = fmaxf(0.0f, worldSpan.x), exit = fminf(best, worldSpan.y);
  if (exit < enter) { *treeOut = -1; return {best, -1.0f}; }
  v3 pos = ro + rd * (enter + 0.0002f);
  int ix = (int)floorf(pos.x / kSpacing + (float)kSide * 0.5f + 0.5f);
  int iz = (int)floorf(pos.z / kSpacing + (float)kSide * 0.5f + 0.5f);
  int sx = rd.x >= 0.0f ? 1 : -1, sz = rd.z >= 0.0f ? 1 : -1;
  float invx = invSafe(rd.x), invz = invSafe(rd.z);
  float bx = ((float)(ix - kSide / 2) + (sx > 0 ? 0.5f : -0.5f)) * kSpacing;
  float bz = ((float)(iz - kSide / 2) + (sz > 0 ? 0.5f : -0.5f)) * kSpacing;
  float tx = (bx - ro.x) * invx, tz = (bz - ro.z) * invz;
  float dx = kSpacing * fabsf(invx), dz = kSpacing * fabsf(invz);
  for (int step = 0; step < 48; step++) {
    if (ix < 0 || iz < 0 || ix >= kSide || iz >= kSide) break;
    int tree = iz * kSide + ix;
    const TreeHead hd = treeHead(trees, tree);
    float sc = hd.sc;
    v3 centre = mk(hd.px, 0.0f, hd.pz) + mk(0.0f, 4.4f * sc, 0.0f);
    float broad = sphereT(ro, rd, centre, 6.6f * sc);
    if (broad < best) {
      const TreeInst ti = treeRecord(trees, tree, hd);
      v3 lr = toLocalPoint(ro, ti), ld = toLocalDir(rd, ti);
      v2 h = traceTree<W>(P, Nodes4, top, Order, lr, ld, cone / sc, time + ti.phase, wind, best / sc, -1, shadows);
      if (h.y >= 0.0f && h.x * sc < best) { best = h.x * sc; found = h.y; foundTree = tree; }
    }
    float next = tx < tz ? tx : tz;
    if (next > exit || next > best) break;
    if (tx < tz) { ix += sx; tx += dx; } else { iz += sz; tz += dz; }
  }
  *treeOut = foundTree;
  return {best, found};
}

// ---- surface normal (Arbor.cu:231-234) ----
__device__ inline v3 objectNormal(const Prim& pr, v3 p, float footprint, float time, float wind) {
  int type = (int)pr.type();
  if (type == 0) {
    float e = fmaxf(0.000035f, fminf(0.0005f, footprint * 0.18f));
    v3 dx = mk(e, 0.0f, 0.0f), dy = mk(0.0f, e, 0.0f), dz = mk(0.0f, 0.0f, e);
    const WoodFrame wf = woodFrame(pr, footprint);
    return unit(mk(woodField(wf, p + dx) - woodField(wf, p - dx), woodField(wf, p + dy) - woodField(wf, p - dy),
                   woodField(wf, p + dz) - woodField(wf, p - dz)));
  }
  if (type == 1) {
    LeafFrame fr = leafFrame(pr, time, wind);
    v3 u = fr.u, v = pr.c(), n = fr.n;
    v3 q = p - pr.a();
    float x = dot(q, u) - pr.f16(), y = dot(q, v);
    if (x < 0.0f) {
      v3 centre = pr.a() + u * (x + pr.f16());
      return unit(p - centre);
    }
    const float e = 0.0002f;
    float len = pr.ra(), hw = pr.rb(), curl = pr.f17(), twist = pr.f18();
    float a = (leafHeight(x + e, y, len, hw, curl, twist, footprint) - leafHeight(x - e, y, len, hw, curl, twist, footprint)) /
              (2.0f * e);
    float c = (leafHeight(x, y + e, len, hw, curl, twist, footprint) - leafHeight(x, y - e, len, hw, curl, twist, footprint)) /
              (2.0f * e);
    return unit(n - u * a - v * c);
  }
  v3 q = p - pr.a(), h = pr.b();
  return unit(mk(q.x / (h.x * h.x), q.y / (h.y * h.y), q.z / (h.z * h.z)));
}

}  // namespace arbor
