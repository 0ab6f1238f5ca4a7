// Wave-scheduled primary-ray tracing for gfx950.
//
// traceForest + traceTree + objectT (Arbor.cu:186-251) are split into phases
// of a per-lane state machine:
//   RAY    take the next pixel from the block's pool and set up its DDA
//   DDA    one forest cell: sphere test, and on a hit enter the tree
//   BVH    one heap-BVH step (children test / pop / leaf reached / tree done)
//   LEAF   an object test's prologue: wood and stone complete here; a leaf
//          runs its rejection tests, petiole tube and slab span
//   NEWTON the leaf's 7-step Newton solve and acceptance
// Each iteration the wave executes ONE phase, chosen by ballot: expensive
// phases wait until enough lanes are ready for them, so they run with most of
// the wave active, while the cheap traversal phases fill the gaps. A lane
// whose ray is finished writes Hit and pulls the next ray from the block's
// LDS pool, so no lane idles until the pool is empty.
//
// Every lane still performs exactly the reference's sequence of operations
// for its own ray (a leaf's result reaches that lane's best before its walk
// continues), so Hit is bit-identical to the per-lane walk; only the
// interleaving of lanes changes.
#pragma once
#include "arbor_trace.hpp"

namespace arbor {

enum : int { kPhRay = 0, kPhDda = 1, kPhBvh = 2, kPhLeaf = 3, kPhNewton = 4, kPhIdle = 5 };

struct LaneRay {
  int phase;
  int pixel;
  // world ray and result
  v3 rd;
  float best, found;
  int foundTree;
  // DDA
  int ix, iz, sx, sz, step;
  float tx, tz, ddx, ddz, exit;
  // current tree
  int tree;
  v3 lr, ld, inv;
  float sc, cone, time, lbest;
  int lfound;
  uint32_t node, pending;
  v2 span;
  bool live;
  // object test in flight
  int leafId;
  float fp;
  // Newton state (leaf frame)
  float nx, ny, nz, ndx, ndy, ndz, nt, spanLo, spanHi, pbest;
};

// Scheduling thresholds (lanes of 64) for the expensive phases.
struct WaveSchedule {
  int leafMin, newtonMin;
};

struct ViewCtx {
  View view;
  float cone, time, wind;
  int width, height, tilesX;
};

// Pixel of ray index r: rays are numbered tile-major (64 per 8x8 tile) so
// consecutive rays are neighbours.
AD void rayPixel(int r, int tilesX, int& x, int& y) {
  int tile = r >> 6, lane = r & 63;
  x = (tile % tilesX) * 8 + (lane & 7);
  y = (tile / tilesX) * 8 + (lane >> 3);
}

AD void laneFinishRay(LaneRay& s, float4* Hit, const ViewCtx& vc, float scale) {
  if (s.pixel >= 0)
    Hit[s.pixel] = make_float4(s.best, s.found, (float)s.foundTree, scale * s.best / (float)vc.height);
  s.phase = kPhRay;
  s.pixel = -1;
}

// DDA advance after a cell (Arbor.cu:248-249).
AD bool laneAdvance(LaneRay& s) {
  float next = s.tx < s.tz ? s.tx : s.tz;
  if (next > s.exit || next > s.best) return false;
  if (s.tx < s.tz) { s.ix += s.sx; s.tx += s.ddx; } else { s.iz += s.sz; s.tz += s.ddz; }
  s.step++;
  return s.step < 48;
}

// RAY phase: set up pixel (x, y) as traceForest does before its loop.
AD void laneBegin(LaneRay& s, int x, int y, const ViewCtx& vc) {
  v3 ro = vc.view.eye, rd = viewRay(vc.view, x, y, vc.width, vc.height);
  float ground = rd.y < -0.00001f ? -ro.y / rd.y : kFar;
  s.rd = rd;
  s.best = ground > 0.0f ? ground : kFar;
  s.found = -1.0f;
  s.foundTree = -1;
  s.tree = -1;
  s.leafId = -1;
  const float half = (float)kSide * kSpacing * 0.5f;
  v2 worldSpan = boxSpan(ro, rd, mk(-half, -1.2f, -half), mk(half, 13.5f, half));
  float enter = fmaxf(0.0f, worldSpan.x);
  s.exit = fminf(s.best, worldSpan.y);
  if (s.exit < enter) { s.phase = kPhIdle; return; }  // finished: write on the next RAY phase
  v3 pos = ro + rd * (enter + 0.0002f);
  s.ix = (int)floorf(pos.x / kSpacing + (float)kSide * 0.5f + 0.5f);
  s.iz = (int)floorf(pos.z / kSpacing + (float)kSide * 0.5f + 0.5f);
  s.sx = rd.x >= 0.0f ? 1 : -1;
  s.sz = rd.z >= 0.0f ? 1 : -1;
  float invx = invSafe(rd.x), invz = invSafe(rd.z);
  float bx = ((float)(s.ix - kSide / 2) + (s.sx > 0 ? 0.5f : -0.5f)) * kSpacing;
  float bz = ((float)(s.iz - kSide / 2) + (s.sz > 0 ? 0.5f : -0.5f)) * kSpacing;
  s.tx = (bx - ro.x) * invx;
  s.tz = (bz - ro.z) * invz;
  s.ddx = kSpacing * fabsf(invx);
  s.ddz = kSpacing * fabsf(invz);
  s.step = 0;
  s.phase = kPhDda;
}

// DDA phase: one cell (Arbor.cu:245-247 up to entering traceTree).
AD void laneDda(LaneRay& s, const float4* Nodes4, const TreeInst* trees, const ViewCtx& vc) {
  if (s.ix < 0 || s.iz < 0 || s.ix >= kSide || s.iz >= kSide) { s.phase = kPhIdle; return; }
  int tree = s.iz * kSide + s.ix;
  const TreeInst& ti = trees[tree];
  float sc = ti.sc;
  v3 ro = vc.view.eye;
  v3 centre = mk(ti.px, 0.0f, ti.pz) + mk(0.0f, 4.4f * sc, 0.0f);
  float broad = sphereT(ro, s.rd, centre, 6.6f * sc);
  if (!(broad < s.best)) {
    if (!laneAdvance(s)) s.phase = kPhIdle;
    return;
  }
  s.tree = tree;
  s.sc = sc;
  s.lr = toLocalPoint(ro, ti);
  s.ld = toLocalDir(s.rd, ti);
  s.inv = invDir(s.ld);
  s.cone = vc.cone / sc;
  s.time = vc.time + ti.phase;
  s.lbest = s.best / sc;
  s.lfound = -1;
  s.node = 1;
  s.pending = 0;
  float4 rl = Nodes4[2], rh = Nodes4[3];
  s.span = slabSpan(s.lr * s.inv, s.inv, mk(rl.x, rl.y, rl.z), mk(rh.x, rh.y, rh.z));
  s.live = spanHit(s.span, rl.w, s.lbest);
  s.phase = kPhBvh;
}

// BVH phase: one step of the walk of traceTree (stack as one bit per depth).
AD void laneBvh(LaneRay& s, const float4* Nodes4, const uint32_t* Order) {
  if (s.live) {
    if (s.node < (uint32_t)kCapacity) {
      const float4* c = Nodes4 + s.node * 4;
      float4 l0 = c[0], l1 = c[1], r0 = c[2], r1 = c[3];
      v2 ls = slabSpan(s.lr * s.inv, s.inv, mk(l0.x, l0.y, l0.z), mk(l1.x, l1.y, l1.z));
      v2 rs = slabSpan(s.lr * s.inv, s.inv, mk(r0.x, r0.y, r0.z), mk(r1.x, r1.y, r1.z));
      bool lh = spanHit(ls, l0.w, s.lbest), rhit = spanHit(rs, r0.w, s.lbest);
      if (lh || rhit) {
        bool takeLeft = lh && (!rhit || ls.x < rs.x);
        if (lh && rhit) s.pending |= 1u << (heapDepth(s.node) + 1);
        s.node = s.node * 2 + (takeLeft ? 0u : 1u);
        s.span = takeLeft ? ls : rs;
        return;
      }
    } else {
      // a live leaf: its object test runs (LEAF phase) before the walk goes on
      s.leafId = (int)Order[s.node - kCapacity];
      s.fp = fmaxf(0.000005f, s.cone * fmaxf(0.05f, s.span.x));
      s.live = false;
      s.phase = kPhLeaf;
      return;
    }
  }
  if (s.pending != 0) {
    int d = 31 - __builtin_clz(s.pending);
    s.pending &= ~(1u << d);
    s.node = (s.node >> (heapDepth(s.node) - d)) ^ 1u;
    float4 b0 = Nodes4[s.node * 2], b1 = Nodes4[s.node * 2 + 1];
    s.span = slabSpan(s.lr * s.inv, s.inv, mk(b0.x, b0.y, b0.z), mk(b1.x, b1.y, b1.z));
    s.live = spanHit(s.span, b0.w, s.lbest);
    return;
  }
  // tree finished: the update after traceTree returns (Arbor.cu:247)
  if (s.lfound >= 0 && s.lbest * s.sc < s.best) {
    s.best = s.lbest * s.sc;
    s.found = (float)s.lfound;
    s.foundTree = s.tree;
  }
  s.tree = -1;
  s.phase = laneAdvance(s) ? kPhDda : kPhIdle;
}

AD void laneObjectResult(LaneRay& s, float t) {
  if (t < s.lbest) { s.lbest = t; s.lfound = s.leafId; }
  s.phase = kPhBvh;
}

// LEAF phase: objectT (Arbor.cu:219) up to the leaf's Newton solve.
AD void laneLeaf(LaneRay& s, const float* P, float wind) {
  Prim pr = loadPrim(P, s.leafId);
  int type = (int)pr.type();
  v3 ro = s.lr, rd = s.ld;
  if (type != 1) {
    laneObjectResult(s, objectT(pr, ro, rd, s.fp, s.time, wind, true, true));
    return;
  }
  // leafT (Arbor.cu:203-218), prologue
  v3 origin = pr.a();
  LeafFrame fr = leafFrame(pr, s.time, wind);
  v3 u = fr.u, v = pr.c(), n = fr.n;
  float stem = pr.f16(), len = pr.ra(), width = pr.rb();
  bool petiole = !sphereMiss(ro, rd, origin + u * (stem * 0.5f), stem * 0.5f + 0.0011f, kDiscriminantPad);
  float hx = len * 0.5f + 0.001f;
  bool slab = !sphereMiss(ro, rd, origin + u * (stem + len * 0.5f) + n * 0.015f,
                          sqrtf(hx * hx + width * width + 0.05f * 0.05f), kSlabPad);
  float best = petiole ? tubeT(ro, rd, origin, origin + u * stem, 0.0011f, 0.00065f) : kFar;
  if (!slab) { laneObjectResult(s, best); return; }
  v3 q = ro - origin;
  float x = dot(q, u) - stem, y = dot(q, v), z = dot(q, n);
  float dx = dot(rd, u), dy = dot(rd, v), dz = dot(rd, n);
  v2 span = boxSpan(mk(x, y, z), mk(dx, dy, dz), mk(-0.001f, -width, -0.035f), mk(len + 0.001f, width, 0.065f));
  if (span.y < fmaxf(0.0001f, span.x)) { laneObjectResult(s, best); return; }
  float t = fabsf(dz) > 0.000001f ? -z / dz : (span.x + span.y) * 0.5f;
  s.nt = clampf(t, fmaxf(0.0001f, span.x), span.y);
  s.nx = x; s.ny = y; s.nz = z; s.ndx = dx; s.ndy = dy; s.ndz = dz;
  s.spanLo = span.x;
  s.spanHi = span.y;
  s.pbest = best;
  s.phase = kPhNewton;
}

// NEWTON phase: the rest of leafT.
AD void laneNewton(LaneRay& s, const float* P) {
  const float* r = P + s.leafId * kPS;
  float len = r[7], width = r[11], curl = r[17], twist = r[18];
  const float x = s.nx, y = s.ny, z = s.nz, dx = s.ndx, dy = s.ndy, dz = s.ndz;
  float t = s.nt;
  LeafLocal L{x, y, z, dx, dy, dz, len, width, curl, twist, s.fp};
  float best = fminf(s.pbest, leafNewton(L, t, s.spanLo, s.spanHi));
  laneObjectResult(s, best);
}

// One wave's loop over the rays [*pool, end) of its block; `pool` is an LDS
// counter shared by the block's waves.
AD void waveTrace(int* pool, int end, const float* P, const float4* Nodes4, const uint32_t* Order,
                  const TreeInst* trees, const ViewCtx& vc, float4* Hit, float scale, WaveSchedule ws) {
  LaneRay s;
  s.phase = kPhRay;
  s.pixel = -1;
  const int lane = threadIdx.x & 63;
  bool poolEmpty = false;
  while (true) {
    // lanes whose ray finished write it and want a new one
    if (s.phase == kPhIdle) laneFinishRay(s, Hit, vc, scale);
    uint64_t want = __ballot(s.phase == kPhRay);
    if (want != 0 && !poolEmpty) {
      int n = __popcll(want);
      int base = 0;
      if (lane == __builtin_ctzll(want)) base = atomicAdd(pool, n);
      base = __shfl(base, __builtin_ctzll(want), 64);
      if (base >= end) poolEmpty = true;
      if (s.phase == kPhRay) {
        int r = base + __popcll(want & ((1ull << lane) - 1ull));
        if (r < end) {
          int x, y;
          rayPixel(r, vc.tilesX, x, y);
          if (x < vc.width && y < vc.height) {
            s.pixel = y * vc.width + x;
            laneBegin(s, x, y, vc);
          }
        }
      }
      continue;
    }
    uint64_t trav = __ballot(s.phase == kPhDda || s.phase == kPhBvh);
    uint64_t leaf = __ballot(s.phase == kPhLeaf);
    uint64_t newton = __ballot(s.phase == kPhNewton);
    int nt = __popcll(trav), nl = __popcll(leaf), nn = __popcll(newton);
    if (nt + nl + nn == 0) break;
    if (nn > 0 && (nn >= ws.newtonMin || nt + nl == 0)) {
      if (s.phase == kPhNewton) laneNewton(s, P);
    } else if (nl > 0 && (nl >= ws.leafMin || nt == 0)) {
      if (s.phase == kPhLeaf) laneLeaf(s, P, vc.wind);
    } else {
      if (s.phase == kPhBvh) laneBvh(s, Nodes4, Order);
      else if (s.phase == kPhDda) laneDda(s, Nodes4, trees, vc);
    }
  }
}

}  // namespace arbor
