// Wave-cooperative forest traversal for primary rays on gfx950.
//
// traceForest + traceTree (Arbor.cu:220-251) flattened into one per-lane
// state machine: a lane's "step" is either one DDA cell (sphere test and, on
// a hit, entering the tree) or one BVH step of the current tree. Leaf tests
// are batched across the wave: every lane steps until it holds one leaf
// candidate (or has finished its ray), then all lanes holding a candidate run
// objectT together. Each lane still performs exactly the reference's sequence
// of operations for its own ray (a leaf's result updates that lane's best
// before its traversal continues), so Hit is identical to the per-lane walk;
// only the interleaving across lanes changes, which keeps the expensive
// intersection code running with most of the wave active.
#pragma once
#include "arbor_trace.hpp"

namespace arbor {

struct ForestRay {
  // world ray and result
  v3 ro, rd;
  float best;
  float found;
  int foundTree;
  // DDA state
  int ix, iz, sx, sz;
  float tx, tz, ddx, ddz, exit;
  int step;
  // current tree (tree < 0: between trees)
  int tree;
  v3 lr, ld, inv;
  float sc, cone, time, lbest;
  int lfound;
  uint32_t node, pending;
  v2 span;
  bool live;
  // pending leaf test
  int leafId;
  float leafFp;
  bool done;
};

// Sets up the DDA exactly as traceForest does before its loop.
AD void forestBegin(ForestRay& s, v3 ro, v3 rd, float best) {
  s.ro = ro;
  s.rd = rd;
  s.best = best;
  s.found = -1.0f;
  s.foundTree = -1;
  s.tree = -1;
  s.leafId = -1;
  s.done = false;
  const float half = (float)kSide * kSpacing * 0.5f;
  v2 worldSpan = boxSpan(ro, rd, mk(-half, -1.2f, -half), mk(half, 13.5f, half));
  float enter = fmaxf(0.0f, worldSpan.x);
  s.exit = fminf(best, worldSpan.y);
  if (s.exit < enter) { s.done = true; return; }
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
}

// The end of one DDA iteration (after the cell's tree, if any): break tests
// and the advance to the next cell.
AD void forestAdvance(ForestRay& s) {
  float next = s.tx < s.tz ? s.tx : s.tz;
  if (next > s.exit || next > s.best) { s.done = true; return; }
  if (s.tx < s.tz) { s.ix += s.sx; s.tx += s.ddx; } else { s.iz += s.sz; s.tz += s.ddz; }
  s.step++;
  if (s.step >= 48) s.done = true;
}

// One step of the lane's walk. Returns with either done, a pending leaf, or
// more steps to take.
AD void forestStep(ForestRay& s, const float4* Nodes4, const uint32_t* Order, const TreeInst* trees, float cone,
                   float time, float wind, bool shadows, int ignore) {
  if (s.tree < 0) {
    // DDA cell (Arbor.cu:245-247)
    if (s.ix < 0 || s.iz < 0 || s.ix >= kSide || s.iz >= kSide) { s.done = true; return; }
    int tree = s.iz * kSide + s.ix;
    const TreeInst& ti = trees[tree];
    float sc = ti.sc;
    v3 centre = mk(ti.px, 0.0f, ti.pz) + mk(0.0f, 4.4f * sc, 0.0f);
    float broad = sphereT(s.ro, s.rd, centre, 6.6f * sc);
    if (!(broad < s.best)) { forestAdvance(s); return; }
    s.tree = tree;
    s.sc = sc;
    s.lr = toLocalPoint(s.ro, ti);
    s.ld = toLocalDir(s.rd, ti);
    s.inv = invDir(s.ld);
    s.cone = cone / sc;
    s.time = time + ti.phase;
    s.lbest = s.best / sc;
    s.lfound = -1;
    s.node = 1;
    s.pending = 0;
    float4 rl = Nodes4[2], rh = Nodes4[3];
    s.span = boxSpanInv(s.lr, s.inv, mk(rl.x, rl.y, rl.z), mk(rh.x, rh.y, rh.z));
    s.live = spanHit(s.span, rl.w, s.lbest);
    return;
  }
  // BVH step (Arbor.cu:221-229), stack as one bit per heap depth
  if (s.live) {
    if (s.node < (uint32_t)kCapacity) {
      const float4* c = Nodes4 + s.node * 4;
      float4 l0 = c[0], l1 = c[1], r0 = c[2], r1 = c[3];
      v2 ls = boxSpanInv(s.lr, s.inv, mk(l0.x, l0.y, l0.z), mk(l1.x, l1.y, l1.z));
      v2 rs = boxSpanInv(s.lr, s.inv, mk(r0.x, r0.y, r0.z), mk(r1.x, r1.y, r1.z));
      bool lh = spanHit(ls, l0.w, s.lbest), rhit = spanHit(rs, r0.w, s.lbest);
      if (lh || rhit) {
        bool takeLeft = lh && (!rhit || ls.x < rs.x);
        if (lh && rhit) s.pending |= 1u << (heapDepth(s.node) + 1);
        s.node = s.node * 2 + (takeLeft ? 0u : 1u);
        s.span = takeLeft ? ls : rs;
        return;
      }
    } else {
      int id = (int)Order[s.node - kCapacity];
      s.live = false;  // the leaf is consumed by the test below
      if (id != ignore) {
        s.leafId = id;
        s.leafFp = fmaxf(0.000005f, s.cone * fmaxf(0.05f, s.span.x));
        return;
      }
    }
  }
  if (s.pending != 0) {
    int d = 31 - __builtin_clz(s.pending);
    s.pending &= ~(1u << d);
    s.node = (s.node >> (heapDepth(s.node) - d)) ^ 1u;
    float4 b0 = Nodes4[s.node * 2], b1 = Nodes4[s.node * 2 + 1];
    s.span = boxSpanInv(s.lr, s.inv, mk(b0.x, b0.y, b0.z), mk(b1.x, b1.y, b1.z));
    s.live = spanHit(s.span, b0.w, s.lbest);
    return;
  }
  // tree finished: the traceForest update after traceTree (Arbor.cu:247)
  if (s.lfound >= 0 && s.lbest * s.sc < s.best) {
    s.best = s.lbest * s.sc;
    s.found = (float)s.lfound;
    s.foundTree = s.tree;
  }
  s.tree = -1;
  forestAdvance(s);
}

// Whole-wave loop: traversal phases alternate with batched leaf tests.
AD void forestTraceBatched(ForestRay& s, const float* P, const float4* Nodes4, const uint32_t* Order,
                           const TreeInst* trees, float cone, float time, float wind, bool shadows) {
  while (true) {
    while (!s.done && s.leafId < 0) forestStep(s, Nodes4, Order, trees, cone, time, wind, shadows, -1);
    bool test = s.leafId >= 0;
    if (!__builtin_amdgcn_ballot_w64(test) ) {
      if (!__builtin_amdgcn_ballot_w64(!s.done)) break;
      continue;
    }
    if (test) {
      Prim pr = loadPrim(P, s.leafId);
      float t = objectT(pr, s.lr, s.ld, s.leafFp, s.time, wind, !shadows);
      if (t < s.lbest) { s.lbest = t; s.lfound = s.leafId; }
      s.leafId = -1;
    }
  }
}

}  // namespace arbor
