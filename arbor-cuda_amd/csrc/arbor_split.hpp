// The stage-split wavefront form of the SAH-walk trace (tracePixels,
// Arbor.cu:235-252, the forest DDA around the per-tree walk).
//
// The single-kernel trace runs a whole ray per lane: a lane whose ray has
// left the forest, or is stepping cells between trees, idles while the other
// lanes of its tile walk trees. Here the work is split into launches over
// compacted queues of rays held in structure-of-arrays buffers in HBM:
//   start   every pixel: camera ray, forest box, DDA to the first cell whose
//           tree's bounding sphere is entered before the best; rays with
//           such a cell are appended to queue 0, the others write Hit;
//   round r each ray of queue r walks that one tree (the SAH walk, the
//           tree-local best bounded by the ray's best), keeps a nearer hit,
//           and steps the DDA on to its next such cell: appended to queue
//           r + 1, or Hit written;
//   finish  the rays of the last queue walk and step to the end in place.
// A ray walks its trees one at a time in its DDA order with the best of the
// trees before, exactly as the single-kernel trace does, so Hit is bit-equal
// for every pixel. Appends are wave-aggregated (one atomic per wave), so a
// queue keeps each wave's rays together and the waves roughly in pixel order.
#pragma once
#include "arbor_trace.hpp"

namespace arbor {

// Ray state between launches, n rays each (structure of arrays).
struct SplitBuffers {
  uint32_t* cell;    // step | (ix + 1) << 8 | (iz + 1) << 16 of the ray's current cell
  float* tx;         // DDA boundary crossings
  float* tz;
  float* best;       // world-space best t so far
  float* found;      // its primitive id (-1: none)
  int* tree;         // its tree (-1: none)
  uint32_t* queue[2];
  uint32_t* counts;  // ray count of queue r (r = 0..rounds), then round r's chunk cursor at kSplitCursors + r;
                     // kSplitWords words, zeroed by the clear launch
};
constexpr int kSplitMaxRounds = 30;
constexpr int kSplitCursors = 32;
constexpr int kSplitWords = 64;

struct DdaCell {
  int ix, iz, step;
  float tx, tz;
};
AD uint32_t packCell(const DdaCell& s) {
  return (uint32_t)s.step | ((uint32_t)(s.ix + 1) << 8) | ((uint32_t)(s.iz + 1) << 16);
}
AD DdaCell unpackCell(uint32_t w, float tx, float tz) {
  return {(int)((w >> 8) & 0xff) - 1, (int)((w >> 16) & 0xff) - 1, (int)(w & 0xff), tx, tz};
}

// The pixel's camera ray and the two ends of its march: the first best
// (ground or far) and the forest box exit it is clipped to.
struct SplitRay {
  v3 ro, rd;
  float best0, enter, exit;
};
AD SplitRay splitRay(const View& view, int x, int y, int width, int height) {
  SplitRay r;
  r.ro = view.eye;
  r.rd = viewRay(view, x, y, width, height);
  const float ground = r.rd.y < -0.00001f ? -r.ro.y / r.rd.y : kFar;
  r.best0 = ground > 0.0f ? ground : kFar;
  const float half = (float)kSide * kSpacing * 0.5f;
  const v2 worldSpan = boxSpan(r.ro, r.rd, mk(-half, -1.2f, -half), mk(half, 13.5f, half));
  r.enter = fmaxf(0.0f, worldSpan.x);
  r.exit = fminf(r.best0, worldSpan.y);
  return r;
}
// The DDA's entry cell and first crossings (Arbor.cu:238-244).
AD DdaCell ddaEnter(const SplitRay& r) {
  const v3 ro = r.ro, rd = r.rd;
  v3 pos = ro + rd * (r.enter + 0.0002f);
  DdaCell s;
  s.ix = (int)floorf(pos.x / kSpacing + (float)kSide * 0.5f + 0.5f);
  s.iz = (int)floorf(pos.z / kSpacing + (float)kSide * 0.5f + 0.5f);
  s.step = 0;
  const int sx = rd.x >= 0.0f ? 1 : -1, sz = rd.z >= 0.0f ? 1 : -1;
  const float invx = invSafe(rd.x), invz = invSafe(rd.z);
  const float bx = ((float)(s.ix - kSide / 2) + (sx > 0 ? 0.5f : -0.5f)) * kSpacing;
  const float bz = ((float)(s.iz - kSide / 2) + (sz > 0 ? 0.5f : -0.5f)) * kSpacing;
  s.tx = (bx - ro.x) * invx;
  s.tz = (bz - ro.z) * invz;
  return s;
}
// Steps the DDA (Arbor.cu:245-251) to the next cell whose tree's bounding
// sphere is entered before best: true with s at that cell, false when the
// march ends. With resume the current cell's tree has been walked already,
// so the march first steps past it.
template <class Trees>
AD bool ddaSeek(Trees trees, const SplitRay& r, float best, DdaCell& s, bool resume) {
  const v3 ro = r.ro, rd = r.rd;
  const int sx = rd.x >= 0.0f ? 1 : -1, sz = rd.z >= 0.0f ? 1 : -1;
  const float invx = invSafe(rd.x), invz = invSafe(rd.z);
  const float dx = kSpacing * fabsf(invx), dz = kSpacing * fabsf(invz);
  while (true) {
    if (!resume) {
      if (s.step >= 48 || s.ix < 0 || s.iz < 0 || s.ix >= kSide || s.iz >= kSide) return false;
      const TreeHead hd = treeHead(trees, s.iz * kSide + s.ix);
      const float sc = hd.sc;
      const v3 centre = mk(hd.px, 0.0f, hd.pz) + mk(0.0f, 4.4f * sc, 0.0f);
      if (sphereT(ro, rd, centre, 6.6f * sc) < best) return true;
    }
    resume = false;
    const float next = s.tx < s.tz ? s.tx : s.tz;
    if (next > r.exit || next > best) return false;
    if (s.tx < s.tz) { s.ix += sx; s.tx += dx; } else { s.iz += sz; s.tz += dz; }
    s.step++;
  }
}
// The walk of the tree at the ray's cell; a nearer hit replaces the best.
template <class W, class Trees>
AD void walkCell(const float* P, const float4* Nodes4, const float4* accel, const uint32_t* Order, Trees trees,
                 const SplitRay& r, const DdaCell& s, float cone, float time, float wind, float& best, float& found,
                 int& foundTree) {
  const int tree = s.iz * kSide + s.ix;
  const TreeHead hd = treeHead(trees, tree);
  const float sc = hd.sc;
  const TreeInst ti = treeRecord(trees, tree, hd);
  const v3 lr = toLocalPoint(r.ro, ti), ld = toLocalDir(r.rd, ti);
  const v2 h = traceTree<W>(P, Nodes4, accel, Order, lr, ld, cone / sc, time + ti.phase, wind, best / sc, -1, false);
  if (h.y >= 0.0f && h.x * sc < best) { best = h.x * sc; found = h.y; foundTree = tree; }
}

// Appends item for the lanes with want: one atomic per wave on count, the
// lanes' slots in lane order. Every active lane of the wave calls it.
AD void splitAppend(bool want, uint32_t item, uint32_t* queue, uint32_t* count) {
  const uint64_t m = __ballot(want);
  if (m == 0) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
  base = __shfl(base, leader);
  if (want) queue[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = item;
}

AD void splitHit(float4* Hit, const float* C, int height, int pixel, float best, float found, int tree) {
  Hit[pixel] = make_float4(best, found, (float)tree, C[14] * best / (float)height);
}
AD void splitSave(const SplitBuffers& b, uint32_t ray, const DdaCell& s, float best, float found, int tree) {
  b.cell[ray] = packCell(s);
  b.tx[ray] = s.tx;
  b.tz[ray] = s.tz;
  b.best[ray] = best;
  b.found[ray] = found;
  b.tree[ray] = tree;
}

__global__ void k_splitClear(uint32_t* counts, int n) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) counts[i] = 0;
}

// The start launch: one wave per 8x8 tile, pixels in tile order.
template <class Trees>
__device__ __forceinline__ void splitStart(Trees trees, const float* C, float4* Hit, int width, int height, int tilesX,
                                           int tiles, SplitBuffers b) {
  const View view = makeView(C);
  const int lane = threadIdx.x & 63;
  for (int t = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); t < tiles; t += gridDim.x * (blockDim.x >> 6)) {
    const int x = (t % tilesX) * 8 + (lane & 7), y = (t / tilesX) * 8 + (lane >> 3);
    const bool in = x < width && y < height;
    bool queued = false;
    const int pixel = in ? y * width + x : 0;
    if (in) {
      const SplitRay r = splitRay(view, x, y, width, height);
      if (r.exit < r.enter) {
        splitHit(Hit, C, height, pixel, r.best0, -1.0f, -1);
      } else {
        DdaCell s = ddaEnter(r);
        if (ddaSeek(trees, r, r.best0, s, false)) {
          splitSave(b, (uint32_t)pixel, s, r.best0, -1.0f, -1);
          queued = true;
        } else {
          splitHit(Hit, C, height, pixel, r.best0, -1.0f, -1);
        }
      }
    }
    splitAppend(queued, (uint32_t)pixel, b.queue[0], b.counts);
  }
}

// Round r (kFinish: the last launch, which walks each ray to its end).
template <class W, bool kFinish, class Trees>
__device__ __forceinline__ void splitRound(const float* P, const float4* Nodes4, const uint32_t* Order, Trees trees,
                                           const float* C, float4* Hit, int width, int height, const float4* accel,
                                           SplitBuffers b, int round) {
  const uint32_t count = b.counts[round];
  const uint32_t* in = b.queue[round & 1];
  uint32_t* out = b.queue[(round + 1) & 1];
  const View view = makeView(C);
  const float cone = C[14] / (float)height, time = C[8], wind = C[11];
  const int lane = threadIdx.x & 63;
  // waves claim chunks of 64 queued rays from the round's cursor until the
  // queue is used up: a wave that drew cheap chunks takes more of them
  uint32_t* cursor = b.counts + kSplitCursors + round;
  while (true) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(cursor, 64u);
    base = __shfl(base, 0);
    if (base >= count) break;
    const bool active = base + (uint32_t)lane < count;
    const uint32_t ray = active ? in[base + lane] : 0u;
    bool queued = false;
    if (active) {
      const int x = (int)(ray % (uint32_t)width), y = (int)(ray / (uint32_t)width);
      const SplitRay r = splitRay(view, x, y, width, height);
      DdaCell s = unpackCell(b.cell[ray], b.tx[ray], b.tz[ray]);
      float best = b.best[ray], found = b.found[ray];
      int tree = b.tree[ray];
      bool more = true;
      do {
        walkCell<W>(P, Nodes4, accel, Order, trees, r, s, cone, time, wind, best, found, tree);
        more = ddaSeek(trees, r, best, s, true);
      } while (kFinish && more);
      if (more) splitSave(b, ray, s, best, found, tree);
      else splitHit(Hit, C, height, (int)ray, best, found, tree);
      queued = more;
    }
    if constexpr (!kFinish) splitAppend(queued, ray, out, b.counts + round + 1);
  }
}

}  // namespace arbor
