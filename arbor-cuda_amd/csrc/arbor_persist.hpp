// Persistent wavefront form of tracePixels over the SAH hierarchy
// (tracePixels / traceForest / traceTree, Arbor.cu:220-252).
//
// The grid is as many waves as the GPU holds at once. Each lane owns one
// primary ray at a time, with its whole state in registers: the forest DDA
// (cell, next boundary t, step count), the world best, and while it walks a
// tree, that tree's local ray and SAH walk (pending siblings in LDS and
// scratch as in traceTreeSah). A lane whose ray ends writes its Hit and takes
// the next ray at once: the wave keeps a queue of the 64 rays of its current
// 8x8 tile and fetches the next tile (in the tile schedule's order) from a
// global counter when the queue runs dry. So a wave stays full until the
// launch runs out of tiles, instead of waiting on the slowest ray of a tile.
//
// Within a wave the interior steps and object tests are gathered as in
// traceTreeSah with kGather: lanes keep stepping through interior nodes until
// at least g/8 of the lanes inside a tree wait at a leaf, then those lanes
// run their object tests together. Every ray still performs exactly the steps
// of traceForest + traceTreeSah in their order (its DDA cells, its tree walks,
// each walk's node visits, pops and object tests), so Hit is bit-identical to
// the other SAH-walk kernels; only which rays share a wave, and when, changes.
//
// With a tile schedule, tileCost[t] receives the work of tile t's rays
// (interior steps, object tests and DDA cells, summed): the wave that takes a
// tile zeroes its entry first, its rays add to it when they end. counters[0]
// hands out tiles, counters[1] counts the waves that have finished; the last
// wave resets both, so they are zero again for the next launch.
#pragma once
#include "arbor_trace.hpp"

namespace arbor {

// One lane's walk of one tree over the SAH hierarchy: the state and steps of
// traceTreeSah, kept across the persistent kernel's loop iterations.
template <class W>
struct SahLane {
  v3 ro, rd, inv, roi;  // tree-local ray
  float cone, time;     // tree-local cone and wind time
  float best;
  int found;
  int ref;
  v2 span;
  bool live;
  int top;
  uint64_t cached;
  bool haveCached;

  __device__ __forceinline__ void begin(const float4* accel, v3 lro, v3 lrd, float lcone, float ltime, float lbest) {
    ro = lro;
    rd = lrd;
    inv = invDir(rd);
    roi = ro * inv;
    cone = lcone;
    time = ltime;
    best = lbest;
    found = -1;
    const float4 h0 = accel[0], h1 = accel[1];
    span = slabSpan(roi, inv, mk(h0.x, h0.y, h0.z), mk(h1.x, h1.y, h1.z));
    ref = __float_as_int(h0.w);
    live = spanHit(span, 1.0f, best);
    top = 0;
    cached = 0;
    haveCached = false;
  }
  // the deepest pending sibling still entered before the best
  __device__ __forceinline__ bool popNext(uint64_t* lds, uint64_t* stack) {
    constexpr int kL = W::ldsStack;
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
  }
  // one interior step: false when neither child is entered
  __device__ __forceinline__ bool stepNode(const float4* accel, uint64_t* lds, uint64_t* stack) {
    constexpr int kL = W::ldsStack;
    const float4* nd = accel + kAccelHeader + 4 * ref;
    const float4 k0 = nd[0], k1 = nd[1], k2 = nd[2], k3 = nd[3];
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
  }
  // the object test at the leaf `ref` (primary rays: refined, no ignore)
  __device__ __forceinline__ void testLeaf(const float* P, const float4* Nodes4, const uint32_t* slotOf, float wind) {
    const int id = ~ref;
    const float footprint = fmaxf(0.000005f, cone * fmaxf(0.05f, span.x));
    const Prim pr = loadPrim(P, id);
    const float t = objectT(pr, ro, rd, footprint, time, wind, true, W::cull, nextUp(best));
    if (t < best) {
      best = t;
      found = id;
    } else if (t == best && found >= 0 && heapLeafFirst(Nodes4, roi, inv, slotOf[id], slotOf[found])) {
      found = id;
    }
  }
};

// The persistent trace kernel. W is a SAH walk with kGather > 0.
template <int kWaves, class W, int kMinWaves = 1>
__global__ __launch_bounds__(kWaves * 64) __attribute__((amdgpu_waves_per_eu(kMinWaves))) void k_tracePersistent(
    const float* P, const float4* Nodes4, const float* C, float4* Hit, int width, int height, int tilesX, int tiles,
    const uint32_t* tileOrder, uint32_t* tileCost, const float4* accel, uint32_t* counters) {
  static_assert(W::sah && W::gather > 0, "the persistent trace gathers object tests over the SAH walk");
  __shared__ TreeInst trees[kTrees];
  for (int i = threadIdx.x; i < kTrees; i += blockDim.x) trees[i] = makeTree(i, (int)C[31]);
  __syncthreads();  // the last block-wide barrier: from here on the waves run independently
  const View view = makeView(C);
  const v3 eye = view.eye;
  const float cone = C[14] / (float)height, time = C[8], wind = C[11], scale = C[14];
  const uint32_t* slotOf = reinterpret_cast<const uint32_t*>(accel + kAccelSlots);
  const int lane = threadIdx.x & 63;
  constexpr int kL = W::ldsStack;
  uint64_t* lds = kL > 0 ? sahLdsStack + (size_t)(threadIdx.x >> 6) * kL * 64 + lane : nullptr;
  uint64_t stack[kSahDepth];
  asm volatile("" : : "v"(stack) : "memory");  // keeps the stack in scratch, not registers

  // ray state
  int pix = -1, tile = 0;
  uint32_t work = 0;
  v3 rd = mk(0.0f, 0.0f, 1.0f);
  float best = kFar, found = -1.0f, exitT = 0.0f, tx = 0.0f, tz = 0.0f, invx = 0.0f, invz = 0.0f;
  int foundTree = -1, ix = 0, iz = 0, step = 0, tree = 0;
  bool inTree = false, cellPending = false;
  SahLane<W> walk;
  walk.live = false;
  walk.ref = 0;

  // wave-uniform tile queue
  int curTile = 0, nextRay = 64;
  bool exhausted = false;
  const float half = (float)kSide * kSpacing * 0.5f;

  while (true) {
    // Refill idle lanes and run the DDA of every lane outside a tree, until
    // no lane is idle or no tile is left. Afterwards every lane with a ray
    // is inside a tree.
    while (true) {
      uint64_t idle = __ballot(pix < 0);
      while (idle != 0 && !exhausted) {
        if (nextRay >= 64) {
          uint32_t t = 0;
          if (lane == 0) t = atomicAdd(&counters[0], 1u);
          t = __shfl(t, 0);
          if (t >= (uint32_t)tiles) {
            exhausted = true;
            break;
          }
          curTile = tileOrder ? (int)tileOrder[t] : (int)t;
          nextRay = 0;
          if (tileCost && lane == 0) tileCost[curTile] = 0u;
        }
        const int rank =
            (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
        const int avail = 64 - nextRay;
        if (pix < 0 && rank < avail) {
          const int r = nextRay + rank;
          const int x = (curTile % tilesX) * 8 + (r & 7), y = (curTile / tilesX) * 8 + (r >> 3);
          if (x < width && y < height) {
            // tracePixels' ray and initial best (Arbor.cu:252), traceForest's
            // world box and DDA set-up (Arbor.cu:236-240)
            rd = viewRay(view, x, y, width, height);
            const float ground = rd.y < -0.00001f ? -eye.y / rd.y : kFar;
            best = ground > 0.0f ? ground : kFar;
            found = -1.0f;
            foundTree = -1;
            inTree = false;
            const v2 worldSpan = boxSpan(eye, rd, mk(-half, -1.2f, -half), mk(half, 13.5f, half));
            const float enter = fmaxf(0.0f, worldSpan.x);
            exitT = fminf(best, worldSpan.y);
            step = 0;
            cellPending = true;
            if (exitT < enter) {  // the ray misses the forest: it ends here, the lane stays idle
              Hit[y * width + x] = make_float4(best, -1.0f, -1.0f, scale * best / (float)height);
              if (tileCost) atomicAdd(&tileCost[curTile], 1u);
            } else {
              pix = y * width + x;
              tile = curTile;
              work = 1;
              const v3 pos = eye + rd * (enter + 0.0002f);
              ix = (int)floorf(pos.x / kSpacing + (float)kSide * 0.5f + 0.5f);
              iz = (int)floorf(pos.z / kSpacing + (float)kSide * 0.5f + 0.5f);
              const int sx = rd.x >= 0.0f ? 1 : -1, sz = rd.z >= 0.0f ? 1 : -1;
              invx = invSafe(rd.x);
              invz = invSafe(rd.z);
              const float bx = ((float)(ix - kSide / 2) + (sx > 0 ? 0.5f : -0.5f)) * kSpacing;
              const float bz = ((float)(iz - kSide / 2) + (sz > 0 ? 0.5f : -0.5f)) * kSpacing;
              tx = (bx - eye.x) * invx;
              tz = (bz - eye.z) * invz;
            }
          }
        }
        nextRay += min(__popcll(idle), avail);
        idle = __ballot(pix < 0);
      }
      // DDA (Arbor.cu:241-250) of lanes outside a tree: test the current
      // cell's tree, or step to the next cell, until the lane enters a tree
      // or its ray ends.
      if (pix >= 0 && !inTree) {
        while (true) {
          bool done = false;
          if (cellPending) {
            cellPending = false;
            if (step >= 48 || ix < 0 || iz < 0 || ix >= kSide || iz >= kSide) {
              done = true;
            } else {
              tree = iz * kSide + ix;
              const TreeInst& ti = trees[tree];
              const float sc = ti.sc;
              const v3 centre = mk(ti.px, 0.0f, ti.pz) + mk(0.0f, 4.4f * sc, 0.0f);
              const float broad = sphereT(eye, rd, centre, 6.6f * sc);
              work++;
              if (broad < best) {
                walk.begin(accel, toLocalPoint(eye, ti), toLocalDir(rd, ti), cone / sc, time + ti.phase, best / sc);
                inTree = true;
                break;
              }
            }
          }
          if (!done) {
            const float next = tx < tz ? tx : tz;
            if (next > exitT || next > best) {
              done = true;
            } else {
              if (tx < tz) {
                ix += rd.x >= 0.0f ? 1 : -1;
                tx += kSpacing * fabsf(invx);
              } else {
                iz += rd.z >= 0.0f ? 1 : -1;
                tz += kSpacing * fabsf(invz);
              }
              step++;
              cellPending = true;
            }
          }
          if (done) {
            Hit[pix] = make_float4(best, found, (float)foundTree, scale * best / (float)height);
            if (tileCost) atomicAdd(&tileCost[tile], work);
            pix = -1;
            break;
          }
        }
      }
      if (exhausted || __ballot(pix < 0) == 0) break;
    }
    if (__ballot(pix >= 0) == 0) break;  // no ray left anywhere in the wave

    // Gathered walk of the lanes inside a tree (traceTreeSah, kGather).
    while (true) {
      const bool stepping = inTree && walk.live && walk.ref >= 0;
      const uint64_t steppers = __ballot(stepping);
      if (steppers == 0) break;
      const uint64_t waiting = __ballot(inTree && walk.live && walk.ref < 0);
      if (8 * __popcll(waiting) >= W::gather * __popcll(waiting | steppers)) break;
      if (stepping) {
        work++;
        if (!walk.stepNode(accel, lds, stack)) walk.live = walk.popNext(lds, stack);
      }
    }
    if (inTree && walk.live && walk.ref < 0) {
      work++;
      walk.testLeaf(P, Nodes4, slotOf, wind);
      walk.live = walk.popNext(lds, stack);
    }
    // a finished tree walk hands its result to the DDA (Arbor.cu:244)
    if (inTree && !walk.live) {
      const float sc = trees[tree].sc;
      if (walk.found >= 0 && walk.best * sc < best) {
        best = walk.best * sc;
        found = (float)walk.found;
        foundTree = tree;
      }
      inTree = false;
    }
  }
  if (lane == 0) {
    const uint32_t waves = gridDim.x * kWaves;
    if (atomicAdd(&counters[1], 1u) == waves - 1) {  // the last wave: every tile was taken
      atomicExch(&counters[0], 0u);
      atomicExch(&counters[1], 0u);
    }
  }
}

}  // namespace arbor
