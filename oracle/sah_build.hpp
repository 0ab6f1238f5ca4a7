// Oracle (test infrastructure): the host restatement of the device SAH
// builder (arbor-cuda_amd/csrc/accel.hip), recursive and serial, which the
// GPU tests compare the device build against word for word. A binned
// surface-area-heuristic binary tree over the heap BVH's leaf boxes,
// flattened into the device layout the SAH walk reads (arbor_trace.hpp):
//   float4 [0] {root lo, root ref}, [1] {root hi, depth}, [2] {nodes, prims,
//   0, 0}, [3] 0; node j at float4 4 + 4j: {left lo, left ref} {left hi, 0}
//   {right lo, right ref} {right hi, 0}; from float4 4 + 4 (capacity - 1):
//   capacity u32, the heap leaf slot of each primitive id.
// A ref >= 0 is an interior node, ref < 0 the primitive ~ref. The build is
// deterministic, and every leaf lies within maxDepth of the root.
#pragma once

#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

namespace arbor {
namespace sah {

struct Box {
  float lo[3], hi[3];
};
struct Prim {
  Box b;
  float c[3];  // box centre
  int id;
};

inline Box unite(const Box& a, const Box& b) {
  Box r;
  for (int k = 0; k < 3; k++) {
    r.lo[k] = std::min(a.lo[k], b.lo[k]);
    r.hi[k] = std::max(a.hi[k], b.hi[k]);
  }
  return r;
}
inline double area(const Box& b) {
  double d[3];
  for (int k = 0; k < 3; k++) d[k] = std::max(0.0, (double)b.hi[k] - (double)b.lo[k]);
  return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
}
inline uint32_t bits(float f) {
  uint32_t u;
  memcpy(&u, &f, sizeof u);
  return u;
}
inline int ceilLog2(int n) {
  int l = 0;
  while ((1 << l) < n) l++;
  return l;
}

struct Builder {
  static constexpr int kBins = 32;
  int maxDepthAllowed;
  std::vector<Prim> prims;
  std::vector<uint32_t> nodes;  // 16 words per interior node
  int maxDepth = 0;

  explicit Builder(int depthLimit) : maxDepthAllowed(depthLimit) {}

  Box boundsOf(int lo, int hi) const {
    Box b = prims[lo].b;
    for (int i = lo + 1; i < hi; i++) b = unite(b, prims[i].b);
    return b;
  }
  // Splits prims[lo, hi) and returns the first index of the right part.
  int split(int lo, int hi, int depth) {
    const int n = hi - lo;
    float cl[3], ch[3];
    for (int k = 0; k < 3; k++) cl[k] = ch[k] = prims[lo].c[k];
    for (int i = lo + 1; i < hi; i++)
      for (int k = 0; k < 3; k++) {
        cl[k] = std::min(cl[k], prims[i].c[k]);
        ch[k] = std::max(ch[k], prims[i].c[k]);
      }
    int axis = -1, cut = -1;
    // Close to the depth limit the rest is split at the median, which keeps
    // every leaf within maxDepthAllowed of the root.
    if (depth + ceilLog2(n) < maxDepthAllowed) {
      double bestCost = 1e300;
      for (int k = 0; k < 3; k++) {
        const float ext = ch[k] - cl[k];
        if (!(ext > 0.0f)) continue;
        Box bin[kBins];
        int cnt[kBins] = {0};
        for (int i = lo; i < hi; i++) {
          int b = std::min(kBins - 1, (int)((prims[i].c[k] - cl[k]) / ext * kBins));
          bin[b] = cnt[b] ? unite(bin[b], prims[i].b) : prims[i].b;
          cnt[b]++;
        }
        double rightArea[kBins];
        int rightCount[kBins];
        Box acc;
        int accN = 0;
        for (int b = kBins - 1; b > 0; b--) {
          if (cnt[b]) {
            acc = accN ? unite(acc, bin[b]) : bin[b];
            accN += cnt[b];
          }
          rightArea[b] = accN ? area(acc) : 0.0;
          rightCount[b] = accN;
        }
        accN = 0;
        for (int b = 0; b < kBins - 1; b++) {
          if (cnt[b]) {
            acc = accN ? unite(acc, bin[b]) : bin[b];
            accN += cnt[b];
          }
          if (!accN || !rightCount[b + 1]) continue;
          const double cost = area(acc) * accN + rightArea[b + 1] * rightCount[b + 1];
          if (cost < bestCost) {
            bestCost = cost;
            axis = k;
            cut = b;
          }
        }
      }
    }
    if (axis >= 0) {
      const float ext = ch[axis] - cl[axis], base = cl[axis];
      const int a = axis, c = cut;
      auto mid = std::stable_partition(prims.begin() + lo, prims.begin() + hi, [&](const Prim& p) {
        return std::min(kBins - 1, (int)((p.c[a] - base) / ext * kBins)) <= c;
      });
      const int m = (int)(mid - prims.begin());
      if (m > lo && m < hi) return m;
    }
    // median along the widest centre extent (ties by primitive id)
    int k = 0;
    for (int j = 1; j < 3; j++)
      if (ch[j] - cl[j] > ch[k] - cl[k]) k = j;
    const int m = lo + n / 2;
    std::nth_element(prims.begin() + lo, prims.begin() + m, prims.begin() + hi, [&](const Prim& p, const Prim& q) {
      return p.c[k] < q.c[k] || (p.c[k] == q.c[k] && p.id < q.id);
    });
    return m;
  }
  static void put(uint32_t* w, const Box& b, int ref) {
    w[0] = bits(b.lo[0]); w[1] = bits(b.lo[1]); w[2] = bits(b.lo[2]); w[3] = (uint32_t)ref;
    w[4] = bits(b.hi[0]); w[5] = bits(b.hi[1]); w[6] = bits(b.hi[2]); w[7] = 0u;
  }
  // Builds prims[lo, hi) at `depth`; returns its ref (interior index or ~id).
  int build(int lo, int hi, int depth) {
    maxDepth = std::max(maxDepth, depth);
    if (hi - lo == 1) return ~prims[lo].id;
    const int me = (int)(nodes.size() / 16);
    nodes.resize(nodes.size() + 16);
    const int m = split(lo, hi, depth);
    const Box lb = boundsOf(lo, m), rb = boundsOf(m, hi);
    const int l = build(lo, m, depth + 1);
    const int r = build(m, hi, depth + 1);
    put(&nodes[16 * (size_t)me], lb, l);
    put(&nodes[16 * (size_t)me + 8], rb, r);
    return me;
  }
};

// The hierarchy over the leaves of a heap BVH: leaves[capacity x 8] are the
// heap's leaf records {lo, count, hi, id} (node capacity + slot), order the
// slot -> primitive id map. Returns the device words, or an empty vector if
// the depth limit could not be met.
inline std::vector<uint32_t> buildAccel(const float* leaves, const unsigned* order, int capacity, int depthLimit,
                                        float emptyFar) {
  Builder b(depthLimit);
  for (int slot = 0; slot < capacity; slot++) {
    const float* q = leaves + (size_t)slot * 8;
    if (!(q[3] > 0.0f)) continue;  // unused slot (count 0)
    Prim p;
    for (int k = 0; k < 3; k++) {
      p.b.lo[k] = q[k];
      p.b.hi[k] = q[4 + k];
      p.c[k] = 0.5f * (q[k] + q[4 + k]);
    }
    p.id = (int)order[slot];
    b.prims.push_back(p);
  }
  const size_t slotsAt = 16 + 16 * (size_t)(capacity - 1);
  std::vector<uint32_t> out(slotsAt + capacity, 0u);
  Box root;
  int rootRef = 0;
  if (b.prims.empty()) {
    for (int k = 0; k < 3; k++) {
      root.lo[k] = emptyFar;
      root.hi[k] = -emptyFar;
    }
  } else {
    root = b.boundsOf(0, (int)b.prims.size());
    rootRef = b.build(0, (int)b.prims.size(), 0);
  }
  if (b.maxDepth > depthLimit) return {};
  Builder::put(&out[0], root, rootRef);
  out[7] = bits((float)b.maxDepth);
  out[8] = bits((float)(b.nodes.size() / 16));
  out[9] = bits((float)b.prims.size());
  std::copy(b.nodes.begin(), b.nodes.end(), out.begin() + 16);
  for (int slot = 0; slot < capacity; slot++)
    if (leaves[(size_t)slot * 8 + 3] > 0.0f && order[slot] < (unsigned)capacity) out[slotsAt + order[slot]] = slot;
  return out;
}

}  // namespace sah
}  // namespace arbor
