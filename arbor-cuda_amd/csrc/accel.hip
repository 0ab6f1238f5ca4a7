// SAH hierarchy over the heap BVH's leaf boxes, for the trace entry points
// that take an `accel` (arbor.h). Built on the host once per scene: the leaf
// boxes and the slot -> primitive order are copied from the device, a binned
// surface-area-heuristic tree is built over them and the node array is
// uploaded. The device walk and the layout are in arbor_trace.hpp
// (traceTreeSah). The build is deterministic: the same Nodes and Order give
// the same bytes.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "arbor_trace.hpp"
#include "launch.hpp"

namespace arbor {
namespace {

struct HBox {
  float lo[3], hi[3];
};
struct HPrim {
  HBox b;
  float c[3];  // box centre
  int id;
};

constexpr int kBins = 32;

HBox unite(const HBox& a, const HBox& b) {
  HBox r;
  for (int k = 0; k < 3; k++) {
    r.lo[k] = std::min(a.lo[k], b.lo[k]);
    r.hi[k] = std::max(a.hi[k], b.hi[k]);
  }
  return r;
}
double area(const HBox& b) {
  double d[3];
  for (int k = 0; k < 3; k++) d[k] = std::max(0.0, (double)b.hi[k] - (double)b.lo[k]);
  return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
}
float refBits(int ref) {
  float f;
  memcpy(&f, &ref, sizeof f);
  return f;
}
int ceilLog2(int n) {
  int l = 0;
  while ((1 << l) < n) l++;
  return l;
}

struct Builder {
  std::vector<HPrim> prims;
  std::vector<float4> nodes;  // 4 per interior node
  int maxDepth = 0;

  HBox boundsOf(int lo, int hi) const {
    HBox b = prims[lo].b;
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
    // every leaf within kSahDepth of the root.
    if (depth + ceilLog2(n) < kSahDepth) {
      double bestCost = 1e300;
      for (int k = 0; k < 3; k++) {
        const float ext = ch[k] - cl[k];
        if (!(ext > 0.0f)) continue;
        HBox bin[kBins];
        int cnt[kBins] = {0};
        for (int i = lo; i < hi; i++) {
          int b = std::min(kBins - 1, (int)((prims[i].c[k] - cl[k]) / ext * kBins));
          bin[b] = cnt[b] ? unite(bin[b], prims[i].b) : prims[i].b;
          cnt[b]++;
        }
        double rightArea[kBins];
        int rightCount[kBins];
        HBox acc;
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
      auto mid = std::stable_partition(prims.begin() + lo, prims.begin() + hi, [&](const HPrim& p) {
        return std::min(kBins - 1, (int)((p.c[a] - base) / ext * kBins)) <= c;
      });
      int m = (int)(mid - prims.begin());
      if (m > lo && m < hi) return m;
    }
    // median along the widest centre extent (ties by primitive id)
    int k = 0;
    for (int j = 1; j < 3; j++)
      if (ch[j] - cl[j] > ch[k] - cl[k]) k = j;
    const int m = lo + n / 2;
    std::nth_element(prims.begin() + lo, prims.begin() + m, prims.begin() + hi, [&](const HPrim& p, const HPrim& q) {
      return p.c[k] < q.c[k] || (p.c[k] == q.c[k] && p.id < q.id);
    });
    return m;
  }
  // Builds prims[lo, hi) at `depth`; returns its ref (interior index or ~id).
  int build(int lo, int hi, int depth) {
    maxDepth = std::max(maxDepth, depth);
    if (hi - lo == 1) return ~prims[lo].id;
    const int me = (int)(nodes.size() / 4);
    nodes.resize(nodes.size() + 4);
    const int m = split(lo, hi, depth);
    const HBox lb = boundsOf(lo, m), rb = boundsOf(m, hi);
    const int l = build(lo, m, depth + 1);
    const int r = build(m, hi, depth + 1);
    float4* q = &nodes[4 * (size_t)me];
    q[0] = make_float4(lb.lo[0], lb.lo[1], lb.lo[2], refBits(l));
    q[1] = make_float4(lb.hi[0], lb.hi[1], lb.hi[2], 0.0f);
    q[2] = make_float4(rb.lo[0], rb.lo[1], rb.lo[2], refBits(r));
    q[3] = make_float4(rb.hi[0], rb.hi[1], rb.hi[2], 0.0f);
    return me;
  }
};

}  // namespace
}  // namespace arbor

using namespace arbor;

extern "C" {

long long arbor_trace_accel_bytes(void) { return (long long)kAccelSlots * 16 + 4LL * kCapacity; }

int arbor_buildTraceAccel(const float* Nodes, const unsigned* Order, void* accel, hipStream_t stream) {
  ARBOR_REQUIRE(Nodes && Order && accel);
  std::vector<float> leaves((size_t)kCapacity * 8);
  std::vector<unsigned> order(kCapacity);
  if (hipMemcpyAsync(leaves.data(), Nodes + (size_t)kCapacity * 8, leaves.size() * sizeof(float),
                     hipMemcpyDeviceToHost, stream) != hipSuccess ||
      hipMemcpyAsync(order.data(), Order, order.size() * sizeof(unsigned), hipMemcpyDeviceToHost, stream) !=
          hipSuccess ||
      hipStreamSynchronize(stream) != hipSuccess)
    return ARBOR_ERR_DEVICE;
  Builder b;
  for (int slot = 0; slot < kCapacity; slot++) {
    const float* q = &leaves[(size_t)slot * 8];
    if (!(q[3] > 0.0f)) continue;  // unused slot (count 0)
    HPrim p;
    for (int k = 0; k < 3; k++) {
      p.b.lo[k] = q[k];
      p.b.hi[k] = q[4 + k];
      p.c[k] = 0.5f * (q[k] + q[4 + k]);
    }
    p.id = (int)order[slot];
    b.prims.push_back(p);
  }
  std::vector<float4> out(kAccelHeader, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
  int rootRef = 0;
  HBox root;
  if (b.prims.empty()) {
    for (int k = 0; k < 3; k++) {
      root.lo[k] = kFar;
      root.hi[k] = -kFar;
    }
  } else {
    root = b.boundsOf(0, (int)b.prims.size());
    rootRef = b.build(0, (int)b.prims.size(), 0);
  }
  if (b.maxDepth > kSahDepth) return ARBOR_ERR_STATE;  // cannot happen: the split keeps the depth bound
  out[0] = make_float4(root.lo[0], root.lo[1], root.lo[2], refBits(rootRef));
  out[1] = make_float4(root.hi[0], root.hi[1], root.hi[2], (float)b.maxDepth);
  out[2] = make_float4((float)(b.nodes.size() / 4), (float)b.prims.size(), 0.0f, 0.0f);
  out.insert(out.end(), b.nodes.begin(), b.nodes.end());
  // heap leaf slot of each primitive id (the tie order of the walk)
  std::vector<unsigned> slotOf(kCapacity, 0u);
  for (int slot = 0; slot < kCapacity; slot++)
    if (leaves[(size_t)slot * 8 + 3] > 0.0f && order[slot] < (unsigned)kCapacity) slotOf[order[slot]] = (unsigned)slot;
  char* dst = static_cast<char*>(accel);
  if (hipMemcpyAsync(dst, out.data(), out.size() * sizeof(float4), hipMemcpyHostToDevice, stream) != hipSuccess ||
      hipMemcpyAsync(dst + (size_t)kAccelSlots * 16, slotOf.data(), slotOf.size() * sizeof(unsigned),
                     hipMemcpyHostToDevice, stream) != hipSuccess ||
      hipStreamSynchronize(stream) != hipSuccess)
    return ARBOR_ERR_DEVICE;
  return ARBOR_OK;
}

}  // extern "C"
