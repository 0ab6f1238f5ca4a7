// SAH hierarchy over the heap BVH's leaf boxes, for the trace entry points
// that take an `accel` (arbor.h), built on the device: no copy to the host,
// no synchronisation. The build is a binned (32 bins per axis) surface-area
// heuristic split of the primitives' leaf boxes, top-down, one kernel per
// tree level and one workgroup per node of that level; near the depth limit
// (and when no bin boundary separates the primitives) a node is split at
// the median along its widest centroid extent, ties by primitive id, which
// keeps every leaf within the limit. Every choice depends only on the set
// of primitives under a node (min/max reductions, counts, costs in double in
// a fixed order), so the bytes are deterministic. Interior nodes are
// numbered in depth-first pre-order: a node's left child follows it, and its
// right child follows the left subtree's interior nodes. The device walk is
// in arbor_trace.hpp (traceTreeSah).
#include <hip/hip_runtime.h>

#include "arbor_trace.hpp"
#include "launch.hpp"

using namespace arbor;

static_assert(kAccelSlots * 4 == 16 + 16 * (kCapacity - 1), "SAH layout: slot table offset");

namespace {

constexpr int kBins = 32;
constexpr int kBuildBlock = 256;
constexpr int kLevelBlocks = 1024;  // workgroups of a level launch, striding over its nodes

// One primitive of the build: its leaf box, box centre and id.
struct GPrim {
  float lo[3], hi[3], c[3];
  int id;
};
// One node to split: primitives [lo, hi) of the level's array, its depth and
// its pre-order index.
struct Task {
  int lo, hi, depth, index;
};
// Build workspace behind the hierarchy's words in the same allocation.
struct Workspace {
  GPrim prims[2][kCapacity];           // a level reads one, writes the other
  Task tasks[2][kCapacity / 2 + 1];    // nodes of the level and of the next
  int count[kSahDepth + 2];            // nodes per level
  int maxDepth;
  int pad[3];
};
constexpr size_t kWordsBytes = (size_t)kAccelSlots * 16 + 4 * (size_t)kCapacity;
constexpr size_t kWorkspaceAt = (kWordsBytes + 255) / 256 * 256;

__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }
__device__ __forceinline__ int ceilLog2(int n) {
  int l = 0;
  while ((1 << l) < n) l++;
  return l;
}
// Box surface area in double, as the cost needs it.
__device__ __forceinline__ double area(const float* lo, const float* hi) {
  double d[3];
  for (int k = 0; k < 3; k++) d[k] = fmax(0.0, (double)hi[k] - (double)lo[k]);
  return 2.0 * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
}
__device__ __forceinline__ int binOf(float c, float lo, float ext) {
  return min(kBins - 1, (int)((c - lo) / ext * (float)kBins));
}
// Node record words: {lo, ref} {hi, 0}.
__device__ __forceinline__ void putBox(uint32_t* w, const float* lo, const float* hi, int ref) {
  w[0] = fbits(lo[0]); w[1] = fbits(lo[1]); w[2] = fbits(lo[2]); w[3] = (uint32_t)ref;
  w[4] = fbits(hi[0]); w[5] = fbits(hi[1]); w[6] = fbits(hi[2]); w[7] = 0u;
}

// Block-wide min (isMax false) or max of one float per thread.
template <bool kMax>
__device__ float blockReduce(float v, float* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = kBuildBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const float o = red[threadIdx.x + s];
      red[threadIdx.x] = kMax ? fmaxf(red[threadIdx.x], o) : fminf(red[threadIdx.x], o);
    }
    __syncthreads();
  }
  const float r = red[0];
  __syncthreads();
  return r;
}
__device__ int blockSum(int v, int* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = kBuildBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const int r = red[0];
  __syncthreads();
  return r;
}
// Exclusive prefix sum of one int per thread (block order); returns the
// thread's prefix, *total the block's sum.
__device__ int blockScan(int v, int* red, int* total) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int d = 1; d < kBuildBlock; d <<= 1) {
    const int add = (int)threadIdx.x >= d ? red[threadIdx.x - d] : 0;
    __syncthreads();
    red[threadIdx.x] += add;
    __syncthreads();
  }
  const int incl = red[threadIdx.x];
  *total = red[kBuildBlock - 1];
  __syncthreads();
  return incl - v;
}
// Bounds of prims [lo, hi): box (which != 0) or centroid (which == 0).
__device__ void rangeBounds(const GPrim* p, int lo, int hi, bool centres, float* outLo, float* outHi, float* red) {
  for (int k = 0; k < 3; k++) {
    float a = 3.402823466e38f, b = -3.402823466e38f;
    for (int i = lo + (int)threadIdx.x; i < hi; i += kBuildBlock) {
      a = fminf(a, centres ? p[i].c[k] : p[i].lo[k]);
      b = fmaxf(b, centres ? p[i].c[k] : p[i].hi[k]);
    }
    outLo[k] = blockReduce<false>(a, red);
    outHi[k] = blockReduce<true>(b, red);
  }
}

// Zeroes the hierarchy's words and the workspace counters.
__global__ void k_sahClear(uint32_t* words, size_t nwords, Workspace* ws) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nwords; i += (size_t)gridDim.x * blockDim.x)
    words[i] = 0u;
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < kSahDepth + 2; i += blockDim.x) ws->count[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) ws->maxDepth = 0;
}

// The primitives of the used leaf slots in slot order, the slot table, the
// root record, the header counts and the first level's node.
__global__ __launch_bounds__(kBuildBlock) void k_sahSetup(const float* Nodes, const uint32_t* Order, uint32_t* words,
                                                          Workspace* ws, float emptyFar) {
  __shared__ int iscan[kBuildBlock];
  __shared__ float red[kBuildBlock];
  const float* leaves = Nodes + (size_t)kCapacity * 8;
  constexpr int kPer = kCapacity / kBuildBlock;
  const int s0 = (int)threadIdx.x * kPer;
  int mine = 0;
  for (int s = s0; s < s0 + kPer; s++) mine += leaves[(size_t)s * 8 + 3] > 0.0f ? 1 : 0;
  int total = 0;
  int at = blockScan(mine, iscan, &total);
  const uint32_t slotsAt = 16u + 16u * (uint32_t)(kCapacity - 1);
  for (int s = s0; s < s0 + kPer; s++) {
    const float* q = leaves + (size_t)s * 8;
    if (!(q[3] > 0.0f)) continue;
    GPrim p;
    for (int k = 0; k < 3; k++) {
      p.lo[k] = q[k];
      p.hi[k] = q[4 + k];
      p.c[k] = 0.5f * (q[k] + q[4 + k]);
    }
    p.id = (int)Order[s];
    ws->prims[0][at++] = p;
    if (Order[s] < (uint32_t)kCapacity) words[slotsAt + Order[s]] = (uint32_t)s;
  }
  __syncthreads();
  const int n = total;
  float lo[3], hi[3];
  if (n == 0) {
    for (int k = 0; k < 3; k++) {
      lo[k] = emptyFar;
      hi[k] = -emptyFar;
    }
  } else {
    rangeBounds(ws->prims[0], 0, n, false, lo, hi, red);
  }
  if (threadIdx.x == 0) {
    putBox(words, lo, hi, n == 1 ? ~ws->prims[0][0].id : 0);
    words[8] = fbits((float)(n > 1 ? n - 1 : 0));
    words[9] = fbits((float)n);
    if (n >= 2) {
      ws->tasks[0][0] = Task{0, n, 0, 0};
      ws->count[0] = 1;
    }
  }
}

// One level: every node of depth `level` split into its two children. A
// workgroup takes a node at a time.
__global__ __launch_bounds__(kBuildBlock) void k_sahLevel(uint32_t* words, Workspace* ws, int level, int depthLimit) {
  __shared__ float red[kBuildBlock];
  __shared__ int ired[kBuildBlock];
  __shared__ float binLo[3 * kBins][3], binHi[3 * kBins][3];
  __shared__ int binCnt[3 * kBins];
  __shared__ int choice[2];  // axis, cut
  __shared__ float stage[kBuildBlock][9];  // a chunk of primitives: centre, box
  __shared__ int stageId[kBuildBlock];
  const int ntasks = ws->count[level];
  const GPrim* in = ws->prims[level & 1];
  GPrim* out = ws->prims[(level + 1) & 1];
  const Task* tasks = ws->tasks[level & 1];
  Task* next = ws->tasks[(level + 1) & 1];
  for (int t = blockIdx.x; t < ntasks; t += gridDim.x) {
    const Task tk = tasks[t];
    const int lo = tk.lo, hi = tk.hi, n = hi - lo;
    float cl[3], ch[3];
    rangeBounds(in, lo, hi, true, cl, ch, red);
    // binned SAH over the three axes (thread k*32 + b gathers bin b of axis k)
    int axis = -1, cut = -1;
    if (tk.depth + ceilLog2(n) < depthLimit) {
      // the node's primitives pass through LDS a chunk at a time; thread
      // k * 32 + b (< 96) gathers bin b of axis k from each chunk
      const int k = threadIdx.x / kBins, b = threadIdx.x % kBins;
      const float ext = k < 3 ? ch[k] - cl[k] : 0.0f, base = k < 3 ? cl[k] : 0.0f;
      float blo[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
      float bhi[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
      int cnt = 0;
      for (int c0 = lo; c0 < hi; c0 += kBuildBlock) {
        const int i = c0 + (int)threadIdx.x;
        if (i < hi) {
          const GPrim& p = in[i];
          for (int j = 0; j < 3; j++) {
            stage[threadIdx.x][j] = p.c[j];
            stage[threadIdx.x][3 + j] = p.lo[j];
            stage[threadIdx.x][6 + j] = p.hi[j];
          }
        }
        __syncthreads();
        if (threadIdx.x < 3 * kBins && ext > 0.0f) {
          const int m = min(kBuildBlock, hi - c0);
          for (int q = 0; q < m; q++) {
            if (binOf(stage[q][k], base, ext) != b) continue;
            for (int j = 0; j < 3; j++) {
              blo[j] = fminf(blo[j], stage[q][3 + j]);
              bhi[j] = fmaxf(bhi[j], stage[q][6 + j]);
            }
            cnt++;
          }
        }
        __syncthreads();
      }
      if (threadIdx.x < 3 * kBins) {
        for (int j = 0; j < 3; j++) {
          binLo[threadIdx.x][j] = blo[j];
          binHi[threadIdx.x][j] = bhi[j];
        }
        binCnt[threadIdx.x] = cnt;
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        // the costs in a fixed order: axes 0-2, cuts 0-30, strictly lower wins
        double bestCost = 1e300;
        for (int k = 0; k < 3; k++) {
          if (!(ch[k] - cl[k] > 0.0f)) continue;
          double rightArea[kBins];
          int rightCount[kBins];
          float alo[3], ahi[3];
          int accN = 0;
          for (int b = kBins - 1; b > 0; b--) {
            const int i = k * kBins + b;
            if (binCnt[i]) {
              for (int j = 0; j < 3; j++) {
                alo[j] = accN ? fminf(alo[j], binLo[i][j]) : binLo[i][j];
                ahi[j] = accN ? fmaxf(ahi[j], binHi[i][j]) : binHi[i][j];
              }
              accN += binCnt[i];
            }
            rightArea[b] = accN ? area(alo, ahi) : 0.0;
            rightCount[b] = accN;
          }
          accN = 0;
          for (int b = 0; b < kBins - 1; b++) {
            const int i = k * kBins + b;
            if (binCnt[i]) {
              for (int j = 0; j < 3; j++) {
                alo[j] = accN ? fminf(alo[j], binLo[i][j]) : binLo[i][j];
                ahi[j] = accN ? fmaxf(ahi[j], binHi[i][j]) : binHi[i][j];
              }
              accN += binCnt[i];
            }
            if (!accN || !rightCount[b + 1]) continue;
            const double cost = area(alo, ahi) * accN + rightArea[b + 1] * rightCount[b + 1];
            if (cost < bestCost) {
              bestCost = cost;
              axis = k;
              cut = b;
            }
          }
        }
        choice[0] = axis;
        choice[1] = cut;
      }
      __syncthreads();
      axis = choice[0];
      cut = choice[1];
      __syncthreads();
    }
    int m = -1;
    if (axis >= 0) {
      // stable partition: bin <= cut to the left
      const float ext = ch[axis] - cl[axis], base = cl[axis];
      int left = 0;
      for (int i = lo + (int)threadIdx.x; i < hi; i += kBuildBlock) left += binOf(in[i].c[axis], base, ext) <= cut;
      const int nLeft = blockSum(left, ired);
      if (nLeft > 0 && nLeft < n) {
        m = lo + nLeft;
        int doneL = 0, doneR = 0;
        for (int c0 = lo; c0 < hi; c0 += kBuildBlock) {
          const int i = c0 + (int)threadIdx.x;
          const bool valid = i < hi;
          const bool isLeft = valid && binOf(in[i].c[axis], base, ext) <= cut;
          int chunkLeft = 0;
          const int posL = blockScan(isLeft ? 1 : 0, ired, &chunkLeft);
          const int chunkN = min(kBuildBlock, hi - c0);
          const int posR = ((int)threadIdx.x - posL);  // right elements before this one in the chunk
          if (valid) {
            if (isLeft) out[lo + doneL + posL] = in[i];
            else out[m + doneR + posR] = in[i];
          }
          doneL += chunkLeft;
          doneR += chunkN - chunkLeft;
        }
      }
    }
    if (m < 0) {
      // the median along the widest centroid extent, ties by primitive id:
      // every primitive goes to its rank in (centre, id) order (the keys
      // pass through LDS a chunk at a time)
      int k = 0;
      for (int j = 1; j < 3; j++)
        if (ch[j] - cl[j] > ch[k] - cl[k]) k = j;
      m = lo + n / 2;
      for (int ib = lo; ib < hi; ib += kBuildBlock) {
        const int i = ib + (int)threadIdx.x;
        const float ci = i < hi ? in[i].c[k] : 0.0f;
        const int idi = i < hi ? in[i].id : 0;
        int rank = 0;
        for (int jb = lo; jb < hi; jb += kBuildBlock) {
          const int j = jb + (int)threadIdx.x;
          if (j < hi) {
            stage[threadIdx.x][0] = in[j].c[k];
            stageId[threadIdx.x] = in[j].id;
          }
          __syncthreads();
          const int cntj = min(kBuildBlock, hi - jb);
          for (int q = 0; q < cntj; q++) {
            const float cj = stage[q][0];
            rank += (cj < ci || (cj == ci && stageId[q] < idi)) ? 1 : 0;
          }
          __syncthreads();
        }
        if (i < hi) out[lo + rank] = in[i];
      }
    }
    __syncthreads();
    float llo[3], lhi[3], rlo[3], rhi[3];
    rangeBounds(out, lo, m, false, llo, lhi, red);
    rangeBounds(out, m, hi, false, rlo, rhi, red);
    if (threadIdx.x == 0) {
      const int me = tk.index, nl = m - lo, nr = hi - m;
      uint32_t* w = words + 16 + 16 * (size_t)me;
      putBox(w, llo, lhi, nl == 1 ? ~out[lo].id : me + 1);
      putBox(w + 8, rlo, rhi, nr == 1 ? ~out[m].id : me + nl);
      if (nl >= 2) next[atomicAdd(&ws->count[level + 1], 1)] = Task{lo, m, tk.depth + 1, me + 1};
      if (nr >= 2) next[atomicAdd(&ws->count[level + 1], 1)] = Task{m, hi, tk.depth + 1, me + nl};
      if (nl == 1 || nr == 1) atomicMax(&ws->maxDepth, tk.depth + 1);
    }
    __syncthreads();
  }
}

// Header: the depth of the deepest leaf.
__global__ void k_sahFinish(uint32_t* words, const Workspace* ws) {
  if (threadIdx.x == 0 && blockIdx.x == 0) words[7] = fbits((float)ws->maxDepth);
}

int buildAccel(const float* Nodes, const unsigned* Order, void* accel, int depthLimit, hipStream_t stream) {
  uint32_t* words = static_cast<uint32_t*>(accel);
  Workspace* ws = reinterpret_cast<Workspace*>(static_cast<char*>(accel) + kWorkspaceAt);
  hipLaunchKernelGGL(k_sahClear, dim3(512), dim3(256), 0, stream, words, kWordsBytes / 4, ws);
  int rc = launch_status();
  if (rc != ARBOR_OK) return rc;
  hipLaunchKernelGGL(k_sahSetup, dim3(1), dim3(kBuildBlock), 0, stream, Nodes, Order, words, ws, kFar);
  if ((rc = launch_status()) != ARBOR_OK) return rc;
  for (int level = 0; level < depthLimit; level++) {
    hipLaunchKernelGGL(k_sahLevel, dim3(kLevelBlocks), dim3(kBuildBlock), 0, stream, words, ws, level, depthLimit);
    if ((rc = launch_status()) != ARBOR_OK) return rc;
  }
  hipLaunchKernelGGL(k_sahFinish, dim3(1), dim3(64), 0, stream, words, ws);
  return launch_status();
}

}  // namespace

static_assert(kCapacity % kBuildBlock == 0, "setup: slots per thread");
static_assert(kWorkspaceAt % alignof(Workspace) == 0, "workspace alignment");

extern "C" {

long long arbor_trace_accel_bytes(void) { return (long long)(kWorkspaceAt + sizeof(Workspace)); }

int arbor_buildTraceAccel(const float* Nodes, const unsigned* Order, void* accel, hipStream_t stream) {
  ARBOR_REQUIRE(Nodes && Order && accel);
  return buildAccel(Nodes, Order, accel, kSahDepth, stream);
}

int arbor_buildTraceAccelLimit(const float* Nodes, const unsigned* Order, void* accel, int depth_limit,
                               hipStream_t stream) {
  ARBOR_REQUIRE(Nodes && Order && accel);
  if (depth_limit < 15 || depth_limit > kSahDepth) return ARBOR_ERR_ARG;
  return buildAccel(Nodes, Order, accel, depth_limit, stream);
}

}  // extern "C"
