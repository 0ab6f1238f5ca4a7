// Per-frame kernels of the ARBOR render path on gfx950: primary visibility,
// shading, the same-surface lighting filter and the progressive resolve.
//
// Visibility and shading run one wave64 per 8x8 pixel tile; a block of waves
// builds the 324-tree instancing table (and, per walk policy, the top heap
// levels) in LDS once and then takes tiles in a grid-stride loop. The trace
// kernel is instantiated for several walk policies and block sizes, all with
// identical results; the tuning key "trace_variant" selects one. Filter and
// resolve are HBM-bound per-pixel kernels over float4 rows.
#include <limits.h>
#include <stdio.h>
#include <math.h>
#include <string.h>

#include <hip/hip_cooperative_groups.h>

#include <atomic>
#include <mutex>

#include "arbor_camera.hpp"
#include "arbor_features.hpp"
#include "arbor_filter.hpp"
#include "arbor_persist.hpp"
#include "arbor_shade.hpp"
#include "arbor_split.hpp"
#include "arbor_wavefront.hpp"
#include "engine_passes.hpp"
#include "launch.hpp"
#include "tuning.hpp"

namespace arbor {

constexpr int kTileBlock = 256;  // 4 waves, one 8x8 tile each (shade)
constexpr int kWavesPerBlock = kTileBlock / 64;

__device__ __forceinline__ void fillTrees(TreeInst* trees, int ws) {
  for (int i = threadIdx.x; i < kTrees; i += blockDim.x) trees[i] = makeTree(i, ws);
}
// The tree records a kernel reads: its LDS copy, or the device-memory ones.
template <bool kRecTrees>
__device__ __forceinline__ auto treeSource(const TreeInst* lds, const float4* recs) {
  if constexpr (kRecTrees) return TreeRecs{recs};
  else return lds;
}
template <class W>
__device__ __forceinline__ void fillTop(float4* top, const float4* Nodes4) {
  if (W::ldsTop)
    for (int i = threadIdx.x; i < 2 * kTopNodes; i += blockDim.x) top[i] = Nodes4[i];
}

// tracePixels (Arbor.cu:252) for every pixel of the view. With a tile
// schedule, work item i is tile tileOrder[i] and each wave records how long its
// tile took (wall clock ticks, >= 1) in tileCost; the order changes only which
// wave traces which tile, never a pixel's result.
// With a SAH walk policy, accel is the hierarchy of arbor_buildTraceAccel.
// kMinWaves > 1 caps the registers at that many waves per SIMD. With
// kRecTrees the tree records come from treeRecs (device memory, the tile
// schedule's copy) instead of LDS, which then holds only the walk's stack.
template <int kWaves, class W, int kMinWaves = 1, bool kRecTrees = false>
__global__ __launch_bounds__(kWaves * 64) __attribute__((amdgpu_waves_per_eu(kMinWaves))) void k_tracePixels(const float* P, const float4* Nodes4, const uint32_t* Order,
                                                             const float* C, float4* Hit, int width, int height,
                                                             int tilesX, int tiles, const uint32_t* tileOrder,
                                                             uint32_t* tileCost, const float4* accel,
                                                             const float4* treeRecs) {
  __shared__ TreeInst ldsTrees[kRecTrees ? 1 : kTrees];
  __shared__ float4 top[W::ldsTop ? 2 * kTopNodes : 1];
  if constexpr (!kRecTrees) fillTrees(ldsTrees, (int)C[31]);
  fillTop<W>(top, Nodes4);
  __syncthreads();
  const auto trees = treeSource<kRecTrees>(ldsTrees, treeRecs);
  const View view = makeView(C);
  const float cone = C[14] / (float)height, time = C[8], wind = C[11];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = blockIdx.x * kWaves + wave; i < tiles; i += gridDim.x * kWaves) {
    const int t = tileOrder ? (int)tileOrder[i] : i;
    const long long start = tileCost ? wall_clock64() : 0;
    int x = (t % tilesX) * 8 + (lane & 7), y = (t / tilesX) * 8 + (lane >> 3);
    if (x < width && y < height) {
      v3 ro = view.eye, rd = viewRay(view, x, y, width, height);
      float ground = rd.y < -0.00001f ? -ro.y / rd.y : kFar;
      int tree = -1;
      v2 h = traceForest<W>(P, Nodes4, W::sah ? accel : top, Order, trees, ro, rd, cone, time, wind,
                            ground > 0.0f ? ground : kFar, &tree, false);
      Hit[y * width + x] = make_float4(h.x, h.y, (float)tree, C[14] * h.x / (float)height);
    }
    if (tileCost && lane == 0) {
      long long ticks = wall_clock64() - start;
      tileCost[t] = ticks < 1 ? 1u : (ticks > 0x7fffffffLL ? 0x7fffffffu : (uint32_t)ticks);
    }
  }
}

// Sun visibility of every pixel (the shadow traversals of shadePixels,
// Arbor.cu:271-274) as a pass of its own: Vis[p] = {transmission, 1}. It
// runs the trace kernel's launch shape (one wave64 per 8x8 tile, the walk
// inlined, optional tile schedule), so shading's register budget no longer
// limits how many shadow traversals are in flight.
template <int kWaves, class W, int kMinWaves = 1, bool kRecTrees = false>
__global__ __launch_bounds__(kWaves * 64) __attribute__((amdgpu_waves_per_eu(kMinWaves))) void k_sunVisibility(const float* P, const float4* Nodes4,
                                                               const uint32_t* Order, const float* C, const float4* Hit,
                                                               float4* Vis, int width, int height, int tilesX,
                                                               int tiles, const uint32_t* tileOrder,
                                                               uint32_t* tileCost, const float4* accel,
                                                               const float4* treeRecs) {
  __shared__ TreeInst ldsTrees[kRecTrees ? 1 : kTrees];
  __shared__ float4 top[W::ldsTop ? 2 * kTopNodes : 1];
  if (!(C[22] > 0.5f)) return;  // visibility off: shading never reads Vis
  if constexpr (!kRecTrees) fillTrees(ldsTrees, (int)C[31]);
  fillTop<W>(top, Nodes4);
  __syncthreads();
  const auto trees = treeSource<kRecTrees>(ldsTrees, treeRecs);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = blockIdx.x * kWaves + wave; i < tiles; i += gridDim.x * kWaves) {
    const int t = tileOrder ? (int)tileOrder[i] : i;
    const long long start = tileCost ? wall_clock64() : 0;
    int x = (t % tilesX) * 8 + (lane & 7), y = (t / tilesX) * 8 + (lane >> 3);
    if (x < width && y < height) {
      const int pixel = y * width + x;
      // per tile from C: the context is dead during the traversal instead of live across the loop
      const ShadeCtx sc = makeShadeCtx(C, width, height);
      v3 v = sunVisibilityOne<W>(P, Nodes4, W::sah ? accel : top, Order, trees, sc, x, y, Hit[pixel]);
      Vis[pixel] = make_float4(v.x, v.y, v.z, 1.0f);
    }
    if (tileCost && lane == 0) {
      long long ticks = wall_clock64() - start;
      tileCost[t] = ticks < 1 ? 1u : (ticks > 0x7fffffffLL ? 0x7fffffffu : (uint32_t)ticks);
    }
  }
}

// Tile schedule (longest predicted first) for the next tracePixels launch. A
// tile's predicted cost is the largest recorded cost in the 3x3 tiles around
// where its centre ray fell in the previous frame (the direction projected
// through the previous camera; eye motion between frames is ignored); tiles
// with no prediction (first frame, or directions outside the previous view)
// go first. Tiles are bucketed by log2(cost) with two fraction bits and laid
// out bucket by bucket, highest first: keys and a bucket histogram, then
// (after one grid barrier) the bucket starts, which every block derives from
// the histogram, and the scatter. A bucket's 64-bit accumulator holds its
// tile count in the low word and the scatter cursor in the high word; calls
// alternate between two accumulator sets, each call zeroing the set the next
// call uses, so nothing waits for a reset. The call keeps its camera for the
// next frame's reprojection.
constexpr int kOrderBuckets = 128;
constexpr int kOrderBlock = 256;
struct TileSchedule {
  float prevC[64];
  uint32_t valid;
  uint32_t parity;      // accumulator set of the next call
  uint32_t persist[2];  // tile counter and finished-wave count of the persistent trace (zero between launches)
  uint32_t pad[60];
  unsigned long long acc[2][kOrderBuckets];  // low word: tiles in the bucket; high word: scatter cursor
  float4 trees[kTreeRecF4 * kTrees];         // tree records of the world seed C[31] (storeTreeRec)
  // followed by cost[tiles], order[tiles], key[tiles]
};
static_assert(sizeof(TileSchedule) == 2560 + 16 * kTreeRecF4 * kTrees, "schedule header");
__device__ inline int tileKey(const View& cur, const View& prev, bool valid, const uint32_t* cost, int t, int tilesX,
                              int tilesY, int width, int height) {
  if (!valid) return kOrderBuckets - 1;
  int x = (t % tilesX) * 8 + 4, y = (t / tilesX) * 8 + 4;
  v3 d = viewRay(cur, x, y, width, height);
  float fz = dot(d, prev.f);
  if (fz <= 0.000001f) return kOrderBuckets - 1;
  float u = dot(d, prev.r) / (fz * prev.scale), w = dot(d, prev.up) / (fz * prev.scale);
  float px = u * (float)height + (float)width * 0.5f, py = (float)height * 0.5f - w * (float)height;
  if (!(px >= 0.0f && py >= 0.0f && px < (float)width && py < (float)height)) return kOrderBuckets - 1;
  int tx = (int)(px * 0.125f), ty = (int)(py * 0.125f);
  uint32_t c = 1;
  for (int j = max(0, ty - 1); j <= min(tilesY - 1, ty + 1); j++)
    for (int i = max(0, tx - 1); i <= min(tilesX - 1, tx + 1); i++) c = max(c, cost[j * tilesX + i]);
  int k = 31 - __clz((int)c);
  int sub = k >= 2 ? (int)((c >> (k - 2)) & 3u) : 0;
  return min(kOrderBuckets - 2, 4 * k + sub);
}
// Ordering, before the grid barrier: every tile's bucket key and the bucket
// histogram into accumulator set `par`; block 0 zeroes the other set.
// `count` is a kOrderBuckets LDS array of the block; cam the current camera.
__device__ void orderKeys(const float* cam, TileSchedule* s, int par, int tilesX, int tilesY, int width, int height,
                          uint32_t* count) {
  const int tiles = tilesX * tilesY;
  const uint32_t* cost = reinterpret_cast<const uint32_t*>(s + 1);
  uint32_t* key = reinterpret_cast<uint32_t*>(s + 1) + 2 * (size_t)tiles;
  for (int b = threadIdx.x; b < kOrderBuckets; b += blockDim.x) count[b] = 0;
  if (blockIdx.x == 0)
    for (int b = threadIdx.x; b < kOrderBuckets; b += blockDim.x) s->acc[par ^ 1][b] = 0ull;
  __syncthreads();
  const View cur = makeView(cam), prev = makeView(s->prevC);
  const bool valid = s->valid == 1u;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < tiles; t += gridDim.x * blockDim.x) {
    int k = tileKey(cur, prev, valid, cost, t, tilesX, tilesY, width, height);
    key[t] = (uint32_t)k;
    atomicAdd(&count[k], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < kOrderBuckets; b += blockDim.x)
    if (count[b]) atomicAdd(&s->acc[par][b], (unsigned long long)count[b]);
  __syncthreads();
}
// Ordering, after the grid barrier: the bucket starts (highest bucket first)
// from the histogram, then the tiles scattered bucket by bucket into the
// order, a chunk of blockDim tiles at a time (the chunk loop's bound is
// uniform per block). start, count and base are kOrderBuckets LDS arrays.
__device__ void orderScatter(TileSchedule* s, int par, int tiles, uint32_t* start, uint32_t* count,
                             uint32_t* base) {
  const uint32_t* key = reinterpret_cast<const uint32_t*>(s + 1) + 2 * (size_t)tiles;
  uint32_t* order = reinterpret_cast<uint32_t*>(s + 1) + tiles;
  unsigned long long* acc = s->acc[par];
  // start[b] = number of tiles in buckets above b: a suffix scan of the counts
  for (int b = threadIdx.x; b < kOrderBuckets; b += blockDim.x) start[b] = (uint32_t)(acc[b] & 0xffffffffull);
  __syncthreads();
  for (int d = 1; d < kOrderBuckets; d <<= 1) {
    uint32_t v = 0;
    const bool mine = threadIdx.x < kOrderBuckets;
    if (mine) v = start[threadIdx.x] + (threadIdx.x + d < kOrderBuckets ? start[threadIdx.x + d] : 0u);
    __syncthreads();
    if (mine) start[threadIdx.x] = v;
    __syncthreads();
  }
  for (int b = threadIdx.x; b < kOrderBuckets; b += blockDim.x)
    start[b] -= (uint32_t)(acc[b] & 0xffffffffull);  // inclusive -> exclusive
  __syncthreads();
  for (int c0 = blockIdx.x * blockDim.x; c0 < tiles; c0 += gridDim.x * blockDim.x) {
    for (int b = threadIdx.x; b < kOrderBuckets; b += blockDim.x) count[b] = 0;
    __syncthreads();
    const int t = c0 + threadIdx.x;
    uint32_t k = 0, local = 0;
    if (t < tiles) {
      k = key[t];
      local = atomicAdd(&count[k], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kOrderBuckets; b += blockDim.x)
      base[b] = count[b] ? start[b] + (uint32_t)(atomicAdd(&acc[b], (unsigned long long)count[b] << 32) >> 32)
                         : 0u;
    __syncthreads();
    const uint32_t at = base[k] + local;
    if (t < tiles && at < (uint32_t)tiles) order[at] = (uint32_t)t;  // always true for consistent inputs
    __syncthreads();
  }
}
// Frame prologue: optionally updateCamera (I != nullptr) and the per-sample
// camera slots (sample >= 0: C[9] = sample, C[10] = sample == 0, and C[22] =
// light unless light < 0), then the tile order of one or two schedules (s1
// may be null), in one cooperative launch with one grid barrier. Every block
// computes the new camera itself in LDS from the old C (the same bits in
// every block); block 0 stores it after the barrier, when no block reads the
// old one any more, and keeps it in each schedule for the next reprojection.
// Each schedule also gets the tree records of the world seed C[31], which
// the kRecTrees trace and visibility kernels read.
__global__ __launch_bounds__(kOrderBlock) void k_framePrologue(float* C, const float* I, float delta, int sample,
                                                               int light, int width, int height, TileSchedule* s0,
                                                               TileSchedule* s1) {
  __shared__ float cam[64];
  __shared__ uint32_t la[kOrderBuckets], lb[kOrderBuckets], lc[kOrderBuckets];
  __shared__ int par[2];
  const int tilesX = (width + 7) / 8, tilesY = (height + 7) / 8, tiles = tilesX * tilesY;
  if (threadIdx.x < 64) cam[threadIdx.x] = C[threadIdx.x];
  if (threadIdx.x == 0) {
    par[0] = (int)(s0->parity & 1u);
    par[1] = s1 ? (int)(s1->parity & 1u) : 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (I) updateCameraState(cam, I, delta, width, height);
    if (sample >= 0) {
      cam[9] = (float)sample;
      cam[10] = sample == 0 ? 1.0f : 0.0f;
      if (light >= 0) cam[22] = light ? 1.0f : 0.0f;
    }
  }
  __syncthreads();
  // the tree records of the world seed, a few trees per block
  for (int i = threadIdx.x * gridDim.x + blockIdx.x; i < kTrees; i += blockDim.x * gridDim.x) {
    const TreeInst t = makeTree(i, (int)cam[31]);
    storeTreeRec(s0->trees, i, t);
    if (s1) storeTreeRec(s1->trees, i, t);
  }
  const int p0 = par[0], p1 = par[1];
  orderKeys(cam, s0, p0, tilesX, tilesY, width, height, la);
  if (s1) orderKeys(cam, s1, p1, tilesX, tilesY, width, height, la);
  cooperative_groups::this_grid().sync();
  orderScatter(s0, p0, tiles, la, lb, lc);
  if (s1) orderScatter(s1, p1, tiles, la, lb, lc);
  if (blockIdx.x == 0) {
    if (threadIdx.x < 64) {
      const float c = cam[threadIdx.x];
      if (I || sample >= 0) C[threadIdx.x] = c;
      s0->prevC[threadIdx.x] = c;
      if (s1) s1->prevC[threadIdx.x] = c;
    }
    if (threadIdx.x == 0) {
      s0->valid = 1u;
      s0->parity = (uint32_t)(p0 ^ 1);
      if (s1) {
        s1->valid = 1u;
        s1->parity = (uint32_t)(p1 ^ 1);
      }
    }
  }
}

// Single-specimen primary rays: traceTree in tree-local space (no instancing).
template <int kWaves, class W>
__global__ __launch_bounds__(kWaves * 64) void k_traceTreePixels(const float* P, const float4* Nodes4,
                                                                 const uint32_t* Order, const float* C, float4* Hit,
                                                                 int width, int height, int tilesX, int tiles,
                                                                 const uint32_t*, uint32_t*, const float4* accel,
                                                                 const float4*) {
  __shared__ float4 top[W::ldsTop ? 2 * kTopNodes : 1];
  fillTop<W>(top, Nodes4);
  __syncthreads();
  const View view = makeView(C);
  const float cone = C[14] / (float)height, time = C[8], wind = C[11];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int t = blockIdx.x * kWaves + wave; t < tiles; t += gridDim.x * kWaves) {
    int x = (t % tilesX) * 8 + (lane & 7), y = (t / tilesX) * 8 + (lane >> 3);
    if (x >= width || y >= height) continue;
    v3 ro = view.eye, rd = viewRay(view, x, y, width, height);
    float ground = rd.y < -0.00001f ? -ro.y / rd.y : kFar;
    v2 h = traceTree<W>(P, Nodes4, W::sah ? accel : top, Order, ro, rd, cone, time, wind,
                        ground > 0.0f ? ground : kFar, -1, false);
    Hit[y * width + x] = make_float4(h.x, h.y, -1.0f, C[14] * h.x / (float)height);
  }
}

// tracePixels with the wave-scheduled lane state machine: each block owns
// `raysPerBlock` consecutive rays (tile-major) that its waves pull from an
// LDS counter.
template <int kWaves>
__global__ __launch_bounds__(kWaves * 64) void k_tracePixelsSched(const float* P, const float4* Nodes4,
                                                                  const uint32_t* Order, const float* C, float4* Hit,
                                                                  int width, int height, int tilesX, int tiles,
                                                                  int raysPerBlock, WaveSchedule ws) {
  __shared__ TreeInst trees[kTrees];
  __shared__ int pool;
  fillTrees(trees, (int)C[31]);
  const int first = blockIdx.x * raysPerBlock;
  if (threadIdx.x == 0) pool = first;
  __syncthreads();
  const int end = min(tiles * 64, first + raysPerBlock);
  ViewCtx vc;
  vc.view = makeView(C);
  vc.cone = C[14] / (float)height;
  vc.time = C[8];
  vc.wind = C[11];
  vc.width = width;
  vc.height = height;
  vc.tilesX = tilesX;
  waveTrace(&pool, end, P, Nodes4, Order, trees, vc, Hit, C[14], ws);
}

// The trace configurations the tuning key "trace_variant" selects from.
using WalkPlain = Walk<false, false, false>;
using WalkCull = Walk<true, false, false>;
using WalkCullPrefetch = Walk<true, true, false>;
using WalkCullTop = Walk<true, false, true>;
using WalkAll = Walk<true, true, true>;
using WalkShade = WalkCull;
using WalkSah = Walk<true, false, false, true>;  // the accel entry points
template <int K>
using WalkSahL = Walk<true, false, false, true, K>;  // stack entries [0, K) in LDS
template <int K, int G>
using WalkSahG = Walk<true, false, false, true, K, G>;  // the same, object tests gathered per wave
template <int K, int G, int S>
using WalkSahGS = Walk<true, false, false, true, K, G, S>;  // the same, S interior steps per gathered round
constexpr int sahLdsPerWave(int k) { return k * 64 * 8; }

// The stage-split trace's launches (arbor_split.hpp): one-wave blocks, the
// tree records from the tile schedule, the walk's stack entries in LDS.
__global__ __launch_bounds__(64) void k_splitStart(const float* C, float4* Hit, int width, int height, int tilesX,
                                                   int tiles, const float4* treeRecs, SplitBuffers b) {
  splitStart(TreeRecs{treeRecs}, C, Hit, width, height, tilesX, tiles, b);
}
template <class W, bool kFinish>
__global__ __launch_bounds__(64) void k_splitRound(const float* P, const float4* Nodes4, const uint32_t* Order,
                                                   const float* C, float4* Hit, int width, int height,
                                                   const float4* accel, const float4* treeRecs, SplitBuffers b,
                                                   int round) {
  splitRound<W, kFinish>(P, Nodes4, Order, TreeRecs{treeRecs}, C, Hit, width, height, accel, b, round);
}

// shadePixels (Arbor.cu:277-289).
// kTraceVis = false: Vis is given whenever sun visibility is on (the shadow
// traversals are compiled out, and with them the call's register cost).
// With Feat, each pixel's compact denoiser features (compactFeatures of its
// Linear, Guide and Hit distance) are written in place of Linear and Guide.
template <bool kTraceVis>
__global__ __launch_bounds__(kTileBlock) __attribute__((amdgpu_waves_per_eu(kTraceVis ? 1 : 3))) void k_shadePixels(const float* P, const float4* Nodes4, const uint32_t* Order,
                                                            const float* C, const float4* Hit, float4* Linear,
                                                            float4* Guide, int width, int height, int tilesX,
                                                            int tiles, float4* Normal, const float4* Vis,
                                                            featx8* Feat) {
  __shared__ TreeInst trees[kTrees];
  fillTrees(trees, (int)C[31]);
  __syncthreads();
  const ShadeCtx sc = makeShadeCtx(C, width, height);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int t = blockIdx.x * kWavesPerBlock + wave; t < tiles; t += gridDim.x * kWavesPerBlock) {
    int x = (t % tilesX) * 8 + (lane & 7), y = (t / tilesX) * 8 + (lane >> 3);
    if (x >= width || y >= height) continue;
    int pixel = y * width + x;
    float4 l, g;
    v3 n;
    const float4 hit = Hit[pixel];
    shadeOne<WalkShade, kTraceVis>(P, Nodes4, Order, trees, sc, x, y, hit, l, g, Normal ? &n : nullptr, Vis);
    if (Feat) {
      Feat[pixel] = compactFeatures(l, g, hit.x);
      continue;
    }
    Linear[pixel] = l;
    Guide[pixel] = g;
    if (Normal) Normal[pixel] = make_float4(n.x, n.y, n.z, 0.0f);
  }
}

// filterLighting (Arbor.cu:292-302); with kResolve also resolvePixels.
template <bool kResolve>
__global__ __launch_bounds__(kFilterW * kFilterH) void k_filterLighting(const float4* Linear, const float4* Guide,
                                                                       const float4* Hit, float4* Filtered,
                                                                       const float* C, int width, int height,
                                                                       FilterWeights fw, float4* History,
                                                                       uint32_t* Pixels) {
  filterBlock<kResolve>(Linear, Guide, Hit, Filtered, C, width, height, fw, History, Pixels, blockIdx.x, blockIdx.y);
}

// resolvePixels (Arbor.cu:304).
__global__ __launch_bounds__(256) void k_resolvePixels(const float4* Filtered, float4* History, uint32_t* Pixels,
                                                       const float* C, int width, int height) {
  int x = (int)(blockIdx.x * blockDim.x + threadIdx.x), y = (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (x >= width || y >= height) return;
  int i = y * width + x;
  float4 f = Filtered[i];
  resolveOne(mk(f.x, f.y, f.z), History, Pixels, C, x, y, width, height, i);
}

}  // namespace arbor

using namespace arbor;

namespace arbor {
// Compute units per device, read once per device and kept in atomics (any
// host thread may launch).
int deviceCUs() {
  static std::atomic<int> cus[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  int n = cus[dev].load(std::memory_order_relaxed);
  if (n <= 0) {
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cus[dev].store(n, std::memory_order_relaxed);
  }
  return n;
}
}  // namespace arbor

namespace {
// Grid for the tile kernels: one wave per 8x8 tile up to the cap (a tuning
// knob); beyond it waves take further tiles in a grid-stride loop.
inline int tileGrid(int tiles, int wavesPerBlock) {
  int blocks = (tiles + wavesPerBlock - 1) / wavesPerBlock;
  int cap = tuning().grid_cap;
  return blocks < cap ? blocks : cap;
}

typedef void (*TraceKernel)(const float*, const float4*, const uint32_t*, const float*, float4*, int, int, int, int,
                            const uint32_t*, uint32_t*, const float4*, const float4*);
typedef void (*PersistKernel)(const float*, const float4*, const float*, float4*, int, int, int, int,
                              const uint32_t*, uint32_t*, const float4*, uint32_t*);
struct TraceVariant {
  TraceKernel forest, tree;
  int waves;
  int ldsPerWave = 0;              // dynamic LDS bytes per wave (the SAH walk's stack entries)
  PersistKernel persist = nullptr;  // the persistent form (forest and tree unused)
  bool recTrees = false;            // reads the schedule's tree records (scheduled calls only)
  int splitRounds = 0;              // the stage-split form with this many queue rounds (forest and tree unused)
};
// trace_variant index -> kernels. Identical results for every entry.
const TraceVariant kTraceVariants[] = {
    {k_tracePixels<4, WalkPlain>, k_traceTreePixels<4, WalkPlain>, 4},
    {k_tracePixels<4, WalkCull>, k_traceTreePixels<4, WalkCull>, 4},
    {k_tracePixels<4, WalkCullPrefetch>, k_traceTreePixels<4, WalkCullPrefetch>, 4},
    {k_tracePixels<4, WalkCullTop>, k_traceTreePixels<4, WalkCullTop>, 4},
    {k_tracePixels<4, WalkAll>, k_traceTreePixels<4, WalkAll>, 4},
    {k_tracePixels<1, WalkCull>, k_traceTreePixels<1, WalkCull>, 1},
    {k_tracePixels<1, WalkCullPrefetch>, k_traceTreePixels<1, WalkCullPrefetch>, 1},
    {k_tracePixels<2, WalkCullPrefetch>, k_traceTreePixels<2, WalkCullPrefetch>, 2},
};
constexpr int kTraceVariantCount = sizeof(kTraceVariants) / sizeof(kTraceVariants[0]);

int launchTrace(bool forest, const float* P, const float* Nodes, const unsigned* Order, const float* C, float* Hit,
                int width, int height, hipStream_t stream, const uint32_t* tileOrder = nullptr,
                uint32_t* tileCost = nullptr) {
  const Tuning& tu = tuning();
  int tx = (width + 7) / 8, tiles = tx * ((height + 7) / 8);
  if (forest && tu.trace_variant >= kTraceVariantCount) {
    // wave-scheduled kernels: one block per `sched_tiles` tiles
    int rays = tu.sched_tiles * 64;
    int blocks = (tiles * 64 + rays - 1) / rays;
    WaveSchedule ws{tu.sched_leaf_min, tu.sched_newton_min};
    int waves = tu.trace_variant == kTraceVariantCount ? 4 : 1;
    auto kernel = waves == 4 ? k_tracePixelsSched<4> : k_tracePixelsSched<1>;
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(waves * 64), 0, stream, P, reinterpret_cast<const float4*>(Nodes),
                       Order, C, reinterpret_cast<float4*>(Hit), width, height, tx, tiles, rays, ws);
    return launch_status();
  }
  const TraceVariant& v = kTraceVariants[tu.trace_variant < kTraceVariantCount ? tu.trace_variant : 5];
  hipLaunchKernelGGL(forest ? v.forest : v.tree, dim3(tileGrid(tiles, v.waves)), dim3(v.waves * 64), 0, stream, P,
                     reinterpret_cast<const float4*>(Nodes), Order, C, reinterpret_cast<float4*>(Hit), width, height,
                     tx, tiles, tileOrder, tileCost, nullptr, nullptr);
  return launch_status();
}

// accel_variant index -> primary-ray kernel over the SAH hierarchy. 8 runs a
// wave's object tests once 6/8 of its live lanes wait at a leaf: 3.26 ms for
// a scheduled 1080p walk frame against 3.37 ms with the if-if order of 3; a
// threshold of 2/8 took 3.80 ms, 8/8 3.46 ms. The record-tree shapes (13,
// 14, 16-21) read the 324 tree records from the tile schedule instead of
// building a copy in LDS per block, so LDS holds the stack. 16 (one-wave
// blocks, 20 LDS entries, 3 waves per SIMD) traces a
// 1080p walk frame in 3.22 ms against 3.31 for 8 and, for a hierarchy of
// depth <= 21, keeps every stack entry out of scratch (PMC WRITE_SIZE about
// Hit's bytes). 19 and 20 (4 waves per SIMD, registers spilled to scratch)
// are as fast at 1080p and 12% faster at 4K (7.5-7.7 ms against 8.8), at
// four times the write traffic. 24 is 16 with three interior steps per
// gathered round and gathering at 5/8: 3.05 ms against 3.23; 28 and 29 are
// 19 and 20 with the same, the autotuner's choice at 4K. 26 (the default)
// takes up to four steps: with the walk's wave priorities 2.92-2.93 ms
// against 2.95-2.97 for 24.
const TraceVariant kAccelTraceVariants[] = {
    {k_tracePixels<1, WalkSah>, k_traceTreePixels<1, WalkSah>, 1},
    {k_tracePixels<2, WalkSah>, k_traceTreePixels<2, WalkSah>, 2},
    {k_tracePixels<4, WalkSah>, k_traceTreePixels<4, WalkSah>, 4},
    {k_tracePixels<2, WalkSahL<8>>, k_traceTreePixels<2, WalkSahL<8>>, 2, sahLdsPerWave(8)},
    {k_tracePixels<1, WalkSahL<4>>, k_traceTreePixels<1, WalkSahL<4>>, 1, sahLdsPerWave(4)},
    {k_tracePixels<2, WalkSahL<12>>, k_traceTreePixels<2, WalkSahL<12>>, 2, sahLdsPerWave(12)},
    {k_tracePixels<4, WalkSahL<6>>, k_traceTreePixels<4, WalkSahL<6>>, 4, sahLdsPerWave(6)},
    {k_tracePixels<2, WalkSahL<8>, 4>, k_traceTreePixels<2, WalkSahL<8>>, 2, sahLdsPerWave(8)},
    {k_tracePixels<2, WalkSahG<8, 6>>, k_traceTreePixels<2, WalkSahG<8, 6>>, 2, sahLdsPerWave(8)},
    {k_tracePixels<2, WalkSahG<8, 5>, 4>, k_traceTreePixels<2, WalkSahG<8, 5>>, 2, sahLdsPerWave(8)},
    {nullptr, nullptr, 2, sahLdsPerWave(8), k_tracePersistent<2, WalkSahG<8, 6>>},
    {nullptr, nullptr, 1, sahLdsPerWave(8), k_tracePersistent<1, WalkSahG<8, 6>>},
    {nullptr, nullptr, 2, sahLdsPerWave(8), k_tracePersistent<2, WalkSahG<8, 5>>},
    // 13-15: the tree records in device memory (the schedule's copy), LDS
    // free for 20 or 24 stack entries per lane; 15 keeps them in LDS with 14
    {k_tracePixels<2, WalkSahG<20, 6>, 1, true>, k_traceTreePixels<2, WalkSahG<20, 6>>, 2, sahLdsPerWave(20),
     nullptr, true},
    {k_tracePixels<2, WalkSahG<24, 6>, 1, true>, k_traceTreePixels<2, WalkSahG<24, 6>>, 2, sahLdsPerWave(24),
     nullptr, true},
    {k_tracePixels<2, WalkSahG<14, 6>>, k_traceTreePixels<2, WalkSahG<14, 6>>, 2, sahLdsPerWave(14)},
    // 16-18: more record-tree shapes (1-wave blocks; 16 entries; 4 waves per SIMD)
    {k_tracePixels<1, WalkSahG<20, 6>, 1, true>, k_traceTreePixels<1, WalkSahG<20, 6>>, 1, sahLdsPerWave(20),
     nullptr, true},
    {k_tracePixels<2, WalkSahG<16, 6>, 1, true>, k_traceTreePixels<2, WalkSahG<16, 6>>, 2, sahLdsPerWave(16),
     nullptr, true},
    {k_tracePixels<2, WalkSahG<16, 5>, 4, true>, k_traceTreePixels<2, WalkSahG<16, 5>>, 2, sahLdsPerWave(16),
     nullptr, true},
    // 19-21: record-tree shapes with 1-wave blocks
    {k_tracePixels<1, WalkSahG<16, 5>, 4, true>, k_traceTreePixels<1, WalkSahG<16, 5>>, 1, sahLdsPerWave(16),
     nullptr, true},
    {k_tracePixels<1, WalkSahG<20, 5>, 4, true>, k_traceTreePixels<1, WalkSahG<20, 5>>, 1, sahLdsPerWave(20),
     nullptr, true},
    {k_tracePixels<1, WalkSahG<24, 6>, 1, true>, k_traceTreePixels<1, WalkSahG<24, 6>>, 1, sahLdsPerWave(24),
     nullptr, true},
    // 22-27: 16 with several interior steps per gathered round (steps S,
    // gathering at G/8): S2 G6, S3 G6, S3 G5, S3 G4, S4 G5, S2 G5
    {k_tracePixels<1, WalkSahGS<20, 6, 2>, 1, true>, k_traceTreePixels<1, WalkSahGS<20, 6, 2>>, 1,
     sahLdsPerWave(20), nullptr, true},
    {k_tracePixels<1, WalkSahGS<20, 6, 3>, 1, true>, k_traceTreePixels<1, WalkSahGS<20, 6, 3>>, 1,
     sahLdsPerWave(20), nullptr, true},
    {k_tracePixels<1, WalkSahGS<20, 5, 3>, 1, true>, k_traceTreePixels<1, WalkSahGS<20, 5, 3>>, 1,
     sahLdsPerWave(20), nullptr, true},
    {k_tracePixels<1, WalkSahGS<20, 4, 3>, 1, true>, k_traceTreePixels<1, WalkSahGS<20, 4, 3>>, 1,
     sahLdsPerWave(20), nullptr, true},
    {k_tracePixels<1, WalkSahGS<20, 5, 4>, 1, true>, k_traceTreePixels<1, WalkSahGS<20, 5, 4>>, 1,
     sahLdsPerWave(20), nullptr, true},
    {k_tracePixels<1, WalkSahGS<20, 5, 2>, 1, true>, k_traceTreePixels<1, WalkSahGS<20, 5, 2>>, 1,
     sahLdsPerWave(20), nullptr, true},
    // 28-29: 19 and 20 (4 waves per SIMD) with three interior steps per round
    {k_tracePixels<1, WalkSahGS<16, 5, 3>, 4, true>, k_traceTreePixels<1, WalkSahGS<16, 5, 3>>, 1,
     sahLdsPerWave(16), nullptr, true},
    {k_tracePixels<1, WalkSahGS<20, 5, 3>, 4, true>, k_traceTreePixels<1, WalkSahGS<20, 5, 3>>, 1,
     sahLdsPerWave(20), nullptr, true},
    // 30-31: 24 with five and six interior steps per round
    {k_tracePixels<1, WalkSahGS<20, 5, 5>, 1, true>, k_traceTreePixels<1, WalkSahGS<20, 5, 5>>, 1,
     sahLdsPerWave(20), nullptr, true},
    {k_tracePixels<1, WalkSahGS<20, 5, 6>, 1, true>, k_traceTreePixels<1, WalkSahGS<20, 5, 6>>, 1,
     sahLdsPerWave(20), nullptr, true},
    // 32-34: the stage-split form (start, 3/6/12 queue rounds, finish) with
    // 26's walk
    {nullptr, nullptr, 1, sahLdsPerWave(20), nullptr, true, 3},
    {nullptr, nullptr, 1, sahLdsPerWave(20), nullptr, true, 6},
    {nullptr, nullptr, 1, sahLdsPerWave(20), nullptr, true, 12},
};
using WalkSplit = WalkSahGS<20, 5, 4>;
// the shape an unscheduled call runs in place of a record-tree one
constexpr int kAccelUnscheduled = 8;
constexpr int kAccelTraceVariantCount = sizeof(kAccelTraceVariants) / sizeof(kAccelTraceVariants[0]);

// Counters of persistent launches without a tile schedule: one pair per
// device in a device array (zero at module load; each launch leaves them at
// zero), its device address from hipGetSymbolAddress once per device.
static __device__ uint32_t persistCounters[2];
uint32_t* unscheduledCounters() {
  static std::mutex lock;
  static void* ptr[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  std::lock_guard<std::mutex> g(lock);
  if (!ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(persistCounters)) != hipSuccess || !p) return nullptr;
    ptr[dev] = p;
  }
  return static_cast<uint32_t*>(ptr[dev]);
}
// Blocks of a persistent launch: as many as the device holds at once.
int residentBlocks(PersistKernel k, int waves, int ldsBytes) {
  int perCU = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, reinterpret_cast<const void*>(k), waves * 64, ldsBytes) !=
          hipSuccess ||
      perCU < 1)
    perCU = 1;
  return perCU * deviceCUs();
}

void scheduleArrays(void* schedule, int width, int height, const uint32_t** order, uint32_t** cost);

// The stage-split trace's ray state and queues: 32 bytes per ray plus the
// round counts, allocated on a variant's first use on a device and grown to
// the largest view traced (the device is synchronised before a smaller
// buffer is freed). Not for callers that cannot allocate: the default shapes
// allocate nothing.
int splitBuffers(size_t rays, SplitBuffers* b) {
  static std::mutex lock;
  static void* base[64] = {};
  static size_t have[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return ARBOR_ERR_DEVICE;
  std::lock_guard<std::mutex> g(lock);
  if (have[dev] < rays) {
    if (base[dev]) {
      if (hipDeviceSynchronize() != hipSuccess) return ARBOR_ERR_DEVICE;
      (void)hipFree(base[dev]);
      base[dev] = nullptr;
      have[dev] = 0;
    }
    void* p = nullptr;
    if (hipMalloc(&p, rays * 8 * sizeof(uint32_t) + kSplitWords * sizeof(uint32_t)) != hipSuccess)
      return ARBOR_ERR_DEVICE;
    base[dev] = p;
    have[dev] = rays;
  }
  uint32_t* w = static_cast<uint32_t*>(base[dev]);
  b->cell = w;
  b->tx = reinterpret_cast<float*>(w + rays);
  b->tz = reinterpret_cast<float*>(w + 2 * rays);
  b->best = reinterpret_cast<float*>(w + 3 * rays);
  b->found = reinterpret_cast<float*>(w + 4 * rays);
  b->tree = reinterpret_cast<int*>(w + 5 * rays);
  b->queue[0] = w + 6 * rays;
  b->queue[1] = w + 7 * rays;
  b->counts = w + 8 * rays;
  return ARBOR_OK;
}
// The stage-split trace: clear, start, `rounds` queue rounds and the finish.
int launchSplit(const float* P, const float* Nodes, const unsigned* Order, const float4* accel, const float* C,
                float* Hit, int width, int height, hipStream_t stream, const float4* treeRecs, int rounds) {
  if (rounds < 1 || rounds > kSplitMaxRounds) return ARBOR_ERR_ARG;
  SplitBuffers b;
  int rc = splitBuffers((size_t)width * (size_t)height, &b);
  if (rc != ARBOR_OK) return rc;
  const int tx = (width + 7) / 8, tiles = tx * ((height + 7) / 8);
  float4* hit = reinterpret_cast<float4*>(Hit);
  const float4* nodes = reinterpret_cast<const float4*>(Nodes);
  const int lds = sahLdsPerWave(20);
  hipLaunchKernelGGL(k_splitClear, dim3(1), dim3(kSplitWords), 0, stream, b.counts, kSplitWords);
  hipLaunchKernelGGL(k_splitStart, dim3(tileGrid(tiles, 1)), dim3(64), 0, stream, C, hit, width, height, tx, tiles,
                     treeRecs, b);
  static std::atomic<int> resident{0};
  int blocks = resident.load(std::memory_order_relaxed);
  if (blocks <= 0) {
    int perCU = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &perCU, reinterpret_cast<const void*>(k_splitRound<WalkSplit, false>), 64, lds) != hipSuccess ||
        perCU < 1)
      perCU = 1;
    blocks = perCU * deviceCUs();
    resident.store(blocks, std::memory_order_relaxed);
  }
  for (int r = 0; r < rounds; r++)
    hipLaunchKernelGGL((k_splitRound<WalkSplit, false>), dim3(blocks), dim3(64), lds, stream, P, nodes, Order, C,
                       hit, width, height, accel, treeRecs, b, r);
  hipLaunchKernelGGL((k_splitRound<WalkSplit, true>), dim3(blocks), dim3(64), lds, stream, P, nodes, Order, C, hit,
                     width, height, accel, treeRecs, b, rounds);
  return launch_status();
}

// The SAH-walk trace. schedule: an ordered tile schedule (its order and cost
// arrays, and the persistent kernels' counters), or null.
int launchTraceAccel(const float* P, const float* Nodes, const unsigned* Order, const void* accel, const float* C,
                     float* Hit, int width, int height, hipStream_t stream, void* schedule) {
  int tx = (width + 7) / 8, tiles = tx * ((height + 7) / 8);
  const uint32_t* tileOrder = nullptr;
  uint32_t* tileCost = nullptr;
  if (schedule) scheduleArrays(schedule, width, height, &tileOrder, &tileCost);
  const int tuned = tuning().accel_variant;
  // an unscheduled call has no tree records: it runs the default shape instead
  const TraceVariant& v = kAccelTraceVariants[kAccelTraceVariants[tuned].recTrees && !schedule ? kAccelUnscheduled : tuned];
  if (v.splitRounds)
    return launchSplit(P, Nodes, Order, static_cast<const float4*>(accel), C, Hit, width, height, stream,
                       static_cast<const TileSchedule*>(schedule)->trees, v.splitRounds);
  if (v.persist) {
    uint32_t* counters = schedule ? static_cast<TileSchedule*>(schedule)->persist : unscheduledCounters();
    if (!counters) return ARBOR_ERR_DEVICE;
    const int lds = v.ldsPerWave * v.waves;
    hipLaunchKernelGGL(v.persist, dim3(residentBlocks(v.persist, v.waves, lds)), dim3(v.waves * 64), lds, stream, P,
                       reinterpret_cast<const float4*>(Nodes), C, reinterpret_cast<float4*>(Hit), width, height, tx,
                       tiles, tileOrder, tileCost, static_cast<const float4*>(accel), counters);
    return launch_status();
  }
  hipLaunchKernelGGL(v.forest, dim3(tileGrid(tiles, v.waves)), dim3(v.waves * 64), v.ldsPerWave * v.waves, stream, P,
                     reinterpret_cast<const float4*>(Nodes), Order, C, reinterpret_cast<float4*>(Hit), width, height,
                     tx, tiles, tileOrder, tileCost, static_cast<const float4*>(accel),
                     schedule ? static_cast<const TileSchedule*>(schedule)->trees : nullptr);
  return launch_status();
}
}  // namespace

extern "C" {

int arbor_tracePixels(const float* P, const float* Nodes, const unsigned* Order, const float* C, float* Hit,
                      int width, int height, hipStream_t stream) {
  ARBOR_REQUIRE(P && Nodes && Order && C && Hit);
  if (!dims_ok(width, height)) return ARBOR_ERR_ARG;
  return launchTrace(true, P, Nodes, Order, C, Hit, width, height, stream);
}

}  // extern "C"

namespace {
// Launches the frame prologue (k_framePrologue) cooperatively: every block
// resident at once, so its grid barriers complete.
int launchPrologue(float* C, const float* I, float delta, int sample, int light, int width, int height,
                   void* s0, void* s1, hipStream_t stream) {
  const int tiles = ((width + 7) / 8) * ((height + 7) / 8);
  int blocks = (tiles + kOrderBlock - 1) / kOrderBlock;
  const int cus = deviceCUs();
  if (blocks > cus) blocks = cus;
  TileSchedule* a = static_cast<TileSchedule*>(s0);
  TileSchedule* b = static_cast<TileSchedule*>(s1);
  void* args[] = {&C, &I, &delta, &sample, &light, &width, &height, &a, &b};
  if (hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_framePrologue), dim3(blocks), dim3(kOrderBlock),
                                 args, 0, stream) != hipSuccess) {
    (void)hipGetLastError();
    return ARBOR_ERR_LAUNCH;
  }
  return launch_status();
}
// Cost and order arrays of a schedule for a width x height view.
void scheduleArrays(void* schedule, int width, int height, const uint32_t** order, uint32_t** cost) {
  const int tiles = ((width + 7) / 8) * ((height + 7) / 8);
  *cost = reinterpret_cast<uint32_t*>(static_cast<TileSchedule*>(schedule) + 1);
  *order = *cost + (size_t)tiles;
}
// The ordering of one schedule (one launch); its cost and order arrays out.
int orderTiles(const float* C, void* schedule, int width, int height, hipStream_t stream, const uint32_t** order,
               uint32_t** cost) {
  scheduleArrays(schedule, width, height, order, cost);
  return launchPrologue(const_cast<float*>(C), nullptr, 0.0f, -1, -1, width, height, schedule, nullptr, stream);
}
}  // namespace

extern "C" {

long long arbor_tile_schedule_bytes(int width, int height) {
  if (!dims_ok(width, height)) return ARBOR_ERR_ARG;
  long long tiles = (long long)((width + 7) / 8) * ((height + 7) / 8);
  return (long long)sizeof(TileSchedule) + 12 * tiles;
}

int arbor_tracePixelsScheduled(const float* P, const float* Nodes, const unsigned* Order, const float* C, float* Hit,
                               int width, int height, void* schedule, hipStream_t stream) {
  ARBOR_REQUIRE(P && Nodes && Order && C && Hit && schedule);
  if (!dims_ok(width, height)) return ARBOR_ERR_ARG;
  if (tuning().trace_variant >= kTraceVariantCount)  // lane-scheduled kernels keep their own order
    return launchTrace(true, P, Nodes, Order, C, Hit, width, height, stream);
  uint32_t* cost = nullptr;
  const uint32_t* order = nullptr;
  int rc = orderTiles(C, schedule, width, height, stream, &order, &cost);
  if (rc != ARBOR_OK) return rc;
  return launchTrace(true, P, Nodes, Order, C, Hit, width, height, stream, order, cost);
}

int arbor_traceTreePixels(const float* P, const float* Nodes, const unsigned* Order, const float* C, float* Hit,
                          int width, int height, hipStream_t stream) {
  ARBOR_REQUIRE(P && Nodes && Order && C && Hit);
  if (!dims_ok(width, height)) return ARBOR_ERR_ARG;
  return launchTrace(false, P, Nodes, Order, C, Hit, width, height, stream);
}

// Blocks of a shading launch: perPerCU blocks per compute unit, striding over
// the tiles. Shading a tile is light next to each block's fill of the tree
// table in LDS, so a grid the size of the resident blocks pays that fill
// once per block: k_shadePixels<false> at 1920x1080 took 168 us with 3 blocks
// per CU against 200 us with one block per 4 tiles (k_shadePixels<true>: 186
// against 235 us with 2 per CU). Results do not depend on the grid.
inline int shadeGrid(int tiles, int perCU) {
  const int cus = deviceCUs();
  const int g = tileGrid(tiles, kWavesPerBlock), cap = perCU * cus;
  return g < cap ? g : cap;
}

int arbor_shadePixels(const float* P, const float* Nodes, const unsigned* Order, const float* C, const float* Hit,
                      float* Linear, float* Guide, int width, int height, hipStream_t stream) {
  ARBOR_REQUIRE(P && Nodes && Order && C && Hit && Linear && Guide);
  if (!dims_ok(width, height)) return ARBOR_ERR_ARG;
  int tx = (width + 7) / 8, tiles = tx * ((height + 7) / 8);
  hipLaunchKernelGGL(k_shadePixels<true>, dim3(shadeGrid(tiles, 2)), dim3(kTileBlock), 0, stream, P,
                     reinterpret_cast<const float4*>(Nodes), Order, C, reinterpret_cast<const float4*>(Hit),
                     reinterpret_cast<float4*>(Linear), reinterpret_cast<float4*>(Guide), width, height, tx, tiles,
                     nullptr, nullptr, nullptr);
  return launch_status();
}

int arbor_shadePixelsVis(const float* P, const float* Nodes, const unsigned* Order, const float* C, const float* Hit,
                         const float* Vis, float* Linear, float* Guide, int width, int height, hipStream_t stream) {
  ARBOR_REQUIRE(P && Nodes && Order && C && Hit && Vis && Linear && Guide);
  if (!dims_ok(width, height)) return ARBOR_ERR_ARG;
  int tx = (width + 7) / 8, tiles = tx * ((height + 7) / 8);
  hipLaunchKernelGGL(k_shadePixels<false>, dim3(shadeGrid(tiles, 3)), dim3(kTileBlock), 0, stream, P,
                     reinterpret_cast<const float4*>(Nodes), Order, C, reinterpret_cast<const float4*>(Hit),
                     reinterpret_cast<float4*>(Linear), reinterpret_cast<float4*>(Guide), width, height, tx, tiles,
                     nullptr, reinterpret_cast<const float4*>(Vis), nullptr);
  return launch_status();
}

namespace {
typedef void (*VisKernel)(const float*, const float4*, const uint32_t*, const float*, const float4*, float4*, int, int,
                          int, int, const uint32_t*, uint32_t*, const float4*, const float4*);
struct VisVariant {
  VisKernel kernel;
  int waves;
  int ldsPerWave = 0;     // dynamic LDS bytes per wave (the SAH walk's stack entries)
  bool recTrees = false;  // reads the schedule's tree records (scheduled calls only)
};
// vis_variant index -> kernel. Identical results for every entry. The
// kernel keeps the 324-entry tree table (11.7 KB) in LDS per block, so
// one-wave blocks stop at 14 waves per CU; two-wave blocks with a 96-VGPR
// budget (entry 6, the default) reach 5 waves per SIMD.
const VisVariant kVisVariants[] = {
    {k_sunVisibility<1, WalkCull>, 1},         {k_sunVisibility<1, WalkCullPrefetch>, 1},
    {k_sunVisibility<1, WalkCullTop>, 1},      {k_sunVisibility<2, WalkCull>, 2},
    {k_sunVisibility<4, WalkCull>, 4},         {k_sunVisibility<1, WalkPlain>, 1},
    {k_sunVisibility<2, WalkCull, 5>, 2},      {k_sunVisibility<4, WalkCull, 5>, 4},
};
constexpr int kVisVariantCount = sizeof(kVisVariants) / sizeof(kVisVariants[0]);

int launchVisibility(const float* P, const float* Nodes, const unsigned* Order, const float* C, const float* Hit,
                     float* Vis, int width, int height, const uint32_t* tileOrder, uint32_t* tileCost,
                     hipStream_t stream) {
  int tx = (width + 7) / 8, tiles = tx * ((height + 7) / 8);
  const VisVariant& v = kVisVariants[tuning().vis_variant];
  hipLaunchKernelGGL(v.kernel, dim3(tileGrid(tiles, v.waves)), dim3(v.waves * 64), 0, stream, P,
                     reinterpret_cast<const float4*>(Nodes), Order, C, reinterpret_cast<const float4*>(Hit),
                     reinterpret_cast<float4*>(Vis), width, height, tx, tiles, tileOrder, tileCost, nullptr, nullptr);
  return launch_status();
}

// accel_vis_variant index -> sun-visibility kernel over the SAH hierarchy. 8
// gathers object tests at 5/8 of the live lanes: 7.8 ms for 1080p with
// shadows against 8.8 ms for 0, the same kernel without it. 15 reads the
// tree records from the schedule and keeps 14 stack
// entries per lane in LDS, one-wave blocks at 5 waves per SIMD: 118-120 ms
// for a 4-sample 4K frame with sun visibility against 131 ms with 8. 23
// (the default) is 15 with three interior steps per gathered round: 114.2
// against 118.0 ms.
const VisVariant kAccelVisVariants[] = {
    {k_sunVisibility<2, WalkSah, 5>, 2},
    {k_sunVisibility<1, WalkSah>, 1},
    {k_sunVisibility<2, WalkSah>, 2},
    {k_sunVisibility<4, WalkSah, 5>, 4},
    {k_sunVisibility<2, WalkSah, 4>, 2},
    {k_sunVisibility<2, WalkSahL<4>, 5>, 2, sahLdsPerWave(4)},
    {k_sunVisibility<2, WalkSahL<8>, 4>, 2, sahLdsPerWave(8)},
    {k_sunVisibility<4, WalkSahL<4>, 5>, 4, sahLdsPerWave(4)},
    {k_sunVisibility<2, WalkSahG<0, 5>, 5>, 2},
    {k_sunVisibility<2, WalkSahG<8, 5>, 4>, 2, sahLdsPerWave(8)},
    // 10-12: the tree records in device memory (the schedule's copy), LDS
    // free for the stack entries
    {k_sunVisibility<2, WalkSahG<14, 5>, 5, true>, 2, sahLdsPerWave(14), true},
    {k_sunVisibility<2, WalkSahG<18, 5>, 4, true>, 2, sahLdsPerWave(18), true},
    {k_sunVisibility<2, WalkSahG<20, 5>, 1, true>, 2, sahLdsPerWave(20), true},
    // 13-17: more record-tree shapes (16 entries; 6 waves per SIMD; 1-wave
    // blocks; gathering at 4/8 and 6/8)
    {k_sunVisibility<2, WalkSahG<16, 5>, 5, true>, 2, sahLdsPerWave(16), true},
    {k_sunVisibility<2, WalkSahG<12, 5>, 6, true>, 2, sahLdsPerWave(12), true},
    {k_sunVisibility<1, WalkSahG<14, 5>, 5, true>, 1, sahLdsPerWave(14), true},
    {k_sunVisibility<2, WalkSahG<14, 4>, 5, true>, 2, sahLdsPerWave(14), true},
    {k_sunVisibility<2, WalkSahG<14, 6>, 5, true>, 2, sahLdsPerWave(14), true},
    // 18-21: record-tree shapes with 1-wave blocks
    {k_sunVisibility<1, WalkSahG<16, 5>, 5, true>, 1, sahLdsPerWave(16), true},
    {k_sunVisibility<1, WalkSahG<12, 5>, 6, true>, 1, sahLdsPerWave(12), true},
    {k_sunVisibility<1, WalkSahG<14, 6>, 5, true>, 1, sahLdsPerWave(14), true},
    {k_sunVisibility<1, WalkSahG<20, 5>, 4, true>, 1, sahLdsPerWave(20), true},
    // 22-25: 15 with several interior steps per gathered round: S2 G5,
    // S3 G5, S3 G4, S3 G6
    {k_sunVisibility<1, WalkSahGS<14, 5, 2>, 5, true>, 1, sahLdsPerWave(14), true},
    {k_sunVisibility<1, WalkSahGS<14, 5, 3>, 5, true>, 1, sahLdsPerWave(14), true},
    {k_sunVisibility<1, WalkSahGS<14, 4, 3>, 5, true>, 1, sahLdsPerWave(14), true},
    {k_sunVisibility<1, WalkSahGS<14, 6, 3>, 5, true>, 1, sahLdsPerWave(14), true},
    // 26-27: 23 with four and five interior steps per round
    {k_sunVisibility<1, WalkSahGS<14, 5, 4>, 5, true>, 1, sahLdsPerWave(14), true},
    {k_sunVisibility<1, WalkSahGS<14, 5, 5>, 5, true>, 1, sahLdsPerWave(14), true},
};
// the shape an unscheduled call runs in place of a record-tree one
constexpr int kAccelVisUnscheduled = 8;
constexpr int kAccelVisVariantCount = sizeof(kAccelVisVariants) / sizeof(kAccelVisVariants[0]);

// schedule: the ordered tile schedule whose order, cost and tree records the
// pass uses, or null.
int launchVisibilityAccel(const float* P, const float* Nodes, const unsigned* Order, const void* accel,
                          const float* C, const float* Hit, float* Vis, int width, int height, void* schedule,
                          hipStream_t stream) {
  int tx = (width + 7) / 8, tiles = tx * ((height + 7) / 8);
  const uint32_t* tileOrder = nullptr;
  uint32_t* tileCost = nullptr;
  if (schedule) scheduleArrays(schedule, width, height, &tileOrder, &tileCost);
  const int tuned = tuning().accel_vis_variant;
  // an unscheduled call has no tree records: it runs the default shape instead
  const VisVariant& v = kAccelVisVariants[kAccelVisVariants[tuned].recTrees && !schedule ? kAccelVisUnscheduled : tuned];
  hipLaunchKernelGGL(v.kernel, dim3(tileGrid(tiles, v.waves)), dim3(v.waves * 64), v.ldsPerWave * v.waves, stream, P,
                     reinterpret_cast<const float4*>(Nodes), Order, C, reinterpret_cast<const float4*>(Hit),
                     reinterpret_cast<float4*>(Vis), width, height, tx, tiles, tileOrder, tileCost,
                     static_cast<const float4*>(accel),
                     schedule ? static_cast<const TileSchedule*>(schedule)->trees : nullptr);
  return launch_status();
}
}  // namespace

int arbor_sunVisibility(const float* P, const float* Nodes, const unsigned* Order, const float* C, const float* Hit,
                        float* Vis, int width, int height, hipStream_t stream) {
  ARBOR_REQUIRE(P && Nodes && Order && C && Hit && Vis);
  if (!dims_ok(width, height)) return ARBOR_ERR_ARG;
  return launchVisibility(P, Nodes, Order, C, Hit, Vis, width, height, nullptr, nullptr, stream);
}

int arbor_sunVisibilityScheduled(const float* P, const float* Nodes, const unsigned* Order, const float* C,
                                 const float* Hit, float* Vis, int width, int height, void* schedule,
                                 hipStream_t stream) {
  ARBOR_REQUIRE(P && Nodes && Order && C && Hit && Vis && schedule);
  if (!dims_ok(width, height)) return ARBOR_ERR_ARG;
  uint32_t* cost = nullptr;
  const uint32_t* order = nullptr;
  int rc = orderTiles(C, schedule, width, height, stream, &order, &cost);
  if (rc != ARBOR_OK) return rc;
  return launchVisibility(P, Nodes, Order, C, Hit, Vis, width, height, order, cost, stream);
}

int arbor_tracePixelsAccel(const float* P, const float* Nodes, const unsigned* Order, const void* accel,
                           const float* C, float* Hit, int width, int height, void* schedule, hipStream_t stream) {
  ARBOR_REQUIRE(P && Nodes && Order && accel && C && Hit);
  if (!dims_ok(width, height)) return ARBOR_ERR_ARG;
  uint32_t* cost = nullptr;
  const uint32_t* order = nullptr;
  if (schedule) {
    int rc = orderTiles(C, schedule, width, height, stream, &order, &cost);
    if (rc != ARBOR_OK) return rc;
  }
  return launchTraceAccel(P, Nodes, Order, accel, C, Hit, width, height, stream, schedule);
}

int arbor_sunVisibilityAccel(const float* P, const float* Nodes, const unsigned* Order, const void* accel,
                             const float* C, const float* Hit, float* Vis, int width, int height, void* schedule,
                             hipStream_t stream) {
  ARBOR_REQUIRE(P && Nodes && Order && accel && C && Hit && Vis);
  if (!dims_ok(width, height)) return ARBOR_ERR_ARG;
  uint32_t* cost = nullptr;
  const uint32_t* order = nullptr;
  if (schedule) {
    int rc = orderTiles(C, schedule, width, height, stream, &order, &cost);
    if (rc != ARBOR_OK) return rc;
  }
  return launchVisibilityAccel(P, Nodes, Order, accel, C, Hit, Vis, width, height, schedule, stream);
}

}  // extern "C"

namespace {
// Tile schedule of arbor_tracePixelsAccelNoAlloc: one per device, in a
// file-scope device array (zero at module load, as a schedule must start),
// for views of up to kNoAllocTiles 8x8 tiles. In host code `scratch` only
// names the symbol; its device address comes from hipGetSymbolAddress, once
// per device with that device current.
constexpr int kNoAllocTiles = 512 * 512;  // 4096 x 4096 pixels
constexpr size_t kNoAllocBytes = sizeof(TileSchedule) + 12 * (size_t)kNoAllocTiles;  // 3,163,840 bytes
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

void* scratchOnDevice() {
  static std::mutex lock;
  static void* ptr[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  std::lock_guard<std::mutex> g(lock);
  if (!ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
    ptr[dev] = p;
  }
  return ptr[dev];
}
}  // namespace

extern "C" {

long long arbor_no_alloc_scratch_bytes(void) { return (long long)kNoAllocBytes; }

int arbor_tracePixelsAccelNoAlloc(const float* P, const float* Nodes, const unsigned* Order, const void* accel,
                                  const float* C, float* Hit, int width, int height, hipStream_t stream) {
  ARBOR_REQUIRE(P && Nodes && Order && accel && C && Hit);
  if (!dims_ok(width, height)) return ARBOR_ERR_ARG;
  if ((long long)((width + 7) / 8) * ((height + 7) / 8) > kNoAllocTiles) return ARBOR_ERR_ARG;
  void* schedule = scratchOnDevice();
  if (!schedule) return ARBOR_ERR_DEVICE;
  uint32_t* cost = nullptr;
  const uint32_t* order = nullptr;
  int rc = orderTiles(C, schedule, width, height, stream, &order, &cost);
  if (rc != ARBOR_OK) return rc;
  return launchTraceAccel(P, Nodes, Order, accel, C, Hit, width, height, stream, schedule);
}

int arbor_shadePixelsAOV(const float* P, const float* Nodes, const unsigned* Order, const float* C, const float* Hit,
                         float* Linear, float* Guide, float* Normal, int width, int height, hipStream_t stream) {
  ARBOR_REQUIRE(P && Nodes && Order && C && Hit && Linear && Guide && Normal);
  if (!dims_ok(width, height)) return ARBOR_ERR_ARG;
  int tx = (width + 7) / 8, tiles = tx * ((height + 7) / 8);
  hipLaunchKernelGGL(k_shadePixels<true>, dim3(shadeGrid(tiles, 2)), dim3(kTileBlock), 0, stream, P,
                     reinterpret_cast<const float4*>(Nodes), Order, C, reinterpret_cast<const float4*>(Hit),
                     reinterpret_cast<float4*>(Linear), reinterpret_cast<float4*>(Guide), width, height, tx, tiles,
                     reinterpret_cast<float4*>(Normal), nullptr, nullptr);
  return launch_status();
}

int arbor_shadePixelsFeatures(const float* P, const float* Nodes, const unsigned* Order, const float* C,
                              const float* Hit, void* Feat, int width, int height, hipStream_t stream) {
  ARBOR_REQUIRE(P && Nodes && Order && C && Hit && Feat);
  if (!dims_ok(width, height)) return ARBOR_ERR_ARG;
  int tx = (width + 7) / 8, tiles = tx * ((height + 7) / 8);
  hipLaunchKernelGGL(k_shadePixels<true>, dim3(shadeGrid(tiles, 2)), dim3(kTileBlock), 0, stream, P,
                     reinterpret_cast<const float4*>(Nodes), Order, C, reinterpret_cast<const float4*>(Hit),
                     nullptr, nullptr, width, height, tx, tiles, nullptr, nullptr, reinterpret_cast<featx8*>(Feat));
  return launch_status();
}

int arbor_filterLighting(const float* Linear, const float* Guide, const float* Hit, float* Filtered, const float* C,
                         int width, int height, hipStream_t stream) {
  ARBOR_REQUIRE(Linear && Guide && Hit && Filtered && C);
  if (!dims_ok(width, height)) return ARBOR_ERR_ARG;
  dim3 block(kFilterW, kFilterH), grid((width + kFilterW - 1) / kFilterW, (height + kFilterH - 1) / kFilterH);
  hipLaunchKernelGGL(k_filterLighting<false>, grid, block, 0, stream, reinterpret_cast<const float4*>(Linear),
                     reinterpret_cast<const float4*>(Guide), reinterpret_cast<const float4*>(Hit),
                     reinterpret_cast<float4*>(Filtered), C, width, height, filterWeights(), nullptr, nullptr);
  return launch_status();
}

int arbor_resolvePixels(const float* Filtered, float* History, unsigned* Pixels, const float* C, int width, int height,
                        hipStream_t stream) {
  ARBOR_REQUIRE(Filtered && History && Pixels && C);
  if (!dims_ok(width, height)) return ARBOR_ERR_ARG;
  dim3 block(64, 4), grid((width + 63) / 64, (height + 3) / 4);
  hipLaunchKernelGGL(k_resolvePixels, grid, block, 0, stream, reinterpret_cast<const float4*>(Filtered),
                     reinterpret_cast<float4*>(History), Pixels, C, width, height);
  return launch_status();
}

const char* arbor_version(void) { return "arbor-mi355x 0.1 gfx950"; }

int arbor_tune_set(const char* key, int value) {
  if (!key) return ARBOR_ERR_NULL;
  Tuning& t = tuning();
  struct { const char* name; int* slot; int lo, hi; } keys[] = {
      {"trace_variant", &t.trace_variant, 0, kTraceVariantCount + 1},
      {"grid_cap", &t.grid_cap, 1, 1 << 20},
      {"sched_tiles", &t.sched_tiles, 1, 4096},
      {"sched_leaf_min", &t.sched_leaf_min, 1, 64},
      {"sched_newton_min", &t.sched_newton_min, 1, 64},
      {"batch_variant", &t.batch_variant, 0, batchVariantCount() - 1},
      {"vis_variant", &t.vis_variant, 0, kVisVariantCount - 1},
      {"accel_variant", &t.accel_variant, 0, kAccelTraceVariantCount - 1},
      {"accel_vis_variant", &t.accel_vis_variant, 0, kAccelVisVariantCount - 1}};
  for (auto& k : keys)
    if (!strcmp(key, k.name)) {
      if (value < k.lo || value > k.hi) return ARBOR_ERR_ARG;
      *k.slot = value;
      return ARBOR_OK;
    }
  return ARBOR_ERR_ARG;
}

int arbor_tune_get(const char* key) {
  if (!key) return ARBOR_ERR_NULL;
  const Tuning& t = tuning();
  if (!strcmp(key, "trace_variant")) return t.trace_variant;
  if (!strcmp(key, "grid_cap")) return t.grid_cap;
  if (!strcmp(key, "sched_tiles")) return t.sched_tiles;
  if (!strcmp(key, "sched_leaf_min")) return t.sched_leaf_min;
  if (!strcmp(key, "sched_newton_min")) return t.sched_newton_min;
  if (!strcmp(key, "batch_variant")) return t.batch_variant;
  if (!strcmp(key, "vis_variant")) return t.vis_variant;
  if (!strcmp(key, "accel_variant")) return t.accel_variant;
  if (!strcmp(key, "accel_vis_variant")) return t.accel_vis_variant;
  return ARBOR_ERR_ARG;
}

static const char* const kTuneKeys[] = {"trace_variant",    "grid_cap",      "sched_tiles",
                                        "sched_leaf_min",   "sched_newton_min", "batch_variant",
                                        "vis_variant",      "accel_variant", "accel_vis_variant"};

// Tuning cache: one "key value" line per key; unknown keys and out-of-range
// values in a file are rejected as a whole.
int arbor_tune_save(const char* path) {
  if (!path) return ARBOR_ERR_NULL;
  FILE* f = fopen(path, "w");
  if (!f) return ARBOR_ERR_ARG;
  fprintf(f, "# arbor tuning (%s)\n", arbor_version());
  for (const char* k : kTuneKeys) fprintf(f, "%s %d\n", k, arbor_tune_get(k));
  return fclose(f) == 0 ? ARBOR_OK : ARBOR_ERR_ARG;
}

int arbor_tune_load(const char* path) {
  if (!path) return ARBOR_ERR_NULL;
  FILE* f = fopen(path, "r");
  if (!f) return ARBOR_ERR_ARG;
  const Tuning saved = tuning();
  char line[256];
  int rc = ARBOR_OK;
  while (rc == ARBOR_OK && fgets(line, sizeof line, f)) {
    if (line[0] == '#' || line[0] == '\n') continue;
    char key[64];
    int value;
    if (sscanf(line, "%63s %d", key, &value) != 2) rc = ARBOR_ERR_ARG;
    else rc = arbor_tune_set(key, value);
  }
  fclose(f);
  if (rc != ARBOR_OK) tuning() = saved;
  return rc;
}

// Autotuner for single-view tracing: times every trace configuration on the
// caller's buffers (reps launches each after a warm-up, HIP events on
// `stream`), keeps the fastest in the process tuning, writes the timings as
// JSON {"best": i, "ms": [...]}. The wave-scheduled kernels are tried with
// the current sched_* settings.
int arbor_autotune_trace(const float* P, const float* Nodes, const unsigned* Order, const float* C, float* Hit,
                         int width, int height, int reps, hipStream_t stream, char* json, int cap) {
  ARBOR_REQUIRE(P && Nodes && Order && C && Hit);
  if (!dims_ok(width, height) || reps < 1 || reps > 1000) return ARBOR_ERR_ARG;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return ARBOR_ERR_DEVICE;
  Tuning& t = tuning();
  const int saved = t.trace_variant;
  const int n = kTraceVariantCount + 2;
  float ms[kTraceVariantCount + 2] = {};
  int best = saved, rc = ARBOR_OK;
  for (int v = 0; v < n && rc == ARBOR_OK; v++) {
    t.trace_variant = v;
    rc = launchTrace(true, P, Nodes, Order, C, Hit, width, height, stream);
    (void)hipEventRecord(e0, stream);
    for (int r = 0; r < reps && rc == ARBOR_OK; r++) rc = launchTrace(true, P, Nodes, Order, C, Hit, width, height, stream);
    (void)hipEventRecord(e1, stream);
    if (hipEventSynchronize(e1) != hipSuccess) rc = ARBOR_ERR_DEVICE;
    (void)hipEventElapsedTime(&ms[v], e0, e1);
    ms[v] /= (float)reps;
    if (v == 0 || ms[v] < ms[best]) best = v;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  t.trace_variant = rc == ARBOR_OK ? best : saved;
  if (rc != ARBOR_OK) return rc;
  if (json && cap > 0) {
    int used = snprintf(json, (size_t)cap, "{\"best\":%d,\"ms\":[", best);
    for (int v = 0; v < n && used < cap; v++) used += snprintf(json + used, (size_t)(cap - used), "%s%.4f", v ? "," : "", ms[v]);
    if (used < cap) used += snprintf(json + used, (size_t)(cap - used), "]}");
    if (used >= cap) return ARBOR_ERR_ARG;
  }
  return ARBOR_OK;
}

int arbor_autotune_trace_accel(const float* P, const float* Nodes, const unsigned* Order, const void* accel,
                               const float* C, float* Hit, int width, int height, int reps, hipStream_t stream,
                               char* json, int cap) {
  ARBOR_REQUIRE(P && Nodes && Order && accel && C && Hit);
  if (!dims_ok(width, height) || reps < 1 || reps > 1000) return ARBOR_ERR_ARG;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return ARBOR_ERR_DEVICE;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return ARBOR_ERR_DEVICE;
  }
  Tuning& t = tuning();
  const int saved = t.accel_variant;
  float ms[kAccelTraceVariantCount] = {};
  int best = 0, rc = ARBOR_OK;
  // Every shape traces in the same tile order, the one the engine would use
  // for this view: a schedule of its own, ordered twice around one trace of
  // shape kAccelUnscheduled (the second ordering from that trace's costs).
  const long long schedBytes = arbor_tile_schedule_bytes(width, height);
  void* sched = nullptr;
  if (hipMalloc(&sched, (size_t)schedBytes) != hipSuccess) rc = ARBOR_ERR_DEVICE;
  if (rc == ARBOR_OK && hipMemsetAsync(sched, 0, (size_t)schedBytes, stream) != hipSuccess) rc = ARBOR_ERR_DEVICE;
  const uint32_t* order = nullptr;
  uint32_t* cost = nullptr;
  if (rc == ARBOR_OK) rc = orderTiles(C, sched, width, height, stream, &order, &cost);
  if (rc == ARBOR_OK) {
    t.accel_variant = kAccelUnscheduled;
    rc = launchTraceAccel(P, Nodes, Order, accel, C, Hit, width, height, stream, sched);
  }
  if (rc == ARBOR_OK) rc = orderTiles(C, sched, width, height, stream, &order, &cost);
  for (int v = 0; v < kAccelTraceVariantCount && rc == ARBOR_OK; v++) {
    t.accel_variant = v;
    rc = launchTraceAccel(P, Nodes, Order, accel, C, Hit, width, height, stream, sched);
    (void)hipEventRecord(e0, stream);
    for (int r = 0; r < reps && rc == ARBOR_OK; r++)
      rc = launchTraceAccel(P, Nodes, Order, accel, C, Hit, width, height, stream, sched);
    (void)hipEventRecord(e1, stream);
    if (hipEventSynchronize(e1) != hipSuccess) rc = ARBOR_ERR_DEVICE;
    (void)hipEventElapsedTime(&ms[v], e0, e1);
    ms[v] /= (float)reps;
    if (ms[v] < ms[best]) best = v;
  }
  (void)hipStreamSynchronize(stream);
  if (sched) (void)hipFree(sched);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  t.accel_variant = rc == ARBOR_OK ? best : saved;
  if (rc != ARBOR_OK) return rc;
  if (json && cap > 0) {
    int used = snprintf(json, (size_t)cap, "{\"best\":%d,\"ms\":[", best);
    for (int v = 0; v < kAccelTraceVariantCount && used < cap; v++)
      used += snprintf(json + used, (size_t)(cap - used), "%s%.4f", v ? "," : "", ms[v]);
    if (used < cap) used += snprintf(json + used, (size_t)(cap - used), "]}");
    if (used >= cap) return ARBOR_ERR_ARG;
  }
  return ARBOR_OK;
}

}  // extern "C"

namespace arbor {

int framePrologue(float* C, const float* I, float delta, int sample, int light, int width, int height,
                  void* traceSchedule, void* visSchedule, hipStream_t stream) {
  if (!C || !traceSchedule) return ARBOR_ERR_NULL;
  if (!dims_ok(width, height)) return ARBOR_ERR_ARG;
  return launchPrologue(C, I, delta, sample, light, width, height, traceSchedule, visSchedule, stream);
}

int traceOrdered(const float* P, const float* Nodes, const unsigned* Order, const void* accel, const float* C,
                 float* Hit, int width, int height, void* schedule, hipStream_t stream) {
  if (!P || !Nodes || !Order || !C || !Hit || !schedule) return ARBOR_ERR_NULL;
  if (!dims_ok(width, height)) return ARBOR_ERR_ARG;
  const uint32_t* order = nullptr;
  uint32_t* cost = nullptr;
  scheduleArrays(schedule, width, height, &order, &cost);
  if (accel) return launchTraceAccel(P, Nodes, Order, accel, C, Hit, width, height, stream, schedule);
  return launchTrace(true, P, Nodes, Order, C, Hit, width, height, stream, order, cost);
}

int visibilityOrdered(const float* P, const float* Nodes, const unsigned* Order, const void* accel, const float* C,
                      const float* Hit, float* Vis, int width, int height, void* schedule, hipStream_t stream) {
  if (!P || !Nodes || !Order || !C || !Hit || !Vis || !schedule) return ARBOR_ERR_NULL;
  if (!dims_ok(width, height)) return ARBOR_ERR_ARG;
  const uint32_t* order = nullptr;
  uint32_t* cost = nullptr;
  scheduleArrays(schedule, width, height, &order, &cost);
  if (accel) return launchVisibilityAccel(P, Nodes, Order, accel, C, Hit, Vis, width, height, schedule, stream);
  return launchVisibility(P, Nodes, Order, C, Hit, Vis, width, height, order, cost, stream);
}

int shadeVisNormal(const float* P, const float* Nodes, const unsigned* Order, const float* C, const float* Hit,
                   const float* Vis, float* Linear, float* Guide, float* Normal, int width, int height,
                   hipStream_t stream) {
  if (!P || !Nodes || !Order || !C || !Hit || !Vis || !Linear || !Guide || !Normal) return ARBOR_ERR_NULL;
  if (!dims_ok(width, height)) return ARBOR_ERR_ARG;
  int tx = (width + 7) / 8, tiles = tx * ((height + 7) / 8);
  hipLaunchKernelGGL(k_shadePixels<false>, dim3(shadeGrid(tiles, 3)), dim3(kTileBlock), 0, stream, P,
                     reinterpret_cast<const float4*>(Nodes), Order, C, reinterpret_cast<const float4*>(Hit),
                     reinterpret_cast<float4*>(Linear), reinterpret_cast<float4*>(Guide), width, height, tx, tiles,
                     reinterpret_cast<float4*>(Normal), reinterpret_cast<const float4*>(Vis), nullptr);
  return launch_status();
}

int shadeVisFeatures(const float* P, const float* Nodes, const unsigned* Order, const float* C, const float* Hit,
                     const float* Vis, void* Feat, int width, int height, hipStream_t stream) {
  if (!P || !Nodes || !Order || !C || !Hit || !Vis || !Feat) return ARBOR_ERR_NULL;
  if (!dims_ok(width, height)) return ARBOR_ERR_ARG;
  int tx = (width + 7) / 8, tiles = tx * ((height + 7) / 8);
  hipLaunchKernelGGL(k_shadePixels<false>, dim3(shadeGrid(tiles, 3)), dim3(kTileBlock), 0, stream, P,
                     reinterpret_cast<const float4*>(Nodes), Order, C, reinterpret_cast<const float4*>(Hit), nullptr,
                     nullptr, width, height, tx, tiles, nullptr, reinterpret_cast<const float4*>(Vis),
                     reinterpret_cast<featx8*>(Feat));
  return launch_status();
}

int filterResolve(const float* Linear, const float* Guide, const float* Hit, float* Filtered, float* History,
                  unsigned* Pixels, const float* C, int width, int height, hipStream_t stream) {
  if (!Linear || !Guide || !Hit || !Filtered || !History || !Pixels || !C) return ARBOR_ERR_NULL;
  if (!dims_ok(width, height)) return ARBOR_ERR_ARG;
  dim3 block(kFilterW, kFilterH), grid((width + kFilterW - 1) / kFilterW, (height + kFilterH - 1) / kFilterH);
  hipLaunchKernelGGL(k_filterLighting<true>, grid, block, 0, stream, reinterpret_cast<const float4*>(Linear),
                     reinterpret_cast<const float4*>(Guide), reinterpret_cast<const float4*>(Hit),
                     reinterpret_cast<float4*>(Filtered), C, width, height, filterWeights(),
                     reinterpret_cast<float4*>(History), Pixels);
  return launch_status();
}

}  // namespace arbor
