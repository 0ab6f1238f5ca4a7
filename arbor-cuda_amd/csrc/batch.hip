// Batched multi-view rendering on gfx950: many camera views of uneven
// resolution traced and shaded in ONE launch, then filtered and resolved in
// one more over every view's pixels, the whole sequence captured once into a
// hipGraph and replayed.
//
// Layout: a view table {width, height, tilesX, tileBase, pixelBase} and one
// C[64] per view in device memory; Hit/Linear/Guide/Filtered/History are
// pooled float4 arrays and Pixels a pooled u32 array, view v's pixels
// starting at pixelBase. The fused kernel numbers all 8x8 tiles of all views
// consecutively; each wave takes a tile, finds its view in the LDS copy of
// the table, traces the 64 rays (traceForest) and shades them (shadeOne)
// without a round trip through HBM. Per pixel the arithmetic is that of
// tracePixels + shadePixels, so every view equals its single-view render.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "arbor_batch.h"
#include "arbor_filter.hpp"
#include "arbor_shade.hpp"
#include "launch.hpp"
#include "tuning.hpp"

namespace arbor {

struct ViewDesc {
  int width, height, tilesX, tileBase, pixelBase;
  int filterBase, filterX, pad;  // first 32x8 filter block of the view, blocks per row
};
constexpr int kMaxViews = ARBOR_BATCH_MAX_VIEWS;

__device__ __forceinline__ int findView(const ViewDesc* sv, int n, int tile) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {  // last view with tileBase <= tile
    int mid = (lo + hi + 1) >> 1;
    if (sv[mid].tileBase <= tile) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int kWaves, class W>
__global__ __launch_bounds__(kWaves * 64) void k_traceShadeViews(const float* P, const float4* Nodes4,
                                                                 const uint32_t* Order, const float* Cs,
                                                                 const ViewDesc* views, int nviews, int totalTiles,
                                                                 float4* Hit, float4* Linear, float4* Guide,
                                                                 const float4* accel, const float4*) {
  __shared__ TreeInst trees[kTrees];
  __shared__ ViewDesc sv[kMaxViews];
  for (int i = threadIdx.x; i < kTrees; i += blockDim.x) trees[i] = makeTree(i, (int)Cs[31]);
  for (int i = threadIdx.x; i < nviews; i += blockDim.x) sv[i] = views[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int t = blockIdx.x * kWaves + wave; t < totalTiles; t += gridDim.x * kWaves) {
    const int v = findView(sv, nviews, t);
    const ViewDesc d = sv[v];
    const float* C = Cs + 64 * v;
    const int lt = t - d.tileBase;
    const int x = (lt % d.tilesX) * 8 + (lane & 7), y = (lt / d.tilesX) * 8 + (lane >> 3);
    if (x >= d.width || y >= d.height) continue;
    const ShadeCtx sc = makeShadeCtx(C, d.width, d.height);
    v3 ro = sc.view.eye, rd = viewRay(sc.view, x, y, d.width, d.height);
    float ground = rd.y < -0.00001f ? -ro.y / rd.y : kFar;
    int tree = -1;
    v2 h = traceForest<W>(P, Nodes4, W::sah ? accel : nullptr, Order, trees, ro, rd, C[14] / (float)d.height, C[8],
                          C[11], ground > 0.0f ? ground : kFar, &tree, false);
    const int pixel = d.pixelBase + y * d.width + x;
    const float4 hit = make_float4(h.x, h.y, (float)tree, C[14] * h.x / (float)d.height);
    Hit[pixel] = hit;
    float4 l, g;
    shadeOne<Walk<true, false, false>>(P, Nodes4, Order, trees, sc, x, y, hit, l, g);
    Linear[pixel] = l;
    Guide[pixel] = g;
  }
}

// The split form of the same pass: every view's tiles traced by one launch
// (the single-view trace kernel's register budget), their sun visibility by
// a second (views with it on), then shaded by a third from Hit and Vis (16 B
// per pixel each through HBM). Per pixel the arithmetic is the fused
// kernel's, so Hit, Linear and Guide are identical.
template <int kWaves, class W>
__global__ __launch_bounds__(kWaves * 64) void k_traceViews(const float* P, const float4* Nodes4, const uint32_t* Order,
                                                            const float* Cs, const ViewDesc* views, int nviews,
                                                            int totalTiles, float4* Hit, float4*, float4*,
                                                            const float4* accel, const float4*) {
  __shared__ TreeInst trees[kTrees];
  __shared__ ViewDesc sv[kMaxViews];
  for (int i = threadIdx.x; i < kTrees; i += blockDim.x) trees[i] = makeTree(i, (int)Cs[31]);
  for (int i = threadIdx.x; i < nviews; i += blockDim.x) sv[i] = views[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int t = blockIdx.x * kWaves + wave; t < totalTiles; t += gridDim.x * kWaves) {
    const int v = findView(sv, nviews, t);
    const ViewDesc d = sv[v];
    const float* C = Cs + 64 * v;
    const int lt = t - d.tileBase;
    const int x = (lt % d.tilesX) * 8 + (lane & 7), y = (lt / d.tilesX) * 8 + (lane >> 3);
    if (x >= d.width || y >= d.height) continue;
    const View view = makeView(C);
    v3 ro = view.eye, rd = viewRay(view, x, y, d.width, d.height);
    float ground = rd.y < -0.00001f ? -ro.y / rd.y : kFar;
    int tree = -1;
    v2 h = traceForest<W>(P, Nodes4, W::sah ? accel : nullptr, Order, trees, ro, rd, C[14] / (float)d.height, C[8],
                          C[11], ground > 0.0f ? ground : kFar, &tree, false);
    Hit[d.pixelBase + y * d.width + x] = make_float4(h.x, h.y, (float)tree, C[14] * h.x / (float)d.height);
  }
}
// The split trace over the SAH hierarchy with the tree records in device
// memory (the batch's copy of the world seed's records): the block's dynamic
// LDS holds kWaves x W::ldsStack stack entries per lane, then the view
// table. Hit equals k_traceViews'.
template <int kWaves, class W>
__global__ __launch_bounds__(kWaves * 64) void k_traceViewsRec(const float* P, const float4* Nodes4,
                                                               const uint32_t* Order, const float* Cs,
                                                               const ViewDesc* views, int nviews, int totalTiles,
                                                               float4* Hit, float4*, float4*, const float4* accel,
                                                               const float4* treeRecs) {
  static_assert(W::sah && W::ldsStack > 0, "a record-tree batch trace walks the SAH hierarchy with an LDS stack");
  ViewDesc* sv = reinterpret_cast<ViewDesc*>(sahLdsStack + (size_t)kWaves * W::ldsStack * 64);
  for (int i = threadIdx.x; i < nviews; i += blockDim.x) sv[i] = views[i];
  __syncthreads();
  const TreeRecs trees{treeRecs};
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int t = blockIdx.x * kWaves + wave; t < totalTiles; t += gridDim.x * kWaves) {
    const int v = findView(sv, nviews, t);
    const ViewDesc d = sv[v];
    const float* C = Cs + 64 * v;
    const int lt = t - d.tileBase;
    const int x = (lt % d.tilesX) * 8 + (lane & 7), y = (lt / d.tilesX) * 8 + (lane >> 3);
    if (x >= d.width || y >= d.height) continue;
    const View view = makeView(C);
    v3 ro = view.eye, rd = viewRay(view, x, y, d.width, d.height);
    float ground = rd.y < -0.00001f ? -ro.y / rd.y : kFar;
    int tree = -1;
    v2 h = traceForest<W>(P, Nodes4, accel, Order, trees, ro, rd, C[14] / (float)d.height, C[8], C[11],
                          ground > 0.0f ? ground : kFar, &tree, false);
    Hit[d.pixelBase + y * d.width + x] = make_float4(h.x, h.y, (float)tree, C[14] * h.x / (float)d.height);
  }
}
// Sun visibility of every view's pixels (the visibility pass of frame.hip,
// per view: a view with C[22] off gets nothing written, and its shading
// reads nothing).
template <class W>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(5))) void k_sunVisibilityViews(
    const float* P, const float4* Nodes4, const uint32_t* Order, const float* Cs, const ViewDesc* views, int nviews,
    int totalTiles, const float4* Hit, float4* Vis, const float4* accel) {
  __shared__ TreeInst trees[kTrees];
  __shared__ ViewDesc sv[kMaxViews];
  for (int i = threadIdx.x; i < kTrees; i += blockDim.x) trees[i] = makeTree(i, (int)Cs[31]);
  for (int i = threadIdx.x; i < nviews; i += blockDim.x) sv[i] = views[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int t = blockIdx.x * 2 + wave; t < totalTiles; t += gridDim.x * 2) {
    const int v = findView(sv, nviews, t);
    const ViewDesc d = sv[v];
    const float* C = Cs + 64 * v;
    if (!(C[22] > 0.5f)) continue;
    const int lt = t - d.tileBase;
    const int x = (lt % d.tilesX) * 8 + (lane & 7), y = (lt / d.tilesX) * 8 + (lane >> 3);
    if (x >= d.width || y >= d.height) continue;
    const ShadeCtx sc = makeShadeCtx(C, d.width, d.height);
    const int pixel = d.pixelBase + y * d.width + x;
    v3 vis = sunVisibilityOne<W>(P, Nodes4, W::sah ? accel : nullptr, Order, trees, sc, x, y, Hit[pixel]);
    Vis[pixel] = make_float4(vis.x, vis.y, vis.z, 1.0f);
  }
}

// The same over the SAH hierarchy with the batch's tree records, one wave
// per block; dynamic LDS: W::ldsStack stack entries per lane, then the view
// table.
template <class W>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(5))) void k_sunVisibilityViewsRec(
    const float* P, const float4* Nodes4, const uint32_t* Order, const float* Cs, const ViewDesc* views, int nviews,
    int totalTiles, const float4* Hit, float4* Vis, const float4* accel, const float4* treeRecs) {
  static_assert(W::sah && W::ldsStack > 0, "a record-tree batch pass walks the SAH hierarchy with an LDS stack");
  ViewDesc* sv = reinterpret_cast<ViewDesc*>(sahLdsStack + (size_t)W::ldsStack * 64);
  for (int i = threadIdx.x; i < nviews; i += blockDim.x) sv[i] = views[i];
  __syncthreads();
  const TreeRecs trees{treeRecs};
  const int lane = threadIdx.x;
  for (int t = blockIdx.x; t < totalTiles; t += gridDim.x) {
    const int v = findView(sv, nviews, t);
    const ViewDesc d = sv[v];
    const float* C = Cs + 64 * v;
    if (!(C[22] > 0.5f)) continue;
    const int lt = t - d.tileBase;
    const int x = (lt % d.tilesX) * 8 + (lane & 7), y = (lt / d.tilesX) * 8 + (lane >> 3);
    if (x >= d.width || y >= d.height) continue;
    const ShadeCtx sc = makeShadeCtx(C, d.width, d.height);
    const int pixel = d.pixelBase + y * d.width + x;
    v3 vis = sunVisibilityOne<W>(P, Nodes4, accel, Order, trees, sc, x, y, Hit[pixel]);
    Vis[pixel] = make_float4(vis.x, vis.y, vis.z, 1.0f);
  }
}
// The world seed's tree records for the record-tree kernels.
__global__ void k_batchTreeRecs(const float* Cs, float4* recs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kTrees) storeTreeRec(recs, i, makeTree(i, (int)Cs[31]));
}

// Shading of every view's pixels from Hit; with Vis, the sun visibility comes
// from k_sunVisibilityViews (the shadow traversals compiled out).
template <bool kTraceVis>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kTraceVis ? 1 : 3))) void k_shadeViews(
    const float* P, const float4* Nodes4, const uint32_t* Order, const float* Cs, const ViewDesc* views, int nviews,
    int totalTiles, const float4* Hit, float4* Linear, float4* Guide, const float4* Vis) {
  __shared__ TreeInst trees[kTrees];
  __shared__ ViewDesc sv[kMaxViews];
  for (int i = threadIdx.x; i < kTrees; i += blockDim.x) trees[i] = makeTree(i, (int)Cs[31]);
  for (int i = threadIdx.x; i < nviews; i += blockDim.x) sv[i] = views[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int t = blockIdx.x * 4 + wave; t < totalTiles; t += gridDim.x * 4) {
    const int v = findView(sv, nviews, t);
    const ViewDesc d = sv[v];
    const int lt = t - d.tileBase;
    const int x = (lt % d.tilesX) * 8 + (lane & 7), y = (lt / d.tilesX) * 8 + (lane >> 3);
    if (x >= d.width || y >= d.height) continue;
    const ShadeCtx sc = makeShadeCtx(Cs + 64 * v, d.width, d.height);
    const int pixel = d.pixelBase + y * d.width + x;
    float4 l, g;
    // shadeOne indexes Vis by the pixel within its view
    shadeOne<Walk<true, false, false>, kTraceVis>(P, Nodes4, Order, trees, sc, x, y, Hit[pixel], l, g, nullptr,
                                                  Vis ? Vis + d.pixelBase : nullptr);
    Linear[pixel] = l;
    Guide[pixel] = g;
  }
}

// filterLighting (and with kResolve resolvePixels) over every view's pixels
// in one launch: one 32x8 block of one view per workgroup, found by a binary
// search of the view table on the block index (uniform per block); each
// view's pixels get the single-view kernels' arithmetic.
__device__ __forceinline__ int findFilterView(const ViewDesc* views, int n, int block) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {  // last view with filterBase <= block
    int mid = (lo + hi + 1) >> 1;
    if (views[mid].filterBase <= block) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
template <bool kResolve>
__global__ __launch_bounds__(kFilterW * kFilterH) void k_filterViews(const float4* Linear, const float4* Guide,
                                                                    const float4* Hit, float4* Filtered,
                                                                    const float* Cs, const ViewDesc* views,
                                                                    int nviews, FilterWeights fw, float4* History,
                                                                    uint32_t* Pixels) {
  const int v = findFilterView(views, nviews, (int)blockIdx.x);
  const ViewDesc d = views[v];
  const int lb = (int)blockIdx.x - d.filterBase;
  const size_t o = (size_t)d.pixelBase;
  filterBlock<kResolve>(Linear + o, Guide + o, Hit + o, Filtered + o, Cs + 64 * v, d.width, d.height, fw,
                        History + o, Pixels + o, lb % d.filterX, lb / d.filterX);
}
// resolvePixels alone over every view's pixels, in the same blocks.
__global__ __launch_bounds__(kFilterW * kFilterH) void k_resolveViews(const float4* Filtered, float4* History,
                                                                     uint32_t* Pixels, const float* Cs,
                                                                     const ViewDesc* views, int nviews) {
  const int v = findFilterView(views, nviews, (int)blockIdx.x);
  const ViewDesc d = views[v];
  const int lb = (int)blockIdx.x - d.filterBase;
  const int x = (lb % d.filterX) * kFilterW + threadIdx.x, y = (lb / d.filterX) * kFilterH + threadIdx.y;
  if (x >= d.width || y >= d.height) return;
  const size_t o = (size_t)d.pixelBase;
  const int i = y * d.width + x;
  const float4 f = Filtered[o + i];
  resolveOne(mk(f.x, f.y, f.z), History + o, Pixels + o, Cs + 64 * v, x, y, d.width, d.height, i);
}

typedef void (*BatchKernel)(const float*, const float4*, const uint32_t*, const float*, const ViewDesc*, int, int,
                            float4*, float4*, float4*, const float4*, const float4*);
struct BatchVariant {
  BatchKernel kernel;  // fused trace + shade, or (split) the trace followed by k_shadeViews
  bool split;
  int waves;
  const char* name;
  int ldsPerWave = 0;     // dynamic LDS bytes per wave (stack entries) of a record-tree trace
  bool recTrees = false;  // tree records from the batch's copy; stack and view table in dynamic LDS
};
const BatchVariant kBatchVariants[] = {
    {k_traceShadeViews<1, Walk<true, false, false>>, false, 1, "1 wave/block, culled walk"},
    {k_traceShadeViews<2, Walk<true, false, false>>, false, 2, "2 waves/block, culled walk"},
    {k_traceShadeViews<4, Walk<true, false, false>>, false, 4, "4 waves/block, culled walk"},
    {k_traceShadeViews<1, Walk<false, false, false>>, false, 1, "1 wave/block, plain walk"},
    {k_traceShadeViews<4, Walk<false, false, false>>, false, 4, "4 waves/block, plain walk"},
    {k_traceViews<1, Walk<true, false, false>>, true, 1, "split: trace 1 wave/block culled, visibility, shade"},
    {k_traceViews<2, Walk<true, false, false>>, true, 2, "split: trace 2 waves/block culled, visibility, shade"},
    // over the heap these are 6 and 5; over the SAH hierarchy their traces
    // read the tree records from device memory and keep the stack in LDS
    {k_traceViews<2, Walk<true, false, false>>, true, 2,
     "split as 6 (SAH walk: 1 wave/block, record trees, 20-entry LDS stack)"},
    {k_traceViews<1, Walk<true, false, false>>, true, 1,
     "split as 5 (SAH walk: 2 waves/block, record trees, 16-entry LDS stack)"},
    {k_traceViews<2, Walk<true, false, false>>, true, 2,
     "split as 6 (SAH walk: 1 wave/block, record trees, 20-entry LDS stack, three interior steps per gathered "
     "round)"},
};
constexpr int kBatchVariantCount = sizeof(kBatchVariants) / sizeof(kBatchVariants[0]);
int batchVariantCount() { return kBatchVariantCount; }
// The same launch shapes over the SAH hierarchy (arbor_batch_set_accel).
// SAH walk, whole stack in scratch (LDS holds the view and tree tables);
// object tests gathered at 6/8 of a wave's live lanes for the primary rays
// and 5/8 for the sun rays, the thresholds that timed best for the frame
// kernels.
using WalkSahB = Walk<true, false, false, true, 0, 6>;
using WalkSahBVis = Walk<true, false, false, true, 0, 5>;
// The record-tree split passes: the trace with 20 (1-wave blocks) or 16
// (2-wave blocks) LDS stack entries per lane, the visibility pass with 14
// (1-wave blocks, 5 waves per SIMD, three interior steps per gathered round).
using WalkSahBRec1 = Walk<true, false, false, true, 20, 6>;
using WalkSahBRec3 = Walk<true, false, false, true, 20, 5, 3>;  // three interior steps per gathered round
using WalkSahBRec2 = Walk<true, false, false, true, 16, 6>;
using WalkSahBVisRec = Walk<true, false, false, true, 14, 5, 3>;
constexpr int kBatchLds(int k) { return k * 64 * 8; }
const BatchVariant kBatchVariantsSah[] = {
    {k_traceShadeViews<1, WalkSahB>, false, 1, ""}, {k_traceShadeViews<2, WalkSahB>, false, 2, ""},
    {k_traceShadeViews<4, WalkSahB>, false, 4, ""}, {k_traceShadeViews<1, WalkSahB>, false, 1, ""},
    {k_traceShadeViews<4, WalkSahB>, false, 4, ""}, {k_traceViews<1, WalkSahB>, true, 1, ""},
    {k_traceViews<2, WalkSahB>, true, 2, ""},
    {k_traceViewsRec<1, WalkSahBRec1>, true, 1, "", kBatchLds(20), true},
    {k_traceViewsRec<2, WalkSahBRec2>, true, 2, "", kBatchLds(16), true},
    {k_traceViewsRec<1, WalkSahBRec3>, true, 1, "", kBatchLds(20), true},
};
static_assert(sizeof(kBatchVariantsSah) == sizeof(kBatchVariants), "one SAH entry per batch variant");

}  // namespace arbor

using namespace arbor;

struct arbor_batch {
  int maxViews = 0;
  long long maxPixels = 0;
  int nviews = 0;
  long long npixels = 0;
  int totalTiles = 0;
  int filterBlocks = 0;  // 32x8 filter blocks over every view
  std::vector<ViewDesc> host;
  ViewDesc* dViews = nullptr;
  float* dC = nullptr;
  float4 *Hit = nullptr, *Linear = nullptr, *Guide = nullptr, *Filtered = nullptr, *History = nullptr;
  float4* Vis = nullptr;  // sun visibility of the split pass
  uint32_t* Pixels = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const float4* accel = nullptr;  // SAH hierarchy (arbor_batch_set_accel), or NULL for the heap walk
  float4* trees = nullptr;        // the world seed's tree records (set_views), for the record-tree kernels
};

namespace {
template <class T>
int dalloc(T** p, size_t bytes) {
  return hipMalloc(reinterpret_cast<void**>(p), bytes) == hipSuccess ? ARBOR_OK : ARBOR_ERR_DEVICE;
}
template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}
void dropGraph(arbor_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// The launches of one batched render on `stream` (also what gets captured).
int issue(arbor_batch* b, const float* P, const float* Nodes, const unsigned* Order, int stages, hipStream_t s) {
  if (stages & ARBOR_BATCH_TRACE_SHADE) {
    const BatchVariant& v = (b->accel ? kBatchVariantsSah : kBatchVariants)[tuning().batch_variant];
    const size_t table = sizeof(ViewDesc) * (size_t)b->nviews;
    int blocks = (b->totalTiles + v.waves - 1) / v.waves;
    if (blocks > tuning().grid_cap) blocks = tuning().grid_cap;
    hipLaunchKernelGGL(v.kernel, dim3(blocks), dim3(v.waves * 64), v.recTrees ? v.waves * v.ldsPerWave + table : 0, s,
                       P, reinterpret_cast<const float4*>(Nodes), Order, b->dC, b->dViews, b->nviews, b->totalTiles,
                       b->Hit, b->Linear, b->Guide, b->accel, b->trees);
    int rc = launch_status();
    if (rc != ARBOR_OK) return rc;
    if (v.split) {
      int vblocks = (b->totalTiles + 1) / 2, sblocks = (b->totalTiles + 3) / 4;
      if (vblocks > tuning().grid_cap) vblocks = tuning().grid_cap;
      if (sblocks > tuning().grid_cap) sblocks = tuning().grid_cap;
      if (b->accel && v.recTrees) {
        int rblocks = b->totalTiles < tuning().grid_cap ? b->totalTiles : tuning().grid_cap;
        hipLaunchKernelGGL(k_sunVisibilityViewsRec<WalkSahBVisRec>, dim3(rblocks), dim3(64),
                           kBatchLds(WalkSahBVisRec::ldsStack) + table, s, P, reinterpret_cast<const float4*>(Nodes),
                           Order, b->dC, b->dViews, b->nviews, b->totalTiles, b->Hit, b->Vis, b->accel, b->trees);
      } else {
        typedef void (*VisViewsKernel)(const float*, const float4*, const uint32_t*, const float*, const ViewDesc*,
                                       int, int, const float4*, float4*, const float4*);
        const VisViewsKernel vk =
            b->accel ? k_sunVisibilityViews<WalkSahBVis> : k_sunVisibilityViews<Walk<true, false, false>>;
        hipLaunchKernelGGL(vk, dim3(vblocks), dim3(128), 0, s, P, reinterpret_cast<const float4*>(Nodes), Order,
                           b->dC, b->dViews, b->nviews, b->totalTiles, b->Hit, b->Vis, b->accel);
      }
      if ((rc = launch_status()) != ARBOR_OK) return rc;
      hipLaunchKernelGGL(k_shadeViews<false>, dim3(sblocks), dim3(256), 0, s, P,
                         reinterpret_cast<const float4*>(Nodes), Order, b->dC, b->dViews, b->nviews, b->totalTiles,
                         b->Hit, b->Linear, b->Guide, b->Vis);
      if ((rc = launch_status()) != ARBOR_OK) return rc;
    }
  }
  const bool filter = (stages & ARBOR_BATCH_FILTER) != 0, resolve = (stages & ARBOR_BATCH_RESOLVE) != 0;
  if (filter) {
    hipLaunchKernelGGL(resolve ? k_filterViews<true> : k_filterViews<false>, dim3(b->filterBlocks),
                       dim3(kFilterW, kFilterH), 0, s, b->Linear, b->Guide, b->Hit, b->Filtered, b->dC, b->dViews,
                       b->nviews, filterWeights(), b->History, b->Pixels);
    int rc = launch_status();
    if (rc != ARBOR_OK) return rc;
  } else if (resolve) {
    hipLaunchKernelGGL(k_resolveViews, dim3(b->filterBlocks), dim3(kFilterW, kFilterH), 0, s, b->Filtered,
                       b->History, b->Pixels, b->dC, b->dViews, b->nviews);
    int rc = launch_status();
    if (rc != ARBOR_OK) return rc;
  }
  return ARBOR_OK;
}
}  // namespace

extern "C" {

int arbor_batch_create(arbor_batch** out, int max_views, long long max_pixels) {
  if (!out) return ARBOR_ERR_NULL;
  *out = nullptr;
  if (max_views < 1 || max_views > kMaxViews || max_pixels < 1 || max_pixels > (1ll << 31) / 16)
    return ARBOR_ERR_ARG;
  arbor_batch* b = new arbor_batch();
  b->maxViews = max_views;
  b->maxPixels = max_pixels;
  size_t n = (size_t)max_pixels;
  int rc = ARBOR_OK;
  rc |= dalloc(&b->dViews, sizeof(ViewDesc) * max_views);
  rc |= dalloc(&b->dC, sizeof(float) * 64 * max_views);
  rc |= dalloc(&b->Hit, 16 * n);
  rc |= dalloc(&b->Linear, 16 * n);
  rc |= dalloc(&b->Guide, 16 * n);
  rc |= dalloc(&b->Filtered, 16 * n);
  rc |= dalloc(&b->History, 16 * n);
  rc |= dalloc(&b->Vis, 16 * n);
  rc |= dalloc(&b->Pixels, 4 * n);
  rc |= dalloc(&b->trees, sizeof(float4) * kTreeRecF4 * kTrees);
  if (rc != ARBOR_OK || hipMemset(b->History, 0, 16 * n) != hipSuccess) {
    arbor_batch_destroy(b);
    return ARBOR_ERR_DEVICE;
  }
  *out = b;
  return ARBOR_OK;
}

int arbor_batch_set_accel(arbor_batch* b, const void* accel) {
  if (!b) return ARBOR_ERR_NULL;
  b->accel = static_cast<const float4*>(accel);
  dropGraph(b);  // a captured graph holds the previous walk's launches
  return ARBOR_OK;
}

int arbor_batch_destroy(arbor_batch* b) {
  if (!b) return ARBOR_ERR_NULL;
  (void)hipDeviceSynchronize();
  dropGraph(b);
  dfree(b->dViews); dfree(b->dC); dfree(b->Hit); dfree(b->Linear); dfree(b->Guide); dfree(b->Filtered);
  dfree(b->History); dfree(b->Pixels); dfree(b->Vis); dfree(b->trees);
  delete b;
  return ARBOR_OK;
}

int arbor_batch_set_views(arbor_batch* b, int nviews, const float* cameras, const int* widths, const int* heights) {
  if (!b || !cameras || !widths || !heights) return ARBOR_ERR_NULL;
  if (nviews < 1 || nviews > b->maxViews) return ARBOR_ERR_ARG;
  std::vector<ViewDesc> h(nviews);
  long long pix = 0;
  int tiles = 0, fblocks = 0;
  for (int i = 0; i < nviews; i++) {
    int w = widths[i], hh = heights[i];
    if (!dims_ok(w, hh)) return ARBOR_ERR_ARG;
    // all views see one forest: the instancing seed C[31] must agree
    if (cameras[64 * i + 31] != cameras[31]) return ARBOR_ERR_ARG;
    ViewDesc d{};
    d.width = w;
    d.height = hh;
    d.tilesX = (w + 7) / 8;
    d.tileBase = tiles;
    d.pixelBase = (int)pix;
    d.filterBase = fblocks;
    d.filterX = (w + kFilterW - 1) / kFilterW;
    fblocks += d.filterX * ((hh + kFilterH - 1) / kFilterH);
    tiles += d.tilesX * ((hh + 7) / 8);
    pix += (long long)w * hh;
    h[i] = d;
  }
  if (pix > b->maxPixels) return ARBOR_ERR_ARG;
  if (hipMemcpy(b->dViews, h.data(), sizeof(ViewDesc) * nviews, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->dC, cameras, sizeof(float) * 64 * nviews, hipMemcpyHostToDevice) != hipSuccess)
    return ARBOR_ERR_DEVICE;
  // the tree records of the views' world seed (synchronously, as the copies)
  hipLaunchKernelGGL(k_batchTreeRecs, dim3((kTrees + 255) / 256), dim3(256), 0, nullptr, b->dC, b->trees);
  if (launch_status() != ARBOR_OK || hipStreamSynchronize(nullptr) != hipSuccess) return ARBOR_ERR_DEVICE;
  b->host = h;
  b->nviews = nviews;
  b->npixels = pix;
  b->totalTiles = tiles;
  b->filterBlocks = fblocks;
  dropGraph(b);  // the captured sequence depends on the layout
  return ARBOR_OK;
}

int arbor_batch_render(arbor_batch* b, const float* P, const float* Nodes, const unsigned* Order, int stages,
                       hipStream_t stream) {
  if (!b || !P || !Nodes || !Order) return ARBOR_ERR_NULL;
  if (b->nviews < 1) return ARBOR_ERR_STATE;
  if (stages <= 0 || stages > ARBOR_BATCH_ALL) return ARBOR_ERR_ARG;
  return issue(b, P, Nodes, Order, stages, stream);
}

int arbor_batch_capture(arbor_batch* b, const float* P, const float* Nodes, const unsigned* Order, int stages,
                        hipStream_t stream) {
  if (!b || !P || !Nodes || !Order) return ARBOR_ERR_NULL;
  if (b->nviews < 1) return ARBOR_ERR_STATE;
  if (stages <= 0 || stages > ARBOR_BATCH_ALL) return ARBOR_ERR_ARG;
  if (!stream) return ARBOR_ERR_ARG;  // capture needs a created stream, not the legacy default stream
  dropGraph(b);
  if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return ARBOR_ERR_DEVICE;
  int rc = issue(b, P, Nodes, Order, stages, stream);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(stream, &g);
  if (rc != ARBOR_OK || e != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    return rc != ARBOR_OK ? rc : ARBOR_ERR_DEVICE;
  }
  hipGraphExec_t x = nullptr;
  if (hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(g);
    return ARBOR_ERR_DEVICE;
  }
  b->graph = g;
  b->exec = x;
  return ARBOR_OK;
}

int arbor_batch_launch(arbor_batch* b, hipStream_t stream) {
  if (!b) return ARBOR_ERR_NULL;
  if (!b->exec) return ARBOR_ERR_STATE;
  return hipGraphLaunch(b->exec, stream) == hipSuccess ? ARBOR_OK : ARBOR_ERR_LAUNCH;
}

void* arbor_batch_buffer(arbor_batch* b, int view, const char* name) {
  if (!b || !name || view < 0 || view >= b->nviews) return nullptr;
  size_t o = (size_t)b->host[view].pixelBase;
  if (!strcmp(name, "Hit")) return b->Hit + o;
  if (!strcmp(name, "Linear")) return b->Linear + o;
  if (!strcmp(name, "Guide")) return b->Guide + o;
  if (!strcmp(name, "Filtered")) return b->Filtered + o;
  if (!strcmp(name, "History")) return b->History + o;
  if (!strcmp(name, "Pixels")) return b->Pixels + o;
  if (!strcmp(name, "C")) return b->dC + 64 * view;
  return nullptr;
}

long long arbor_batch_pixels(const arbor_batch* b) { return b ? b->npixels : -1; }
int arbor_batch_tiles(const arbor_batch* b) { return b ? b->totalTiles : -1; }

// Times every batched-kernel variant on the batch's current views (trace +
// shade stage, `reps` launches each after one warm-up, HIP events on
// `stream`), keeps the fastest in the process tuning ("batch_variant"), and
// writes {"best": i, "ms": [...], "names": [...]} to json.
int arbor_autotune_batch(arbor_batch* b, const float* P, const float* Nodes, const unsigned* Order, int reps,
                         hipStream_t stream, char* json, int cap) {
  if (!b || !P || !Nodes || !Order) return ARBOR_ERR_NULL;
  if (b->nviews < 1) return ARBOR_ERR_STATE;
  if (reps < 1 || reps > 1000) return ARBOR_ERR_ARG;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return ARBOR_ERR_DEVICE;
  Tuning& t = tuning();
  const int saved = t.batch_variant;
  std::vector<float> ms(kBatchVariantCount, 0.0f);
  int best = 0, rc = ARBOR_OK;
  for (int v = 0; v < kBatchVariantCount && rc == ARBOR_OK; v++) {
    t.batch_variant = v;
    rc = issue(b, P, Nodes, Order, ARBOR_BATCH_TRACE_SHADE, stream);
    (void)hipEventRecord(e0, stream);
    for (int r = 0; r < reps && rc == ARBOR_OK; r++) rc = issue(b, P, Nodes, Order, ARBOR_BATCH_TRACE_SHADE, stream);
    (void)hipEventRecord(e1, stream);
    if (hipEventSynchronize(e1) != hipSuccess) rc = ARBOR_ERR_DEVICE;
    (void)hipEventElapsedTime(&ms[v], e0, e1);
    ms[v] /= (float)reps;
    if (ms[v] < ms[best]) best = v;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  t.batch_variant = rc == ARBOR_OK ? best : saved;
  dropGraph(b);
  if (rc != ARBOR_OK) return rc;
  if (json && cap > 0) {
    std::string s = "{\"best\":" + std::to_string(best) + ",\"ms\":[";
    for (int v = 0; v < kBatchVariantCount; v++) s += (v ? "," : "") + std::to_string(ms[v]);
    s += "],\"names\":[";
    for (int v = 0; v < kBatchVariantCount; v++) s += std::string(v ? "," : "") + "\"" + kBatchVariants[v].name + "\"";
    s += "]}";
    if ((int)s.size() >= cap) return ARBOR_ERR_ARG;
    memcpy(json, s.c_str(), s.size() + 1);
  }
  return ARBOR_OK;
}

}  // extern "C"
