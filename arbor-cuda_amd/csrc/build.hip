// One-time scene build on gfx950: tree growth, primitive emission, Morton keys,
// LDS radix sort and heap-BVH emission, plus the camera state kernels.
//
// growTree / makePrimitives / makeKeys / sortPass / makeBounds / reduceBounds /
// initCamera / updateCamera keep the reference's per-thread semantics
// (Arbor.cu:76-142,181-182). Their transcendentals are evaluated in double and
// rounded once (sin_cr, cos_cr, pow_cr) so the float results agree with the
// reference harness's glibc calls. arbor_sortKeys and arbor_buildBVH replace
// the 120 bitonic stages and 15 reduction launches with one workgroup.
#include "arbor_camera.hpp"
#include "arbor_trace.hpp"
#include "launch.hpp"

namespace arbor {

// ---------------------------------------------------------------- growTree
__device__ void storeCurve(float* B, int id, v3 a, v3 b, v3 c, v3 d, float ra, float rb, float arc, int seed) {
  int o = id * kCS;
  st3(B, o, a); st3(B, o + 3, b); st3(B, o + 6, c); st3(B, o + 9, d);
  B[o + 12] = ra; B[o + 13] = rb; B[o + 14] = arc; B[o + 15] = (float)seed; B[o + 16] = length(d - a);
}

// Arbor.cu:76-102. Launched with exactly the thread count of the level
// (9/12/72/288/1728, native_harness.cpp:46).
__global__ void k_growTree(float* B, int level, int seed, int count) {
  int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= count) return;
  if (level == 0) {
    if (i == 0)
      storeCurve(B, 0, mk(0.0f, -0.15f, 0.0f), mk(-0.22f, 1.1f, 0.16f), mk(0.23f, 2.9f, -0.19f), mk(0.04f, 4.2f, 0.05f),
                 0.68f, 0.23f, 0.0f, seed);
    if (i < 8) {
      float a = (float)i * kPi * 0.25f + (rnd(i, seed) - 0.5f) * 0.5f;
      v3 d = mk(cos_cr(a), 0.0f, sin_cr(a));
      float len = 2.0f + rnd(i + 10, seed) * 1.4f;
      v3 p = d * 0.32f + mk(0.0f, 0.27f, 0.0f);
      storeCurve(B, i + 1, p, d * 0.8f + mk(0.0f, 0.10f, 0.0f), d * (len * 0.68f) + mk(0.0f, -0.06f, 0.0f),
                 d * len + mk(0.0f, -0.16f, 0.0f), 0.235f, 0.012f, 0.0f, seed + i * 37);
    }
    return;
  }
  if (level == 1) {
    if (i >= 12) return;
    float t = 0.31f + 0.65f * (float)i / 11.0f;
    float a = (float)i * 2.399963f + 0.31f;
    float r = rnd(i + 37, seed);
    float len = 3.2f + 2.1f * r;
    v3 base = curveAt(B, 0, t);
    v3 d = mk(cos_cr(a), 0.0f, sin_cr(a));
    v3 side = mk(-d.z, 0.0f, d.x);
    v3 end = base + d * (len * (0.87f - 0.30f * t)) + mk(0.0f, len * (0.60f + 0.40f * t), 0.0f) +
             side * ((r - 0.5f) * 1.2f);
    float ra = curveRadius(B, 0, t) * (0.46f + 0.13f * r);
    storeCurve(B, 9 + i, base, base + d * (len * 0.29f) + mk(0.0f, len * 0.12f, 0.0f),
               end - d * (len * 0.17f) - mk(0.0f, len * 0.26f, 0.0f), end, ra, ra * 0.14f, B[14] + t * 4.2f,
               seed + i * 19);
    return;
  }
  int children = level == 2 ? 6 : (level == 3 ? 4 : 6);
  int offset = level == 2 ? 21 : (level == 3 ? 93 : 381);
  int parentOffset = level == 2 ? 9 : (level == 3 ? 21 : 93);
  int parent = parentOffset + i / children, child = i % children;
  float h = rnd(i + level * 730, seed);
  float t = 0.23f + 0.72f * ((float)child + 0.35f) / (float)children;
  v3 p = curveAt(B, parent, t);
  v3 tangent = curveTangent(B, parent, t);
  v3 u = basisU(tangent);
  v3 v = cross(tangent, u);
  float angle = (float)child * 2.399963f + h * 0.8f;
  v3 lateral = u * cos_cr(angle) + v * sin_cr(angle);
  v3 d = unit(tangent * 0.52f + lateral * (level == 4 ? 0.64f : 0.80f) + mk(0.0f, 0.26f, 0.0f));
  float len = level == 2 ? 1.65f + 1.10f * h : (level == 3 ? 0.85f + 0.65f * h : 0.43f + 0.37f * h);
  len *= 0.84f + 0.25f * (1.0f - t);
  v3 end = p + d * len + mk(0.0f, len * 0.16f, 0.0f);
  float radius = curveRadius(B, parent, t) * (level == 4 ? 0.30f : 0.40f);
  radius = fmaxf(level == 4 ? 0.0035f : 0.008f, radius);
  storeCurve(B, offset + i, p, p + tangent * (len * 0.25f) + lateral * (len * 0.11f),
             end - d * (len * 0.28f) + mk(0.0f, len * 0.05f, 0.0f), end, radius, radius * (level == 4 ? 0.30f : 0.22f),
             B[parent * kCS + 14] + t * B[parent * kCS + 16], seed + (offset + i) * 31);
}

// ---------------------------------------------------------- makePrimitives
// Arbor.cu:103-131, one record per thread.
__global__ void k_makePrimitives(const float* B, float* P, int seed) {
  int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= kCapacity) return;
  int b = i * kPS;
  float r[kPS];
#pragma unroll
  for (int j = 0; j < kPS; j++) r[j] = 0.0f;
  r[3] = -1.0f;
  if (i < kObjects) {
    if (i < kWood) {
      int c, segment, segments;
      if (i < 12) { c = 0; segment = i; segments = 12; }
      else if (i < 60) { c = 1 + (i - 12) / 6; segment = (i - 12) % 6; segments = 6; }
      else if (i < 156) { c = 9 + (i - 60) / 8; segment = (i - 60) % 8; segments = 8; }
      else if (i < 516) { c = 21 + (i - 156) / 5; segment = (i - 156) % 5; segments = 5; }
      else if (i < 1668) { c = 93 + (i - 516) / 4; segment = (i - 516) % 4; segments = 4; }
      else { c = 381 + (i - 1668) / 2; segment = (i - 1668) % 2; segments = 2; }
      float t0 = (float)segment / (float)segments, t1 = (float)(segment + 1) / (float)segments;
      v3 p0 = curveAt(B, c, t0), p1 = curveAt(B, c, t1);
      r[0] = p0.x; r[1] = p0.y; r[2] = p0.z; r[3] = 0.0f;
      r[4] = p1.x; r[5] = p1.y; r[6] = p1.z;
      r[7] = curveRadius(B, c, t0);
      r[11] = curveRadius(B, c, t1);
      r[15] = B[c * kCS + 15];
      float arclength = B[c * kCS + 14];
      for (int j = 0; j < 12; j++)
        if (j < segment)
          arclength += length(curveAt(B, c, (float)(j + 1) / (float)segments) - curveAt(B, c, (float)j / (float)segments));
      r[16] = arclength;
      r[17] = (float)c;
      r[18] = B[c * kCS + 12];
      v3 ax = basisU(curveTangent(B, c, 0.0f));
      r[8] = ax.x; r[9] = ax.y; r[10] = ax.z;
    } else if (i < kWood + kLeaves) {
      int leaf = i - kWood, twig = leaf / 14, k = leaf % 14, c = 381 + twig;
      float h = rnd(leaf + 13, seed);
      float t = 0.16f + 0.81f * ((float)k + 0.4f) / 14.0f;
      v3 anchor = curveAt(B, c, t);
      v3 axis = curveTangent(B, c, t);
      v3 bu = basisU(axis);
      v3 bv = cross(axis, bu);
      float angle = (float)k * 2.399963f + rnd(twig, seed) * 6.28f;
      v3 out = unit(bu * cos_cr(angle) + bv * sin_cr(angle));
      v3 u = unit(out * 0.82f + axis * 0.35f + mk(0.0f, 0.20f, 0.0f));
      v3 normal = unit(mk(0.0f, 1.0f, 0.0f) + out * (0.20f + h * 0.38f));
      v3 v = unit(cross(normal, u));
      normal = unit(cross(u, v));
      r[0] = anchor.x; r[1] = anchor.y; r[2] = anchor.z; r[3] = 1.0f;
      r[4] = u.x; r[5] = u.y; r[6] = u.z;
      r[7] = 0.145f + 0.075f * h;
      r[8] = v.x; r[9] = v.y; r[10] = v.z;
      r[11] = r[7] * (0.34f + 0.07f * rnd(leaf + 37, seed));
      r[12] = normal.x; r[13] = normal.y; r[14] = normal.z;
      r[15] = (float)(leaf + seed * 11);
      r[16] = 0.023f + 0.024f * rnd(leaf + 149, seed);
      r[17] = 0.013f + 0.016f * h;
      r[18] = (rnd(leaf + 777, seed) - 0.5f) * 0.10f;
      r[19] = (float)twig;
    } else {
      int rock = i - kWood - kLeaves;
      float a = rnd(rock * 3 + 1, seed) * 6.2831853f;
      float rr = 0.8f + 4.5f * sqrtf(rnd(rock * 3 + 2, seed));
      float size = 0.045f + 0.19f * pow_cr(rnd(rock + 371, seed), 3.0f);
      r[0] = cos_cr(a) * rr; r[1] = size * 0.30f - 0.015f; r[2] = sin_cr(a) * rr; r[3] = 2.0f;
      r[4] = size; r[5] = size * 0.55f; r[6] = size * (0.7f + rnd(rock + 16, seed) * 0.5f);
      r[15] = (float)(rock + seed);
    }
  }
  float4* out = reinterpret_cast<float4*>(P + b);
#pragma unroll
  for (int j = 0; j < kPS / 4; j++) out[j] = make_float4(r[4 * j], r[4 * j + 1], r[4 * j + 2], r[4 * j + 3]);
}

// ------------------------------------------------------- keys and bounds
__device__ __forceinline__ uint32_t spread10(uint32_t x) {
  x &= 1023u;
  x = (x | (x << 16)) & 50331903u;
  x = (x | (x << 8)) & 50393103u;
  x = (x | (x << 4)) & 51130563u;
  x = (x | (x << 2)) & 153391689u;
  return x;
}
// makeKeys for one record (Arbor.cu:139).
__device__ __forceinline__ uint32_t mortonKey(const float* P, int i) {
  Prim pr = loadPrim(P, i);
  if (pr.type() < 0.0f) return 4294967295u;
  v3 lo, hi;
  primBounds(pr, lo, hi);
  v3 p = (lo + hi) * 0.5f;
  uint32_t x = (uint32_t)(sat((p.x + 12.0f) / 24.0f) * 1023.0f);
  uint32_t y = (uint32_t)(sat((p.y + 1.0f) / 18.0f) * 1023.0f);
  uint32_t z = (uint32_t)(sat((p.z + 12.0f) / 24.0f) * 1023.0f);
  return spread10(x) | (spread10(y) << 1) | (spread10(z) << 2);
}
__global__ void k_makeKeys(const float* P, uint32_t* Keys, uint32_t* Order) {
  int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= kCapacity) return;
  Order[i] = (uint32_t)i;
  Keys[i] = mortonKey(P, i);
}
// Arbor.cu:140: one bitonic stage; only the lower index of a pair writes.
__global__ void k_sortPass(uint32_t* Keys, uint32_t* Order, int distance, int span) {
  int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= kCapacity) return;
  int j = i ^ distance;
  if (j <= i) return;
  uint32_t a = Keys[i], b = Keys[j], ia = Order[i], ib = Order[j];
  bool greater = a > b || (a == b && ia > ib);
  bool ascending = (i & span) == 0;
  if (greater == ascending) { Keys[i] = b; Keys[j] = a; Order[i] = ib; Order[j] = ia; }
}
// Leaf node CAPACITY+i (Arbor.cu:141).
__device__ __forceinline__ void writeLeaf(const float* P, float* Nodes, int i, uint32_t id) {
  Prim pr = loadPrim(P, (int)id);
  v3 lo, hi;
  primBounds(pr, lo, hi);
  float4* n = reinterpret_cast<float4*>(Nodes + (size_t)(kCapacity + i) * 8);
  n[0] = make_float4(lo.x, lo.y, lo.z, pr.type() >= 0.0f ? 1.0f : 0.0f);
  n[1] = make_float4(hi.x, hi.y, hi.z, (float)id);
}
__global__ void k_makeBounds(const float* P, const uint32_t* Order, float* Nodes) {
  int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= kCapacity) return;
  writeLeaf(P, Nodes, i, Order[i]);
}
// Interior node first+i (Arbor.cu:142).
__device__ __forceinline__ void writeInterior(float* Nodes, int node) {
  const float4* c = reinterpret_cast<const float4*>(Nodes) + (size_t)node * 4;
  float4 a0 = c[0], a1 = c[1], b0 = c[2], b1 = c[3];
  float4* n = reinterpret_cast<float4*>(Nodes) + (size_t)node * 2;
  n[0] = make_float4(fminf(a0.x, b0.x), fminf(a0.y, b0.y), fminf(a0.z, b0.z), a0.w + b0.w);
  n[1] = make_float4(fmaxf(a1.x, b1.x), fmaxf(a1.y, b1.y), fmaxf(a1.z, b1.z), -1.0f);
}
__global__ void k_reduceBounds(float* Nodes, int first) {
  int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= first) return;
  writeInterior(Nodes, first + i);
}

// ------------------------------------------------- single-workgroup sort
// 1024 threads = 16 wave64s, 32 elements per thread. Element e lives in wave
// e/2048, item (e%2048)/64, lane e%64; that order is the stable order. Each
// 8-bit LSD pass ranks digits per wave with ballots (the peers of a lane's
// digit are found bit by bit), turns per-wave counts into global offsets and
// scatters through one 128 KiB LDS array. The payload carried with each key is
// its 15-bit input slot, packed with the element's rank in one register, so a
// pass holds 64 live values per thread; the caller gathers by slot at the end.
constexpr int kSortThreads = 1024;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kItems = kCapacity / kSortThreads;

struct SortShared {
  uint32_t stage[kCapacity];       // 128 KiB scatter buffer
  uint32_t hist[kSortWaves][256];  // per-wave digit counts, then offsets
  uint32_t total[256];
};

__device__ __forceinline__ uint64_t lanesBelow(int lane) { return lane == 0 ? 0ull : (~0ull >> (64 - lane)); }
__device__ __forceinline__ int sortHome() { return (threadIdx.x >> 6) * (kItems * 64) + (threadIdx.x & 63); }

// Stable permutation of (key[], pk[]) by the 8-bit digit of key at `shift`.
// pk[i] low 16 bits: slot; high 16 bits: scratch for the rank.
__device__ __forceinline__ void radixPass(SortShared& sh, uint32_t (&key)[kItems], uint32_t (&pk)[kItems], int shift) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  uint32_t* hist = &sh.hist[0][0];
  for (int i = tid; i < kSortWaves * 256; i += kSortThreads) hist[i] = 0;
  __syncthreads();
  const uint64_t below = lanesBelow(lane);
#pragma unroll
  for (int it = 0; it < kItems; it++) {
    uint32_t d = (key[it] >> shift) & 255u;
    uint64_t peers = ~0ull;
#pragma unroll
    for (int bit = 0; bit < 8; bit++) {
      uint32_t set = (d >> bit) & 1u;
      uint64_t bal = __ballot(set);
      peers &= set ? bal : ~bal;
    }
    uint32_t rank = (uint32_t)__popcll(peers & below);
    uint32_t base = sh.hist[wave][d];
    __builtin_amdgcn_wave_barrier();
    if (rank == 0) sh.hist[wave][d] = base + (uint32_t)__popcll(peers);
    __builtin_amdgcn_wave_barrier();
    pk[it] = (pk[it] & 0xFFFFu) | ((base + rank) << 16);
  }
  __syncthreads();
  if (tid < 256) {
    uint32_t run = 0;
    for (int w = 0; w < kSortWaves; w++) {
      uint32_t c = sh.hist[w][tid];
      sh.hist[w][tid] = run;
      run += c;
    }
    sh.total[tid] = run;
  }
  __syncthreads();
  if (tid < 64) {  // exclusive scan of the 256 digit totals by one wave
    uint32_t v0 = sh.total[tid * 4], v1 = sh.total[tid * 4 + 1], v2 = sh.total[tid * 4 + 2],
             v3_ = sh.total[tid * 4 + 3];
    uint32_t s4 = v0 + v1 + v2 + v3_, incl = s4;
    for (int off = 1; off < 64; off <<= 1) {
      uint32_t o = __shfl_up(incl, off, 64);
      if (lane >= off) incl += o;
    }
    uint32_t excl = incl - s4;
    sh.total[tid * 4] = excl;
    sh.total[tid * 4 + 1] = excl + v0;
    sh.total[tid * 4 + 2] = excl + v0 + v1;
    sh.total[tid * 4 + 3] = excl + v0 + v1 + v2;
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < kItems; it++) {
    uint32_t d = (key[it] >> shift) & 255u;
    uint32_t dest = sh.total[d] + sh.hist[wave][d] + (pk[it] >> 16);
    pk[it] = (pk[it] & 0xFFFFu) | (dest << 16);
  }
  const int home = sortHome();
#pragma unroll
  for (int it = 0; it < kItems; it++) sh.stage[pk[it] >> 16] = key[it];
  __syncthreads();
#pragma unroll
  for (int it = 0; it < kItems; it++) key[it] = sh.stage[home + it * 64];
  __syncthreads();
#pragma unroll
  for (int it = 0; it < kItems; it++) sh.stage[pk[it] >> 16] = pk[it] & 0xFFFFu;
  __syncthreads();
#pragma unroll
  for (int it = 0; it < kItems; it++) pk[it] = sh.stage[home + it * 64];
  __syncthreads();
}

// Four passes: ascending by all 32 key bits, stable in slot order.
__device__ __forceinline__ void sortByKey(SortShared& sh, uint32_t (&key)[kItems], uint32_t (&pk)[kItems]) {
#pragma unroll 1
  for (int shift = 0; shift < 32; shift += 8) radixPass(sh, key, pk, shift);
}

// In-place sort of (Keys, Order) into ascending (key, order) order, the
// order the bitonic network's comparison (a > b || (a == b && ia > ib))
// produces. When Order is already strictly increasing (as makeKeys writes
// it), a stable sort by key is that order; otherwise LSD passes on the order
// value run first.
__global__ __launch_bounds__(kSortThreads) void k_sortKeys(uint32_t* Keys, uint32_t* Order) {
  __shared__ SortShared sh;
  __shared__ int increasing, wide;
  const int home = sortHome();
  if (threadIdx.x == 0) { increasing = 1; wide = 0; }
  __syncthreads();
  uint32_t key[kItems], pk[kItems];
  bool ok = true, big = false;
#pragma unroll
  for (int it = 0; it < kItems; it++) {
    int e = home + it * 64;
    uint32_t o = Order[e];
    if (e > 0 && Order[e - 1] >= o) ok = false;
    big |= o > 0xFFFFu;
    key[it] = o;
    pk[it] = (uint32_t)e;
  }
  if (!ok) increasing = 0;
  if (big) wide = 1;
  __syncthreads();
  if (!increasing) {
    const int bits = wide ? 32 : 16;
#pragma unroll 1
    for (int shift = 0; shift < bits; shift += 8) radixPass(sh, key, pk, shift);
  }
#pragma unroll
  for (int it = 0; it < kItems; it++) key[it] = Keys[pk[it] & 0xFFFFu];
  sortByKey(sh, key, pk);
  uint32_t ord[kItems];
#pragma unroll
  for (int it = 0; it < kItems; it++) ord[it] = Order[pk[it] & 0xFFFFu];
  __syncthreads();
#pragma unroll
  for (int it = 0; it < kItems; it++) {
    Keys[home + it * 64] = key[it];
    Order[home + it * 64] = ord[it];
  }
}

// makeKeys + sort + makeBounds + reduceBounds(16384..1) in one workgroup.
// Order starts as the identity (makeKeys), so the slot is the order value.
__global__ __launch_bounds__(kSortThreads) void k_buildBVH(const float* P, uint32_t* Keys, uint32_t* Order,
                                                          float* Nodes) {
  __shared__ SortShared sh;
  const int tid = threadIdx.x, home = sortHome();
  uint32_t key[kItems], pk[kItems];
#pragma unroll
  for (int it = 0; it < kItems; it++) {
    int e = home + it * 64;
    key[it] = mortonKey(P, e);
    pk[it] = (uint32_t)e;
  }
  sortByKey(sh, key, pk);
#pragma unroll
  for (int it = 0; it < kItems; it++) {
    int e = home + it * 64;
    uint32_t id = pk[it] & 0xFFFFu;
    Keys[e] = key[it];
    Order[e] = id;
    writeLeaf(P, Nodes, e, id);
  }
  __syncthreads();
  for (int first = kCapacity / 2; first >= 1; first >>= 1) {
    for (int i = tid; i < first; i += kSortThreads) writeInterior(Nodes, first + i);
    __syncthreads();
  }
}

// ------------------------------------------------------------------ camera
// Arbor.cu:181.
__global__ void k_initCamera(float* C, int seed) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  initCameraState(C, seed);
}
// Arbor.cu:182.
__global__ void k_updateCamera(float* C, const float* I, float delta, int width, int height) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  updateCameraState(C, I, delta, width, height);
}

__global__ void k_fill(float4* p, size_t n4, float v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n4; i += stride) p[i] = make_float4(v, v, v, v);
}

}  // namespace arbor

using namespace arbor;

static const int kGrowCounts[5] = {9, 12, 72, 288, 1728};

extern "C" {

int arbor_growTree(float* B, int level, int seed, hipStream_t stream) {
  ARBOR_REQUIRE(B);
  if (level < 0 || level > 4 || seed < 0 || seed > 1000000) return ARBOR_ERR_ARG;
  int n = kGrowCounts[level];
  hipLaunchKernelGGL(k_growTree, dim3((n + 63) / 64), dim3(64), 0, stream, B, level, seed, n);
  return launch_status();
}
int arbor_makePrimitives(const float* B, float* P, int seed, hipStream_t stream) {
  ARBOR_REQUIRE(B && P);
  if (seed < 0 || seed > 1000000) return ARBOR_ERR_ARG;
  hipLaunchKernelGGL(k_makePrimitives, dim3(kCapacity / 64), dim3(64), 0, stream, B, P, seed);
  return launch_status();
}
int arbor_makeKeys(const float* P, unsigned* Keys, unsigned* Order, hipStream_t stream) {
  ARBOR_REQUIRE(P && Keys && Order);
  hipLaunchKernelGGL(k_makeKeys, dim3(kCapacity / 256), dim3(256), 0, stream, P, Keys, Order);
  return launch_status();
}
int arbor_sortPass(unsigned* Keys, unsigned* Order, int distance, int span, hipStream_t stream) {
  ARBOR_REQUIRE(Keys && Order);
  if (!is_pow2(distance) || !is_pow2(span) || distance >= span || span > kCapacity) return ARBOR_ERR_ARG;
  hipLaunchKernelGGL(k_sortPass, dim3(kCapacity / 256), dim3(256), 0, stream, Keys, Order, distance, span);
  return launch_status();
}
int arbor_makeBounds(const float* P, const unsigned* Order, float* Nodes, hipStream_t stream) {
  ARBOR_REQUIRE(P && Order && Nodes);
  hipLaunchKernelGGL(k_makeBounds, dim3(kCapacity / 256), dim3(256), 0, stream, P, Order, Nodes);
  return launch_status();
}
int arbor_reduceBounds(float* Nodes, int first, hipStream_t stream) {
  ARBOR_REQUIRE(Nodes);
  if (!is_pow2(first) || first > kCapacity / 2) return ARBOR_ERR_ARG;
  hipLaunchKernelGGL(k_reduceBounds, dim3((first + 255) / 256), dim3(256), 0, stream, Nodes, first);
  return launch_status();
}
int arbor_sortKeys(unsigned* Keys, unsigned* Order, hipStream_t stream) {
  ARBOR_REQUIRE(Keys && Order);
  hipLaunchKernelGGL(k_sortKeys, dim3(1), dim3(kSortThreads), 0, stream, Keys, Order);
  return launch_status();
}
int arbor_buildBVH(const float* P, unsigned* Keys, unsigned* Order, float* Nodes, hipStream_t stream) {
  ARBOR_REQUIRE(P && Keys && Order && Nodes);
  hipLaunchKernelGGL(k_buildBVH, dim3(1), dim3(kSortThreads), 0, stream, P, Keys, Order, Nodes);
  return launch_status();
}
int arbor_initCamera(float* C, const float* P, int seed, hipStream_t stream) {
  (void)P;
  ARBOR_REQUIRE(C);
  if (seed < 0 || seed > 1000000) return ARBOR_ERR_ARG;
  hipLaunchKernelGGL(k_initCamera, dim3(1), dim3(1), 0, stream, C, seed);
  return launch_status();
}
int arbor_updateCamera(float* C, const float* I, const float* P, float delta, int width, int height,
                       hipStream_t stream) {
  (void)P;
  ARBOR_REQUIRE(C && I);
  if (!dims_ok(width, height)) return ARBOR_ERR_ARG;
  hipLaunchKernelGGL(k_updateCamera, dim3(1), dim3(1), 0, stream, C, I, delta, width, height);
  return launch_status();
}
int arbor_generate(float* B, float* P, unsigned* Keys, unsigned* Order, float* Nodes, float* C, int seed,
                   float* History, int width, int height, hipStream_t stream) {
  ARBOR_REQUIRE(B && P && Keys && Order && Nodes && C);
  if (seed < 0 || seed > 1000000) return ARBOR_ERR_ARG;
  if (History && !dims_ok(width, height)) return ARBOR_ERR_ARG;
  int rc;
  for (int level = 0; level < 5; level++)
    if ((rc = arbor_growTree(B, level, seed, stream)) != ARBOR_OK) return rc;
  if ((rc = arbor_makePrimitives(B, P, seed, stream)) != ARBOR_OK) return rc;
  if ((rc = arbor_buildBVH(P, Keys, Order, Nodes, stream)) != ARBOR_OK) return rc;
  if ((rc = arbor_initCamera(C, P, seed, stream)) != ARBOR_OK) return rc;
  if (History) {
    size_t n4 = (size_t)width * height;
    hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, stream, reinterpret_cast<float4*>(History), n4, 0.0f);
    if ((rc = launch_status()) != ARBOR_OK) return rc;
  }
  return ARBOR_OK;
}

}  // extern "C"
