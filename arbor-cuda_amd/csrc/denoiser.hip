// The learned denoiser's network: an L-level 3x3-convolution encoder-decoder
// (U-Net, L = 1..5; by default 5 levels of 32/48/64/96/128 channels), BF16
// activations and weights, FP32 accumulation, master weights and Adam state,
// trained with an L1 loss on log(1 + radiance), and used for inference on
// whole frames.
//
// Input per pixel (16 channels, 10 used): log(1 + 1-spp radiance) rgb,
// albedo rgb (Guide), shading normal xyz, and depth log(1 + min(t, 1000))/7
// (1 for sky). Output: log(1 + radiance) rgb.
//
// Graph of the 5-level net (s = ReLU conv pairs kept as skips, p = 2x2
// average pool, u = nearest 2x upsample, c = channel concatenation [u, s]):
//   in -> L0 -> L1 = s0 -> p -> L2 -> L3 = s1 -> p -> L4 -> L5 = s2 -> p ->
//   L6 -> L7 = s3 -> p -> L8 -> L9 = e4 -> u,c(s3) -> L10 -> L11 -> u,c(s2) ->
//   L12 -> L13 -> u,c(s1) -> L14 -> L15 -> u,c(s0) -> L16 -> L17 -> L18 (linear)
// With L levels: encoder layers 0..2L-1 (two per level), decoder layers
// 2L..4L-3 (two per level above the bottom), output layer 4L-2.
//
// The lean network (levels = 0): the compact input features (8 channels:
// log(1 + radiance) rgb, albedo rgb, depth, zero), one 3x3 convolution of 16
// or 32 channels with ReLU and a linear 1x1 output layer; layer 0 and
// layer 1. Its one-pass kernel reads the 16-byte compact features the
// shading pass can write directly (arbor_shadePixelsFeatures).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <random>
#include <string>
#include <vector>

#include "arbor_conv1x1.hpp"
#include "arbor_denoise.h"
#include "arbor_features.hpp"
#include "arbor_filter.hpp"
#include "engine_passes.hpp"
#include "launch.hpp"

namespace {

using bf16 = __bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kInC = 16, kOutC = 16, kMaxLevels = 5;
constexpr int kLeanInC = 8;  // the compact features (arbor::compactFeatures)
constexpr int kDefaultWidths[kMaxLevels] = {32, 48, 64, 96, 128};

// ---- data kernels ----
using arbor::depthFeature;
using arbor::expm1Fast;
using arbor::log1pFast;
// The ten used input features of one pixel from its Linear, Guide, Normal
// and Hit values.
__device__ __forceinline__ void featuresOf(const float4 l, const float4 g, const float4 n, const float4 t, float* f) {
  f[0] = log1pFast(fmaxf(l.x, 0.0f));
  f[1] = log1pFast(fmaxf(l.y, 0.0f));
  f[2] = log1pFast(fmaxf(l.z, 0.0f));
  f[3] = g.x;
  f[4] = g.y;
  f[5] = g.z;
  f[6] = n.x;
  f[7] = n.y;
  f[8] = n.z;
  f[9] = depthFeature(t.x);
}
// The same of pixel i of the frame buffers.
__device__ __forceinline__ void pixelFeatures(const float4* __restrict__ Linear, const float4* __restrict__ Guide,
                                              const float4* __restrict__ Normal, const float4* __restrict__ Hit,
                                              int i, float* f) {
  featuresOf(Linear[i], Guide[i], Normal[i], Hit[i], f);
}
// Input slot (H x W) from a frame (w x h; zero outside): the ten features in
// 16 channels, or with kCompact the compact set in 8 (Normal not read).
template <bool kCompact>
__global__ void k_features(const float4* __restrict__ Linear, const float4* __restrict__ Guide,
                           const float4* __restrict__ Normal, const float4* __restrict__ Hit, int w, int h,
                           bf16* __restrict__ X, int H, int W) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= W) return;
  if constexpr (kCompact) {
    arbor::featx8 f;
#pragma unroll
    for (int c = 0; c < 8; c++) f[c] = (bf16)0.0f;
    if (x < w && y < h) {
      const int i = y * w + x;
      f = arbor::compactFeatures(Linear[i], Guide[i], Hit[i].x);
    }
    *reinterpret_cast<arbor::featx8*>(X + ((size_t)y * W + x) * kLeanInC) = f;
    return;
  }
  float f[kInC];
  for (int c = 0; c < kInC; c++) f[c] = 0.0f;
  if (x < w && y < h) pixelFeatures(Linear, Guide, Normal, Hit, y * w + x, f);
  bf16* o = X + ((size_t)y * W + x) * kInC;
  for (int c = 0; c < kInC; c++) o[c] = (bf16)f[c];
}
// Target slot: radiance rgb and a validity mask in w.
__global__ void k_target(const float4* __restrict__ R, int w, int h, float4* __restrict__ T, int H, int W) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= W) return;
  float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (x < w && y < h) {
    const float4 r = R[y * w + x];
    v = make_float4(log1pf(fmaxf(r.x, 0.0f)), log1pf(fmaxf(r.y, 0.0f)), log1pf(fmaxf(r.z, 0.0f)), 1.0f);
  }
  T[(size_t)y * W + x] = v;
}
// L1 in log space over the masked pixels: gradient sign(o - t) * scale into G
// (channels 0..2; the rest 0); per-block partial sums of |o - t| into part.
__global__ __launch_bounds__(256) void k_l1Loss(const bf16* __restrict__ Out, const float4* __restrict__ T, long long n,
                                                float scale, bf16* __restrict__ G, float* __restrict__ part) {
  __shared__ float red[256];
  float s = 0.0f;
  for (long long p = blockIdx.x * 256LL + threadIdx.x; p < n; p += (long long)gridDim.x * 256) {
    const float4 t = T[p];
    const float tv[3] = {t.x, t.y, t.z};
    for (int c = 0; c < kOutC; c++) {
      float g = 0.0f;
      if (c < 3 && t.w > 0.0f) {
        const float d = (float)Out[p * kOutC + c] - tv[c];
        s += fabsf(d);
        g = d > 0.0f ? scale : (d < 0.0f ? -scale : 0.0f);
      }
      G[p * kOutC + c] = (bf16)g;
    }
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ void k_sumTo(const float* __restrict__ part, int n, float scale, float* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    float s = 0.0f;
    for (int i = 0; i < n; i++) s += part[i];
    *out = s * scale;
  }
}
// ---- a 1-level network of 16 channels in one pass ----
// Input features, the three convolutions and the output of one frame for a
// 32x16-pixel tile per block, the intermediate tensors in LDS: each layer is
// computed on the tile plus the halo the later layers read (38x22 features,
// then 36x20 and 34x18 activations, 32 B per pixel). Every convolution issues
// the v_mfma_f32_16x16x32_bf16 sequence of arbor_dn_conv3x3 for 16 input
// channels (five instructions per 16 output channels x 16 pixels, taps paired
// in K: taps 0-1, 2-3, 4-5, 6-7, then tap 8 with zero weights), adds the
// bias, applies ReLU and rounds to bf16 as it does, and positions outside
// the slot are zero as its padding is: the output has the layered network's
// bits. With kResolve each pixel is also resolved (resolvePixels).
namespace fused {
constexpr int kTW = 32, kTH = 16, kC = 16;
constexpr int kFW = kTW + 6, kFH = kTH + 6;  // features
constexpr int kAW = kTW + 4, kAH = kTH + 4;  // first convolution
constexpr int kBW = kTW + 2, kBH = kTH + 2;  // second convolution
constexpr int kFPix = kFW * kFH, kAPix = kAW * kAH, kBPix = kBW * kBH;
constexpr int kThreads = 256;
constexpr int kXcds = 8;  // accelerator complex dies, each with its own L2
static_assert(kBPix <= kFPix, "the second convolution's region reuses the features' LDS");

__device__ __forceinline__ bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int j = 0; j < 8; j++) z[j] = (bf16)0.0f;
  return z;
}
// The tap of K-slot group g in paired instruction k (9: none).
__device__ __forceinline__ int pairTap(int k, int g) { return 2 * k + (g >> 1); }
// The five paired instructions' weight fragments of lane (row = output
// channel, g) for a 16 -> 16 convolution [9][16][16]: tap pairTap(k, g),
// channels 8 (g & 1) .. +7; zero for the slot without a tap.
__device__ __forceinline__ void loadTaps(const bf16* __restrict__ Wt, int row, int g, bf16x8* w) {
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const int tap = pairTap(k, g);
    w[k] = tap < 9 ? *reinterpret_cast<const bf16x8*>(Wt + (tap * kC + row) * kC + 8 * (g & 1)) : zero8();
  }
}
// The 3x3 convolutions of target pixels p0 and p1 (region wd wide) from the
// source region (wd + 2 wide) at src: lane (row, g) gets channels 4g..4g+3
// of the pixel with index p0 (p1) of its 16-pixel group (row = its place in
// the group). Every lane reads its paired tap's 8 channels (tap 8 again for
// the slot without a tap, whose weights are zero); the two groups' products
// alternate, so neither waits on its own previous one.
__device__ __forceinline__ void convPair(const bf16* src, int wd, int p0, int p1, int g, const bf16x8* w,
                                         f32x4& acc0, f32x4& acc1) {
  const int ws = wd + 2;
  const int ty0 = p0 / wd, tx0 = p0 - ty0 * wd, ty1 = p1 / wd, tx1 = p1 - ty1 * wd;
  const bf16* b0 = src + (ty0 * ws + tx0) * kC + 8 * (g & 1);
  const bf16* b1 = src + (ty1 * ws + tx1) * kC + 8 * (g & 1);
  bf16x8 a0[5], a1[5];
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const int tap = min(pairTap(k, g), 8), off = ((tap / 3) * ws + tap - 3 * (tap / 3)) * kC;
    a0[k] = *reinterpret_cast<const bf16x8*>(b0 + off);
    a1[k] = *reinterpret_cast<const bf16x8*>(b1 + off);
  }
  acc0 = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  acc1 = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int k = 0; k < 5; k++) {
    acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[k], a0[k], acc0, 0, 0, 0);
    acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[k], a1[k], acc1, 0, 0, 0);
  }
}
// One ReLU convolution of a whole region: dst (wd x hd, origin (yo, xo) in
// slot coordinates) from src; two 16-pixel groups per wave at a time.
__device__ __forceinline__ void convRegion(const bf16* src, bf16* dst, int wd, int hd, int yo, int xo,
                                           const bf16* __restrict__ Wt, const float* __restrict__ bias, int H,
                                           int W) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, row = lane & 15, g = lane >> 4;
  bf16x8 w[5];
  loadTaps(Wt, row, g, w);
  float bv[4];
#pragma unroll
  for (int i = 0; i < 4; i++) bv[i] = bias[4 * g + i];
  const int n = wd * hd;
  // a region wholly inside the slot needs no per-pixel padding test
  const bool interior = yo >= 0 && xo >= 0 && yo + hd <= H && xo + wd <= W;
  for (int base = 32 * wave; base < n; base += 32 * (kThreads / 64)) {
    const int p0 = min(base + row, n - 1), p1 = min(base + 16 + row, n - 1);
    f32x4 a0, a1;
    convPair(src, wd, p0, p1, g, w, a0, a1);
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const int q = base + 16 * k + row;
      if (q >= n) continue;
      const f32x4 acc = k ? a1 : a0;
      f32x4 v;
#pragma unroll
      for (int i = 0; i < 4; i++) v[i] = fmaxf(acc[i] + bv[i], 0.0f);
      bf16x4 o = __builtin_convertvector(v, bf16x4);
      if (!interior) {
        const int ty = q / wd, tx = q - ty * wd, y = yo + ty, x = xo + tx;
        if (!(y >= 0 && y < H && x >= 0 && x < W)) o = __builtin_convertvector(f32x4{0.0f, 0.0f, 0.0f, 0.0f}, bf16x4);
      }
      *reinterpret_cast<bf16x4*>(dst + q * kC + 4 * g) = o;
    }
  }
}
}  // namespace fused

template <bool kResolve>
__global__ __launch_bounds__(fused::kThreads) void k_denoise1(
    const float4* __restrict__ Linear, const float4* __restrict__ Guide, const float4* __restrict__ Normal,
    const float4* __restrict__ Hit, int w, int h, int H, int W, const bf16* __restrict__ W0,
    const float* __restrict__ B0, const bf16* __restrict__ W1, const float* __restrict__ B1,
    const bf16* __restrict__ W2, const float* __restrict__ B2, float4* __restrict__ Out, const float* C,
    float4* History, uint32_t* Pixels) {
  using namespace fused;
  __shared__ __attribute__((aligned(16))) bf16 featB[kFPix * kC];  // features, then the second convolution
  __shared__ __attribute__((aligned(16))) bf16 actA[kAPix * kC];   // the first convolution
  // Blocks are dealt to the 8 XCDs in turn: block b takes tile t so that
  // each XCD walks a contiguous run of tiles (row-major), and the halo a
  // tile shares with its neighbours is read through that XCD's L2.
  const int tilesX = (w + kTW - 1) / kTW, tiles = tilesX * ((h + kTH - 1) / kTH);
  const int per = (tiles + kXcds - 1) / kXcds, t = (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds);
  if (t >= tiles) return;
  const int x0 = (t % tilesX) * kTW, y0 = (t / tilesX) * kTH;
  // Features: every load of the thread's pixels is issued before the first
  // is used (pixel 0 stands in outside the image).
  constexpr int kIter = (kFPix + kThreads - 1) / kThreads;
  float4 lq[kIter], gq[kIter], nq[kIter], tq[kIter];
  bool in[kIter];
#pragma unroll
  for (int s = 0; s < kIter; s++) {
    const int p = threadIdx.x + s * kThreads, fy = p / kFW, y = y0 - 3 + fy, x = x0 - 3 + (p - fy * kFW);
    in[s] = p < kFPix && y >= 0 && y < h && x >= 0 && x < w;
    const int i = in[s] ? y * w + x : 0;
    lq[s] = Linear[i];
    gq[s] = Guide[i];
    nq[s] = Normal[i];
    tq[s] = Hit[i];
  }
#pragma unroll
  for (int s = 0; s < kIter; s++) {
    const int p = threadIdx.x + s * kThreads;
    if (p >= kFPix) continue;
    float f[10];
#pragma unroll
    for (int c = 0; c < 10; c++) f[c] = 0.0f;
    if (in[s]) featuresOf(lq[s], gq[s], nq[s], tq[s], f);
    bf16x8 lo, hi = zero8();
#pragma unroll
    for (int c = 0; c < 8; c++) lo[c] = (bf16)f[c];
    hi[0] = (bf16)f[8];
    hi[1] = (bf16)f[9];
    *reinterpret_cast<bf16x8*>(featB + p * kC) = lo;
    *reinterpret_cast<bf16x8*>(featB + p * kC + 8) = hi;
  }
  __syncthreads();
  convRegion(featB, actA, kAW, kAH, y0 - 2, x0 - 2, W0, B0, H, W);
  __syncthreads();
  convRegion(actA, featB, kBW, kBH, y0 - 1, x0 - 1, W1, B1, H, W);
  __syncthreads();
  // The linear output convolution. Lanes of group 0 hold channels 0-3 of
  // their pixel; they leave the three used ones, rounded to bf16 as the
  // layered output is, in LDS (the first convolution's, no longer read), and
  // every thread then finishes two pixels: expm1, Filtered, the resolve.
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, row = lane & 15, g = lane >> 4;
  bf16x8 wt[5];
  loadTaps(W2, row, g, wt);
  float bv[4];
#pragma unroll
  for (int i = 0; i < 4; i++) bv[i] = B2[4 * g + i];
  bf16* rgb = actA;  // [kTW * kTH][4]
  for (int base = 32 * wave; base < kTW * kTH; base += 32 * (kThreads / 64)) {
    f32x4 a0, a1;
    convPair(featB, kTW, base + row, base + 16 + row, g, wt, a0, a1);
    if (g == 0) {
#pragma unroll
      for (int k = 0; k < 2; k++) {
        const f32x4 acc = k ? a1 : a0;
        f32x4 v;
#pragma unroll
        for (int i = 0; i < 4; i++) v[i] = acc[i] + bv[i];
        *reinterpret_cast<bf16x4*>(rgb + (base + 16 * k + row) * 4) = __builtin_convertvector(v, bf16x4);
      }
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < kTW * kTH; q += kThreads) {
    const int ty = q / kTW, y = y0 + ty, x = x0 + (q - ty * kTW);
    if (y >= h || x >= w) continue;
    const bf16x4 o = *reinterpret_cast<const bf16x4*>(rgb + q * 4);
    const float r = fmaxf(expm1Fast((float)o[0]), 0.0f);
    const float gg = fmaxf(expm1Fast((float)o[1]), 0.0f);
    const float b = fmaxf(expm1Fast((float)o[2]), 0.0f);
    const int i = y * w + x;
    Out[i] = make_float4(r, gg, b, 1.0f);
    if constexpr (kResolve) arbor::resolveOne(arbor::mk(r, gg, b), History, Pixels, C, x, y, w, h, i);
  }
}

// ---- the lean network in one pass ----
// A block computes a 32x16-pixel tile: it stages the compact features of the
// tile plus a 1-pixel halo (34x18 pixels, 16 bytes each) in LDS, and each
// wave takes eight 16-pixel groups. Per group, the 3x3 convolution is three
// v_mfma_f32_16x16x32_bf16 per 16 output channels (four taps per
// instruction: K slots 8g..8g+7 hold tap 4k+g's 8 channels; the third holds
// tap 8 and zero weights), as arbor_dn_conv3x3 issues them for 8 input
// channels, the bias the accumulator's initial value; ReLU and the rounding
// to bf16 leave the activations where the 1x1 output layer takes its B
// operand (arbor_conv1x1.hpp), so that layer is one more instruction on
// registers. The output, rounded to bf16
// as the layered network's is, goes through LDS so that every thread then
// finishes two pixels (expm1, Out, and with kResolve resolvePixels). Feat
// holds the features of a w x h frame with row stride `stride` (zero outside
// the frame, as the layered network's padding): the output has the layered
// runner's bits.
namespace lean {
constexpr int kTW = 32, kTH = 16, kRW = kTW + 2, kRH = kTH + 2, kRPix = kRW * kRH;
constexpr int kThreads = 256, kGroupsPerWave = kTW * kTH / 16 / (kThreads / 64);
}  // namespace lean

template <int J, bool kResolve>
__global__ __launch_bounds__(lean::kThreads) void k_denoiseLean(
    const arbor::featx8* __restrict__ Feat, int stride, int w, int h, const bf16* __restrict__ W0,
    const float* __restrict__ B0, const bf16* __restrict__ W1, const float* __restrict__ B1, float4* __restrict__ Out,
    const float* C, float4* History, uint32_t* Pixels) {
  using namespace lean;
  using Op = arbor::dn1::Op1x1<J>;
  __shared__ arbor::featx8 feat[kRPix];
  __shared__ bf16x4 rgb[kTW * kTH];
  // XCD-contiguous tiles, as k_denoise1 takes them
  const int tilesX = (w + kTW - 1) / kTW, tiles = tilesX * ((h + kTH - 1) / kTH);
  const int per = (tiles + fused::kXcds - 1) / fused::kXcds;
  const int t = (int)(blockIdx.x % fused::kXcds) * per + (int)(blockIdx.x / fused::kXcds);
  if (t >= tiles) return;
  const int x0 = (t % tilesX) * kTW, y0 = (t / tilesX) * kTH;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, row = lane & 15, g = lane >> 4;
  // stage: every load issued before the first store (pixel 0 outside the frame)
  constexpr int kIter = (kRPix + kThreads - 1) / kThreads;
  arbor::featx8 v[kIter];
  bool in[kIter];
#pragma unroll
  for (int s = 0; s < kIter; s++) {
    const int p = threadIdx.x + s * kThreads, fy = p / kRW, y = y0 - 1 + fy, x = x0 - 1 + (p - fy * kRW);
    in[s] = p < kRPix && y >= 0 && y < h && x >= 0 && x < w;
    v[s] = Feat[in[s] ? y * stride + x : 0];  // frames stay below 2^31 pixels
  }
  // weight fragments: instruction k, output group j (lane row = channel
  // 16j + row, K slots 8g.. = tap 4k+g), the 1x1 layer's, and the biases
  bf16x8 w0[3][J];
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int j = 0; j < J; j++) {
      const int tap = 4 * k + g;
      w0[k][j] = tap < 9 ? *reinterpret_cast<const bf16x8*>(W0 + ((size_t)tap * 16 * J + 16 * j + row) * kLeanInC)
                         : fused::zero8();
    }
  const typename Op::T w1 = Op::load(W1 + (size_t)row * 16 * J, g);
  float b0[J][4], b1[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
#pragma unroll
    for (int j = 0; j < J; j++) b0[j][i] = B0[16 * j + 4 * g + i];
    b1[i] = B1[4 * g + i];
  }
#pragma unroll
  for (int s = 0; s < kIter; s++) {
    const int p = threadIdx.x + s * kThreads;
    if (p < kRPix) {
      if (!in[s]) {
#pragma unroll
        for (int c = 0; c < 8; c++) v[s][c] = (bf16)0.0f;
      }
      feat[p] = v[s];
    }
  }
  __syncthreads();
  // Group s of the wave covers tile row 4 * wave + s / 2, columns
  // 16 * (s % 2) + row: each lane's three tap positions are fixed offsets
  // from its first group's, so the loop's LDS addresses differ by constants.
  const int wv = __builtin_amdgcn_readfirstlane(wave);
  int tap0[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int tl = min(4 * k + g, 8), dy = tl / 3, dx = tl - 3 * (tl / 3);
    tap0[k] = (4 * wv + dy) * kRW + row + dx;
  }
  const int out0 = 4 * wv * kTW + row;
#pragma unroll
  for (int s = 0; s < kGroupsPerWave; s++) {
    const int step = (s >> 1) * kRW + (s & 1) * 16;
    f32x4 acc[J];  // the biases are the accumulators' initial values
#pragma unroll
    for (int j = 0; j < J; j++) acc[j] = f32x4{b0[j][0], b0[j][1], b0[j][2], b0[j][3]};
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const bf16x8 a = feat[tap0[k] + step];
#pragma unroll
      for (int j = 0; j < J; j++) acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w0[k][j], a, acc[j], 0, 0, 0);
    }
    bf16x4 hid[J];
#pragma unroll
    for (int j = 0; j < J; j++) hid[j] = arbor::dn1::reluToBf16(acc[j]);
    const f32x4 o = Op::mma(w1, Op::fromGroups(hid), f32x4{b1[0], b1[1], b1[2], b1[3]});
    if (g == 0) rgb[out0 + (s >> 1) * kTW + (s & 1) * 16] = arbor::dn1::toBf16(o);
  }
  __syncthreads();
#pragma unroll
  for (int q = threadIdx.x; q < kTW * kTH; q += kThreads) {
    const int ty = q / kTW, y = y0 + ty, x = x0 + (q - ty * kTW);
    if (y >= h || x >= w) continue;
    const bf16x4 o = rgb[q];
    const float r = fmaxf(expm1Fast((float)o[0]), 0.0f);
    const float gg = fmaxf(expm1Fast((float)o[1]), 0.0f);
    const float b = fmaxf(expm1Fast((float)o[2]), 0.0f);
    const int i = y * w + x;
    Out[i] = make_float4(r, gg, b, 1.0f);
    if constexpr (kResolve) arbor::resolveOne(arbor::mk(r, gg, b), History, Pixels, C, x, y, w, h, i);
  }
}

// Output slot -> radiance (expm1 of the prediction, clamped at 0), alpha 1.
__global__ void k_output(const bf16* __restrict__ Y, int H, int W, float4* __restrict__ Out, int w, int h) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= w || y >= h) return;
  const bf16* o = Y + ((size_t)y * W + x) * kOutC;
  Out[y * w + x] = make_float4(fmaxf(expm1Fast((float)o[0]), 0.0f), fmaxf(expm1Fast((float)o[1]), 0.0f),
                               fmaxf(expm1Fast((float)o[2]), 0.0f), 1.0f);
}
// Adam on fp32 master values; also writes the bf16 copy when w16 != NULL.
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       long long n, float lr, float b1, float b2, float eps, bf16* __restrict__ w16) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi, vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * mi / (sqrtf(vi) + eps);
    w[i] = wi;
    if (w16) w16[i] = (bf16)wi;
  }
}
__global__ void k_toBf16(const float* __restrict__ a, long long n, bf16* __restrict__ b) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    b[i] = (bf16)a[i];
}

// Zeroes every tap but the centre one (4) of a [9][cout][cin] gradient
// (per = cin * cout elements per tap).
__global__ void k_keepCentre(float* __restrict__ g, long long per) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < 9 * per; i += (long long)gridDim.x * blockDim.x)
    if (i / per != 4) g[i] = 0.0f;
}

int grid1(long long n) {
  long long b = (n + 255) / 256;
  return (int)(b < 1 ? 1 : (b > 65535 ? 65535 : b));
}

struct Conv {
  int cin, cout, level;
  bool relu;
  bool k1 = false;  // 1x1: only the centre tap (4) of the stored [9][cout][cin] weights is used, the rest stay 0
  float *w = nullptr, *b = nullptr, *gw = nullptr, *gb = nullptr, *mw = nullptr, *vw = nullptr, *mb = nullptr,
        *vb = nullptr;
  uint16_t *w16 = nullptr, *wflip = nullptr;
  long long nw() const { return 9LL * cin * cout; }
};
struct Act {
  int c, level;
  uint16_t* p = nullptr;  // activation
  uint16_t* g = nullptr;  // gradient (training)
};

}  // namespace

struct arbor_denoiser {
  int maxBatch = 0, H = 0, W = 0;
  bool train = false;
  int step = 0;
  int levels = kMaxLevels;
  int widths[kMaxLevels] = {};
  std::vector<Conv> conv;
  // activations (level k): a = first conv of level k, s = skip (second conv
  // of level k < L-1), e4 = bottleneck (second conv of level L-1), p = pooled
  // skip, u / c = upsampled / concatenated decoder input, dA / d = the
  // decoder's two convs
  Act in, a[kMaxLevels], s[kMaxLevels - 1], p[kMaxLevels - 1], e4, u[kMaxLevels - 1], c[kMaxLevels - 1],
      dA[kMaxLevels - 1], d[kMaxLevels - 1], out;
  uint16_t* gpool = nullptr;  // pool-gradient temporary (largest: level 0 x 32)
  float* target = nullptr;    // [maxBatch][H][W] float4
  float* partial = nullptr;
  long long partialFloats = 0;
  float* lossPart = nullptr;
  float* loss = nullptr;
  std::vector<void*> owned;

  // levels == 0: the lean network (widths[0] channels)
  bool lean() const { return levels == 0; }
  int inC() const { return lean() ? kLeanInC : kInC; }
  int hl(int level) const { return H >> level; }
  int wl(int level) const { return W >> level; }
  long long count(const Act& t, int n) const { return (long long)n * hl(t.level) * wl(t.level) * t.c; }
  template <class T>
  int alloc(T** ptr, size_t bytes) {
    void* q = nullptr;
    if (hipMalloc(&q, bytes) != hipSuccess) return ARBOR_ERR_DEVICE;
    if (hipMemset(q, 0, bytes) != hipSuccess) return ARBOR_ERR_DEVICE;
    owned.push_back(q);
    *ptr = static_cast<T*>(q);
    return ARBOR_OK;
  }
  int allocAct(Act& t, int c, int level) {
    t.c = c;
    t.level = level;
    size_t bytes = (size_t)maxBatch * hl(level) * wl(level) * c * 2;
    int rc = alloc(&t.p, bytes);
    if (rc == ARBOR_OK && train) rc = alloc(&t.g, bytes);
    return rc;
  }
};

namespace {

// Layer table: (cin, cout, level, relu) in graph order: the encoder pairs,
// the decoder pairs (deepest first), the linear output layer.
void buildLayers(arbor_denoiser* d) {
  const int L = d->levels, *c = d->widths;
  auto add = [&](int cin, int cout, int level, bool relu) {
    Conv cv;
    cv.cin = cin;
    cv.cout = cout;
    cv.level = level;
    cv.relu = relu;
    d->conv.push_back(cv);
  };
  if (d->lean()) {
    add(kLeanInC, c[0], 0, true);
    add(c[0], kOutC, 0, false);
    d->conv.back().k1 = true;
    return;
  }
  for (int k = 0; k < L; k++) {
    add(k == 0 ? kInC : c[k - 1], c[k], k, true);
    add(c[k], c[k], k, true);
  }
  for (int k = L - 2; k >= 0; k--) {
    add(c[k + 1] + c[k], c[k], k, true);
    add(c[k], c[k], k, true);
  }
  add(c[0], kOutC, 0, false);
}

int initWeights(arbor_denoiser* d, int seed) {
  std::mt19937 rng((uint32_t)seed * 2654435761u + 12345u);
  for (size_t li = 0; li < d->conv.size(); li++) {
    Conv& cv = d->conv[li];
    const bool last = li + 1 == d->conv.size();
    const int used = d->lean() ? 7 : 10;  // input channels in use
    const int fanIn = (cv.k1 ? 1 : 9) * (li == 0 ? used : cv.cin);
    std::normal_distribution<float> nd(0.0f, last ? sqrtf(1.0f / fanIn) : sqrtf(2.0f / fanIn));
    std::vector<float> w((size_t)cv.nw());
    for (int tap = 0; tap < 9; tap++)
      for (int co = 0; co < cv.cout; co++)
        for (int ci = 0; ci < cv.cin; ci++) {
          float v = nd(rng);
          // unused input / output channels, and the taps a 1x1 layer has not
          if ((li == 0 && ci >= used) || (last && co >= 3) || (cv.k1 && tap != 4)) v = 0.0f;
          w[((size_t)tap * cv.cout + co) * cv.cin + ci] = v;
        }
    if (hipMemcpy(cv.w, w.data(), w.size() * 4, hipMemcpyHostToDevice) != hipSuccess) return ARBOR_ERR_DEVICE;
  }
  return ARBOR_OK;
}

int refreshBf16(arbor_denoiser* d, hipStream_t s) {
  for (Conv& cv : d->conv) {
    hipLaunchKernelGGL(k_toBf16, dim3(grid1(cv.nw())), dim3(256), 0, s, cv.w, cv.nw(), reinterpret_cast<bf16*>(cv.w16));
    int rc = arbor_dn_flip_weights(cv.w16, cv.cin, cv.cout, cv.wflip, s);
    if (rc != ARBOR_OK) return rc;
  }
  return arbor::launch_status();
}

#define DN_TRY(x)                      \
  do {                                 \
    int rc_ = (x);                     \
    if (rc_ != ARBOR_OK) return rc_;   \
  } while (0)

int runConv(arbor_denoiser* d, int li, const Act& x, Act& y, int n, hipStream_t s) {
  const Conv& cv = d->conv[li];
  if (cv.k1)
    return arbor_dn_conv1x1(x.p, (long long)n * d->hl(cv.level) * d->wl(cv.level), cv.cin,
                            cv.w16 + 4LL * cv.cout * cv.cin, cv.b, cv.relu ? 1 : 0, y.p, cv.cout, s);
  return arbor_dn_conv3x3(x.p, n, d->hl(cv.level), d->wl(cv.level), cv.cin, cv.w16, cv.b, cv.relu ? 1 : 0, y.p,
                          cv.cout, s);
}

// The decoder's input at level k (the coarser result it upsamples).
const Act& decoderLow(const arbor_denoiser* d, int k) { return k + 1 < d->levels - 1 ? d->d[k + 1] : d->e4; }
// The layer index of decoder level k's first convolution.
int decoderLayer(const arbor_denoiser* d, int k) { return 2 * d->levels + 2 * (d->levels - 2 - k); }

int forward(arbor_denoiser* d, int n, hipStream_t s) {
  if (d->lean()) {
    DN_TRY(runConv(d, 0, d->in, d->a[0], n, s));
    return runConv(d, 1, d->a[0], d->out, n, s);
  }
  const int L = d->levels;
  // encoder
  const Act* x = &d->in;
  for (int k = 0; k < L; k++) {
    Act& first = d->a[k];
    Act& second = k < L - 1 ? d->s[k] : d->e4;
    DN_TRY(runConv(d, 2 * k, *x, first, n, s));
    DN_TRY(runConv(d, 2 * k + 1, first, second, n, s));
    if (k < L - 1) {
      DN_TRY(arbor_dn_pool2(second.p, n, d->hl(k), d->wl(k), second.c, d->p[k].p, s));
      x = &d->p[k];
    }
  }
  // decoder: level k = L-2..0
  for (int k = L - 2; k >= 0; k--) {
    const Act& low = decoderLow(d, k);
    const int li = decoderLayer(d, k);
    if (d->train) {  // the weight gradient needs the concatenated input in memory
      DN_TRY(arbor_dn_up2(low.p, n, d->hl(k + 1), d->wl(k + 1), low.c, d->u[k].p, s));
      DN_TRY(arbor_dn_concat(d->u[k].p, d->u[k].c, d->s[k].p, d->s[k].c, (long long)n * d->hl(k) * d->wl(k),
                             d->c[k].p, s));
      DN_TRY(runConv(d, li, d->c[k], d->dA[k], n, s));
    } else {  // upsampling and concatenation inside the convolution's staging
      const Conv& cv = d->conv[li];
      DN_TRY(arbor_dn_conv3x3_cat(low.p, low.c, d->s[k].p, d->s[k].c, n, d->hl(k), d->wl(k), cv.w16, cv.b, 1,
                                  d->dA[k].p, cv.cout, s));
    }
    DN_TRY(runConv(d, li + 1, d->dA[k], d->d[k], n, s));
  }
  return runConv(d, 4 * L - 2, L > 1 ? d->d[0] : d->e4, d->out, n, s);
}

// dX of conv li from dY (into gx; gx may be NULL for the input layer), and
// the layer's weight and bias gradients from its input x.
int backConv(arbor_denoiser* d, int li, const Act& x, const Act& y, uint16_t* gx, int n, hipStream_t s) {
  const Conv& cv = d->conv[li];
  const int h = d->hl(cv.level), w = d->wl(cv.level);
  if (cv.relu) DN_TRY(arbor_dn_relu_bwd(y.g, y.p, (long long)n * h * w * cv.cout, s));
  DN_TRY(arbor_dn_conv3x3_wgrad(x.p, y.g, n, h, w, cv.cin, cv.cout, d->partial, d->partialFloats, cv.gw, cv.gb, s));
  if (cv.k1) {  // a 1x1 layer's gradient is the centre tap's; its other taps stay 0
    hipLaunchKernelGGL(k_keepCentre, dim3(grid1(cv.nw())), dim3(256), 0, s, cv.gw, (long long)cv.cin * cv.cout);
    DN_TRY(arbor::launch_status());
  }
  // (for a 1x1 layer the flipped weights are zero outside the centre tap, so
  // this 3x3 convolution is the 1x1 layer's transpose)
  if (gx) DN_TRY(arbor_dn_conv3x3(y.g, n, h, w, cv.cout, cv.wflip, nullptr, 0, gx, cv.cin, s));
  return ARBOR_OK;
}

int backward(arbor_denoiser* d, int n, hipStream_t s) {
  if (d->lean()) {
    DN_TRY(backConv(d, 1, d->a[0], d->out, d->a[0].g, n, s));
    return backConv(d, 0, d->in, d->a[0], nullptr, n, s);
  }
  const int L = d->levels;
  // output layer and decoder
  Act& top = L > 1 ? d->d[0] : d->e4;
  DN_TRY(backConv(d, 4 * L - 2, top, d->out, top.g, n, s));
  for (int k = 0; k <= L - 2; k++) {
    const int li = decoderLayer(d, k);
    DN_TRY(backConv(d, li + 1, d->dA[k], d->d[k], d->dA[k].g, n, s));
    DN_TRY(backConv(d, li, d->c[k], d->dA[k], d->c[k].g, n, s));
    const long long px = (long long)n * d->hl(k) * d->wl(k);
    DN_TRY(arbor_dn_split(d->c[k].g, d->u[k].c, d->s[k].c, px, d->u[k].g, d->s[k].g, s));
    const Act& low = decoderLow(d, k);
    DN_TRY(arbor_dn_up2_bwd(d->u[k].g, n, d->hl(k + 1), d->wl(k + 1), low.c, low.g, s));
  }
  // bottleneck and encoder, deepest first
  for (int k = L - 1; k >= 0; k--) {
    Act& second = k < L - 1 ? d->s[k] : d->e4;
    const Act& input = k == 0 ? d->in : d->p[k - 1];
    DN_TRY(backConv(d, 2 * k + 1, d->a[k], second, d->a[k].g, n, s));
    DN_TRY(backConv(d, 2 * k, input, d->a[k], k == 0 ? nullptr : d->p[k - 1].g, n, s));
    if (k > 0) {  // the skip below receives the pooled path's gradient too
      Act& skip = d->s[k - 1];
      DN_TRY(arbor_dn_pool2_bwd(d->p[k - 1].g, n, d->hl(k - 1), d->wl(k - 1), skip.c, d->gpool, s));
      DN_TRY(arbor_dn_add(skip.g, d->gpool, (long long)n * d->hl(k - 1) * d->wl(k - 1) * skip.c, s));
    }
  }
  return ARBOR_OK;
}

// Checkpoint kinds: "ARBORDN1" files hold a U-Net; "ARBORDN2" files name
// their kind after the magic.
constexpr int32_t kKindUNet = 0, kKindLean = 1;
// Reads the magic, the kind and the layer count.
bool readHeader(FILE* f, int32_t* kind, int32_t* layers) {
  char magic[8];
  if (fread(magic, 1, 8, f) != 8) return false;
  if (memcmp(magic, "ARBORDN1", 8) == 0)
    *kind = kKindUNet;
  else if (memcmp(magic, "ARBORDN2", 8) != 0 || fread(kind, 4, 1, f) != 1 || *kind != kKindLean)
    return false;
  return fread(layers, 4, 1, f) == 1;
}

bool slotOk(const arbor_denoiser* d, int index, int w, int h) {
  return index >= 0 && index < d->maxBatch && w >= 1 && h >= 1 && w <= d->W && h <= d->H;
}

}  // namespace

extern "C" {

int arbor_denoiser_create_net(arbor_denoiser** out, int max_batch, int height, int width, int train, int seed,
                              int levels, const int* widths) {
  if (!out || !widths) return ARBOR_ERR_NULL;
  *out = nullptr;
  if (max_batch < 1 || max_batch > 1024 || height < 16 || width < 16 || height % 16 || width % 16 ||
      height > ARBOR_MAX_DIM || width > ARBOR_MAX_DIM || levels < 0 || levels > kMaxLevels)
    return ARBOR_ERR_ARG;
  if (levels == 0 && widths[0] != 16 && widths[0] != 32) return ARBOR_ERR_ARG;
  for (int k = 0; k < levels; k++) {
    const int below = k + 1 < levels ? widths[k + 1] : 0;  // the decoder concatenates the level below
    if (widths[k] < 16 || widths[k] % 16 || widths[k] + below > ARBOR_DN_MAX_CHANNELS) return ARBOR_ERR_ARG;
  }
  arbor_denoiser* d = new arbor_denoiser();
  d->maxBatch = max_batch;
  d->H = height;
  d->W = width;
  d->train = train != 0;
  d->levels = levels;
  for (int k = 0; k < (levels ? levels : 1); k++) d->widths[k] = widths[k];
  buildLayers(d);
  int rc = ARBOR_OK;
  for (Conv& cv : d->conv) {
    rc |= d->alloc(&cv.w, cv.nw() * 4);
    rc |= d->alloc(&cv.b, cv.cout * 4);
    rc |= d->alloc(&cv.w16, cv.nw() * 2);
    rc |= d->alloc(&cv.wflip, cv.nw() * 2);
    if (d->train) {
      rc |= d->alloc(&cv.gw, cv.nw() * 4);
      rc |= d->alloc(&cv.gb, cv.cout * 4);
      rc |= d->alloc(&cv.mw, cv.nw() * 4);
      rc |= d->alloc(&cv.vw, cv.nw() * 4);
      rc |= d->alloc(&cv.mb, cv.cout * 4);
      rc |= d->alloc(&cv.vb, cv.cout * 4);
    }
  }
  const int* c = d->widths;
  const int L = levels;
  rc |= d->allocAct(d->in, d->inC(), 0);
  if (d->lean()) rc |= d->allocAct(d->a[0], c[0], 0);
  for (int k = 0; k < L; k++) rc |= d->allocAct(d->a[k], c[k], k);
  for (int k = 0; k < L - 1; k++) {
    rc |= d->allocAct(d->s[k], c[k], k);
    rc |= d->allocAct(d->p[k], c[k], k + 1);
    if (d->train) {
      rc |= d->allocAct(d->u[k], c[k + 1], k);
      rc |= d->allocAct(d->c[k], c[k + 1] + c[k], k);
    } else {  // the fused convolution reads the coarse and skip tensors directly
      d->u[k].c = c[k + 1];
      d->u[k].level = k;
      d->c[k].c = c[k + 1] + c[k];
      d->c[k].level = k;
    }
    rc |= d->allocAct(d->dA[k], c[k], k);
    rc |= d->allocAct(d->d[k], c[k], k);
  }
  if (!d->lean()) rc |= d->allocAct(d->e4, c[L - 1], L - 1);
  rc |= d->allocAct(d->out, kOutC, 0);
  if (d->train) {
    size_t poolElems = 0;  // the largest skip a pooled gradient returns to
    for (int k = 0; k < L - 1; k++) {
      size_t e = (size_t)max_batch * d->hl(k) * d->wl(k) * c[k];
      poolElems = e > poolElems ? e : poolElems;
    }
    if (poolElems) rc |= d->alloc(&d->gpool, poolElems * 2);
    rc |= d->alloc(&d->target, (size_t)max_batch * height * width * 16);
    long long most = 0;
    for (Conv& cv : d->conv) {
      long long need = arbor_dn_wgrad_partial_floats(max_batch, d->hl(cv.level), d->wl(cv.level), cv.cin, cv.cout);
      most = need > most ? need : most;
    }
    d->partialFloats = most;
    rc |= d->alloc(&d->partial, (size_t)most * 4);
    rc |= d->alloc(&d->lossPart, 1024 * 4);
    rc |= d->alloc(&d->loss, 4);
  }
  if (rc == ARBOR_OK) rc = initWeights(d, seed);
  if (rc == ARBOR_OK) rc = refreshBf16(d, nullptr);
  if (rc == ARBOR_OK && hipDeviceSynchronize() != hipSuccess) rc = ARBOR_ERR_DEVICE;
  if (rc != ARBOR_OK) {
    arbor_denoiser_destroy(d);
    return ARBOR_ERR_DEVICE;
  }
  *out = d;
  return ARBOR_OK;
}

int arbor_denoiser_create(arbor_denoiser** out, int max_batch, int height, int width, int train, int seed) {
  return arbor_denoiser_create_net(out, max_batch, height, width, train, seed, kMaxLevels, kDefaultWidths);
}

int arbor_denoiser_shape(const arbor_denoiser* d, int* levels, int* widths) {
  if (!d || !levels || !widths) return ARBOR_ERR_NULL;
  *levels = d->levels;
  for (int k = 0; k < kMaxLevels; k++) widths[k] = k < (d->levels ? d->levels : 1) ? d->widths[k] : 0;
  return ARBOR_OK;
}

// The architecture a checkpoint file was saved from: its layer count gives the
// levels (4L - 1 layers), the encoder layers' output widths the channels.
int arbor_denoiser_checkpoint_shape(const char* path, int* levels, int* widths) {
  if (!path || !levels || !widths) return ARBOR_ERR_NULL;
  FILE* f = fopen(path, "rb");
  if (!f) return ARBOR_ERR_ARG;
  int32_t kind = 0, layers = 0;
  bool ok = readHeader(f, &kind, &layers);
  int w[kMaxLevels] = {};
  if (ok && kind == kKindLean) {  // levels 0: the lean network, width = layer 0's output channels
    int32_t dims[2];
    ok = layers == 2 && fread(dims, 4, 2, f) == 2 && dims[0] == kLeanInC && (dims[1] == 16 || dims[1] == 32);
    fclose(f);
    if (!ok) return ARBOR_ERR_ARG;
    *levels = 0;
    for (int k = 0; k < kMaxLevels; k++) widths[k] = k == 0 ? dims[1] : 0;
    return ARBOR_OK;
  }
  ok = ok && kind == kKindUNet && layers >= 3 && (layers + 1) % 4 == 0 && (layers + 1) / 4 <= kMaxLevels;
  const int L = ok ? (layers + 1) / 4 : 0;
  for (int li = 0; ok && li < 2 * L; li++) {
    int32_t dims[2];
    ok = fread(dims, 4, 2, f) == 2 && dims[0] > 0 && dims[1] > 0 &&
         fseek(f, (long)(9LL * dims[0] * dims[1] + dims[1]) * 4, SEEK_CUR) == 0;
    if (ok && li % 2 == 0) w[li / 2] = dims[1];
  }
  fclose(f);
  if (!ok) return ARBOR_ERR_ARG;
  *levels = L;
  for (int k = 0; k < kMaxLevels; k++) widths[k] = w[k];
  return ARBOR_OK;
}

void* arbor_denoiser_tensor(arbor_denoiser* d, const char* name) {
  if (!d || !name) return nullptr;
  if (strcmp(name, "input") == 0) return d->in.p;
  if (strcmp(name, "output") == 0) return d->out.p;
  if (strcmp(name, "target") == 0) return d->target;
  const char* colon = strchr(name, ':');
  if (!colon) return nullptr;
  char* end = nullptr;
  long li = strtol(colon + 1, &end, 10);
  if (end == colon + 1 || *end != '\0' || li < 0 || li >= (long)d->conv.size()) return nullptr;
  const Conv& cv = d->conv[(size_t)li];
  const std::string kind(name, colon - name);
  if (kind == "weight") return cv.w;
  if (kind == "bias") return cv.b;
  if (kind == "wgrad") return cv.gw;
  if (kind == "bgrad") return cv.gb;
  return nullptr;
}

int arbor_denoiser_destroy(arbor_denoiser* d) {
  if (!d) return ARBOR_ERR_NULL;
  (void)hipDeviceSynchronize();
  for (void* q : d->owned) (void)hipFree(q);
  delete d;
  return ARBOR_OK;
}

int arbor_denoiser_slot(const arbor_denoiser* d, int* max_batch, int* height, int* width) {
  if (!d || !max_batch || !height || !width) return ARBOR_ERR_NULL;
  *max_batch = d->maxBatch;
  *height = d->H;
  *width = d->W;
  return ARBOR_OK;
}

long long arbor_denoiser_parameters(const arbor_denoiser* d) {
  if (!d) return ARBOR_ERR_NULL;
  long long n = 0;
  for (const Conv& cv : d->conv) n += cv.nw() + cv.cout;
  return n;
}

int arbor_denoiser_set_input(arbor_denoiser* d, int index, const float* Linear, const float* Guide,
                             const float* Normal, const float* Hit, int w, int h, hipStream_t stream) {
  ARBOR_REQUIRE(d && Linear && Guide && (Normal || d->lean()) && Hit);
  if (!slotOk(d, index, w, h)) return ARBOR_ERR_ARG;
  auto X = reinterpret_cast<bf16*>(d->in.p) + (size_t)index * d->H * d->W * d->inC();
  hipLaunchKernelGGL(d->lean() ? k_features<true> : k_features<false>, dim3((d->W + 63) / 64, d->H), dim3(64), 0,
                     stream, reinterpret_cast<const float4*>(Linear), reinterpret_cast<const float4*>(Guide),
                     reinterpret_cast<const float4*>(Normal), reinterpret_cast<const float4*>(Hit), w, h, X, d->H,
                     d->W);
  return arbor::launch_status();
}

int arbor_denoiser_set_target(arbor_denoiser* d, int index, const float* Radiance, int w, int h, hipStream_t stream) {
  ARBOR_REQUIRE(d && Radiance);
  if (!d->train) return ARBOR_ERR_STATE;
  if (!slotOk(d, index, w, h)) return ARBOR_ERR_ARG;
  auto T = reinterpret_cast<float4*>(d->target) + (size_t)index * d->H * d->W;
  hipLaunchKernelGGL(k_target, dim3((d->W + 63) / 64, d->H), dim3(64), 0, stream,
                     reinterpret_cast<const float4*>(Radiance), w, h, T, d->H, d->W);
  return arbor::launch_status();
}

int arbor_denoiser_forward(arbor_denoiser* d, int n, hipStream_t stream) {
  ARBOR_REQUIRE(d);
  if (n < 1 || n > d->maxBatch) return ARBOR_ERR_ARG;
  return forward(d, n, stream);
}

}  // extern "C"

namespace {
// Whether the network runs as the one-pass kernel (1 level of 16 channels).
bool fusedShape(const arbor_denoiser* d) { return d->levels == 1 && d->widths[0] == fused::kC; }

// The lean network's one-pass kernel over compact features (w x h, row
// stride `stride`) into Out (and, with History and Pixels, the resolve).
int leanPass(arbor_denoiser* d, const void* Feat, int stride, int w, int h, float* Out, float* History,
             unsigned* Pixels, const float* C, hipStream_t stream) {
  const Conv &c0 = d->conv[0], &c1 = d->conv[1];
  const int tiles = ((w + lean::kTW - 1) / lean::kTW) * ((h + lean::kTH - 1) / lean::kTH);
  const dim3 grid(fused::kXcds * ((tiles + fused::kXcds - 1) / fused::kXcds));
  const bool resolve = History != nullptr;
  auto kernel = d->widths[0] == 32 ? (resolve ? k_denoiseLean<2, true> : k_denoiseLean<2, false>)
                                   : (resolve ? k_denoiseLean<1, true> : k_denoiseLean<1, false>);
  hipLaunchKernelGGL(kernel, grid, dim3(lean::kThreads), 0, stream, reinterpret_cast<const arbor::featx8*>(Feat),
                     stride, w, h, reinterpret_cast<const bf16*>(c0.w16), c0.b,
                     reinterpret_cast<const bf16*>(c1.w16) + 4LL * c1.cout * c1.cin, c1.b,
                     reinterpret_cast<float4*>(Out), C, reinterpret_cast<float4*>(History),
                     reinterpret_cast<uint32_t*>(Pixels));
  return arbor::launch_status();
}

// One frame through the network into Out (and, with History and Pixels, the
// resolve): the one-pass kernel for the 1-level 16-channel shape; for the
// lean network its compact features into slot 0, then its one-pass kernel;
// else the input features, forward and output on slot 0 (then
// resolvePixels).
int denoiseFrame(arbor_denoiser* d, const float* Linear, const float* Guide, const float* Normal, const float* Hit,
                 int w, int h, float* Out, float* History, unsigned* Pixels, const float* C, hipStream_t stream) {
  if (!slotOk(d, 0, w, h)) return ARBOR_ERR_ARG;
  const bool resolve = History != nullptr;
  if (fusedShape(d)) {
    const Conv &c0 = d->conv[0], &c1 = d->conv[1], &c2 = d->conv[2];
    const int tiles = ((w + fused::kTW - 1) / fused::kTW) * ((h + fused::kTH - 1) / fused::kTH);
    const dim3 grid(fused::kXcds * ((tiles + fused::kXcds - 1) / fused::kXcds));
    auto kernel = resolve ? k_denoise1<true> : k_denoise1<false>;
    hipLaunchKernelGGL(kernel, grid, dim3(fused::kThreads), 0, stream, reinterpret_cast<const float4*>(Linear),
                       reinterpret_cast<const float4*>(Guide), reinterpret_cast<const float4*>(Normal),
                       reinterpret_cast<const float4*>(Hit), w, h, d->H, d->W,
                       reinterpret_cast<const bf16*>(c0.w16), c0.b, reinterpret_cast<const bf16*>(c1.w16), c1.b,
                       reinterpret_cast<const bf16*>(c2.w16), c2.b, reinterpret_cast<float4*>(Out), C,
                       reinterpret_cast<float4*>(History), reinterpret_cast<uint32_t*>(Pixels));
    return arbor::launch_status();
  }
  DN_TRY(arbor_denoiser_set_input(d, 0, Linear, Guide, Normal, Hit, w, h, stream));
  if (d->lean()) return leanPass(d, d->in.p, d->W, w, h, Out, History, Pixels, C, stream);
  DN_TRY(arbor_denoiser_forward(d, 1, stream));
  DN_TRY(arbor_denoiser_get_output(d, 0, Out, w, h, stream));
  if (resolve) return arbor_resolvePixels(Out, History, Pixels, C, w, h, stream);
  return ARBOR_OK;
}
}  // namespace

namespace arbor {
int denoiseResolve(arbor_denoiser* d, const float* Linear, const float* Guide, const float* Normal, const float* Hit,
                   float* Filtered, float* History, unsigned* Pixels, const float* C, int w, int h,
                   hipStream_t stream) {
  if (!d || !Linear || !Guide || !Normal || !Hit || !Filtered || !History || !Pixels || !C) return ARBOR_ERR_NULL;
  return denoiseFrame(d, Linear, Guide, Normal, Hit, w, h, Filtered, History, Pixels, C, stream);
}
bool denoiserLean(const arbor_denoiser* d) { return d && d->lean(); }
int denoiseFeaturesResolve(arbor_denoiser* d, const void* Feat, float* Filtered, float* History, unsigned* Pixels,
                           const float* C, int w, int h, hipStream_t stream) {
  if (!d || !Feat || !Filtered || !History || !Pixels || !C) return ARBOR_ERR_NULL;
  if (!d->lean()) return ARBOR_ERR_STATE;
  if (!slotOk(d, 0, w, h)) return ARBOR_ERR_ARG;
  return leanPass(d, Feat, w, w, h, Filtered, History, Pixels, C, stream);
}
}  // namespace arbor

extern "C" {

int arbor_denoiser_denoise(arbor_denoiser* d, const float* Linear, const float* Guide, const float* Normal,
                           const float* Hit, int w, int h, float* Out, hipStream_t stream) {
  ARBOR_REQUIRE(d && Linear && Guide && (Normal || d->lean()) && Hit && Out);
  return denoiseFrame(d, Linear, Guide, Normal, Hit, w, h, Out, nullptr, nullptr, nullptr, stream);
}

int arbor_denoiser_denoise_features(arbor_denoiser* d, const void* Feat, int w, int h, float* Out,
                                    hipStream_t stream) {
  ARBOR_REQUIRE(d && Feat && Out);
  if (!d->lean()) return ARBOR_ERR_STATE;
  if (!slotOk(d, 0, w, h)) return ARBOR_ERR_ARG;
  return leanPass(d, Feat, w, w, h, Out, nullptr, nullptr, nullptr, stream);
}

int arbor_denoiser_get_output(arbor_denoiser* d, int index, float* Out, int w, int h, hipStream_t stream) {
  ARBOR_REQUIRE(d && Out);
  if (!slotOk(d, index, w, h)) return ARBOR_ERR_ARG;
  auto Y = reinterpret_cast<const bf16*>(d->out.p) + (size_t)index * d->H * d->W * kOutC;
  hipLaunchKernelGGL(k_output, dim3((w + 63) / 64, h), dim3(64), 0, stream, Y, d->H, d->W,
                     reinterpret_cast<float4*>(Out), w, h);
  return arbor::launch_status();
}

int arbor_denoiser_train_step(arbor_denoiser* d, int n, float lr, hipStream_t stream) {
  ARBOR_REQUIRE(d);
  if (!d->train) return ARBOR_ERR_STATE;
  if (n < 1 || n > d->maxBatch || !(lr > 0.0f)) return ARBOR_ERR_ARG;
  DN_TRY(forward(d, n, stream));
  const long long px = (long long)n * d->H * d->W;
  const int blocks = (int)((px + 255) / 256 < 1024 ? (px + 255) / 256 : 1024);
  // mean over every pixel and channel of the batch (padding counts as zero loss)
  const float scale = 1.0f / (float)(px * 3);
  hipLaunchKernelGGL(k_l1Loss, dim3(blocks), dim3(256), 0, stream, reinterpret_cast<const bf16*>(d->out.p),
                     reinterpret_cast<const float4*>(d->target), px, scale, reinterpret_cast<bf16*>(d->out.g),
                     d->lossPart);
  hipLaunchKernelGGL(k_sumTo, dim3(1), dim3(64), 0, stream, d->lossPart, blocks, scale, d->loss);
  DN_TRY(arbor::launch_status());
  DN_TRY(backward(d, n, stream));
  d->step++;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float lrT = lr * sqrtf(1.0f - powf(b2, (float)d->step)) / (1.0f - powf(b1, (float)d->step));
  for (Conv& cv : d->conv) {
    hipLaunchKernelGGL(k_adam, dim3(grid1(cv.nw())), dim3(256), 0, stream, cv.w, cv.gw, cv.mw, cv.vw, cv.nw(), lrT, b1,
                       b2, eps, reinterpret_cast<bf16*>(cv.w16));
    hipLaunchKernelGGL(k_adam, dim3(grid1(cv.cout)), dim3(256), 0, stream, cv.b, cv.gb, cv.mb, cv.vb,
                       (long long)cv.cout, lrT, b1, b2, eps, (bf16*)nullptr);
    DN_TRY(arbor_dn_flip_weights(cv.w16, cv.cin, cv.cout, cv.wflip, stream));
  }
  return arbor::launch_status();
}

int arbor_denoiser_loss(arbor_denoiser* d, float* host_loss) {
  ARBOR_REQUIRE(d && host_loss);
  if (!d->train) return ARBOR_ERR_STATE;
  if (hipMemcpy(host_loss, d->loss, 4, hipMemcpyDeviceToHost) != hipSuccess) return ARBOR_ERR_DEVICE;
  return ARBOR_OK;
}

// File: "ARBORDN1" (a U-Net) or "ARBORDN2" and an int32 network kind (1:
// the lean network), then the int32 layer count, per layer int32 cin, cout,
// then fp32 weights [9][cout][cin] and biases [cout].
int arbor_denoiser_save(const arbor_denoiser* d, const char* path) {
  ARBOR_REQUIRE(d && path);
  if (hipDeviceSynchronize() != hipSuccess) return ARBOR_ERR_DEVICE;
  FILE* f = fopen(path, "wb");
  if (!f) return ARBOR_ERR_ARG;
  bool ok = fwrite(d->lean() ? "ARBORDN2" : "ARBORDN1", 1, 8, f) == 8;
  if (d->lean()) ok = ok && fwrite(&kKindLean, 4, 1, f) == 1;
  int32_t layers = (int32_t)d->conv.size();
  ok = ok && fwrite(&layers, 4, 1, f) == 1;
  for (const Conv& cv : d->conv) {
    int32_t dims[2] = {cv.cin, cv.cout};
    std::vector<float> w((size_t)cv.nw()), b((size_t)cv.cout);
    ok = ok && hipMemcpy(w.data(), cv.w, w.size() * 4, hipMemcpyDeviceToHost) == hipSuccess &&
         hipMemcpy(b.data(), cv.b, b.size() * 4, hipMemcpyDeviceToHost) == hipSuccess;
    ok = ok && fwrite(dims, 4, 2, f) == 2 && fwrite(w.data(), 4, w.size(), f) == w.size() &&
         fwrite(b.data(), 4, b.size(), f) == b.size();
  }
  ok = fclose(f) == 0 && ok;
  return ok ? ARBOR_OK : ARBOR_ERR_DEVICE;
}

// The whole file is read and checked (magic, layer count, every layer's shape
// and values, nothing after the last layer) before any device weight changes,
// so a truncated or mismatched checkpoint leaves the network as it was.
int arbor_denoiser_load(arbor_denoiser* d, const char* path) {
  ARBOR_REQUIRE(d && path);
  FILE* f = fopen(path, "rb");
  if (!f) return ARBOR_ERR_ARG;
  int32_t kind = 0, layers = 0;
  bool ok = readHeader(f, &kind, &layers) && kind == (d->lean() ? kKindLean : kKindUNet) &&
            layers == (int32_t)d->conv.size();
  std::vector<std::vector<float>> ws, bs;
  for (size_t li = 0; ok && li < d->conv.size(); li++) {
    const Conv& cv = d->conv[li];
    int32_t dims[2];
    ok = fread(dims, 4, 2, f) == 2 && dims[0] == cv.cin && dims[1] == cv.cout;
    if (!ok) break;
    ws.emplace_back((size_t)cv.nw());
    bs.emplace_back((size_t)cv.cout);
    ok = fread(ws.back().data(), 4, ws.back().size(), f) == ws.back().size() &&
         fread(bs.back().data(), 4, bs.back().size(), f) == bs.back().size();
  }
  char extra;
  ok = ok && fread(&extra, 1, 1, f) == 0;
  fclose(f);
  if (!ok) return ARBOR_ERR_ARG;
  for (size_t li = 0; li < d->conv.size(); li++) {
    Conv& cv = d->conv[li];
    if (hipMemcpy(cv.w, ws[li].data(), ws[li].size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(cv.b, bs[li].data(), bs[li].size() * 4, hipMemcpyHostToDevice) != hipSuccess)
      return ARBOR_ERR_DEVICE;
  }
  int rc = refreshBf16(d, nullptr);
  if (rc == ARBOR_OK && hipDeviceSynchronize() != hipSuccess) rc = ARBOR_ERR_DEVICE;
  return rc;
}

}  // extern "C"
