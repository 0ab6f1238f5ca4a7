// filterLighting and resolvePixels as device code shared by the single-view
// kernels and the batched views (Arbor.cu:292-304): one 32x8 block of a
// view's pixels filtered from an LDS tile with a 2-pixel halo, and one
// pixel's progressive resolve.
#pragma once
#include <limits.h>
#include <math.h>

#include "arbor_device.hpp"

namespace arbor {

// The display mapping of resolvePixels (Arbor.cu:304) on the vector ALU's
// transcendental units: the ACES fit's quotient through the hardware
// reciprocal (1 ulp), the 1/2.2 power as exp2(log2(x) / 2.2) (v_log_f32,
// v_exp_f32: under 1e-6 relative for x in (0, 1]), and the vignette's
// coordinates from reciprocals of the view size. Each value is within a few
// ulp of the correctly rounded operations, far inside the 8-bit output step
// (the reference's glibc powf and divisions are matched within one step);
// the accumulation into History is exact. IEEE powf and divisions cost
// about 575 vector instructions per wave of pixels, this mapping about 100.
AD float aces(float x) {
  return sat(x * (2.51f * x + 0.03f) * __builtin_amdgcn_rcpf(x * (2.43f * x + 0.59f) + 0.14f));
}
AD float gamma22(float x) { return x > 0.0f ? __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * (1.0f / 2.2f)) : 0.0f; }

// One pixel of resolvePixels (Arbor.cu:304) given its Filtered colour.
__device__ __forceinline__ void resolveOne(v3 current, float4* History, uint32_t* Pixels, const float* C, int x,
                                           int y, int width, int height, int i) {
  float count = C[10] > 0.5f ? 0.0f : C[9];
  v3 mean = current;
  if (!(count < 0.5f)) {
    float4 h = History[i];
    mean = mix(mk(h.x, h.y, h.z), current, 1.0f / (count + 1.0f));
  }
  History[i] = make_float4(mean.x, mean.y, mean.z, count + 1.0f);
  v3 col = mean * C[20];
  col = mk(aces(col.x), aces(col.y), aces(col.z));
  col = mk(gamma22(col.x), gamma22(col.y), gamma22(col.z));
  float u = ((float)x * __builtin_amdgcn_rcpf((float)width) - 0.5f) * 2.0f;
  float v = ((float)y * __builtin_amdgcn_rcpf((float)height) - 0.5f) * 2.0f;
  col = col * (1.0f - 0.09f * (u * u + v * v));
  Pixels[i] = packRGB(col);
}

// filterLighting (Arbor.cu:292-302): 5x5 cross-bilateral, same id/tree/depth.
// One block filters a 32x8 tile. The tile plus a 2-pixel halo (36x12) is
// staged in LDS as {t, id, tree, Linear/(Guide+0.02)}: the 25 taps read LDS
// and each pixel's demodulation (three divisions) is computed once instead of
// once per tap; the same operands and summation order give the same result.
// The 6 distinct tap weights expf(-0.3*(j*j+k*k)) arrive precomputed.
constexpr int kFilterW = 32, kFilterH = 8, kHalo = 2;
constexpr int kTileW = kFilterW + 2 * kHalo, kTileH = kFilterH + 2 * kHalo;
// One 32x8 block of pixels (block coordinates bx, by of the view). With
// kResolve the block then resolves each pixel from the colour it just
// filtered (History and Pixels as resolvePixels writes them): the fused
// filter + resolve, with the same bits as the two kernels.
struct FilterWeights { float w[9]; };  // indexed by j*j+k*k (0,1,2,4,5,8 used)
template <bool kResolve>
__device__ __forceinline__ void filterBlock(const float4* Linear, const float4* Guide, const float4* Hit,
                                            float4* Filtered, const float* C, int width, int height,
                                            const FilterWeights& fw, float4* History, uint32_t* Pixels, int bx,
                                            int by) {
  __shared__ float st[kTileH][kTileW];
  __shared__ int sid[kTileH][kTileW], stree[kTileH][kTileW];
  __shared__ float sdr[kTileH][kTileW], sdg[kTileH][kTileW], sdb[kTileH][kTileW];
  const int x0 = bx * kFilterW - kHalo, y0 = by * kFilterH - kHalo;
  const int tid = threadIdx.y * kFilterW + threadIdx.x;
  // Every load of the block is issued before the first is used: the tile's
  // halo pixels (pixel 0 stands in outside the view) and the thread's own
  // pixel, whose raw Linear and Guide the blend needs.
  constexpr int kItems = kTileW * kTileH, kThreads = kFilterW * kFilterH;
  constexpr int kIter = (kItems + kThreads - 1) / kThreads;
  float4 hq[kIter], lq[kIter], gq[kIter];
  bool inView[kIter];
#pragma unroll
  for (int s = 0; s < kIter; s++) {
    const int i = tid + s * kThreads, ty = i / kTileW, tx = i % kTileW, gx = x0 + tx, gy = y0 + ty;
    inView[s] = i < kItems && gx >= 0 && gy >= 0 && gx < width && gy < height;
    const int q = inView[s] ? gy * width + gx : 0;
    hq[s] = Hit[q];
    lq[s] = Linear[q];
    gq[s] = Guide[q];
  }
  const int x = bx * kFilterW + threadIdx.x, y = by * kFilterH + threadIdx.y;
  const bool mine = x < width && y < height;
  const int b = mine ? y * width + x : 0;
  const float4 hb = Hit[b], lb = Linear[b], gb = Guide[b];
#pragma unroll
  for (int s = 0; s < kIter; s++) {
    const int i = tid + s * kThreads, ty = i / kTileW, tx = i % kTileW;
    if (i >= kItems) continue;
    if (!inView[s]) {
      sid[ty][tx] = INT_MIN;  // never equals a pixel's id (ids are >= -1)
      continue;
    }
    st[ty][tx] = hq[s].x;
    sid[ty][tx] = (int)hq[s].y;
    stree[ty][tx] = (int)hq[s].z;
    v3 g = mk(gq[s].x, gq[s].y, gq[s].z) + mk(0.02f, 0.02f, 0.02f);
    sdr[ty][tx] = lq[s].x / g.x;
    sdg[ty][tx] = lq[s].y / g.y;
    sdb[ty][tx] = lq[s].z / g.z;
  }
  __syncthreads();
  if (!mine) return;
  v3 col = mk(lb.x, lb.y, lb.z);
  v3 base = mk(gb.x, gb.y, gb.z) + mk(0.02f, 0.02f, 0.02f);
  v3 sum = mk(0.0f, 0.0f, 0.0f);
  float weight = 0.0f;
  int id = (int)hb.y, tree = (int)hb.z;
  float depth = hb.x;
  if (depth < kFar && C[15] < 0.5f) {
    float tol = fmaxf(0.005f, hb.w * 6.0f);
    const int cx = threadIdx.x + kHalo, cy = threadIdx.y + kHalo;
#pragma unroll
    for (int j = -2; j <= 2; j++)
#pragma unroll
      for (int k = -2; k <= 2; k++) {
        int ty = cy + j, tx = cx + k;
        if (sid[ty][tx] != id || stree[ty][tx] != tree || fabsf(st[ty][tx] - depth) > tol) continue;
        float w = fw.w[j * j + k * k];
        sum = sum + mk(sdr[ty][tx], sdg[ty][tx], sdb[ty][tx]) * w;
        weight += w;
      }
    if (weight > 0.0f) col = mix(col, (sum / weight) * base, 0.86f);
  }
  Filtered[b] = make_float4(col.x, col.y, col.z, 1.0f);
  if (kResolve) resolveOne(col, History, Pixels, C, x, y, width, height, b);
}

// The 6 distinct tap weights expf(-0.3*(j*j+k*k)) from the host's libm, as
// the reference harness computes them.
inline FilterWeights filterWeights() {
  FilterWeights fw;
  for (int i = 0; i < 9; i++) fw.w[i] = expf(-(float)i * 0.30f);
  return fw;
}

}  // namespace arbor
