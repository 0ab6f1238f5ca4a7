// The learned denoiser's per-pixel input features, shared by the network
// runner (which computes them from a frame's Linear, Guide, Normal and Hit)
// and the shading pass that writes the compact set directly, so both give
// the same bits.
#pragma once
#include <hip/hip_runtime.h>

namespace arbor {

typedef __bf16 featx8 __attribute__((ext_vector_type(8)));

// Hit distance of a ray that left the scene (FAR).
constexpr float kFeatFar = 100000.0f;

// log(1 + x) for x >= 0 within about 1e-6 relative: below 1/16 its Taylor
// series to x^6 (truncation x^6 / 7 relative, under 2e-8), above it the
// hardware logarithm of 1 + x (the rounding of 1 + x costs at most 2^-24 /
// log(1 + 1/16), 1e-6 relative). The results are rounded to bf16 (2^-8).
// Both forms are evaluated and one selected (no branch: neighbouring pixels
// take either side).
__device__ __forceinline__ float log1pFast(float x) {
  const float series = x * (1.0f - x * (0.5f - x * (1.0f / 3.0f - x * (0.25f - x * (0.2f - x * (1.0f / 6.0f))))));
  const float hw = __logf(1.0f + x);
  return x < 0.0625f ? series : hw;
}
// exp(y) - 1 within about 2e-6 relative: below |y| = 1/16 its Taylor series
// to y^4 in fused multiply-adds (truncation under y^4 / 120 = 1.3e-7
// relative), above it the hardware exponential (1 ulp) minus 1, which
// magnifies its error at most e^(1/16) / (e^(1/16) - 1) = 16.5 times.
__device__ __forceinline__ float expm1Fast(float y) {
  const float series = y * fmaf(y, fmaf(y, fmaf(y, 1.0f / 24.0f, 1.0f / 6.0f), 0.5f), 1.0f);
  const float hw = __expf(y) - 1.0f;
  return fabsf(y) < 0.0625f ? series : hw;
}
// Depth feature of hit distance t: log(1 + min(t, 1000)) / 7, 1 for sky.
__device__ __forceinline__ float depthFeature(float t) {
  return t < kFeatFar ? log1pFast(fmaxf(fminf(t, 1000.0f), 0.0f)) * (1.0f / 7.0f) : 1.0f;
}
// The compact feature set of one pixel (8 bf16, 16 bytes): log(1 + radiance)
// rgb of its Linear value, albedo rgb (Guide), the depth feature of its Hit
// distance t, and a zero channel.
__device__ __forceinline__ featx8 compactFeatures(const float4 l, const float4 g, float t) {
  featx8 f;
  f[0] = (__bf16)log1pFast(fmaxf(l.x, 0.0f));
  f[1] = (__bf16)log1pFast(fmaxf(l.y, 0.0f));
  f[2] = (__bf16)log1pFast(fmaxf(l.z, 0.0f));
  f[3] = (__bf16)g.x;
  f[4] = (__bf16)g.y;
  f[5] = (__bf16)g.z;
  f[6] = (__bf16)depthFeature(t);
  f[7] = (__bf16)0.0f;
  return f;
}

}  // namespace arbor
