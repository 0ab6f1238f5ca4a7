// Kernels of the learned denoiser on gfx950: a 3x3 convolution encoder-decoder
// in BF16 with FP32 accumulation on the matrix cores, and its training pieces.
//
// Tensors are NHWC bf16 with the channel count a multiple of 8, so a lane's
// 8 consecutive channels are one 16-byte load. A 3x3 convolution is the
// implicit GEMM Y[p, co] = sum_{tap, ci} X[p + off(tap), ci] * W[tap, co, ci]
// (zero outside the image): per tap and 32-channel chunk one
// v_mfma_f32_16x16x32_bf16 per 16x16 output subtile. Operand layout of that
// instruction (checked on the device): A row = lane % 16, B column = lane % 16,
// both with k = 8*(lane/16) + j for their 8 bf16; D column = lane % 16, rows
// 4*(lane/16) + i for its 4 floats.
//
// The gradient with respect to X is the same convolution of dY with the
// weights flipped in the taps and transposed in the channels. The weight
// gradient is a GEMM over pixels: tiles of dY and X are staged transposed in
// LDS so each operand fragment is one 16-byte LDS read; blocks write partial
// sums per pixel chunk, reduced by a second kernel (deterministic).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "arbor_conv1x1.hpp"
#include "arbor_denoise.h"
#include "launch.hpp"

namespace arbor {
namespace dn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(__bf16 v) { return (float)v; }
__device__ __forceinline__ __bf16 f2bf(float v) { return (__bf16)v; }  // round to nearest even
__device__ __forceinline__ bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int j = 0; j < 8; j++) z[j] = (__bf16)0.0f;
  return z;
}

// ---- 3x3 convolution, stride 1, zero padding 1 ----
// A block computes a (4*MSUB)x16-pixel output tile of one image for the
// output channels [co0, co0 + 16*NSUB) of a CoutTotal-channel output. The
// input tile with its 1-pixel halo is staged in LDS 32 channels at a time, so
// the nine taps read their pixel fragments from LDS: each input value leaves
// HBM/L2 about 1.3 times instead of nine. Wave w owns output rows
// MSUB*w .. MSUB*w + MSUB-1 (one 16-pixel subtile each); weight fragments
// come from global memory, shared by every block through the caches, and
// each feeds MSUB matrix instructions.
//
// With Xc non-NULL the input is the channel concatenation [up2(Xc), X] of a
// half-resolution tensor Xc (Cc channels, read at (y/2, x/2)) and X (Cin - Cc
// channels): the decoder's upsampling and skip concatenation happen in the
// staging, never in memory.
constexpr int kTileW = 16, kHaloW = kTileW + 2, kChunk = 32;
// LDS index (in bf16) of the 16-byte channel group grp (0..3) of staged pixel
// pix: 64 bytes per pixel, the group XORed with bits 1-2 of the pixel index.
// Both access patterns are then free of bank conflicts on gfx950 (measured
// with SQ_LDS_BANK_CONFLICT): the staging stores (four lanes per pixel) and
// the tap reads (16 consecutive pixels per quarter-wave, one group each). A
// plain 80-byte pixel stride costs 7.7 conflict cycles per store and 4 per
// read.
__device__ __forceinline__ int ldsSlot(int pix, int grp) { return pix * kChunk + ((grp ^ (pix >> 1)) & 3) * 8; }
// Waves per SIMD the register allocation targets, per output width: the
// highest each width reaches without scratch (5 for 32 channels, 4 for 48, 3
// for 128; 1 leaves the allocation unforced). Over the network's layer shapes
// at 1088x1920 these targets give the lowest total time.
constexpr int convWaves(int nsub) { return nsub == 2 ? 5 : (nsub == 3 ? 4 : (nsub == 8 ? 3 : 1)); }
// kPack taps share a matrix instruction (no concatenation):
//  * 2 (16 -> 16): K slots 0-15 hold tap 2k's 16 channels and 16-31 tap
//    2k+1's; the fifth instruction holds tap 8 and zero weights. Five
//    instructions per output subtile instead of nine half-empty ones.
//  * 4 (8 input channels): K slots 8g..8g+7 hold tap 4k+g's 8 channels; the
//    third instruction holds tap 8 and zero weights. Three instructions
//    instead of nine quarter-full ones. The bias is the accumulator's
//    initial value (no addition after the products), as the lean network's
//    one-pass kernel computes it.
template <int NSUB, int MSUB, int kPack = 1>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(convWaves(NSUB)))) void k_conv3x3(
    const __bf16* __restrict__ X, const __bf16* __restrict__ Xc, int Cc, int N, int H, int W, int Cin,
    const __bf16* __restrict__ Wt, const float* __restrict__ bias, int relu, __bf16* __restrict__ Y, int CoutTotal,
    int co0) {
  constexpr int kTileH = 4 * MSUB, kHaloH = kTileH + 2;
  constexpr int kGroups = kHaloH * kHaloW * 4;   // 16-byte groups staged per chunk
  constexpr int kStage = (kGroups + 255) / 256;  // groups per thread
  __shared__ __attribute__((aligned(16))) __bf16 tile[kHaloH * kHaloW * kChunk];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = lane & 15, g = lane >> 4;
  const int tilesX = (W + kTileW - 1) / kTileW, tilesY = (H + kTileH - 1) / kTileH;
  int b = blockIdx.x;
  const int tx = b % tilesX;
  b /= tilesX;
  const int ty = b % tilesY, n = b / tilesY;
  const int x0 = tx * kTileW, y0 = ty * kTileH;
  const int Cs = Cin - Cc;  // channels read from X
  // The weights are the A operand and the pixels the B operand, so lane
  // (row, g) accumulates output pixel x0 + row, channels co0 + 16j + 4g + i:
  // four consecutive channels of one pixel, stored as one 8-byte write.
  f32x4 acc[MSUB][NSUB];
#pragma unroll
  for (int j = 0; j < NSUB; j++) {
    f32x4 init = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    if (kPack == 4 && bias) {
#pragma unroll
      for (int i = 0; i < 4; i++) init[i] = bias[co0 + 16 * j + 4 * g + i];
    }
#pragma unroll
    for (int m = 0; m < MSUB; m++) acc[m][j] = init;
  }
  for (int c0 = 0; c0 < Cin; c0 += kChunk) {
    // Stage (kHaloH x 18 pixels) x 32 channels as 16-byte groups. Every load
    // is issued before the first store, from an address clamped into the
    // tensor; groups outside the image or past Cin are then stored as zeros.
    bf16x8 sv[kStage];
    bool ok[kStage];
    // the staging addresses are recomputed per chunk rather than held in
    // registers through the matrix loop
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
#pragma unroll
    for (int s = 0; s < kStage; s++) {
      const int i = tid + 256 * s;
      const int pix = i >> 2, grp = i & 3;
      const int yy = y0 - 1 + pix / kHaloW, xx = x0 - 1 + pix % kHaloW, ci = c0 + grp * 8;
      ok[s] = i < kGroups && yy >= 0 && yy < H && xx >= 0 && xx < W && ci < Cin;
      const int yc = min(max(yy, 0), H - 1), xc = min(max(xx, 0), W - 1), cc = min(ci, Cin - 8);
      const __bf16* src =
          cc < Cc ? Xc + ((size_t)(n * (H >> 1) + (yc >> 1)) * (W >> 1) + (xc >> 1)) * Cc + cc
                  : X + ((size_t)(n * H + yc) * W + xc) * Cs + cc - Cc;
      sv[s] = *reinterpret_cast<const bf16x8*>(src);
    }
#pragma unroll
    for (int s = 0; s < kStage; s++) {
      const int i = tid + 256 * s;
      if (i < kGroups) *reinterpret_cast<bf16x8*>(&tile[ldsSlot(i >> 2, i & 3)]) = ok[s] ? sv[s] : zero8();
    }
    __syncthreads();
    if constexpr (kPack == 4) {
      // lane (row, g): tap 4k + g of instruction k, channels 0-7 (staged
      // group 0); every lane reads a staged pixel (tap 8 for the slots
      // without a tap, whose weights are zero)
      auto loadQuad = [&](int k, bf16x8* w8) {
        const int tap = 4 * k + g;
#pragma unroll
        for (int j = 0; j < NSUB; j++)
          w8[j] = tap < 9 ? *reinterpret_cast<const bf16x8*>(Wt + ((size_t)tap * CoutTotal + co0 + j * 16 + row) * Cin)
                          : zero8();
      };
      auto quadTaps = [&](int k, const bf16x8* w8) {
        const int tl = min(4 * k + g, 8), dy = tl / 3, dx = tl - 3 * (tl / 3);
#pragma unroll
        for (int m = 0; m < MSUB; m++) {
          const bf16x8 a =
              *reinterpret_cast<const bf16x8*>(&tile[ldsSlot((MSUB * wave + m + dy) * kHaloW + row + dx, 0)]);
#pragma unroll
          for (int j = 0; j < NSUB; j++)
            acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w8[j], a, acc[m][j], 0, 0, 0);
        }
      };
      bf16x8 w0[NSUB], w1[NSUB], w2[NSUB];
      loadQuad(0, w0);
      loadQuad(1, w1);
      loadQuad(2, w2);
      asm volatile("" ::: "memory");
      quadTaps(0, w0);
      quadTaps(1, w1);
      quadTaps(2, w2);
      __syncthreads();
      continue;
    }
    if constexpr (kPack == 2) {
      // lane (row, g): tap 2k + (g >> 1) of pair k, channels 8 (g & 1) ..;
      // the next pair's weights are loaded while the current one computes
      const int grp = g & 1;
      auto loadPair = [&](int k, bf16x8* w8) {
        const int tap = 2 * k + (g >> 1);
#pragma unroll
        for (int j = 0; j < NSUB; j++)
          w8[j] = tap < 9 ? *reinterpret_cast<const bf16x8*>(Wt + ((size_t)tap * CoutTotal + co0 + j * 16 + row) * Cin +
                                                             8 * grp)
                          : zero8();
      };
      auto pairTaps = [&](int k, const bf16x8* w8) {
        const int tl = min(2 * k + (g >> 1), 8), dy = tl / 3, dx = tl - 3 * (tl / 3);
#pragma unroll
        for (int m = 0; m < MSUB; m++) {
          const bf16x8 a =
              *reinterpret_cast<const bf16x8*>(&tile[ldsSlot((MSUB * wave + m + dy) * kHaloW + row + dx, grp)]);
#pragma unroll
          for (int j = 0; j < NSUB; j++)
            acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w8[j], a, acc[m][j], 0, 0, 0);
        }
      };
      bf16x8 w0[NSUB], w1[NSUB];
      loadPair(0, w0);
      loadPair(1, w1);
      asm volatile("" ::: "memory");
      pairTaps(0, w0);
      loadPair(2, w0);
      asm volatile("" ::: "memory");
      pairTaps(1, w1);
      loadPair(3, w1);
      asm volatile("" ::: "memory");
      pairTaps(2, w0);
      loadPair(4, w0);
      asm volatile("" ::: "memory");
      pairTaps(3, w1);
      pairTaps(4, w0);
      __syncthreads();
      continue;
    }
    // Weight fragments come from global memory (L2); the next tap's are
    // loaded while the current one computes. A lane whose channel group lies
    // past Cin reads group 0's weights instead: its staged pixels are zero,
    // so those products are zero either way.
    const int ci = min(c0 + g * 8, Cin - 8);
    auto loadW = [&](int tap, bf16x8* w8) {
#pragma unroll
      for (int j = 0; j < NSUB; j++)
        w8[j] = *reinterpret_cast<const bf16x8*>(Wt + ((size_t)tap * CoutTotal + co0 + j * 16 + row) * Cin + ci);
    };
    auto taps = [&](int tap, const bf16x8* w8) {
      const int dy = tap / 3, dx = tap - dy * 3;
#pragma unroll
      for (int m = 0; m < MSUB; m++) {
        const bf16x8 a =
            *reinterpret_cast<const bf16x8*>(&tile[ldsSlot((MSUB * wave + m + dy) * kHaloW + row + dx, g)]);
#pragma unroll
        for (int j = 0; j < NSUB; j++) acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w8[j], a, acc[m][j], 0, 0, 0);
      }
    };
    {
      // two weight buffers used in turn, two taps per iteration; each empty
      // asm statement keeps the loads above it from sinking to their use
      bf16x8 w0[NSUB], w1[NSUB];
      loadW(0, w0);
#pragma unroll 1
      for (int tap = 0; tap < 8; tap += 2) {
        loadW(tap + 1, w1);
        asm volatile("" ::: "memory");
        taps(tap, w0);
        loadW(tap + 2, w0);
        asm volatile("" ::: "memory");
        taps(tap + 1, w1);
      }
      taps(8, w0);
    }
    __syncthreads();
  }
  const int x = x0 + row;
  if (x >= W) return;
  // every bias load is issued before the first store
  float bv[NSUB][4];
#pragma unroll
  for (int j = 0; j < NSUB; j++)
#pragma unroll
    for (int i = 0; i < 4; i++) bv[j][i] = bias && kPack != 4 ? bias[co0 + 16 * j + 4 * g + i] : 0.0f;
#pragma unroll
  for (int m = 0; m < MSUB; m++) {
    const int y = y0 + MSUB * wave + m;
    if (y >= H) continue;
    const size_t p = (size_t)(n * H + y) * W + x;
#pragma unroll
    for (int j = 0; j < NSUB; j++) {
      if constexpr (kPack == 4) {  // the lean network's input layer: the one-pass kernel's ReLU and rounding
        const f32x4 a = acc[m][j];
        *reinterpret_cast<bf16x4*>(Y + p * CoutTotal + co0 + 16 * j + 4 * g) =
            relu ? dn1::reluToBf16(a) : dn1::toBf16(a);
        continue;
      }
      bf16x4 o;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        float v = kPack == 4 ? acc[m][j][i] : acc[m][j][i] + bv[j][i];
        if (relu) v = fmaxf(v, 0.0f);
        o[i] = f2bf(v);
      }
      *reinterpret_cast<bf16x4*>(Y + p * CoutTotal + co0 + 16 * j + 4 * g) = o;
    }
  }
}

// ---- 1x1 convolution (the lean network's output layer) ----
// Y[p, co] = act(sum_ci X[p, ci] * W1[co, ci] + bias[co]) over P pixels of
// 16*J channels, W1 [Cout][16*J]: each wave takes 16-pixel groups, one
// matrix instruction per 16 output channels (operands as arbor_conv1x1.hpp
// lays them out) whose accumulator starts at the bias.
template <int J>
__global__ __launch_bounds__(256) void k_conv1x1(const __bf16* __restrict__ X, long long P,
                                                 const __bf16* __restrict__ W1, const float* __restrict__ bias,
                                                 int relu, __bf16* __restrict__ Y, int Cout) {
  using Op = dn1::Op1x1<J>;
  constexpr int Cin = 16 * J;
  const int lane = threadIdx.x & 63, row = lane & 15, g = lane >> 4;
  const long long groups = (P + 15) / 16;
  for (long long grp = (blockIdx.x * 256LL + threadIdx.x) >> 6; grp < groups; grp += gridDim.x * 4LL) {
    const long long p = grp * 16 + row;
    const bool ok = p < P;
    const typename Op::T b = Op::load(X + (ok ? p : P - 1) * Cin, g);
    for (int co0 = 0; co0 < Cout; co0 += 16) {
      f32x4 init = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      if (bias) {
#pragma unroll
        for (int i = 0; i < 4; i++) init[i] = bias[co0 + 4 * g + i];
      }
      const f32x4 acc = Op::mma(Op::load(W1 + (size_t)(co0 + row) * Cin, g), b, init);
      if (ok) *reinterpret_cast<bf16x4*>(Y + p * Cout + co0 + 4 * g) = relu ? dn1::reluToBf16(acc) : dn1::toBf16(acc);
    }
  }
}

// ---- weight and bias gradients ----
// Block (chunk c, tap, 64x64 tile of (co, ci)): sum over the chunk's pixels of
// dY[p, co] * X[p + off(tap), ci] into partial[c][tap][co][ci] (fp32); the
// tap-0 blocks of the first ci tile also sum dY[p, co] into partialB[c][co].
constexpr int kWgTile = 64, kWgK = 32;
__global__ __launch_bounds__(256) void k_conv3x3_wgrad(const __bf16* __restrict__ X, const __bf16* __restrict__ dY,
                                                       int N, int H, int W, int Cin, int Cout, int chunk,
                                                       float* __restrict__ partial, float* __restrict__ partialB) {
  __shared__ __attribute__((aligned(16))) __bf16 sA[kWgTile][kWgK + 8];  // [co][px], rows padded
  __shared__ __attribute__((aligned(16))) __bf16 sB[kWgTile][kWgK + 8];  // [ci][px]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, t = threadIdx.x;
  const int c = blockIdx.x, tap = blockIdx.y;
  const int tilesCi = (Cin + kWgTile - 1) / kWgTile;
  const int co0 = (blockIdx.z / tilesCi) * kWgTile, ci0 = (blockIdx.z % tilesCi) * kWgTile;
  const int M = N * H * W, HW = H * W;
  const int dy = tap / 3 - 1, dx = tap % 3 - 1;
  const int p0 = c * chunk, p1 = min(M, p0 + chunk);
  const bool doBias = tap == 0 && ci0 == 0;
  float biasAcc = 0.0f;  // thread t < 64 sums channel co0 + t
  f32x4 acc[4];
#pragma unroll
  for (int n = 0; n < 4; n++) acc[n] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  const int row = lane & 15, g = lane >> 4;
  // staging: thread t loads 8 channels (group t % 8) of pixel t / 8 of the 32-pixel step
  const int sp = t >> 3, sc = (t & 7) << 3;
  for (int pb = p0; pb < p1; pb += kWgK) {
    const int p = pb + sp;
    bf16x8 vy = zero8(), vx = zero8();
    if (p < p1) {
      if (co0 + sc < Cout) vy = *reinterpret_cast<const bf16x8*>(dY + (size_t)p * Cout + co0 + sc);
      const int n = p / HW, r = p - n * HW, y = r / W, x = r - y * W;
      const int yy = y + dy, xx = x + dx;
      if (ci0 + sc < Cin && yy >= 0 && yy < H && xx >= 0 && xx < W)
        vx = *reinterpret_cast<const bf16x8*>(X + ((size_t)(n * H + yy) * W + xx) * Cin + ci0 + sc);
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
      sA[sc + j][sp] = vy[j];
      sB[sc + j][sp] = vx[j];
    }
    __syncthreads();
    if (doBias && t < kWgTile) {
#pragma unroll 8
      for (int k = 0; k < kWgK; k++) biasAcc += bf2f(sA[t][k]);
    }
    const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sA[wave * 16 + row][8 * g]);
#pragma unroll
    for (int n = 0; n < 4; n++) {
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(&sB[n * 16 + row][8 * g]);
      acc[n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[n], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int n = 0; n < 4; n++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int co = co0 + wave * 16 + 4 * g + i, ci = ci0 + n * 16 + row;
      if (co < Cout && ci < Cin) partial[(((size_t)c * 9 + tap) * Cout + co) * Cin + ci] = acc[n][i];
    }
  if (doBias && t < kWgTile && co0 + t < Cout) partialB[(size_t)c * Cout + co0 + t] = biasAcc;
}

// out[i] = sum over c of part[c * n + i]
__global__ void k_sumParts(const float* __restrict__ part, int chunks, long long n, float* __restrict__ out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    float s = 0.0f;
    for (int c = 0; c < chunks; c++) s += part[c * n + i];
    out[i] = s;
  }
}

// Wf[tap][ci][co] = W[8 - tap][co][ci]
__global__ void k_flipWeights(const __bf16* __restrict__ Wt, int Cin, int Cout, __bf16* __restrict__ Wf) {
  const long long n = 9LL * Cin * Cout;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const int co = (int)(i % Cout);
    const long long r = i / Cout;
    const int ci = (int)(r % Cin), tap = (int)(r / Cin);
    Wf[i] = Wt[((size_t)(8 - tap) * Cout + co) * Cin + ci];
  }
}

// ---- resampling, concatenation, activation gradient ----
__device__ __forceinline__ bf16x8 ld8(const __bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }
__device__ __forceinline__ void st8(__bf16* p, bf16x8 v) { *reinterpret_cast<bf16x8*>(p) = v; }
// Elementwise kernels work on 8-channel groups (16 bytes); C is a multiple of
// 8 and every count fits 32 bits (checked by the host wrappers).
__global__ void k_pool2(const __bf16* __restrict__ X, int N, int H, int W, int C, __bf16* __restrict__ Y) {
  const int Ho = H / 2, Wo = W / 2, G = C >> 3;
  const int n = N * Ho * Wo * G;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int c = (i % G) << 3;
    int r = i / G;
    const int x = r % Wo;
    r /= Wo;
    const int y = r % Ho, b = r / Ho;
    const size_t base = (((size_t)b * H + 2 * y) * W + 2 * x) * C + c, down = (size_t)W * C;
    const bf16x8 p = ld8(X + base), q = ld8(X + base + C), u = ld8(X + base + down), v = ld8(X + base + down + C);
    bf16x8 o;
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = f2bf((bf2f(p[k]) + bf2f(q[k]) + bf2f(u[k]) + bf2f(v[k])) * 0.25f);
    st8(Y + (size_t)i * 8, o);
  }
}
__global__ void k_pool2Bwd(const __bf16* __restrict__ dY, int N, int H, int W, int C, __bf16* __restrict__ dX) {
  const int G = C >> 3, n = N * H * W * G;  // H, W: the input (fine) size
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int c = (i % G) << 3;
    int r = i / G;
    const int x = r % W;
    r /= W;
    const int y = r % H, b = r / H;
    const bf16x8 g = ld8(dY + (((size_t)b * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c);
    bf16x8 o;
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = f2bf(bf2f(g[k]) * 0.25f);
    st8(dX + (size_t)i * 8, o);
  }
}
__global__ void k_up2(const __bf16* __restrict__ X, int N, int H, int W, int C, __bf16* __restrict__ Y) {
  const int G = C >> 3, n = N * 2 * H * 2 * W * G;  // H, W: the input (coarse) size
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int c = (i % G) << 3;
    int r = i / G;
    const int x = r % (2 * W);
    r /= 2 * W;
    const int y = r % (2 * H), b = r / (2 * H);
    st8(Y + (size_t)i * 8, ld8(X + (((size_t)b * H + y / 2) * W + x / 2) * C + c));
  }
}
__global__ void k_up2Bwd(const __bf16* __restrict__ dY, int N, int H, int W, int C, __bf16* __restrict__ dX) {
  const int G = C >> 3, n = N * H * W * G;  // H, W: the coarse size
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int c = (i % G) << 3;
    int r = i / G;
    const int x = r % W;
    r /= W;
    const int y = r % H, b = r / H;
    const size_t base = (((size_t)b * 2 * H + 2 * y) * 2 * W + 2 * x) * C + c, down = (size_t)2 * W * C;
    const bf16x8 p = ld8(dY + base), q = ld8(dY + base + C), u = ld8(dY + base + down), v = ld8(dY + base + down + C);
    bf16x8 o;
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = f2bf(bf2f(p[k]) + bf2f(q[k]) + bf2f(u[k]) + bf2f(v[k]));
    st8(dX + (size_t)i * 8, o);
  }
}
// Y[p] = [A[p] (Ca channels), B[p] (Cb channels)]; the backward splits.
__global__ void k_concat(const __bf16* __restrict__ A, int Ca, const __bf16* __restrict__ B, int Cb, int pixels,
                         __bf16* __restrict__ Y) {
  const int Ga = Ca >> 3, G = (Ca + Cb) >> 3, n = pixels * G;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int p = i / G, k = i - p * G;
    st8(Y + (size_t)i * 8, k < Ga ? ld8(A + ((size_t)p * Ca + 8 * k)) : ld8(B + ((size_t)p * Cb + 8 * (k - Ga))));
  }
}
__global__ void k_split(const __bf16* __restrict__ Y, int Ca, int Cb, int pixels, __bf16* __restrict__ A,
                        __bf16* __restrict__ B) {
  const int Ga = Ca >> 3, G = (Ca + Cb) >> 3, n = pixels * G;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int p = i / G, k = i - p * G;
    const bf16x8 v = ld8(Y + (size_t)i * 8);
    if (k < Ga) st8(A + ((size_t)p * Ca + 8 * k), v);
    else st8(B + ((size_t)p * Cb + 8 * (k - Ga)), v);
  }
}
// G *= (Yact > 0): the gradient through ReLU given its output.
__global__ void k_reluBwd(__bf16* __restrict__ G, const __bf16* __restrict__ Yact, int groups) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < groups; i += gridDim.x * blockDim.x) {
    bf16x8 g = ld8(G + (size_t)i * 8);
    const bf16x8 y = ld8(Yact + (size_t)i * 8);
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (!(bf2f(y[k]) > 0.0f)) g[k] = (__bf16)0.0f;
    st8(G + (size_t)i * 8, g);
  }
}
__global__ void k_add(__bf16* __restrict__ A, const __bf16* __restrict__ B, int groups) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < groups; i += gridDim.x * blockDim.x) {
    bf16x8 a = ld8(A + (size_t)i * 8);
    const bf16x8 b = ld8(B + (size_t)i * 8);
#pragma unroll
    for (int k = 0; k < 8; k++) a[k] = f2bf(bf2f(a[k]) + bf2f(b[k]));
    st8(A + (size_t)i * 8, a);
  }
}

inline int gridFor(long long n, int block = 256) {
  long long b = (n + block - 1) / block;
  return (int)(b < 1 ? 1 : (b > 65535 ? 65535 : b));
}

}  // namespace dn
}  // namespace arbor

using namespace arbor;
using namespace arbor::dn;

namespace {
template <int NSUB>
void launchConv(const __bf16* X, const __bf16* Xc, int Cc, int N, int H, int W, int Cin, const __bf16* Wt,
                const float* bias, int relu, __bf16* Y, int coutTotal, int co0, hipStream_t s) {
  // narrow outputs: 16-row tiles (each weight fragment feeds 4 MFMAs); wide: 8-row tiles
  constexpr int MSUB = NSUB <= 4 ? 4 : 2;
  const long long blocks = (long long)N * ((H + 4 * MSUB - 1) / (4 * MSUB)) * ((W + kTileW - 1) / kTileW);
  // paired taps for 16 -> 16 only: with wider outputs the mixed-tap LDS reads
  // conflict (16 -> 32 took 117 against 66-82 us at 1088x1920)
  if (NSUB == 1 && Cin == 16 && !Xc)
    hipLaunchKernelGGL((k_conv3x3<NSUB, MSUB, 2>), dim3((unsigned)blocks), dim3(256), 0, s, X, Xc, Cc, N, H, W,
                       Cin, Wt, bias, relu, Y, coutTotal, co0);
  else if (Cin == 8 && !Xc)  // four taps per instruction
    hipLaunchKernelGGL((k_conv3x3<NSUB, MSUB, 4>), dim3((unsigned)blocks), dim3(256), 0, s, X, Xc, Cc, N, H, W,
                       Cin, Wt, bias, relu, Y, coutTotal, co0);
  else
    hipLaunchKernelGGL((k_conv3x3<NSUB, MSUB>), dim3((unsigned)blocks), dim3(256), 0, s, X, Xc, Cc, N, H, W, Cin,
                       Wt, bias, relu, Y, coutTotal, co0);
}
int conv(const __bf16* x, const __bf16* xc, int cc, int n, int h, int w, int cin, const __bf16* wt, const float* bias,
         int relu, __bf16* y, int cout, hipStream_t stream) {
  // output channels in launches of at most 128
  for (int co0 = 0; co0 < cout; co0 += 128) {
    const int width = cout - co0 < 128 ? cout - co0 : 128;
    switch (width / 16) {
      case 1: launchConv<1>(x, xc, cc, n, h, w, cin, wt, bias, relu, y, cout, co0, stream); break;
      case 2: launchConv<2>(x, xc, cc, n, h, w, cin, wt, bias, relu, y, cout, co0, stream); break;
      case 3: launchConv<3>(x, xc, cc, n, h, w, cin, wt, bias, relu, y, cout, co0, stream); break;
      case 4: launchConv<4>(x, xc, cc, n, h, w, cin, wt, bias, relu, y, cout, co0, stream); break;
      case 5: launchConv<5>(x, xc, cc, n, h, w, cin, wt, bias, relu, y, cout, co0, stream); break;
      case 6: launchConv<6>(x, xc, cc, n, h, w, cin, wt, bias, relu, y, cout, co0, stream); break;
      case 7: launchConv<7>(x, xc, cc, n, h, w, cin, wt, bias, relu, y, cout, co0, stream); break;
      case 8: launchConv<8>(x, xc, cc, n, h, w, cin, wt, bias, relu, y, cout, co0, stream); break;
    }
  }
  return launch_status();
}
bool dnShapeOk(int n, int h, int w, int c) { return n >= 1 && h >= 1 && w >= 1 && c >= 8 && (c & 7) == 0; }
// element counts of the vectorised kernels stay below 2^31
bool fits32(long long elements) { return elements > 0 && elements < (1LL << 31); }
}  // namespace

extern "C" {

long long arbor_dn_wgrad_partial_floats(int n, int h, int w, int cin, int cout) {
  if (!dnShapeOk(n, h, w, cin) || !dnShapeOk(1, 1, 1, cout)) return ARBOR_ERR_ARG;
  const long long M = (long long)n * h * w;
  const long long chunks = (M + ARBOR_DN_WGRAD_CHUNK - 1) / ARBOR_DN_WGRAD_CHUNK;
  return chunks * (9LL * cin * cout + cout);
}

int arbor_dn_conv3x3(const uint16_t* X, int n, int h, int w, int cin, const uint16_t* Wt, const float* bias, int relu,
                     uint16_t* Y, int cout, hipStream_t stream) {
  ARBOR_REQUIRE(X && Wt && Y);
  if (!dnShapeOk(n, h, w, cin) || cout < 16 || cout % 16 != 0 || cout > ARBOR_DN_MAX_CHANNELS) return ARBOR_ERR_ARG;
  if (cin > ARBOR_DN_MAX_CHANNELS || (long long)n * h * w > (1LL << 31) / 256) return ARBOR_ERR_ARG;
  return conv(reinterpret_cast<const __bf16*>(X), nullptr, 0, n, h, w, cin, reinterpret_cast<const __bf16*>(Wt), bias,
              relu, reinterpret_cast<__bf16*>(Y), cout, stream);
}

int arbor_dn_conv1x1(const uint16_t* X, long long pixels, int cin, const uint16_t* W1, const float* bias, int relu,
                     uint16_t* Y, int cout, hipStream_t stream) {
  ARBOR_REQUIRE(X && W1 && Y);
  if (pixels < 1 || !fits32(pixels * (cin > cout ? cin : cout)) || (cin != 16 && cin != 32) || cout < 16 ||
      cout % 16 != 0 || cout > ARBOR_DN_MAX_CHANNELS)
    return ARBOR_ERR_ARG;
  const long long groups = (pixels + 15) / 16;
  const int blocks = (int)((groups + 3) / 4 < 65535 ? (groups + 3) / 4 : 65535);
  auto x = reinterpret_cast<const __bf16*>(X);
  auto w = reinterpret_cast<const __bf16*>(W1);
  auto y = reinterpret_cast<__bf16*>(Y);
  if (cin == 16)
    hipLaunchKernelGGL(k_conv1x1<1>, dim3(blocks), dim3(256), 0, stream, x, pixels, w, bias, relu, y, cout);
  else
    hipLaunchKernelGGL(k_conv1x1<2>, dim3(blocks), dim3(256), 0, stream, x, pixels, w, bias, relu, y, cout);
  return launch_status();
}

int arbor_dn_conv3x3_cat(const uint16_t* Xc, int cc, const uint16_t* X, int cs, int n, int h, int w, const uint16_t* Wt,
                         const float* bias, int relu, uint16_t* Y, int cout, hipStream_t stream) {
  ARBOR_REQUIRE(Xc && X && Wt && Y);
  const int cin = cc + cs;
  if (!dnShapeOk(n, h, w, cin) || cc < 8 || (cc & 7) || cs < 8 || (cs & 7) || (h & 1) || (w & 1)) return ARBOR_ERR_ARG;
  if (cout < 16 || cout % 16 != 0 || cout > ARBOR_DN_MAX_CHANNELS || cin > ARBOR_DN_MAX_CHANNELS) return ARBOR_ERR_ARG;
  if ((long long)n * h * w > (1LL << 31) / 256) return ARBOR_ERR_ARG;
  return conv(reinterpret_cast<const __bf16*>(X), reinterpret_cast<const __bf16*>(Xc), cc, n, h, w, cin,
              reinterpret_cast<const __bf16*>(Wt), bias, relu, reinterpret_cast<__bf16*>(Y), cout, stream);
}

int arbor_dn_conv3x3_wgrad(const uint16_t* X, const uint16_t* dY, int n, int h, int w, int cin, int cout,
                           float* partial, long long partial_floats, float* dW, float* dB, hipStream_t stream) {
  ARBOR_REQUIRE(X && dY && partial && dW && dB);
  const long long need = arbor_dn_wgrad_partial_floats(n, h, w, cin, cout);
  if (need < 0 || partial_floats < need) return ARBOR_ERR_ARG;
  const long long M = (long long)n * h * w;
  const int chunks = (int)((M + ARBOR_DN_WGRAD_CHUNK - 1) / ARBOR_DN_WGRAD_CHUNK);
  if (chunks > 65535) return ARBOR_ERR_ARG;
  const int tiles = ((cout + kWgTile - 1) / kWgTile) * ((cin + kWgTile - 1) / kWgTile);
  float* partialB = partial + (size_t)chunks * 9 * cin * cout;
  hipLaunchKernelGGL(k_conv3x3_wgrad, dim3(chunks, 9, tiles), dim3(256), 0, stream, reinterpret_cast<const __bf16*>(X),
                     reinterpret_cast<const __bf16*>(dY), n, h, w, cin, cout, ARBOR_DN_WGRAD_CHUNK, partial, partialB);
  const long long nw = 9LL * cin * cout;
  hipLaunchKernelGGL(k_sumParts, dim3(gridFor(nw)), dim3(256), 0, stream, partial, chunks, nw, dW);
  hipLaunchKernelGGL(k_sumParts, dim3(gridFor(cout)), dim3(256), 0, stream, partialB, chunks, (long long)cout, dB);
  return launch_status();
}

int arbor_dn_flip_weights(const uint16_t* Wt, int cin, int cout, uint16_t* Wflip, hipStream_t stream) {
  ARBOR_REQUIRE(Wt && Wflip);
  if (cin < 1 || cout < 1) return ARBOR_ERR_ARG;
  hipLaunchKernelGGL(k_flipWeights, dim3(gridFor(9LL * cin * cout)), dim3(256), 0, stream,
                     reinterpret_cast<const __bf16*>(Wt), cin, cout, reinterpret_cast<__bf16*>(Wflip));
  return launch_status();
}

int arbor_dn_pool2(const uint16_t* X, int n, int h, int w, int c, uint16_t* Y, hipStream_t stream) {
  ARBOR_REQUIRE(X && Y);
  if (!dnShapeOk(n, h, w, c) || h < 2 || w < 2 || (h & 1) || (w & 1) || !fits32((long long)n * h * w * c))
    return ARBOR_ERR_ARG;
  hipLaunchKernelGGL(k_pool2, dim3(gridFor((long long)n * h * w * c / 32)), dim3(256), 0, stream,
                     reinterpret_cast<const __bf16*>(X), n, h, w, c, reinterpret_cast<__bf16*>(Y));
  return launch_status();
}
int arbor_dn_pool2_bwd(const uint16_t* dY, int n, int h, int w, int c, uint16_t* dX, hipStream_t stream) {
  ARBOR_REQUIRE(dY && dX);
  if (!dnShapeOk(n, h, w, c) || h < 2 || w < 2 || (h & 1) || (w & 1) || !fits32((long long)n * h * w * c))
    return ARBOR_ERR_ARG;
  hipLaunchKernelGGL(k_pool2Bwd, dim3(gridFor((long long)n * h * w * c / 8)), dim3(256), 0, stream,
                     reinterpret_cast<const __bf16*>(dY), n, h, w, c, reinterpret_cast<__bf16*>(dX));
  return launch_status();
}
int arbor_dn_up2(const uint16_t* X, int n, int h, int w, int c, uint16_t* Y, hipStream_t stream) {
  ARBOR_REQUIRE(X && Y);
  if (!dnShapeOk(n, h, w, c) || !fits32((long long)n * h * w * c * 4)) return ARBOR_ERR_ARG;
  hipLaunchKernelGGL(k_up2, dim3(gridFor((long long)n * h * w * c / 2)), dim3(256), 0, stream,
                     reinterpret_cast<const __bf16*>(X), n, h, w, c, reinterpret_cast<__bf16*>(Y));
  return launch_status();
}
int arbor_dn_up2_bwd(const uint16_t* dY, int n, int h, int w, int c, uint16_t* dX, hipStream_t stream) {
  ARBOR_REQUIRE(dY && dX);
  if (!dnShapeOk(n, h, w, c) || !fits32((long long)n * h * w * c * 4)) return ARBOR_ERR_ARG;
  hipLaunchKernelGGL(k_up2Bwd, dim3(gridFor((long long)n * h * w * c / 8)), dim3(256), 0, stream,
                     reinterpret_cast<const __bf16*>(dY), n, h, w, c, reinterpret_cast<__bf16*>(dX));
  return launch_status();
}
int arbor_dn_concat(const uint16_t* A, int ca, const uint16_t* B, int cb, long long pixels, uint16_t* Y,
                    hipStream_t stream) {
  ARBOR_REQUIRE(A && B && Y);
  if (ca < 8 || cb < 8 || (ca & 7) || (cb & 7) || !fits32(pixels * (ca + cb))) return ARBOR_ERR_ARG;
  hipLaunchKernelGGL(k_concat, dim3(gridFor(pixels * (ca + cb) / 8)), dim3(256), 0, stream,
                     reinterpret_cast<const __bf16*>(A), ca, reinterpret_cast<const __bf16*>(B), cb, (int)pixels,
                     reinterpret_cast<__bf16*>(Y));
  return launch_status();
}
int arbor_dn_split(const uint16_t* Y, int ca, int cb, long long pixels, uint16_t* A, uint16_t* B, hipStream_t stream) {
  ARBOR_REQUIRE(Y && A && B);
  if (ca < 8 || cb < 8 || (ca & 7) || (cb & 7) || !fits32(pixels * (ca + cb))) return ARBOR_ERR_ARG;
  hipLaunchKernelGGL(k_split, dim3(gridFor(pixels * (ca + cb) / 8)), dim3(256), 0, stream,
                     reinterpret_cast<const __bf16*>(Y), ca, cb, (int)pixels, reinterpret_cast<__bf16*>(A),
                     reinterpret_cast<__bf16*>(B));
  return launch_status();
}
int arbor_dn_relu_bwd(uint16_t* G, const uint16_t* Yact, long long count, hipStream_t stream) {
  ARBOR_REQUIRE(G && Yact);
  if (!fits32(count) || (count & 7)) return ARBOR_ERR_ARG;
  hipLaunchKernelGGL(k_reluBwd, dim3(gridFor(count / 8)), dim3(256), 0, stream, reinterpret_cast<__bf16*>(G),
                     reinterpret_cast<const __bf16*>(Yact), (int)(count / 8));
  return launch_status();
}
int arbor_dn_add(uint16_t* A, const uint16_t* B, long long count, hipStream_t stream) {
  ARBOR_REQUIRE(A && B);
  if (!fits32(count) || (count & 7)) return ARBOR_ERR_ARG;
  hipLaunchKernelGGL(k_add, dim3(gridFor(count / 8)), dim3(256), 0, stream, reinterpret_cast<__bf16*>(A),
                     reinterpret_cast<const __bf16*>(B), (int)(count / 8));
  return launch_status();
}

}  // extern "C"
