// Operands of a 1x1 convolution over 16*J input channels (J = 1 or 2) on the
// matrix cores, shared by the layered kernel and the lean network's fused
// kernel so both issue the same instruction on the same operands.
//
// One instruction yields 16 output channels x 16 pixels. The weights are
// the A operand (row = output channel), the pixels the B operand (column =
// pixel); lane (row, g), g = lane / 16, holds:
//  * J = 1 (v_mfma_f32_16x16x16_bf16): channels 4g..4g+3;
//  * J = 2 (v_mfma_f32_16x16x32_bf16): channels 4g..4g+3 and 16+4g..16+4g+3
//    in its 8 K slots.
// That is where a 3x3 convolution's output lies in the lanes (its D operand:
// channels 4g..4g+3 of each 16-channel group for pixel row), so a fused
// kernel feeds its activations to this layer without moving them. D: lane
// (row, g) gets output channels 4g..4g+3 of pixel row.
#pragma once
#include <hip/hip_runtime.h>

namespace arbor {
namespace dn1 {

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

// Four values rounded to bf16 (round to nearest even), converted in pairs
// (v_cvt_pk_bf16_f32) and joined as two 32-bit words, so no element
// shuffles follow.
__device__ __forceinline__ bf16x4 toBf16(f32x4 v) {
  const bf16x2 lo = __builtin_convertvector(f32x2{v[0], v[1]}, bf16x2);
  const bf16x2 hi = __builtin_convertvector(f32x2{v[2], v[3]}, bf16x2);
  return __builtin_bit_cast(bf16x4, u32x2{__builtin_bit_cast(unsigned, lo), __builtin_bit_cast(unsigned, hi)});
}
// ReLU of four finite values, then rounded to bf16: the sign bit decides (an
// integer maximum of the bit pattern with 0, so -0 becomes +0), one
// instruction per value.
__device__ __forceinline__ bf16x4 reluToBf16(f32x4 v) {
  f32x4 r;
#pragma unroll
  for (int i = 0; i < 4; i++) r[i] = __int_as_float(max(__float_as_int(v[i]), 0));
  return toBf16(r);
}

template <int J>
struct Op1x1;

template <>
struct Op1x1<1> {
  typedef bf16x4 T;
  // lane g's fragment of a 16-channel row at p
  __device__ static __forceinline__ T load(const __bf16* p, int g) { return *reinterpret_cast<const bf16x4*>(p + 4 * g); }
  // the fragment from a 3x3 convolution's output groups (bf16)
  __device__ static __forceinline__ T fromGroups(const bf16x4* h) { return h[0]; }
  __device__ static __forceinline__ f32x4 mma(T a, T b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(s16x4, a), __builtin_bit_cast(s16x4, b), c, 0,
                                                     0, 0);
  }
};

template <>
struct Op1x1<2> {
  typedef bf16x8 T;
  __device__ static __forceinline__ T load(const __bf16* p, int g) {
    const bf16x4 lo = *reinterpret_cast<const bf16x4*>(p + 4 * g), hi = *reinterpret_cast<const bf16x4*>(p + 16 + 4 * g);
    return T{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
  }
  __device__ static __forceinline__ T fromGroups(const bf16x4* h) {
    return T{h[0][0], h[0][1], h[0][2], h[0][3], h[1][0], h[1][1], h[1][2], h[1][3]};
  }
  __device__ static __forceinline__ f32x4 mma(T a, T b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};

}  // namespace dn1
}  // namespace arbor
