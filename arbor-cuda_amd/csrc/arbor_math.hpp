// Float sine/cosine for the bounded arguments of the render path's hot loops,
// and division by a divisor whose correctly rounded reciprocal is known.
//
// sincosBounded: Cody-Waite reduction by pi/2 split into three floats (the
// leading one exact in k*P1 through fma), then the Taylor series of sin and
// cos on [-pi/4, pi/4] through x^9 and x^10 (first omitted terms below 2.4e-9
// and 1.2e-10 relative), evaluated by Horner's rule with fma. Within 1.5 ulp
// of the exact value for |x| up to 2e4 (tests/test_math.py); the callers'
// arguments stay below ~10^4 (leaf outlines |x| < 4, bark |x| < 10^3, wind
// phases < 10^4).
//
// divKnown: a / b given rb = RN(1/b): q = RN(a*rb), r = a - b*q (exact with
// fma), RN(q + r*rb) is the correctly rounded quotient for normal a, b and
// quotients away from overflow and underflow (Markstein's correction step).
//
// The header compiles for the device (HIP) and for the host (a C++ compiler),
// so the same code is checked on the CPU against double precision.
#pragma once
#include <math.h>

#if defined(__HIPCC__)
#define AM __host__ __device__ __forceinline__
#else
#define AM static inline
#endif

namespace arbor {

struct SinCos {
  float s, c;
};

AM SinCos sincosBounded(float x) {
  const float k = rintf(x * 0.636619747f);  // x * 2/pi, nearest integer
  float r = fmaf(-k, 1.57079637f, x);       // pi/2 = P1 + P2 + P3
  r = fmaf(-k, -4.37113883e-08f, r);
  r = fmaf(-k, -1.71512451e-15f, r);
  const float r2 = r * r;
  float ps = fmaf(r2, 2.75573188e-06f, -1.98412701e-04f);  // 1/9!, -1/7!
  ps = fmaf(r2, ps, 8.33333377e-03f);                       // 1/5!
  ps = fmaf(r2, ps, -0.166666672f);                         // -1/3!
  const float s = fmaf(r * r2, ps, r);
  float pc = fmaf(r2, -2.75573200e-07f, 2.48015876e-05f);   // -1/10!, 1/8!
  pc = fmaf(r2, pc, -1.38888892e-03f);                      // -1/6!
  pc = fmaf(r2, pc, 4.16666679e-02f);                       // 1/4!
  pc = fmaf(r2, pc, -0.5f);                                 // -1/2!
  const float c = fmaf(r2, pc, 1.0f);
  const int q = ((int)k) & 3;
  SinCos out;
  out.s = q == 0 ? s : (q == 1 ? c : (q == 2 ? -s : -c));
  out.c = q == 0 ? c : (q == 1 ? -s : (q == 2 ? -c : s));
  return out;
}
AM float sinBounded(float x) { return sincosBounded(x).s; }
AM float cosBounded(float x) { return sincosBounded(x).c; }

AM float divKnown(float a, float b, float rb) {
  const float q = a * rb;
  const float r = fmaf(-b, q, a);
  return r == 0.0f ? q : fmaf(r, rb, q);  // r == 0: q is exact (and keeps the sign of a zero quotient)
}

}  // namespace arbor
