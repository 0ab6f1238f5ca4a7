// Host build of the device math in arbor-cuda_amd/csrc/arbor_math.hpp for the
// CPU accuracy checks (tests/test_math.py): same source, same fma/rint steps.
#include "arbor_math.hpp"

extern "C" {
void mc_sincos(const float* x, float* s, float* c, long n) {
  for (long i = 0; i < n; i++) {
    arbor::SinCos v = arbor::sincosBounded(x[i]);
    s[i] = v.s;
    c[i] = v.c;
  }
}
void mc_div_known(const float* a, const float* b, const float* rb, float* q, long n) {
  for (long i = 0; i < n; i++) q[i] = arbor::divKnown(a[i], b[i], rb[i]);
}
}
