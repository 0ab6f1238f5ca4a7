// Host-side argument checks and launch status shared by the C-ABI wrappers.
#pragma once
#include <hip/hip_runtime.h>

#include "arbor.h"

#define ARBOR_REQUIRE(cond) \
  do {                      \
    if (!(cond)) return ARBOR_ERR_NULL; \
  } while (0)

namespace arbor {

inline bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }
inline bool dims_ok(int w, int h) { return w >= 1 && h >= 1 && w <= ARBOR_MAX_DIM && h <= ARBOR_MAX_DIM; }
// Status of the launch just issued on this thread (hipGetLastError clears it).
inline int launch_status() { return hipGetLastError() == hipSuccess ? ARBOR_OK : ARBOR_ERR_LAUNCH; }

}  // namespace arbor
