// Process-wide launch tuning of the tile kernels, set by the autotuner or the
// caller through arbor_tune_set (arbor.h). Every setting selects between
// implementations that produce identical results; none changes the work done.
#pragma once

namespace arbor {

struct Tuning {
  int trace_variant = 1;  // index into the trace kernel table (frame.hip)
  int grid_cap = 65535;   // maximum blocks of a tile kernel launch
};

inline Tuning& tuning() {
  static Tuning t;
  return t;
}

}  // namespace arbor
