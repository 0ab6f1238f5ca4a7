// Process-wide launch tuning of the tile kernels, set by the autotuner or the
// caller through arbor_tune_set (arbor.h). Every setting selects between
// implementations that produce identical results; none changes the work done.
#pragma once

namespace arbor {

struct Tuning {
  int trace_variant = 5;      // index into the trace kernel table (frame.hip); the
                              // two values past the table select the wave-scheduled
                              // kernel with 4-wave and 1-wave blocks
  int grid_cap = 1 << 20;     // maximum blocks of a tile kernel launch (8192x8192 has 2^20 tiles)
  int sched_tiles = 16;       // wave-scheduled kernel: 8x8 tiles per block pool
  int sched_leaf_min = 24;    // lanes that must be ready before a LEAF phase runs
  int sched_newton_min = 24;  // lanes that must be ready before a NEWTON phase runs
  int batch_variant = 9;      // index into the batched trace+shade kernel table (batch.hip)
  int vis_variant = 6;        // index into the sun-visibility kernel table (frame.hip)
  int accel_variant = 26;      // index into the SAH-walk trace kernel table (frame.hip)
  int accel_vis_variant = 23;  // index into the SAH-walk sun-visibility kernel table (frame.hip)
};
int batchVariantCount();

inline Tuning& tuning() {
  static Tuning t;
  return t;
}

}  // namespace arbor
