// Frame passes the C++ engine chains into its five launches per frame
// (the reference engine's five dispatches, src/engine.js:48-50):
//   1. framePrologue: updateCamera (and the per-sample camera slots) and the
//      tile orders of the trace and of the sun-visibility pass, in one
//      cooperative launch;
//   2. traceOrdered: tracePixels over the schedule's current order;
//   3. visibilityOrdered: the sun-visibility pass over its schedule's order
//      (it returns at once with visibility off);
//   4. arbor_shadePixelsVis;
//   5. filterResolve: filterLighting and resolvePixels in one kernel.
// Internal to the library (C++ linkage); every pass gives the bits of the
// public entry points it replaces.
#pragma once
#include <hip/hip_runtime.h>

#include "arbor_denoise.h"

namespace arbor {

// I may be null (no camera update); sample < 0 leaves C[9], C[10] and C[22];
// light < 0 leaves C[22]; visSchedule may be null.
int framePrologue(float* C, const float* I, float delta, int sample, int light, int width, int height,
                  void* traceSchedule, void* visSchedule, hipStream_t stream);
// accel null: the heap walk (arbor_tracePixelsScheduled's kernels).
int traceOrdered(const float* P, const float* Nodes, const unsigned* Order, const void* accel, const float* C,
                 float* Hit, int width, int height, void* schedule, hipStream_t stream);
int visibilityOrdered(const float* P, const float* Nodes, const unsigned* Order, const void* accel, const float* C,
                      const float* Hit, float* Vis, int width, int height, void* schedule, hipStream_t stream);
// arbor_shadePixelsVis that also writes the shading normal (Normal as
// arbor_shadePixelsAOV writes it); Linear and Guide are the same bits.
int shadeVisNormal(const float* P, const float* Nodes, const unsigned* Order, const float* C, const float* Hit,
                   const float* Vis, float* Linear, float* Guide, float* Normal, int width, int height,
                   hipStream_t stream);
int filterResolve(const float* Linear, const float* Guide, const float* Hit, float* Filtered, float* History,
                  unsigned* Pixels, const float* C, int width, int height, hipStream_t stream);
// The learned denoiser in place of filterResolve: the network's output into
// Filtered, then resolvePixels (one launch for the 1-level 16-channel shape).
int denoiseResolve(arbor_denoiser* d, const float* Linear, const float* Guide, const float* Normal, const float* Hit,
                   float* Filtered, float* History, unsigned* Pixels, const float* C, int w, int h,
                   hipStream_t stream);
// Whether d is the lean network (compact features in, one pass).
bool denoiserLean(const arbor_denoiser* d);
// The lean network's one-pass kernel over compact features (w x h, as
// arbor_shadePixelsFeatures writes them) into Filtered, with resolvePixels.
int denoiseFeaturesResolve(arbor_denoiser* d, const void* Feat, float* Filtered, float* History, unsigned* Pixels,
                           const float* C, int w, int h, hipStream_t stream);
// arbor_shadePixelsVis writing the compact features (as
// arbor_shadePixelsFeatures does) in place of Linear and Guide.
int shadeVisFeatures(const float* P, const float* Nodes, const unsigned* Order, const float* C, const float* Hit,
                     const float* Vis, void* Feat, int width, int height, hipStream_t stream);
// Compute units of the current device (cached per device).
int deviceCUs();

}  // namespace arbor
