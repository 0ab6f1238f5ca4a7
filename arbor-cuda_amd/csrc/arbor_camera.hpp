// Camera state of the render path (Arbor.cu:179-182): the first-person pose,
// the view presets, initCamera and updateCamera as single-thread functions
// over the 64-float C buffer, shared by the stand-alone camera kernels and
// the engine's fused frame prologue. Transcendentals are evaluated in double
// and rounded once (sin_cr, cos_cr), as for the rest of the one-time build.
#pragma once
#include "arbor_device.hpp"

namespace arbor {

// Arbor.cu:179.
__device__ inline void fpsPose(float* C, v3 eye, float yaw, float pitch) {
  C[16] = yaw; C[17] = pitch; C[27] = yaw; C[28] = pitch;
  st3(C, 0, eye);
  float cp = cos_cr(pitch);
  v3 f = unit(mk(sin_cr(yaw) * cp, sin_cr(pitch), cos_cr(yaw) * cp));
  st3(C, 4, eye + f);
  C[3] = 1.0f; C[18] = 1.0f;
}
// Arbor.cu:180.
__device__ inline void viewPreset(float* C, int preset) {
  if (preset == 1) fpsPose(C, mk(2.0f, 1.72f, 18.0f), kPi, 0.0f);
  if (preset == 2) fpsPose(C, mk(-1.4f, 1.72f, 7.0f), kPi * 0.94f, 0.02f);
  if (preset == 3) fpsPose(C, mk(8.0f, 1.72f, -3.0f), kPi * 0.55f, -0.02f);
  if (preset == 4) fpsPose(C, mk(-11.0f, 1.72f, -7.0f), 0.65f, 0.03f);
  if (preset == 5) fpsPose(C, mk(0.0f, 12.0f, 72.0f), kPi, -0.12f);
  C[9] = 0.0f; C[10] = 1.0f;
}
// initCamera (Arbor.cu:181).
__device__ inline void initCameraState(float* C, int seed) {
  for (int i = 0; i < 64; i++) C[i] = 0.0f;
  C[12] = -0.45f; C[13] = 0.63f; C[14] = 0.90f; C[20] = 1.18f; C[22] = 1.0f; C[23] = 1.0f;
  C[31] = (float)seed; C[44] = 1.0f;
  fpsPose(C, mk(2.0f, 1.72f, 18.0f), kPi, 0.0f);
}
// updateCamera (Arbor.cu:182).
__device__ inline void updateCameraState(float* C, const float* I, float delta, int width, int height) {
  C[10] = 0.0f;
  C[24] += 1.0f;
  float dt = clampf(delta, 0.0f, 0.1f);
  v3 old = ld3(C, 0);
  if (I[5] > 0.0f) viewPreset(C, (int)I[5]);
  if (I[7] > 0.0f) { C[11] = C[11] > 0.0f ? 0.0f : 0.60f; C[10] = 1.0f; }
  if (I[9] > 0.0f) { C[15] = (float)(((int)C[15] + 1) % 4); C[10] = 1.0f; }
  if (I[10] > 0.0f) { C[22] = 1.0f - C[22]; C[10] = 1.0f; }
  if (I[8] != 0.0f) { C[12] += I[8] * dt * 0.6f; C[10] = 1.0f; }
  if (I[12] != 0.0f) { C[20] = clampf(C[20] + I[12] * dt, 0.25f, 3.0f); C[10] = 1.0f; }
  if (C[11] > 0.0f) { C[8] += dt; C[10] = 1.0f; }
  C[16] -= I[0] * 0.00225f;
  C[17] = clampf(C[17] - I[1] * 0.0020f, -1.45f, 1.45f);
  float yaw = C[16], cp = cos_cr(C[17]);
  v3 look = unit(mk(sin_cr(yaw) * cp, sin_cr(C[17]), cos_cr(yaw) * cp));
  v3 walkF = unit(mk(sin_cr(yaw), 0.0f, cos_cr(yaw)));
  v3 right = unit(mk(walkF.z, 0.0f, -walkF.x));
  float speed = I[16] > 0.5f ? 10.5f : 4.2f;
  v3 move = walkF * I[15] + right * I[13];
  if (dot(move, move) > 0.0001f) move = unit(move) * speed * dt;
  v3 eye = ld3(C, 0) + move;
  eye.y = 1.72f;
  st3(C, 0, eye);
  st3(C, 4, eye + look);
  C[27] = C[16]; C[28] = C[17]; C[3] = 1.0f; C[18] = 1.0f;
  if (length(eye - old) > 0.000001f || I[0] != 0.0f || I[1] != 0.0f) C[10] = 1.0f;
  if (C[32] != (float)width || C[33] != (float)height) { C[10] = 1.0f; C[32] = (float)width; C[33] = (float)height; }
  C[9] = C[10] > 0.5f ? 0.0f : fminf(255.0f, C[9] + 1.0f);
}

}  // namespace arbor
