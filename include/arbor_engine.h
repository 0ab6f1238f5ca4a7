/* ARBOR host engine over the C-ABI kernels: the generate / resize / frame /
 * inspect contract of the reference host (src/engine.js:6-57) in C++, with
 * device buffers owned by the engine. All calls return ARBOR_OK or a negative
 * status (arbor.h). */
#ifndef ARBOR_ENGINE_H
#define ARBOR_ENGINE_H

#include "arbor.h"
#include "arbor_denoise.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct arbor_engine arbor_engine;

/* engine.js:36 rounding: width up to a multiple of 64 (min 320), height up to
 * a multiple of 8 (min 192). */
void arbor_engine_round_size(int* width, int* height);

/* Engine.init + generate (engine.js:6-24,27-34). round_size != 0 applies the
 * engine.js resize rounding; stream NULL makes the engine create its own. */
int arbor_engine_create(arbor_engine** out, int width, int height, int seed, int round_size, hipStream_t stream);
int arbor_engine_destroy(arbor_engine* e);
/* engine.js:27: seed must be an integer in [0, 1000000]. */
int arbor_engine_generate(arbor_engine* e, int seed);
/* engine.js:35-43. */
int arbor_engine_resize(arbor_engine* e, int width, int height, int round_size);
/* initCamera on the engine's C buffer (the state generate() leaves). */
int arbor_engine_reset_camera(arbor_engine* e);
/* Host copy of the 32-float input slot array used by the next frame. */
int arbor_engine_set_input(arbor_engine* e, const float* input32);
/* Per-stage HIP-event timing of every frame (synchronises each frame). */
int arbor_engine_set_timing(arbor_engine* e, int on);
/* Trace and sun visibility over the SAH hierarchy built at generate
 * (arbor_tracePixelsAccel / arbor_sunVisibilityAccel; on by default) or over
 * the heap BVH (arbor_tracePixelsScheduled / arbor_sunVisibilityScheduled). */
int arbor_engine_set_accel(arbor_engine* e, int on);
/* Denoise with a learned network in place of filterLighting (no counterpart
 * in the reference, which filters only, Arbor.cu:292-302). With d not NULL,
 * every later frame whose size fits d's slot (arbor_denoiser_slot) shades
 * with the normal as well (a Normal buffer, 16 B a pixel, is allocated),
 * runs the network on slot 0 (input features, forward, output into Filtered)
 * and resolves (arbor_denoiser_denoise, then resolvePixels; for a 1-level
 * 16-channel network one kernel does both, and the frame stays five
 * launches); a frame larger than the slot falls back to filterLighting,
 * as does every frame with d NULL (the default). inspect() names the last
 * frame's denoiser. The engine does not own d: the caller loads its weights,
 * keeps it alive while attached and does not use slot 0 of it meanwhile
 * from another stream. */
int arbor_engine_set_denoiser(arbor_engine* e, arbor_denoiser* d);
/* engine.js:45-53: updateCamera, tracePixels, shadePixels, filterLighting,
 * resolvePixels, issued as five kernel launches like the reference's five
 * dispatches (engine.js:48-50, source.test.mjs:6): (1) updateCamera with the
 * tile orders of the trace and of the sun-visibility pass (one cooperative
 * launch), (2) tracePixels, (3) the sun-visibility pass, (4) shadePixels from
 * it, (5) filterLighting and resolvePixels in one kernel. device_input, if
 * not NULL, is a device array of 32 floats used instead of the host input
 * (otherwise the host input is copied to the device first). */
int arbor_engine_frame(arbor_engine* e, float dt, const float* device_input);
/* Last timed frame: {prologue (camera and tile orders), trace, shading
 * (visibility pass and shadePixels), filter + resolve (or the network and
 * resolve), 0, whole frame} in ms,
 * measured with HIP events on the engine's stream. */
int arbor_engine_stage_ms(const arbor_engine* e, float* out6);
int arbor_engine_sync(arbor_engine* e);
/* Device pointer of a named buffer: B P Keys Order Nodes C I Hit Linear Guide
 * Filtered History Pixels Schedule VisSchedule Vis Accel Normal; NULL for an
 * unknown name (and for Normal without a denoiser attached). */
void* arbor_engine_buffer(arbor_engine* e, const char* name);
hipStream_t arbor_engine_stream(arbor_engine* e);
int arbor_engine_size(const arbor_engine* e, int* width, int* height);
/* engine.js:55 inspect(): JSON into out (cap bytes). */
int arbor_engine_inspect(const arbor_engine* e, char* out, int cap);
/* Seeded walk-path input for frame `frame` (32 floats). */
int arbor_walk_input(int seed, int frame, float* input32);
/* The reference harness's render mode (native_harness.cpp:60-63): camera to
 * view preset `view` (0..5, one updateCamera of 1/60 s with input[5] = view),
 * sun visibility C[22] = light, then `samples` (1..512) jittered samples
 * C[9] = i, C[10] = (i == 0), each traced, shaded, filtered and accumulated by
 * resolvePixels (History holds their mean, Pixels the tone-mapped image). */
int arbor_engine_render(arbor_engine* e, int view, int samples, int light);
/* One frame of `samples` (1..512) jittered samples: updateCamera once, then per
 * sample C[9] = i, C[10] = (i == 0), tracePixels, shadePixels, filterLighting,
 * resolvePixels (History: the frame's sample mean), five launches per sample
 * as in arbor_engine_frame. The multi-sample frame of the 4-sample,
 * sun-visibility configuration. */
int arbor_engine_frame_samples(arbor_engine* e, float dt, const float* device_input, int samples);
/* Input recording (no counterpart in the reference, whose inputs come from
 * the browser each frame): with path non-NULL every later frame's dt, sample
 * count and 32 input floats are appended to the file ("ARBORIN1", version,
 * width, height, seed, then 136 bytes per frame); a frame given a device
 * input copies it back first, which waits for the stream. NULL stops. */
int arbor_engine_record(arbor_engine* e, const char* path);
/* Replays a recording: frames first .. first+count-1 (count <= 0: to the end)
 * through arbor_engine_frame / arbor_engine_frame_samples as recorded, with
 * the recorded inputs. *played (may be NULL) = frames run. A replay on an
 * engine of the recording's size and seed, from the same state, reproduces
 * its Pixels bit for bit. ARBOR_ERR_ARG for a file that is not a recording. */
int arbor_engine_replay(arbor_engine* e, const char* path, int first, int count, int* played);
/* native_harness.cpp:53 save(): binary PPM (P6) of the engine's Pixels, RGB
 * from the low three bytes of each packed pixel. */
int arbor_engine_save_ppm(arbor_engine* e, const char* path);
/* The same PPM from host pixels (no GPU). */
int arbor_write_ppm(const char* path, const unsigned* pixels, int width, int height);

#ifdef __cplusplus
}
#endif
#endif /* ARBOR_ENGINE_H */
