// C++ host engine: the generate / resize / frame contract of the reference's
// host (src/engine.js:27-53) over the C-ABI kernels, with owned device buffers,
// per-stage HIP-event timing and a seeded walk-path input generator.
#include <stdio.h>
#include <string.h>

#include <string>

#include <math.h>

#include "arbor_batch.h"
#include "arbor_denoise.h"
#include "arbor_engine.h"
#include "engine_passes.hpp"
#include "launch.hpp"

namespace {

struct Buffers {
  float *B = nullptr, *P = nullptr, *Nodes = nullptr, *C = nullptr, *I = nullptr;
  unsigned *Keys = nullptr, *Order = nullptr;
  float *Hit = nullptr, *Linear = nullptr, *Guide = nullptr, *Filtered = nullptr, *History = nullptr;
  unsigned* Pixels = nullptr;
  void* Schedule = nullptr;     // tile schedule of the trace (arbor_tracePixelsScheduled)
  void* VisSchedule = nullptr;  // tile schedule of the sun-visibility pass
  float* Vis = nullptr;         // sun visibility per pixel
  void* Accel = nullptr;        // SAH trace hierarchy (arbor_buildTraceAccel)
  float* Normal = nullptr;      // shading normal, the denoiser's input, or the lean network's compact features
                                // (only with a denoiser attached)
};

template <class T>
int alloc(T** p, size_t bytes) {
  return hipMalloc(reinterpret_cast<void**>(p), bytes) == hipSuccess ? ARBOR_OK : ARBOR_ERR_DEVICE;
}
template <class T>
void release(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

}  // namespace

struct arbor_engine {
  Buffers b;
  hipStream_t stream = nullptr;
  bool owns_stream = false;
  int width = 0, height = 0, seed = -1;
  long frames = 0;
  bool generated = false;
  hipEvent_t ev[6] = {};
  bool timing = false;
  float stage_ms[6] = {0, 0, 0, 0, 0, 0};  // prologue, trace, shade, filter + resolve, 0, whole frame
  float input[ARBOR_I_FLOATS] = {};
  FILE* rec = nullptr;  // input recording (arbor_engine_record)
  bool accel = true;    // trace and sun visibility over the SAH hierarchy (arbor_engine_set_accel)
  arbor_denoiser* dn = nullptr;  // learned denoiser in place of filterLighting (arbor_engine_set_denoiser)
  bool denoised = false;         // the last frame's denoiser: the network (true) or filterLighting
};

namespace {

size_t pixel_bytes(int w, int h) { return (size_t)w * h * 16; }

void free_pixels(arbor_engine* e) {
  release(e->b.Hit); release(e->b.Linear); release(e->b.Guide); release(e->b.Filtered);
  release(e->b.History); release(e->b.Pixels); release(e->b.Schedule); release(e->b.VisSchedule); release(e->b.Vis);
  release(e->b.Normal);
}

// Whether this frame denoises with the attached network: only when the
// frame fits one of its slots; otherwise filterLighting.
bool useNetwork(const arbor_engine* e) {
  if (!e->dn || !e->b.Normal) return false;
  int n = 0, h = 0, w = 0;
  if (arbor_denoiser_slot(e->dn, &n, &h, &w) != ARBOR_OK) return false;
  return n >= 1 && e->width <= w && e->height <= h;
}

}  // namespace

// Input recording: "ARBORIN1", then int32 version, width, height, seed; then
// per frame float dt, int32 samples, float input[32] (little-endian, 136 B).
constexpr char kRecMagic[8] = {'A', 'R', 'B', 'O', 'R', 'I', 'N', '1'};
constexpr int kRecVersion = 1;
struct RecFrame {
  float dt;
  int32_t samples;
  float input[ARBOR_I_FLOATS];
};
static_assert(sizeof(RecFrame) == 136, "recorded frame");

// Appends the frame about to run to the recording (a device input is copied
// back first, which waits for the stream).
int recordFrame(arbor_engine* e, float dt, const float* device_input, int samples) {
  if (!e->rec) return ARBOR_OK;
  RecFrame r;
  r.dt = dt;
  r.samples = samples;
  if (device_input) {
    if (hipMemcpyAsync(r.input, device_input, sizeof(r.input), hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
        hipStreamSynchronize(e->stream) != hipSuccess)
      return ARBOR_ERR_DEVICE;
  } else {
    memcpy(r.input, e->input, sizeof(r.input));
  }
  if (fwrite(&r, sizeof(r), 1, e->rec) != 1) return ARBOR_ERR_ARG;
  return ARBOR_OK;
}

namespace {

// One sample of the engine's frame in five launches (the reference engine's
// five dispatches, src/engine.js:48-50):
//   1. the prologue: updateCamera when I is not null (delta seconds), the
//      per-sample slots when sample >= 0 (light >= 0 also sets C[22]), and
//      the tile orders of the trace and of the sun-visibility pass;
//   2. tracePixels over the SAH hierarchy (or the heap) in the trace order;
//   3. the sun-visibility pass in its order (it returns at once when C[22]
//      is off);
//   4. shadePixels from the visibility pass;
//   5. filterLighting and resolvePixels as one kernel.
// Every buffer gets the bits the separate entry points give. With `timed`,
// HIP events on the engine's stream bracket the prologue, the trace, the
// shading (3-4) and the filter + resolve.
// With a denoiser attached and a frame that fits its slot, shading also
// writes the normal and step 5 becomes the network's input features, its
// forward pass and output into Filtered (slot 0), then resolvePixels: one
// launch for a 1-level 16-channel network, so the frame stays five. With the
// lean network, shading writes the compact features (in the Normal buffer)
// in place of Linear and Guide, which that frame leaves as they were, and
// step 5 is the network's one-pass kernel from them with the resolve.
int frameLaunches(arbor_engine* e, const float* I, float delta, int sample, int light, bool timed) {
  const int w = e->width, h = e->height;
  const void* accel = e->accel ? e->b.Accel : nullptr;
  int rc;
  if (timed) (void)hipEventRecord(e->ev[0], e->stream);
  if ((rc = arbor::framePrologue(e->b.C, I, delta, sample, light, w, h, e->b.Schedule, e->b.VisSchedule, e->stream)) !=
      ARBOR_OK)
    return rc;
  if (timed) (void)hipEventRecord(e->ev[1], e->stream);
  if ((rc = arbor::traceOrdered(e->b.P, e->b.Nodes, e->b.Order, accel, e->b.C, e->b.Hit, w, h, e->b.Schedule,
                         e->stream)) != ARBOR_OK)
    return rc;
  if (timed) (void)hipEventRecord(e->ev[2], e->stream);
  if ((rc = arbor::visibilityOrdered(e->b.P, e->b.Nodes, e->b.Order, accel, e->b.C, e->b.Hit, e->b.Vis, w, h,
                              e->b.VisSchedule, e->stream)) != ARBOR_OK)
    return rc;
  const bool net = useNetwork(e), lean = net && arbor::denoiserLean(e->dn);
  if (lean)  // the compact features, in the normal's buffer (both 16 B per pixel)
    rc = arbor::shadeVisFeatures(e->b.P, e->b.Nodes, e->b.Order, e->b.C, e->b.Hit, e->b.Vis, e->b.Normal, w, h,
                                 e->stream);
  else if (net)
    rc = arbor::shadeVisNormal(e->b.P, e->b.Nodes, e->b.Order, e->b.C, e->b.Hit, e->b.Vis, e->b.Linear, e->b.Guide,
                               e->b.Normal, w, h, e->stream);
  else
    rc = arbor_shadePixelsVis(e->b.P, e->b.Nodes, e->b.Order, e->b.C, e->b.Hit, e->b.Vis, e->b.Linear, e->b.Guide, w,
                              h, e->stream);
  if (rc != ARBOR_OK) return rc;
  if (timed) (void)hipEventRecord(e->ev[3], e->stream);
  if (lean) {
    if ((rc = arbor::denoiseFeaturesResolve(e->dn, e->b.Normal, e->b.Filtered, e->b.History, e->b.Pixels, e->b.C, w,
                                            h, e->stream)) != ARBOR_OK)
      return rc;
  } else if (net) {
    if ((rc = arbor::denoiseResolve(e->dn, e->b.Linear, e->b.Guide, e->b.Normal, e->b.Hit, e->b.Filtered,
                                    e->b.History, e->b.Pixels, e->b.C, w, h, e->stream)) != ARBOR_OK)
      return rc;
  } else if ((rc = arbor::filterResolve(e->b.Linear, e->b.Guide, e->b.Hit, e->b.Filtered, e->b.History, e->b.Pixels,
                                        e->b.C, w, h, e->stream)) != ARBOR_OK) {
    return rc;
  }
  e->denoised = net;
  if (timed) {
    (void)hipEventRecord(e->ev[4], e->stream);
    if (hipEventSynchronize(e->ev[4]) != hipSuccess) return ARBOR_ERR_DEVICE;
    for (int s = 0; s < 4; s++) (void)hipEventElapsedTime(&e->stage_ms[s], e->ev[s], e->ev[s + 1]);
    e->stage_ms[4] = 0.0f;
    (void)hipEventElapsedTime(&e->stage_ms[5], e->ev[0], e->ev[4]);
  }
  return ARBOR_OK;
}

}  // namespace

extern "C" {

void arbor_engine_round_size(int* width, int* height) {
  // engine.js:36: width up to a multiple of 64 (min 320), height to 8 (min 192).
  int w = ((*width + 63) / 64) * 64, h = ((*height + 7) / 8) * 8;
  *width = w < 320 ? 320 : w;
  *height = h < 192 ? 192 : h;
}

int arbor_engine_create(arbor_engine** out, int width, int height, int seed, int round_size, hipStream_t stream) {
  if (!out) return ARBOR_ERR_NULL;
  *out = nullptr;
  if (seed < 0 || seed > 1000000) return ARBOR_ERR_ARG;
  arbor_engine* e = new arbor_engine();
  if (stream) e->stream = stream;
  else {
    if (hipStreamCreate(&e->stream) != hipSuccess) { delete e; return ARBOR_ERR_DEVICE; }
    e->owns_stream = true;
  }
  int rc = ARBOR_OK;
  rc |= alloc(&e->b.B, (size_t)ARBOR_CURVES * ARBOR_CS * 4);
  rc |= alloc(&e->b.P, (size_t)ARBOR_CAPACITY * ARBOR_PS * 4);
  rc |= alloc(&e->b.Keys, (size_t)ARBOR_CAPACITY * 4);
  rc |= alloc(&e->b.Order, (size_t)ARBOR_CAPACITY * 4);
  rc |= alloc(&e->b.Nodes, (size_t)ARBOR_NODE_COUNT * 8 * 4);
  rc |= alloc(&e->b.C, ARBOR_C_FLOATS * 4);
  rc |= alloc(&e->b.I, ARBOR_I_FLOATS * 4);
  rc |= alloc(&e->b.Accel, (size_t)arbor_trace_accel_bytes());
  for (int i = 0; i < 6 && rc == ARBOR_OK; i++)
    if (hipEventCreate(&e->ev[i]) != hipSuccess) rc = ARBOR_ERR_DEVICE;
  if (rc == ARBOR_OK) {
    // B, Nodes and I start zeroed as WebGPU buffers do (engine.js:22).
    if (hipMemsetAsync(e->b.B, 0, (size_t)ARBOR_CURVES * ARBOR_CS * 4, e->stream) != hipSuccess ||
        hipMemsetAsync(e->b.Nodes, 0, (size_t)ARBOR_NODE_COUNT * 32, e->stream) != hipSuccess ||
        hipMemsetAsync(e->b.I, 0, ARBOR_I_FLOATS * 4, e->stream) != hipSuccess)
      rc = ARBOR_ERR_DEVICE;
  }
  if (rc == ARBOR_OK) rc = arbor_engine_resize(e, width, height, round_size);
  if (rc == ARBOR_OK) rc = arbor_engine_generate(e, seed);
  if (rc != ARBOR_OK) {
    arbor_engine_destroy(e);
    return rc;
  }
  *out = e;
  return ARBOR_OK;
}

int arbor_engine_destroy(arbor_engine* e) {
  if (!e) return ARBOR_ERR_NULL;
  if (e->rec) fclose(e->rec);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  free_pixels(e);
  release(e->b.B); release(e->b.P); release(e->b.Keys); release(e->b.Order); release(e->b.Nodes); release(e->b.Accel);
  release(e->b.C); release(e->b.I);
  for (auto& v : e->ev)
    if (v) (void)hipEventDestroy(v);
  if (e->owns_stream && e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return ARBOR_OK;
}

int arbor_engine_generate(arbor_engine* e, int seed) {
  if (!e) return ARBOR_ERR_NULL;
  if (seed < 0 || seed > 1000000) return ARBOR_ERR_ARG;  // engine.js:28
  int rc = arbor_generate(e->b.B, e->b.P, e->b.Keys, e->b.Order, e->b.Nodes, e->b.C, seed, e->b.History, e->width,
                          e->height, e->stream);
  if (rc == ARBOR_OK) rc = arbor_buildTraceAccel(e->b.Nodes, e->b.Order, e->b.Accel, e->stream);
  if (rc == ARBOR_OK) {
    e->seed = seed;
    e->generated = true;
  }
  return rc;
}

int arbor_engine_resize(arbor_engine* e, int width, int height, int round_size) {
  if (!e) return ARBOR_ERR_NULL;
  if (round_size) arbor_engine_round_size(&width, &height);
  if (!arbor::dims_ok(width, height)) return ARBOR_ERR_ARG;
  if (width == e->width && height == e->height && e->b.Hit) return ARBOR_OK;
  (void)hipStreamSynchronize(e->stream);
  free_pixels(e);
  size_t n = pixel_bytes(width, height);
  int rc = ARBOR_OK;
  rc |= alloc(&e->b.Hit, n);
  rc |= alloc(&e->b.Linear, n);
  rc |= alloc(&e->b.Guide, n);
  rc |= alloc(&e->b.Filtered, n);
  rc |= alloc(&e->b.History, n);
  rc |= alloc(&e->b.Pixels, (size_t)width * height * 4);
  const long long sched = arbor_tile_schedule_bytes(width, height);
  rc |= alloc(&e->b.Schedule, (size_t)sched);
  rc |= alloc(&e->b.VisSchedule, (size_t)sched);
  rc |= alloc(&e->b.Vis, n);
  if (e->dn) rc |= alloc(&e->b.Normal, n);
  if (rc != ARBOR_OK) {
    free_pixels(e);
    e->width = e->height = 0;
    return ARBOR_ERR_DEVICE;
  }
  if (hipMemsetAsync(e->b.History, 0, n, e->stream) != hipSuccess ||
      hipMemsetAsync(e->b.Schedule, 0, (size_t)sched, e->stream) != hipSuccess ||
      hipMemsetAsync(e->b.VisSchedule, 0, (size_t)sched, e->stream) != hipSuccess)
    return ARBOR_ERR_DEVICE;
  e->width = width;
  e->height = height;
  return ARBOR_OK;
}

int arbor_engine_reset_camera(arbor_engine* e) {
  if (!e) return ARBOR_ERR_NULL;
  if (!e->generated) return ARBOR_ERR_STATE;
  return arbor_initCamera(e->b.C, e->b.P, e->seed, e->stream);
}

int arbor_engine_set_input(arbor_engine* e, const float* input32) {
  if (!e || !input32) return ARBOR_ERR_NULL;
  memcpy(e->input, input32, sizeof(e->input));
  return ARBOR_OK;
}

int arbor_engine_set_timing(arbor_engine* e, int on) {
  if (!e) return ARBOR_ERR_NULL;
  e->timing = on != 0;
  return ARBOR_OK;
}

int arbor_engine_set_accel(arbor_engine* e, int on) {
  if (!e) return ARBOR_ERR_NULL;
  e->accel = on != 0;
  return ARBOR_OK;
}

int arbor_engine_set_denoiser(arbor_engine* e, arbor_denoiser* d) {
  if (!e) return ARBOR_ERR_NULL;
  if (d && !e->b.Normal && e->b.Hit) {
    if (alloc(&e->b.Normal, pixel_bytes(e->width, e->height)) != ARBOR_OK) return ARBOR_ERR_DEVICE;
  }
  if (!d && e->b.Normal) {
    (void)hipStreamSynchronize(e->stream);
    release(e->b.Normal);
  }
  e->dn = d;
  return ARBOR_OK;
}

// One frame: the five launches of frameLaunches. With device_input NULL the
// host input set by arbor_engine_set_input is uploaded first.
int arbor_engine_frame(arbor_engine* e, float dt, const float* device_input) {
  if (!e) return ARBOR_ERR_NULL;
  if (!e->generated || !e->b.Hit) return ARBOR_ERR_STATE;
  int rc0 = recordFrame(e, dt, device_input, 0);
  if (rc0 != ARBOR_OK) return rc0;
  const float* I = device_input;
  if (!I) {
    if (hipMemcpyAsync(e->b.I, e->input, sizeof(e->input), hipMemcpyHostToDevice, e->stream) != hipSuccess)
      return ARBOR_ERR_DEVICE;
    I = e->b.I;
  }
  float delta = dt < 0.0f ? 0.0f : (dt > 0.1f ? 0.1f : dt);
  int rc = frameLaunches(e, I, delta, -1, -1, e->timing);
  if (rc != ARBOR_OK) return rc;
  e->frames++;
  return ARBOR_OK;
}

int arbor_engine_stage_ms(const arbor_engine* e, float* out6) {
  if (!e || !out6) return ARBOR_ERR_NULL;
  for (int i = 0; i < 6; i++) out6[i] = e->stage_ms[i];
  return ARBOR_OK;
}

int arbor_engine_sync(arbor_engine* e) {
  if (!e) return ARBOR_ERR_NULL;
  return hipStreamSynchronize(e->stream) == hipSuccess ? ARBOR_OK : ARBOR_ERR_DEVICE;
}

void* arbor_engine_buffer(arbor_engine* e, const char* name) {
  if (!e || !name) return nullptr;
  const Buffers& b = e->b;
  struct { const char* n; void* p; } table[] = {
      {"B", b.B}, {"P", b.P}, {"Keys", b.Keys}, {"Order", b.Order}, {"Nodes", b.Nodes}, {"C", b.C}, {"I", b.I},
      {"Hit", b.Hit}, {"Linear", b.Linear}, {"Guide", b.Guide}, {"Filtered", b.Filtered}, {"History", b.History},
      {"Pixels", b.Pixels}, {"Schedule", b.Schedule}, {"VisSchedule", b.VisSchedule}, {"Vis", b.Vis},
      {"Accel", b.Accel}, {"Normal", b.Normal}};
  for (auto& t : table)
    if (strcmp(t.n, name) == 0) return t.p;
  return nullptr;
}

hipStream_t arbor_engine_stream(arbor_engine* e) { return e ? e->stream : nullptr; }

int arbor_engine_size(const arbor_engine* e, int* width, int* height) {
  if (!e || !width || !height) return ARBOR_ERR_NULL;
  *width = e->width;
  *height = e->height;
  return ARBOR_OK;
}

// inspect() (engine.js:55): JSON with sizes, counters and timings.
int arbor_engine_inspect(const arbor_engine* e, char* out, int cap) {
  if (!e || !out || cap <= 0) return ARBOR_ERR_NULL;
  size_t scene = (size_t)ARBOR_CURVES * ARBOR_CS * 4 + (size_t)ARBOR_CAPACITY * ARBOR_PS * 4 +
                 (size_t)ARBOR_CAPACITY * 8 + (size_t)ARBOR_NODE_COUNT * 32;
  // Hit, Linear, Guide, Filtered, History, Vis (and Normal with a denoiser
  // attached) at 16 B a pixel, Pixels at 4 B, the two tile schedules, and
  // the SAH hierarchy beside the scene.
  const size_t sched = e->width > 0 ? (size_t)arbor_tile_schedule_bytes(e->width, e->height) : 0;
  size_t allocated = scene + ARBOR_C_FLOATS * 4 + ARBOR_I_FLOATS * 4 + (size_t)arbor_trace_accel_bytes() +
                     (e->b.Normal ? 7 : 6) * pixel_bytes(e->width, e->height) + (size_t)e->width * e->height * 4 +
                     2 * sched;
  int n = snprintf(out, (size_t)cap,
                   "{\"width\":%d,\"height\":%d,\"seed\":%d,\"frames\":%ld,\"accel\":%s,\"denoiser\":\"%s\","
                   "\"sceneBytes\":%zu,"
                   "\"allocatedBytes\":%zu,\"timings_ms\":{\"camera\":%.4f,\"visibility\":%.4f,\"lighting\":%.4f,"
                   "\"filter\":%.4f,\"resolve\":%.4f,\"frame\":%.4f},\"abi\":{\"CAPACITY\":%d,\"NODE_COUNT\":%d,"
                   "\"PS\":%d,\"CS\":%d,\"CURVES\":%d,\"WOOD_COUNT\":%d,\"LEAF_COUNT\":%d,\"ROCK_COUNT\":%d,"
                   "\"OBJECT_COUNT\":%d}}",
                   e->width, e->height, e->seed, e->frames, e->accel ? "true" : "false",
                   e->denoised ? "network" : "filterLighting", scene, allocated, e->stage_ms[0], e->stage_ms[1],
                   e->stage_ms[2], e->stage_ms[3], e->stage_ms[4], e->stage_ms[5], ARBOR_CAPACITY, ARBOR_NODE_COUNT, ARBOR_PS,
                   ARBOR_CS, ARBOR_CURVES, ARBOR_WOOD_COUNT, ARBOR_LEAF_COUNT, ARBOR_ROCK_COUNT, ARBOR_OBJECT_COUNT);
  return (n >= 0 && n < cap) ? ARBOR_OK : ARBOR_ERR_ARG;
}

// Seeded first-person walk: per frame, forward/back and strafe keys in
// {-1,0,1} and a mouse-x delta in [-20,20) from the integer hash of
// (frame, seed). Input slots follow src/app.js:10 as updateCamera reads them.
static unsigned walk_hash(unsigned x) {
  x ^= x >> 16; x *= 2146121005u; x ^= x >> 15; x *= 2221713035u; x ^= x >> 16;
  return x;
}
static float walk_unit(unsigned n, unsigned seed) {  // [0,1) from the same hash as rnd
  return (float)(walk_hash(n ^ walk_hash(seed)) & 16777215u) / 16777216.0f;
}

// Host-side camera state: initCamera's defaults (Arbor.cu:181) with fpsPose
// at (eye, yaw, pitch) (Arbor.cu:179), evaluated with the host libm as the
// reference harness does.
int arbor_camera_pose(float* C, int seed, float ex, float ey, float ez, float yaw, float pitch) {
  if (!C) return ARBOR_ERR_NULL;
  if (seed < 0 || seed > 1000000) return ARBOR_ERR_ARG;
  for (int i = 0; i < ARBOR_C_FLOATS; i++) C[i] = 0.0f;
  C[12] = -0.45f; C[13] = 0.63f; C[14] = 0.90f; C[20] = 1.18f; C[22] = 1.0f; C[23] = 1.0f;
  C[31] = (float)seed; C[44] = 1.0f;
  C[16] = yaw; C[17] = pitch; C[27] = yaw; C[28] = pitch;
  C[0] = ex; C[1] = ey; C[2] = ez;
  float cp = cosf(pitch);
  float fx = sinf(yaw) * cp, fy = sinf(pitch), fz = cosf(yaw) * cp;
  float len = sqrtf(fx * fx + fy * fy + fz * fz);
  float d = len > 0.0000001f ? len : 0.0000001f;
  C[4] = ex + fx / d; C[5] = ey + fy / d; C[6] = ez + fz / d;
  C[3] = 1.0f; C[18] = 1.0f;
  C[9] = 0.0f; C[10] = 1.0f;
  return ARBOR_OK;
}

// The batched-views workload: view i stands at a seeded position inside the
// forest (|x|,|z| < 240 m, eye height 1.72 m) with a seeded yaw and a small
// pitch, sun visibility off, at a resolution drawn uniformly from {256x256,
// 640x360, 1280x720, 1920x1080, 2560x1440}.
int arbor_batch_walk_views(int seed, int world_seed, int n, float* cameras, int* widths, int* heights) {
  if (!cameras || !widths || !heights) return ARBOR_ERR_NULL;
  if (n < 1 || n > ARBOR_BATCH_MAX_VIEWS) return ARBOR_ERR_ARG;
  static const int res[5][2] = {{256, 256}, {640, 360}, {1280, 720}, {1920, 1080}, {2560, 1440}};
  for (int i = 0; i < n; i++) {
    unsigned base = (unsigned)seed * 7919u + (unsigned)i * 101u;
    float x = (walk_unit(base + 1, 811) - 0.5f) * 480.0f;
    float z = (walk_unit(base + 2, 811) - 0.5f) * 480.0f;
    float yaw = walk_unit(base + 3, 811) * 6.2831853f;
    float pitch = (walk_unit(base + 4, 811) - 0.5f) * 0.2f;
    int r = (int)(walk_unit(base + 5, 811) * 5.0f);
    if (r > 4) r = 4;
    int rc = arbor_camera_pose(cameras + 64 * i, world_seed, x, 1.72f, z, yaw, pitch);
    if (rc != ARBOR_OK) return rc;
    cameras[64 * i + 22] = 0.0f;
    widths[i] = res[r][0];
    heights[i] = res[r][1];
  }
  return ARBOR_OK;
}
int arbor_engine_render(arbor_engine* e, int view, int samples, int light) {
  if (!e) return ARBOR_ERR_NULL;
  if (!e->generated || !e->b.Hit) return ARBOR_ERR_STATE;
  if (view < 0 || view > 5 || samples < 1 || samples > 512) return ARBOR_ERR_ARG;
  float input[ARBOR_I_FLOATS] = {};
  input[5] = (float)view;  // Renderer::view (native_harness.cpp:50)
  if (hipMemcpyAsync(e->b.I, input, sizeof(input), hipMemcpyHostToDevice, e->stream) != hipSuccess)
    return ARBOR_ERR_DEVICE;
  if (hipStreamSynchronize(e->stream) != hipSuccess) return ARBOR_ERR_DEVICE;  // `input` is on this stack
  // per-sample camera slots of the harness render loop (native_harness.cpp:62)
  int rc = ARBOR_OK;
  for (int i = 0; i < samples && rc == ARBOR_OK; i++)
    rc = frameLaunches(e, i == 0 ? e->b.I : nullptr, 1.0f / 60.0f, i, light ? 1 : 0, false);
  if (rc == ARBOR_OK) e->frames += samples;
  return rc;
}

int arbor_engine_frame_samples(arbor_engine* e, float dt, const float* device_input, int samples) {
  if (!e) return ARBOR_ERR_NULL;
  if (!e->generated || !e->b.Hit) return ARBOR_ERR_STATE;
  if (samples < 1 || samples > 512) return ARBOR_ERR_ARG;
  int rc0 = recordFrame(e, dt, device_input, samples);
  if (rc0 != ARBOR_OK) return rc0;
  const float* I = device_input;
  if (!I) {
    if (hipMemcpyAsync(e->b.I, e->input, sizeof(e->input), hipMemcpyHostToDevice, e->stream) != hipSuccess)
      return ARBOR_ERR_DEVICE;
    I = e->b.I;
  }
  float delta = dt < 0.0f ? 0.0f : (dt > 0.1f ? 0.1f : dt);
  int rc = ARBOR_OK;
  for (int i = 0; i < samples && rc == ARBOR_OK; i++) rc = frameLaunches(e, i == 0 ? I : nullptr, delta, i, -1, false);
  if (rc == ARBOR_OK) e->frames++;
  return rc;
}

int arbor_engine_record(arbor_engine* e, const char* path) {
  if (!e) return ARBOR_ERR_NULL;
  if (e->rec) {
    int bad = fclose(e->rec);
    e->rec = nullptr;
    if (bad) return ARBOR_ERR_ARG;
  }
  if (!path) return ARBOR_OK;
  FILE* f = fopen(path, "wb");
  if (!f) return ARBOR_ERR_ARG;
  const int32_t head[4] = {kRecVersion, e->width, e->height, e->seed};
  if (fwrite(kRecMagic, 8, 1, f) != 1 || fwrite(head, sizeof(head), 1, f) != 1) {
    fclose(f);
    return ARBOR_ERR_ARG;
  }
  e->rec = f;
  return ARBOR_OK;
}

int arbor_engine_replay(arbor_engine* e, const char* path, int first, int count, int* played) {
  if (!e || !path) return ARBOR_ERR_NULL;
  if (played) *played = 0;
  if (first < 0) return ARBOR_ERR_ARG;
  if (!e->generated || !e->b.Hit) return ARBOR_ERR_STATE;
  FILE* f = fopen(path, "rb");
  if (!f) return ARBOR_ERR_ARG;
  char magic[8];
  int32_t head[4];
  int rc = ARBOR_OK;
  if (fread(magic, 8, 1, f) != 1 || memcmp(magic, kRecMagic, 8) != 0 || fread(head, sizeof(head), 1, f) != 1 ||
      head[0] != kRecVersion)
    rc = ARBOR_ERR_ARG;
  if (rc == ARBOR_OK && fseek(f, (long)first * (long)sizeof(RecFrame), SEEK_CUR) != 0) rc = ARBOR_ERR_ARG;
  int n = 0;
  RecFrame r;
  FILE* keep = e->rec;  // a replay is not recorded again
  e->rec = nullptr;
  while (rc == ARBOR_OK && (count <= 0 || n < count) && fread(&r, sizeof(r), 1, f) == 1) {
    if (r.samples < 0 || r.samples > 512 || !(r.dt == r.dt)) {
      rc = ARBOR_ERR_ARG;
      break;
    }
    memcpy(e->input, r.input, sizeof(e->input));
    rc = r.samples == 0 ? arbor_engine_frame(e, r.dt, nullptr) : arbor_engine_frame_samples(e, r.dt, nullptr, r.samples);
    if (rc == ARBOR_OK) n++;
  }
  e->rec = keep;
  fclose(f);
  if (played) *played = n;
  return rc;
}

int arbor_write_ppm(const char* path, const unsigned* pixels, int width, int height) {
  if (!path || !pixels) return ARBOR_ERR_NULL;
  if (width < 1 || height < 1) return ARBOR_ERR_ARG;
  FILE* f = fopen(path, "wb");
  if (!f) return ARBOR_ERR_ARG;
  fprintf(f, "P6\n%d %d\n255\n", width, height);
  std::string row((size_t)width * 3, '\0');
  bool ok = true;
  for (int y = 0; y < height && ok; y++) {
    for (int x = 0; x < width; x++) {
      unsigned p = pixels[(size_t)y * width + x];
      row[3 * x] = (char)(p & 255u);
      row[3 * x + 1] = (char)((p >> 8) & 255u);
      row[3 * x + 2] = (char)((p >> 16) & 255u);
    }
    ok = fwrite(row.data(), 1, row.size(), f) == row.size();
  }
  ok = fclose(f) == 0 && ok;
  return ok ? ARBOR_OK : ARBOR_ERR_DEVICE;
}

int arbor_engine_save_ppm(arbor_engine* e, const char* path) {
  if (!e || !path) return ARBOR_ERR_NULL;
  if (!e->b.Pixels) return ARBOR_ERR_STATE;
  std::string host((size_t)e->width * e->height * 4, '\0');
  if (hipMemcpyAsync(&host[0], e->b.Pixels, host.size(), hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
      hipStreamSynchronize(e->stream) != hipSuccess)
    return ARBOR_ERR_DEVICE;
  return arbor_write_ppm(path, reinterpret_cast<const unsigned*>(host.data()), e->width, e->height);
}

int arbor_walk_input(int seed, int frame, float* input32) {
  if (!input32) return ARBOR_ERR_NULL;
  if (frame < 0) return ARBOR_ERR_ARG;
  for (int i = 0; i < ARBOR_I_FLOATS; i++) input32[i] = 0.0f;
  unsigned h = walk_hash((unsigned)frame * 2654435761u ^ walk_hash((unsigned)seed + 40503u));
  static const float fwd[4] = {1.0f, 1.0f, 0.0f, -1.0f};
  static const float strafe[4] = {0.0f, 1.0f, 0.0f, -1.0f};
  input32[15] = fwd[h & 3u];
  input32[13] = strafe[(h >> 2) & 3u];
  input32[0] = (float)((h >> 8) & 0xFFFFu) / 65536.0f * 40.0f - 20.0f;
  return ARBOR_OK;
}

}  // extern "C"
