/* ARBOR render path on MI355X (gfx950): the C-ABI drop-in boundary.
 *
 * The twelve entry points keep the parameter order, types and buffer layouts
 * of the reference CUDA kernels (Arbor.cu, line cited per entry) and add the
 * HIP stream they are ordered on. Launch geometry is internal. Every pointer
 * is a device pointer the caller allocated with the sizes below; no entry
 * point allocates. Every call returns ARBOR_OK or a negative status and never
 * aborts. One buffer set must not be used from two streams at once.
 *
 * Buffer layouts (flat little-endian f32/u32, docs/ARCHITECTURE.md:9-15):
 *   B      CURVES x CS f32     Bezier curves: ctrl0..3, ra, rb, arc0, seed, chord
 *   P      CAPACITY x PS f32   primitives; P[3] = -1 unused, 0 wood, 1 leaf, 2 stone
 *   Keys   CAPACITY u32        30-bit Morton keys (0xFFFFFFFF for unused)
 *   Order  CAPACITY u32        Morton-sorted slot -> primitive id
 *   Nodes  NODE_COUNT x 8 f32  heap BVH {lo.xyz, count, hi.xyz, id|-1}, root 1,
 *                              leaves CAPACITY+i, node 0 unused
 *   C      64 f32              camera/state; I 32 f32 input
 *   Hit, Linear, Guide, Filtered, History   W*H x 4 f32
 *   Pixels W*H u32             RGBA8
 */
#ifndef ARBOR_H
#define ARBOR_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ABI constants (Arbor.cu:5-15,54-56; src/engine.js:2). */
#define ARBOR_CAPACITY 32768
#define ARBOR_NODE_COUNT 65536
#define ARBOR_PS 24
#define ARBOR_CS 20
#define ARBOR_CURVES 2109
#define ARBOR_WOOD_COUNT 5124
#define ARBOR_LEAF_COUNT 24192
#define ARBOR_ROCK_COUNT 96
#define ARBOR_OBJECT_COUNT 29412
#define ARBOR_FOREST_SIDE 18
#define ARBOR_FOREST_COUNT 324
#define ARBOR_C_FLOATS 64
#define ARBOR_I_FLOATS 32
/* Resolution limits: the reference harness accepts 32..4096 (native_harness.cpp:61);
 * this library accepts 1..8192 per side. */
#define ARBOR_MAX_DIM 8192

/* Status codes. */
#define ARBOR_OK 0
#define ARBOR_ERR_NULL (-1)     /* a required pointer is NULL */
#define ARBOR_ERR_ARG (-2)      /* a scalar argument is out of its documented range */
#define ARBOR_ERR_LAUNCH (-3)   /* the HIP launch failed (hipGetLastError) */
#define ARBOR_ERR_DEVICE (-4)   /* no usable HIP device / runtime call failed */
#define ARBOR_ERR_STATE (-5)    /* engine used after destroy or before generate */

/* ---- the twelve reference entry points ---------------------------------- */

/* Arbor.cu:76  growTree(float* B, int level, int seed): level 0..4 in order;
 * level L reads only curves of level L-1. seed in [0, 1000000] (engine.js:28). */
int arbor_growTree(float* B, int level, int seed, hipStream_t stream);

/* Arbor.cu:103  makePrimitives(const float* B, float* P, int seed). */
int arbor_makePrimitives(const float* B, float* P, int seed, hipStream_t stream);

/* Arbor.cu:139  makeKeys(const float* P, unsigned* Keys, unsigned* Order). */
int arbor_makeKeys(const float* P, unsigned* Keys, unsigned* Order, hipStream_t stream);

/* Arbor.cu:140  sortPass(unsigned* Keys, unsigned* Order, int distance, int span):
 * one bitonic compare-exchange stage; distance and span powers of two with
 * 1 <= distance < span <= CAPACITY. */
int arbor_sortPass(unsigned* Keys, unsigned* Order, int distance, int span, hipStream_t stream);

/* Arbor.cu:141  makeBounds(const float* P, const unsigned* Order, float* Nodes). */
int arbor_makeBounds(const float* P, const unsigned* Order, float* Nodes, hipStream_t stream);

/* Arbor.cu:142  reduceBounds(float* Nodes, int first): first a power of two
 * in [1, CAPACITY/2]. */
int arbor_reduceBounds(float* Nodes, int first, hipStream_t stream);

/* Arbor.cu:181  initCamera(float* C, const float* P, int seed). P is unused by
 * the reference and may be NULL. */
int arbor_initCamera(float* C, const float* P, int seed, hipStream_t stream);

/* Arbor.cu:182  updateCamera(float* C, const float* I, const float* P,
 * float delta, int width, int height). P is unused and may be NULL. */
int arbor_updateCamera(float* C, const float* I, const float* P, float delta, int width, int height,
                       hipStream_t stream);

/* Arbor.cu:252  tracePixels(P, Nodes, Order, C, Hit, width, height):
 * Hit[4p..] = {t, primitive id | -1, tree | -1, C[14]*t/height}. */
int arbor_tracePixels(const float* P, const float* Nodes, const unsigned* Order, const float* C, float* Hit,
                      int width, int height, hipStream_t stream);

/* Arbor.cu:277  shadePixels(P, Nodes, Order, C, Hit, Linear, Guide, width, height). */
int arbor_shadePixels(const float* P, const float* Nodes, const unsigned* Order, const float* C, const float* Hit,
                      float* Linear, float* Guide, int width, int height, hipStream_t stream);

/* Arbor.cu:292  filterLighting(Linear, Guide, Hit, Filtered, C, width, height). */
int arbor_filterLighting(const float* Linear, const float* Guide, const float* Hit, float* Filtered, const float* C,
                         int width, int height, hipStream_t stream);

/* Arbor.cu:304  resolvePixels(Filtered, History, Pixels, C, width, height). */
int arbor_resolvePixels(const float* Filtered, float* History, unsigned* Pixels, const float* C, int width,
                        int height, hipStream_t stream);

/* ---- additions beside the reference interface --------------------------- */

/* Stable radix sort of the CAPACITY (Keys, Order) pairs into ascending
 * (key, order) order: the same result as the full 120-stage sortPass network
 * (engine.js:31) and as std::sort of pairs (native_harness.cpp:47). One
 * workgroup, LDS-staged; no workspace. */
int arbor_sortKeys(unsigned* Keys, unsigned* Order, hipStream_t stream);

/* makeKeys -> sortKeys -> makeBounds -> reduceBounds(16384..1) in one launch:
 * Keys, Order and Nodes bit-identical to the reference build sequence
 * (engine.js:30-32) for the same P. */
int arbor_buildBVH(const float* P, unsigned* Keys, unsigned* Order, float* Nodes, hipStream_t stream);

/* Whole one-time build, engine.js:27-34: growTree 0..4, makePrimitives,
 * buildBVH, initCamera. History (W*H*4 f32, may be NULL) is cleared. */
int arbor_generate(float* B, float* P, unsigned* Keys, unsigned* Order, float* Nodes, float* C, int seed,
                   float* History, int width, int height, hipStream_t stream);

/* shadePixels plus the shading normal per pixel (no counterpart in the
 * reference; an auxiliary output for the learned denoiser): Normal[4p..] =
 * {n, 0}, the normal shadePixels lights with, facing the camera; zero for sky.
 * Linear and Guide are identical to arbor_shadePixels'. */
int arbor_shadePixelsAOV(const float* P, const float* Nodes, const unsigned* Order, const float* C, const float* Hit,
                         float* Linear, float* Guide, float* Normal, int width, int height, hipStream_t stream);

/* tracePixels with a tile schedule kept across frames (no counterpart in the
 * reference; for callers that trace a moving view frame after frame). The
 * schedule is arbor_tile_schedule_bytes(width, height) bytes of device memory
 * the caller allocates and zeroes once; each call orders the view's 8x8 tiles
 * longest-predicted-first (the previous call's per-tile trace times,
 * reprojected through the camera change), traces them in that order and
 * records this call's times. Hit is identical to arbor_tracePixels'. One
 * schedule per view size; concurrent calls must not share it. Each ordering
 * also writes the 324 tree records of the world seed C[31] into it, which
 * the SAH walk's record-tree launch shapes read (an unscheduled call of
 * such a shape runs the default shape instead). */
long long arbor_tile_schedule_bytes(int width, int height);
int arbor_tracePixelsScheduled(const float* P, const float* Nodes, const unsigned* Order, const float* C, float* Hit,
                               int width, int height, void* schedule, hipStream_t stream);

/* Sun visibility as a pass of its own (no counterpart in the reference; the
 * shadow traversals of shadePixels, Arbor.cu:271-274, with the trace's launch
 * shape): Vis[4p..] = {transmission rgb, 1} for every pixel (1 for sky); with
 * C[22] turning tree-to-tree visibility off it writes nothing, as shading
 * then reads nothing. arbor_shadePixelsVis then shades
 * from it; Linear and Guide equal arbor_shadePixels' bit for bit. The
 * scheduled form keeps its own tile schedule (arbor_tile_schedule_bytes). */
int arbor_sunVisibility(const float* P, const float* Nodes, const unsigned* Order, const float* C, const float* Hit,
                        float* Vis, int width, int height, hipStream_t stream);
int arbor_sunVisibilityScheduled(const float* P, const float* Nodes, const unsigned* Order, const float* C,
                                 const float* Hit, float* Vis, int width, int height, void* schedule,
                                 hipStream_t stream);
int arbor_shadePixelsVis(const float* P, const float* Nodes, const unsigned* Order, const float* C, const float* Hit,
                         const float* Vis, float* Linear, float* Guide, int width, int height, hipStream_t stream);

/* SAH trace hierarchy (no counterpart in the reference; for callers that trace
 * one scene many times). arbor_buildTraceAccel builds, on the device and
 * asynchronously on `stream`, a binned surface-area-heuristic hierarchy over
 * the heap BVH's leaf boxes (the Nodes and Order of buildBVH) into `accel`,
 * arbor_trace_accel_bytes() bytes of device memory: the hierarchy's words
 * (16 + 16 (ARBOR_CAPACITY - 1) + ARBOR_CAPACITY u32) first, then the build's
 * workspace. The same Nodes/Order give the same words, and every leaf lies
 * within the walk's 32-entry stack. The Accel trace and visibility entry points walk it in place
 * of the heap (about a third of the box tests and 30% fewer object tests).
 * They take the heap too: objects hit at exactly the same t resolve in the
 * heap walk's order, decided at their common heap ancestor. Results equal
 * arbor_tracePixels' / arbor_sunVisibility's except for rays that hit a
 * swaying leaf outside its rest-pose leaf box, where which of two valid hits
 * the reference reports depends on its walk order (about 1 pixel in 2
 * million at 1920x1080). `schedule` (may be NULL) is a tile schedule as for
 * the Scheduled entry points. `accel` must be the one built from the Nodes
 * and Order passed with it (it is not checked). */
long long arbor_trace_accel_bytes(void);
int arbor_buildTraceAccel(const float* Nodes, const unsigned* Order, void* accel, hipStream_t stream);
/* The same build with every leaf within depth_limit (15..32) of the root:
 * nodes whose primitives would not fit below the limit by a surface-area
 * split are split at the median instead. depth_limit 32 is
 * arbor_buildTraceAccel. */
int arbor_buildTraceAccelLimit(const float* Nodes, const unsigned* Order, void* accel, int depth_limit,
                               hipStream_t stream);
int arbor_tracePixelsAccel(const float* P, const float* Nodes, const unsigned* Order, const void* accel,
                           const float* C, float* Hit, int width, int height, void* schedule, hipStream_t stream);
int arbor_sunVisibilityAccel(const float* P, const float* Nodes, const unsigned* Order, const void* accel,
                             const float* C, const float* Hit, float* Vis, int width, int height, void* schedule,
                             hipStream_t stream);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). arbor_tracePixelsAccel with a tile schedule, the schedule kept
 * in the library's own device array (arbor_no_alloc_scratch_bytes() bytes per
 * GPU, held until the process ends) instead of a caller workspace; it
 * allocates nothing. Views of up to 512 x 512 tiles (4096 x 4096 pixels);
 * larger views return ARBOR_ERR_ARG. Hit equals arbor_tracePixelsAccel's.
 * Every call on one GPU shares its array (the previous call's tile times seed
 * the order, whatever its view), so calls that overlap on one GPU race. */
long long arbor_no_alloc_scratch_bytes(void);
int arbor_tracePixelsAccelNoAlloc(const float* P, const float* Nodes, const unsigned* Order, const void* accel,
                                  const float* C, float* Hit, int width, int height, hipStream_t stream);

/* Single-specimen primary rays: traceTree in tree-local space with the ground
 * plane as initial best and no forest instancing (the single-tree workload of
 * native_harness.cpp:84 per pixel). Hit layout as tracePixels with tree = -1. */
int arbor_traceTreePixels(const float* P, const float* Nodes, const unsigned* Order, const float* C, float* Hit,
                          int width, int height, hipStream_t stream);

/* Library build identification (git-independent). */
const char* arbor_version(void);

/* Launch tuning (no counterpart in the reference, whose launch shapes are
 * fixed by its WebGPU host). Keys: "trace_variant" (0..9: 0-7 the heap
 * walk's launch shape and walk options, 8 and 9 the wave-scheduled kernels),
 * "grid_cap" (max blocks of a tile kernel, 1..1048576).
 * Every setting gives identical results. Process-wide; set before launching. */
int arbor_tune_set(const char* key, int value);
/* Current value of a tuning key, or a negative status for an unknown key.
 * Further keys: "sched_tiles" (1..4096), "sched_leaf_min" and
 * "sched_newton_min" (1..64) for the wave-scheduled trace kernel,
 * "batch_variant" (arbor_batch.h), "vis_variant" (0..7, the
 * sun-visibility kernel's launch shape and register budget), and for the
 * SAH walk "accel_variant" (0..34) and "accel_vis_variant" (0..27): block
 * size, register budget, how many of each lane's stack entries live in LDS,
 * whether a wave holds its object tests until a share of its lanes reach a
 * leaf, how many interior steps a lane takes between those checks, whether
 * the tree records come from the tile schedule, the persistent form
 * (accel_variant 10-12) and the stage-split form over ray queues
 * (accel_variant 32-34, which allocates its queues on first use per device);
 * the defaults are 26 and 23. */
int arbor_tune_get(const char* key);
/* Persisted tuning cache: writes / reads "key value" lines. A file with an
 * unknown key or out-of-range value is rejected whole (tuning unchanged). */
int arbor_tune_save(const char* path);
int arbor_tune_load(const char* path);
/* Autotuner (no counterpart in the reference): times every trace
 * configuration on the caller's buffers and view, keeps the fastest as
 * "trace_variant", writes {"best": i, "ms": [...]} to json. Synchronises. */
int arbor_autotune_trace(const float* P, const float* Nodes, const unsigned* Order, const float* C, float* Hit,
                         int width, int height, int reps, hipStream_t stream, char* json, int cap);
/* The same over the SAH hierarchy: times every "accel_variant" and keeps the
 * fastest (no counterpart in the reference). The shapes trace scheduled, in
 * the tile order the engine would use for the view (a tile schedule the call
 * allocates and frees). */
int arbor_autotune_trace_accel(const float* P, const float* Nodes, const unsigned* Order, const void* accel,
                               const float* C, float* Hit, int width, int height, int reps, hipStream_t stream,
                               char* json, int cap);

#ifdef __cplusplus
}
#endif
#endif /* ARBOR_H */
