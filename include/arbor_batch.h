/* Batched multi-view rendering (no counterpart in the reference, which renders
 * one view per frame): up to ARBOR_BATCH_MAX_VIEWS camera views of uneven
 * resolution sharing one resident scene, traced and shaded in ONE launch,
 * then filtered and resolved per view; the sequence can be captured once
 * into a hipGraph and replayed. Every view's Hit/Linear/Guide/Filtered/
 * History/Pixels equal what tracePixels, shadePixels, filterLighting and
 * resolvePixels give for that view alone. All calls return ARBOR_OK or a
 * negative status (arbor.h). */
#ifndef ARBOR_BATCH_H
#define ARBOR_BATCH_H

#include "arbor.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ARBOR_BATCH_MAX_VIEWS 256
#define ARBOR_BATCH_TRACE_SHADE 1
#define ARBOR_BATCH_FILTER 2
#define ARBOR_BATCH_RESOLVE 4
#define ARBOR_BATCH_ALL 7

typedef struct arbor_batch arbor_batch;

/* Allocates the view table and pooled per-pixel buffers for up to max_views
 * views and max_pixels pixels in total (History cleared). */
int arbor_batch_create(arbor_batch** out, int max_views, long long max_pixels);
int arbor_batch_destroy(arbor_batch* b);
/* Lays out nviews views: cameras is nviews x 64 host floats (each a C state),
 * widths/heights host ints. All views must share the forest seed C[31].
 * Invalidates a captured graph. */
int arbor_batch_set_views(arbor_batch* b, int nviews, const float* cameras, const int* widths, const int* heights);
/* Issues the stages (bitmask of ARBOR_BATCH_*) on `stream`: one fused
 * trace+shade launch over all views' tiles, then per view filter, resolve. */
int arbor_batch_render(arbor_batch* b, const float* P, const float* Nodes, const unsigned* Order, int stages,
                       hipStream_t stream);
/* Captures the same launches into a hipGraph (stream must be a created,
 * non-default stream); arbor_batch_launch replays it. */
int arbor_batch_capture(arbor_batch* b, const float* P, const float* Nodes, const unsigned* Order, int stages,
                        hipStream_t stream);
int arbor_batch_launch(arbor_batch* b, hipStream_t stream);
/* Trace and sun visibility over a SAH hierarchy of arbor_buildTraceAccel
 * (arbor.h) for the scene passed to render/capture, or over the heap BVH
 * with accel NULL (the default). batch_variant keeps selecting the launch
 * shape; over the hierarchy 7, 8 and the default (9) read the views' tree
 * records from a copy set_views writes and keep the walk's stack in LDS.
 * Invalidates a captured graph. No counterpart in the reference. */
int arbor_batch_set_accel(arbor_batch* b, const void* accel);
/* Device pointer of view `view`'s slice of a buffer: Hit Linear Guide
 * Filtered History Pixels C; NULL for an unknown name or view. */
void* arbor_batch_buffer(arbor_batch* b, int view, const char* name);
long long arbor_batch_pixels(const arbor_batch* b);
int arbor_batch_tiles(const arbor_batch* b);
/* Host helpers (no GPU). arbor_camera_pose fills a C state as initCamera
 * (Arbor.cu:181) followed by fpsPose at (eye, yaw, pitch) (Arbor.cu:179).
 * arbor_batch_walk_views generates the batched-views workload: n views at
 * seeded positions inside the forest (|x|,|z| < 240 m, eye height 1.72 m),
 * seeded yaw and small pitch, sun visibility off, resolutions drawn uniformly
 * from {256x256, 640x360, 1280x720, 1920x1080, 2560x1440}. */
int arbor_camera_pose(float* C, int seed, float ex, float ey, float ez, float yaw, float pitch);
int arbor_batch_walk_views(int seed, int world_seed, int n, float* cameras, int* widths, int* heights);
/* Autotuner: times each batched-kernel variant (block waves x walk policy)
 * on the current views, keeps the fastest as the "batch_variant" tuning and
 * writes the timings as JSON. Synchronises `stream`. */
int arbor_autotune_batch(arbor_batch* b, const float* P, const float* Nodes, const unsigned* Order, int reps,
                         hipStream_t stream, char* json, int cap);

#ifdef __cplusplus
}
#endif
#endif /* ARBOR_BATCH_H */
