// SAH hierarchy over the heap BVH's leaf boxes, for the trace entry points
// that take an `accel` (arbor.h). Built on the host once per scene: the leaf
// records and the slot -> primitive order are copied from the device, the
// hierarchy is built by sah_build.hpp and uploaded. The device walk is in
// arbor_trace.hpp (traceTreeSah).
#include <hip/hip_runtime.h>

#include <vector>

#include "arbor_trace.hpp"
#include "launch.hpp"
#include "sah_build.hpp"

using namespace arbor;

static_assert(kAccelSlots * 4 == 16 + 16 * (kCapacity - 1), "SAH layout: slot table offset");

extern "C" {

long long arbor_trace_accel_bytes(void) { return (long long)kAccelSlots * 16 + 4LL * kCapacity; }

int arbor_buildTraceAccel(const float* Nodes, const unsigned* Order, void* accel, hipStream_t stream) {
  ARBOR_REQUIRE(Nodes && Order && accel);
  std::vector<float> leaves((size_t)kCapacity * 8);
  std::vector<unsigned> order(kCapacity);
  if (hipMemcpyAsync(leaves.data(), Nodes + (size_t)kCapacity * 8, leaves.size() * sizeof(float),
                     hipMemcpyDeviceToHost, stream) != hipSuccess ||
      hipMemcpyAsync(order.data(), Order, order.size() * sizeof(unsigned), hipMemcpyDeviceToHost, stream) !=
          hipSuccess ||
      hipStreamSynchronize(stream) != hipSuccess)
    return ARBOR_ERR_DEVICE;
  const std::vector<uint32_t> words = sah::buildAccel(leaves.data(), order.data(), kCapacity, kSahDepth, kFar);
  if (words.empty()) return ARBOR_ERR_STATE;  // cannot happen: the split keeps the depth bound
  if (hipMemcpyAsync(accel, words.data(), words.size() * sizeof(uint32_t), hipMemcpyHostToDevice, stream) !=
          hipSuccess ||
      hipStreamSynchronize(stream) != hipSuccess)
    return ARBOR_ERR_DEVICE;
  return ARBOR_OK;
}

}  // extern "C"
