// batch_static.hip -- the batch run that allocates nothing (dd_batch_run_static).
//
// dd_batch_run keeps a batch's work queues (job parameter blocks, tile
// descriptors, queue bounds, claim counters) in device buffers it allocates
// on its first run.  Callers that cannot allocate device memory -- code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc --
// run the same persistent launches with those queues in `scratch`, a fixed
// array in device memory of DD_BATCH_STATIC_BYTES per GPU, which every
// call on that GPU shares: a call writes its queues there (a host-to-device
// copy on the caller's stream), clears the claim counters with a kernel and
// launches the persistent kernels, all in stream order.  Calls that overlap
// on one GPU (several streams or threads at once) race on the array.
//
// `scratch` is a host-side stand-in in host code; its device address comes
// from hipGetSymbolAddress, once per device with that device current, and
// only that address reaches kernels and copies.
#include <mutex>

#include "batch_internal.hpp"

using namespace ddamd;

namespace {

static __device__ __align__(256) uint8_t scratch[DD_BATCH_STATIC_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
uint8_t *g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up on the
// first call on that device.
dd_status scratch_address(uint8_t **out) {
  int dev = 0;
  DD_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(DD_ERR_UNSUPPORTED, "device index beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[dev]) {
    void *ptr = nullptr;
    DD_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(DD_ERR_HIP, "hipGetSymbolAddress returned no address for the batch scratch array");
    g_scratch[dev] = static_cast<uint8_t *>(ptr);
  }
  *out = g_scratch[dev];
  return DD_OK;
}

} // namespace

extern "C" dd_status dd_batch_run_static(dd_batch *batch, void *stream) {
  if (!batch) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  if (batch_job_count(batch) == 0) return DD_OK;
  uint8_t *dev = nullptr;
  dd_status s = scratch_address(&dev);
  if (s != DD_OK) return s;
  StaticBatchImage &img = batch_image(batch);
  s = batch_static_image(batch, DD_BATCH_STATIC_BYTES, img);
  if (s != DD_OK) return s;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  // the claim counters (the first bytes) are cleared by kernels; the queues
  // after them are copied
  const uint64_t first = (uint64_t)3 * kQueues * kHeadStride * sizeof(uint32_t);
  if (img.bytes > first)
    DD_HIP(hipMemcpyAsync(dev + first, img.host.data() + first, img.bytes - first, hipMemcpyHostToDevice, st));
  return batch_enqueue_static(batch, dev, img, st);
}
