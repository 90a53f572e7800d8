// probe.hip -- a read-bandwidth probe for the measurements (not part of the
// path).  It streams a buffer with coalesced 16-byte loads, four in flight
// per thread, and writes one partial sum per thread: the read bandwidth a
// streaming kernel reaches on the device, the practical ceiling of the
// read-bound kernels beside the nominal HBM peak.
#include "internal.hpp"

#include <string>

namespace ddamd {
namespace {

typedef float probe_f4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) read_probe(const probe_f4 *a, uint64_t n4, float *sink) {
  const uint64_t tid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  probe_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  uint64_t i = tid;
  for (; i + 3 * stride < n4; i += 4 * stride) {
    probe_f4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = __builtin_nontemporal_load(&a[i + u * stride]);
#pragma unroll
    for (int u = 0; u < 4; ++u) acc += v[u];
  }
  for (; i < n4; i += stride) acc += a[i];
  sink[tid] = acc.x + acc.y + acc.z + acc.w;
}

} // namespace
} // namespace ddamd

extern "C" dd_status dd_read_probe(const void *d_data, uint64_t bytes, float *d_sink, void *stream) {
  if (!d_data || !d_sink || bytes % 16 != 0 || (reinterpret_cast<uintptr_t>(d_data) & 15) != 0) {
    ddamd::set_error("dd_read_probe: need a 16-byte aligned buffer of whole 16-byte pieces and a sink");
    return DD_ERR_INVALID_ARGUMENT;
  }
  hipLaunchKernelGGL(ddamd::read_probe, dim3(DD_READ_PROBE_SINK_FLOATS / 256), dim3(256), 0, (hipStream_t)stream,
                     static_cast<const ddamd::probe_f4 *>(d_data), bytes / 16, d_sink);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    ddamd::set_error(std::string("read_probe launch: ") + hipGetErrorString(e));
    return DD_ERR_HIP;
  }
  return DD_OK;
}
