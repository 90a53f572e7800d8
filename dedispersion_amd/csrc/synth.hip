// synth.hip -- seeded synthetic inputs generated on the device.
//
// Every value is a function of (seed, row, column) only, through a
// splitmix64-style counter hash, so the same seed gives the same bytes on any
// launch geometry.  Gaussian values use the Box-Muller transform in double
// precision and are rounded to nearest and clipped to [0, 255].
#include "internal.hpp"

namespace ddamd {
namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void synth_fill(uint8_t *d, uint64_t rows, uint64_t cols, uint64_t pitch, float mean, float std_dev,
                           uint64_t seed, int uniform) {
  const uint64_t col = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t row = blockIdx.y + (uint64_t)blockIdx.z * gridDim.y;
  if (col >= cols || row >= rows) return;
  const uint64_t h = mix64(seed ^ mix64((row << 32) ^ col));
  uint8_t out;
  if (uniform) {
    out = (uint8_t)(h >> 56);
  } else {
    const double u1 = (double)((h >> 11) + 1) * 0x1.0p-53;           // (0, 1]
    const double u2 = (double)(mix64(h) >> 11) * 0x1.0p-53;          // [0, 1)
    const double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    double v = rint((double)mean + (double)std_dev * z);
    v = v < 0.0 ? 0.0 : (v > 255.0 ? 255.0 : v);
    out = (uint8_t)v;
  }
  d[row * pitch + col] = out;
}

__global__ void inject_pulse(uint8_t *d, uint64_t pitch, const float *table, uint32_t channels, float dm_value,
                             uint32_t t0, uint32_t width, uint32_t amplitude, uint32_t row_len) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t ch = blockIdx.y;
  if (k >= width || ch >= channels) return;
  const uint32_t shift = (uint32_t)__fmul_rn(dm_value, table[ch]);  // include/Dedispersion.hpp:81
  const uint64_t t = (uint64_t)t0 + shift + k;
  if (t >= row_len) return;
  uint8_t *p = d + (uint64_t)ch * pitch + t;
  const uint32_t v = (uint32_t)*p + amplitude;
  *p = (uint8_t)(v > 255u ? 255u : v);
}

} // namespace

hipError_t launch_synth_gaussian(uint8_t *d, uint64_t rows, uint64_t cols, uint64_t pitch, float mean,
                                 float std_dev, uint64_t seed, bool uniform, hipStream_t stream) {
  if (rows == 0 || cols == 0) return hipSuccess;
  const uint64_t gy = rows < 65535 ? rows : 65535;
  const uint64_t gz = (rows + gy - 1) / gy;
  dim3 grid((uint32_t)((cols + 255) / 256), (uint32_t)gy, (uint32_t)gz);
  hipLaunchKernelGGL(synth_fill, grid, dim3(256), 0, stream, d, rows, cols, pitch, mean, std_dev, seed,
                     uniform ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_inject_pulse(uint8_t *d, uint64_t pitch, const float *d_table, uint32_t channels,
                               const uint32_t *, float dm_value, uint32_t t0, uint32_t width, uint32_t amplitude,
                               uint32_t row_len, hipStream_t stream) {
  if (width == 0) return hipSuccess;
  dim3 grid((width + 255) / 256, channels);
  hipLaunchKernelGGL(inject_pulse, grid, dim3(256), 0, stream, d, pitch, d_table, channels, dm_value, t0, width,
                     amplitude, row_len);
  return hipGetLastError();
}

// ---- a read-bandwidth probe ----
//
// Streams a buffer with coalesced 16-byte loads, four in flight per thread,
// and writes one partial sum per thread: the read bandwidth a streaming
// kernel reaches on the device, the practical ceiling of the read-bound
// kernels beside the nominal HBM peak.

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
