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

} // namespace ddamd
