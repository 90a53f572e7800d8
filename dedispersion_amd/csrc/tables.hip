// tables.hip -- device-resident shift tables.
//
// The float tables follow src/Shifts.cpp:20-42 of TRASAL/Dedispersion with
// every operation separately rounded in the same order (the square is x * x),
// and the integer tables apply the shift rule of include/Dedispersion.hpp:81,
// :124 and :168: (unsigned)((firstDM + dm * DMStep) * table[row]).  The
// explicit __f*_rn intrinsics keep the compiler from fusing any multiply-add,
// so the device tables equal the host restatement bit for bit.
#include "internal.hpp"

namespace ddamd {
namespace {

__global__ void build_shifts(const ShiftFormula f, float *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= f.padded_rows) return;
  float value = 0.0f;
  if (i + 1 < f.rows) {
    const float inverse_high = __fdiv_rn(1.0f, __fmul_rn(f.max_freq, f.max_freq));   // :21
    const float time_unit = __fmul_rn((float)f.samples_per_batch, f.sampling_time);  // :22
    const float freq = __fadd_rn(f.first_freq, __fmul_rn((float)i, f.freq_step));   // :25
    const float inverse = __fdiv_rn(1.0f, __fmul_rn(freq, freq));
    float delay = __fmul_rn(4148.808f, __fsub_rn(inverse, inverse_high));           // :26
    delay = __fmul_rn(delay, (float)(f.samples_per_batch / f.downsampling));
    value = __fdiv_rn(delay, time_unit);
  }
  out[i] = value;  // rows - 1 and the padding stay 0 (:28)
}

__global__ void relative_shifts(const float *table, uint32_t channels, uint32_t per_subband, uint32_t absolute,
                                float *out) {
  const uint32_t ch = blockIdx.x * blockDim.x + threadIdx.x;
  if (ch >= channels) return;
  const uint32_t top = (ch / per_subband + 1) * per_subband - 1;
  // include/Dedispersion.hpp:124 (the CPU function), or the channel's own
  // delay, as the OpenCL step-one kernel applies it (:595-597, :731, :733)
  out[ch] = absolute ? table[ch] : __fsub_rn(table[ch], table[top]);
}

__global__ void integer_shifts(const float *table, uint32_t rows, uint32_t dms, uint32_t dm_pitch,
                               float first_dm, float dm_step, uint32_t *out) {
  const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t r = blockIdx.y;
  if (d >= dm_pitch || r >= rows) return;
  const uint32_t dm = d < dms ? d : dms - 1;  // padding repeats the last DM
  const float dm_value = __fadd_rn(first_dm, __fmul_rn((float)dm, dm_step));
  const float delay = __fmul_rn(dm_value, table[r]);
  out[(uint64_t)r * dm_pitch + d] = (uint32_t)delay;
}

} // namespace

hipError_t launch_build_shifts(const ShiftFormula &f, float *d_table, hipStream_t stream) {
  const uint32_t blocks = (f.padded_rows + 255) / 256;
  hipLaunchKernelGGL(build_shifts, dim3(blocks), dim3(256), 0, stream, f, d_table);
  return hipGetLastError();
}

hipError_t launch_relative_shifts(const float *d_table, uint32_t channels, uint32_t per_subband, bool absolute,
                                  float *d_relative, hipStream_t stream) {
  const uint32_t blocks = (channels + 255) / 256;
  hipLaunchKernelGGL(relative_shifts, dim3(blocks), dim3(256), 0, stream, d_table, channels, per_subband,
                     absolute ? 1u : 0u, d_relative);
  return hipGetLastError();
}

hipError_t launch_integer_shifts(const float *d_table, uint32_t rows, uint32_t dms, uint32_t dm_pitch,
                                 float first_dm, float dm_step, uint32_t *d_out, hipStream_t stream) {
  dim3 grid((dm_pitch + 255) / 256, rows);
  hipLaunchKernelGGL(integer_shifts, grid, dim3(256), 0, stream, d_table, rows, dms, dm_pitch, first_dm, dm_step,
                     d_out);
  return hipGetLastError();
}

} // namespace ddamd
