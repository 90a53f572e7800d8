// plan_internal.hpp -- the plan object and the launch preparation shared by
// the single-plan entry points (plan.cpp) and batches (batch.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "internal.hpp"

namespace ddamd {

inline uint32_t padto(uint32_t v, uint32_t k) { return k == 0 ? v : ((v + k - 1) / k) * k; }
inline uint32_t roundup(uint32_t v, uint32_t k) { return ((v + k - 1) / k) * k; }

template <typename T> struct DeviceBuffer {
  T *ptr = nullptr;
  size_t count = 0;
  ~DeviceBuffer() { release(); }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    count = 0;
  }
  hipError_t allocate(size_t n) {
    release();
    count = n;
    if (n == 0) return hipSuccess;
    return hipMalloc(reinterpret_cast<void **>(&ptr), n * sizeof(T));
  }
  hipError_t upload(const std::vector<T> &host) {
    hipError_t e = allocate(host.size());
    if (e != hipSuccess || host.empty()) return e;
    return hipMemcpy(ptr, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice);
  }
  hipError_t download(std::vector<T> &host) const {
    host.resize(count);
    if (count == 0) return hipSuccess;
    return hipMemcpy(host.data(), ptr, count * sizeof(T), hipMemcpyDeviceToHost);
  }
};

// Integer shift table of one mode, device copy and host copy.
struct ShiftTable {
  uint32_t rows = 0, dms = 0, pitch = 0;
  DeviceBuffer<uint32_t> dev;
  std::vector<uint32_t> host;
  uint32_t at(uint32_t r, uint32_t d) const { return host[(size_t)r * pitch + d]; }
  uint32_t max_shift() const {
    uint32_t m = 0;
    for (uint32_t r = 0; r < rows; ++r)
      for (uint32_t d = 0; d < dms; ++d) m = std::max(m, at(r, d));
    return m;
  }
  // Largest shift spread inside any DM tile of `tile` consecutive DMs
  // (cached per tile size; the cache is cleared when the table changes).
  uint32_t max_spread(uint32_t tile) const {
    for (auto &c : spread_cache)
      if (c.first == tile) return c.second;
    uint32_t m = 0;
    for (uint32_t r = 0; r < rows; ++r)
      for (uint32_t d0 = 0; d0 < dms; d0 += tile) {
        uint32_t a = at(r, d0), b = at(r, std::min(d0 + tile - 1, pitch - 1));
        m = std::max(m, b - a);
      }
    spread_cache.emplace_back(tile, m);
    return m;
  }
  mutable std::vector<std::pair<uint32_t, uint32_t>> spread_cache;
  // Most values that the shift difference of a row pair (2k, 2k + 1) takes
  // over the DMs of any tile of `tile` consecutive DMs, counted as the width
  // of its range (cached per tile size like max_spread).
  uint32_t pair_range(uint32_t tile) const {
    for (auto &c : pair_cache)
      if (c.first == tile) return c.second;
    uint32_t m = 1;
    for (uint32_t r = 0; r + 1 < rows; r += 2)
      for (uint32_t d0 = 0; d0 < dms; d0 += tile) {
        int64_t lo = INT64_MAX, hi = INT64_MIN;
        for (uint32_t d = d0; d < std::min(d0 + tile, dms); ++d) {
          const int64_t delta = (int64_t)at(r + 1, d) - (int64_t)at(r, d);
          lo = std::min(lo, delta);
          hi = std::max(hi, delta);
        }
        m = (uint32_t)std::max<int64_t>(m, hi - lo + 1);
      }
    pair_cache.emplace_back(tile, m);
    return m;
  }
  mutable std::vector<std::pair<uint32_t, uint32_t>> pair_cache;
};

} // namespace ddamd

struct dd_plan {
  dd_observation obs{};
  uint32_t padding = 0;
  int device = 0;
  size_t max_lds = 65536;
  bool two_step = false;
  bool step_one_absolute = false;  // step one applies each channel's own delay (dd_plan_set_step_one_shifts)
  uint32_t input_bits = 8;  // the reference's inputBits: 1, 2, 4 (packed rows), 8, or 32 (float rows)
  dd_sizes sizes{};
  dd_conf conf[3]{};

  ddamd::DeviceBuffer<float> shifts, shifts2, relative;
  std::vector<float> h_shifts, h_shifts2;
  ddamd::ShiftTable ish_single, ish_one, ish_two;

  std::vector<uint32_t> h_zapped;  // flags [pad(C)]
  ddamd::DeviceBuffer<uint32_t> live_single, live_single_count, live_sub, live_sub_count;
  uint32_t max_live_sub = 0;  // most live channels in one subband
  // channel-block pattern records of the single step (built on first use)
  ddamd::DeviceBuffer<uint32_t> fact_records;
  bool fact_valid = false;
  bool fact_ok = false;        // the pattern kernel can run this plan
  uint32_t fact_max_patterns = 0;
  uint32_t fact_base_spread = 0;   // largest base-shift offset inside a DM tile
  uint32_t fact_delta_spread = 0;  // largest spread of one row's D_i over the patterns
  ddamd::DeviceBuffer<uint32_t> map_single, map_two;
};

namespace ddamd {

dd_status fail(dd_status s, const std::string &msg);
dd_status hip_fail(hipError_t e, const char *what);

#define DD_HIP(call)                                                                                 \
  do {                                                                                               \
    hipError_t e_ = (call);                                                                          \
    if (e_ != hipSuccess) return hip_fail(e_, #call);                                                \
  } while (0)

// Geometry of a mode under a configuration; validates limits.
dd_status geometry_of(const dd_plan *p, dd_mode mode, const dd_conf &c, Geometry &g);

// One prepared launch of a mode: kernel parameters, geometry and grid.
struct U8Launch {
  RowSumU8Params k{};
  Geometry g{};
  dim3 grid;
  bool flush = false;
};
struct F32Launch {
  RowSumF32Params k{};
  Geometry g{};
  dim3 grid;
};

// Validate arguments and fill the launch of one entry point under conf c
// (the checks and parameters of dd_dedisperse, dd_subband_step_one and
// dd_subband_step_two).  n_tiles_x/y/n are set for every geometry.
dd_status prepare_single(dd_plan *plan, const dd_conf &c, const uint8_t *d_input, float *d_output,
                         uint32_t first_sbeam, uint32_t nr_sbeams, U8Launch &L);
dd_status prepare_step_one(dd_plan *plan, const dd_conf &c, const uint8_t *d_input, float *d_intermediate,
                           U8Launch &L);
dd_status prepare_step_two(dd_plan *plan, const dd_conf &c, const float *d_intermediate, float *d_output,
                           uint32_t first_sbeam, uint32_t nr_sbeams, F32Launch &L);
// Step two over the uint16 sums that step one writes with out_u16 (a batch
// that owns its intermediate): the launch of prepare_step_two with `in` read
// as uint16 elements.  Fails with DD_ERR_CONFIGURATION when c does not select
// the staged pipeline or a subband's sum could reach 2^16.  With `pairs`, the
// paired kernel (kernels_u16_pairs.hip) is prepared when the job allows it
// (a pair's sum below 2^16, at most kPairSlotsMax shift differences per pair
// and DM tile, rows 8-element aligned, the windows within LDS); L.g.pairs
// says which was prepared.
dd_status prepare_step_two_u16(dd_plan *plan, const dd_conf &c, const void *d_intermediate, float *d_output,
                               uint32_t first_sbeam, uint32_t nr_sbeams, F32Launch &L, bool pairs = false);
// The single step and step one over float input rows (a plan with
// input_bits 32): the float row-sum kernels over the live-channel lists.
dd_status prepare_single_f32(dd_plan *plan, const dd_conf &c, const float *d_input, float *d_output,
                             uint32_t first_sbeam, uint32_t nr_sbeams, F32Launch &L);
dd_status prepare_step_one_f32(dd_plan *plan, const dd_conf &c, const float *d_input, float *d_intermediate,
                               F32Launch &L);
// Launch a prepared launch (staged or direct kernel).
dd_status run(const U8Launch &L, hipStream_t stream);
dd_status run(const F32Launch &L, hipStream_t stream);

} // namespace ddamd
