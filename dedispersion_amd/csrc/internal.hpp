// internal.hpp -- definitions shared by the HIP sources of the engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "dedispersion_amd.h"

namespace ddamd {

// ---- kernel parameter blocks (passed by value as kernel arguments) ----

// Shifted row sum over uint8 rows: single-step dedispersion
// (include/Dedispersion.hpp:70-109) and step one (include/Dedispersion.hpp:111-153).
// Output element (beam, group, dm, t) = sum over the live rows r of the group
// of in[inbeam(beam, r)][r][t + ishift[r][dm]].
struct RowSumU8Params {
  const uint8_t *in;          // input base
  uint64_t in_bytes;          // bytes addressable from `in`
  float *out;
  const uint32_t *ishift;     // integer shifts [row][dm_pitch]
  uint32_t dm_pitch;
  const uint32_t *live;       // live rows of group g: live[g * live_pitch + k]
  const uint32_t *live_count; // [groups]
  uint32_t live_pitch;
  const uint32_t *map;        // beam mapping [sBeam][map_pitch] or nullptr (beam = z)
  uint32_t map_pitch;
  uint32_t first_sbeam;
  uint64_t in_row_pitch;      // bytes between rows (channels)
  uint64_t in_beam_pitch;     // bytes between input beams
  uint64_t out_dm_pitch;      // floats between DM rows
  uint64_t out_beam_pitch;    // floats between output beams
  uint64_t out_group_pitch;   // floats between groups (subbands) of one DM
  uint32_t n_samples;         // output samples per row
  uint32_t n_dms;
  uint32_t n_groups;          // grid z = beams * n_groups
  uint32_t unroll;            // rows staged per barrier
  uint32_t win_bytes;         // LDS window per row, bytes (multiple of 16)
  uint32_t nontemporal;       // stream the output past the caches
  uint32_t vec_store;         // output rows 16-byte aligned
  // register-window kernels: shifts computed from the float table
  const float *table;         // float delay per row (samples per unit DM)
  float first_dm, dm_step;
  uint32_t n_tiles_x, n_tiles_y, n_tiles;  // 1-D grid decomposition
};

// Shifted row sum over float rows: step two (include/Dedispersion.hpp:155-176).
// Output element (sbeam, coarse, dm, t) = sum over subbands sb = 0..Nsb-1, in
// order, of in[inbeam(sbeam, sb)][coarse][sb][t + ishift[sb][dm]].
struct RowSumF32Params {
  const float *in;
  uint64_t in_floats;         // floats addressable from `in`
  float *out;
  const uint32_t *ishift;     // [sb][dm_pitch]
  uint32_t dm_pitch;
  const uint32_t *map;        // [sBeam][map_pitch]
  uint32_t map_pitch;
  uint32_t first_sbeam;
  uint32_t n_rows;            // subbands
  uint32_t n_coarse;          // grid z = sBeams * n_coarse
  uint64_t in_row_pitch;      // floats between subband rows
  uint64_t in_coarse_pitch;   // floats between coarse DMs
  uint64_t in_beam_pitch;     // floats between beams
  uint64_t out_dm_pitch;      // floats between output DM rows
  uint64_t out_beam_pitch;
  uint32_t n_samples;
  uint32_t n_dms;             // fine DMs
  uint32_t unroll;
  uint32_t win_floats;        // LDS window per row, floats (multiple of 4)
  uint32_t nontemporal;
  uint32_t vec_store;         // output rows 16-byte aligned
  const float *table;         // float delay per subband
  float first_dm, dm_step;
  uint32_t n_tiles_x, n_tiles_y, n_tiles;
};

// Geometry of one launch.
struct Geometry {
  uint32_t threads_d1; // wavefronts per workgroup
  uint32_t items_d0;   // samples per lane
  uint32_t items_d1;   // DMs per wavefront
  uint32_t local;      // LDS staging
  uint32_t win;        // LDS window per row (bytes for uint8 rows, floats for float rows)
  uint32_t lds_bytes;  // dynamic LDS per workgroup
  uint32_t window;     // 1 = register-window kernel
  uint32_t span;       // register-window capacity (largest per-wavefront shift spread)
};

// Launchers (kernels_u8.hip / kernels_f32.hip).  Return hipSuccess or the
// launch error; hipErrorInvalidValue when no kernel matches the geometry.
hipError_t launch_rowsum_u8(const RowSumU8Params &p, const Geometry &g, bool flush, dim3 grid,
                            hipStream_t stream);
hipError_t launch_rowsum_f32(const RowSumF32Params &p, const Geometry &g, dim3 grid,
                             hipStream_t stream);
bool rowsum_u8_supported(uint32_t items_d0, uint32_t items_d1);
bool rowsum_f32_supported(uint32_t items_d0, uint32_t items_d1);
// Register-window kernels (kernels_u8_win.hip / kernels_f32_win.hip).
hipError_t launch_rowsum_u8_win(const RowSumU8Params &p, const Geometry &g, bool flush, hipStream_t stream);
hipError_t launch_rowsum_f32_win(const RowSumF32Params &p, const Geometry &g, hipStream_t stream);
// Smallest supported register-window span >= need, or 0.
uint32_t u8_win_span(uint32_t need);
uint32_t f32_win_span(uint32_t need);
// Geometry constants of the register-window kernels.
constexpr uint32_t kWinSlotsPerWave = 2;   // rows staged per wavefront per group
constexpr uint32_t kU8WinChunks = 2;       // 16-byte staging loads per lane per row (window <= 2 KiB)
inline uint32_t u8_win_reads(uint32_t run, uint32_t span) {  // bytes of a lane window
  const uint32_t emax = run - 1 + span, nwd = emax / 4 + run / 4 + 1;
  return ((nwd * 4 + run - 1) / run) * run;
}
inline uint32_t f32_win_chunks(uint32_t run) { return run / 4 + 1; }  // 16-byte staging loads per lane per row
inline uint32_t f32_win_reads(uint32_t run, uint32_t span) {  // floats of a lane window
  return ((3 + span + run + 3) / 4) * 4;
}

// Table kernels (tables.hip).
struct ShiftFormula {
  float first_freq;            // frequency of row 0
  float freq_step;             // frequency increment per row
  float max_freq;              // reference (highest) frequency
  uint32_t samples_per_batch;  // S
  uint32_t downsampling;
  float sampling_time;
  uint32_t rows;               // channels or subbands
  uint32_t padded_rows;        // table length
};
hipError_t launch_build_shifts(const ShiftFormula &f, float *d_table, hipStream_t stream);
hipError_t launch_relative_shifts(const float *d_table, uint32_t channels, uint32_t per_subband,
                                  float *d_relative, hipStream_t stream);
hipError_t launch_integer_shifts(const float *d_table, uint32_t rows, uint32_t dms, uint32_t dm_pitch,
                                 float first_dm, float dm_step, uint32_t *d_out, hipStream_t stream);

// Synthetic data kernels (synth.hip).
hipError_t launch_synth_gaussian(uint8_t *d, uint64_t rows, uint64_t cols, uint64_t pitch, float mean,
                                 float std_dev, uint64_t seed, bool uniform, hipStream_t stream);
hipError_t launch_inject_pulse(uint8_t *d, uint64_t pitch, const float *d_table, uint32_t channels,
                               const uint32_t *d_live_mask, float dm_value, uint32_t t0, uint32_t width,
                               uint32_t amplitude, uint32_t row_len, hipStream_t stream);

// Thread-local last-error text used by dd_last_error().
void set_error(const std::string &message);

} // namespace ddamd
