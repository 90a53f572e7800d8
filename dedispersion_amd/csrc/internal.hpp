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
  // pipelined kernels: shifts computed from the float table
  const float *table;         // float delay per row (samples per unit DM)
  float first_dm, dm_step;
  uint32_t n_tiles_x, n_tiles_y, n_tiles;  // 1-D grid decomposition
  // Compact output (pipelined kernels): the sums are written as uint16 at
  // the same element offsets (out is then a uint16 array); every sum must
  // be below 2^16 (255 x live rows of any group <= 65535).
  uint32_t out_u16;
  // Channel-block patterns (kernels_u8_factored.hip, single step):
  // kFactoredRecordDwords-dword records per (DM tile of kFactoredDMs, block of
  // four live channels), fact_slots Q rows per phase, fact_qw uint16 elements
  // per Q row (a multiple of 16); win_bytes is then the staged row window.
  const uint32_t *fact_records;
  uint32_t fact_slots;
  uint32_t fact_qw;
  // Packed input (the reference's inputBits < 8, include/Dedispersion.hpp:92-103):
  // log2(8 / inputBits) = 1, 2 or 3 for 4-, 2- and 1-bit samples, 0 for bytes.
  // Sample x of a row is the field at bit (x % (8 / inputBits)) * inputBits
  // of byte x / (8 / inputBits), least significant bit first; in_row_pitch and
  // in_beam_pitch count bytes.  Only the pipelined kernels read packed rows
  // (they unpack while staging into LDS).
  uint32_t in_log2f;
};

// Records of the pattern kernel, one per (DM tile, block of four live channels):
//   [0]      patterns P (<= kFactoredMaxPatterns) | rows in the block << 8
//   [1]      shift of the block's last (base) channel at the tile's first DM
//   [2], [3] per-row minimum / maximum of D_i over the patterns (bytes)
//   [4..19]  patterns: D_i = shift(row i) - shift(base) as bytes
//   [20..35] per DM of the tile, uint16: base-shift offset from the first DM
//            (low byte) | pattern index << 8
constexpr uint32_t kFactoredDMs = 32;
constexpr uint32_t kFactoredMaxPatterns = 16;
constexpr uint32_t kFactoredPatternWord = 4;
constexpr uint32_t kFactoredDMWord = kFactoredPatternWord + kFactoredMaxPatterns;
constexpr uint32_t kFactoredRecordDwords = kFactoredDMWord + kFactoredDMs / 2;

// Shifted row sum over float rows: step two (include/Dedispersion.hpp:155-176),
// and the single step and step one over float input (:70-153 with I = float).
// A launch's slices are z = outer * n_coarse + inner.  Output element
// (outer, inner, dm, t) = the float sum, in order, over the rows r of the
// slice of in[inbeam(outer, r)][inner][r][t + ishift[r][dm]], starting from 0.
// Step two: outer = sBeam, inner = coarse DM, rows = subbands 0..Nsb-1.
// Float single step: outer = sBeam, one inner, rows = the live channels.
// Float step one: outer = beam, inner = subband, rows = its live channels.
struct RowSumF32Params {
  const float *in;
  uint64_t in_floats;         // floats addressable from `in`
  float *out;
  const uint32_t *ishift;     // [sb][dm_pitch]
  uint32_t dm_pitch;
  const uint32_t *map;        // [sBeam][map_pitch]
  uint32_t map_pitch;
  uint32_t first_sbeam;
  uint32_t n_rows;            // rows of every slice without a row list (subbands)
  uint32_t n_coarse;          // inner indices per outer index: grid z = outer * n_coarse
  uint64_t in_row_pitch;      // floats between rows
  uint64_t in_coarse_pitch;   // floats between inner indices (coarse DMs)
  uint64_t in_beam_pitch;     // floats between input beams
  uint64_t out_dm_pitch;      // floats between output DM rows
  uint64_t out_beam_pitch;    // floats between outer indices
  uint32_t n_samples;
  uint32_t n_dms;             // fine DMs
  uint32_t unroll;
  uint32_t win_floats;        // LDS window per row, floats (multiple of 4)
  uint32_t nontemporal;
  uint32_t vec_store;         // output rows 16-byte aligned
  const float *table;         // float delay per subband
  float first_dm, dm_step;
  uint32_t n_tiles_x, n_tiles_y, n_tiles;
  // uint16 rows (kernels_u16_staged.hip): `in` points to uint16 step-one sums
  // with the pitches above counted in uint16 elements (in_floats = elements
  // addressable) and win_floats the window in elements; combine_rows rows
  // (1, 2 or 4) may be added as 16-bit pairs before widening, which needs
  // combine_rows x 255 x (live channels of any subband) <= 65535.
  uint32_t combine_rows;
  // Row lists (float and f32 kernels only): when `rows` is set, slice
  // (outer, inner) sums rows[inner * rows_pitch + k] for k < rows_count[inner],
  // in that order (each < 65536); otherwise rows 0 .. n_rows - 1.  A row's
  // input beam is map[(first_sbeam + outer) * map_pitch + row] when map is
  // set, else outer (each < 65536).
  const uint32_t *rows;
  const uint32_t *rows_count;
  uint32_t rows_pitch;        // also the most rows of a slice
  uint64_t out_inner_pitch;   // floats between the inner indices of one outer index
  // Paired uint16 step two (kernels_u16_pairs.hip): LDS windows per pair of
  // subbands (1 .. kPairSlotsMax; at least the number of values the shift
  // difference of the pair takes over any DM tile); `unroll` then counts
  // pairs per barrier and `combine_rows` pairs added as 16-bit lanes.
  uint32_t pair_slots;
};

// Most LDS windows per pair of subbands in the paired step two.
constexpr uint32_t kPairSlotsMax = 4;

// Geometry of one launch.
struct Geometry {
  uint32_t threads_d1; // wavefronts per workgroup
  uint32_t items_d0;   // samples per lane
  uint32_t items_d1;   // DMs per wavefront
  uint32_t local;      // LDS staging
  uint32_t win;        // LDS window per row (bytes for uint8 rows, floats for float rows)
  uint32_t lds_bytes;  // dynamic LDS per workgroup
  uint32_t staged;     // 1 = pipelined (double-buffered) kernel
  uint32_t packed;     // 1 = the rows may be packed (RowSumU8Params::in_log2f)
  uint32_t unroll;     // rows staged per barrier (the configuration's, or its cap)
  uint32_t pairs;      // 1 = the paired uint16 step two (kernels_u16_pairs.hip)
};

// Launchers (kernels_u8.hip / kernels_f32.hip).  Return hipSuccess or the
// launch error; hipErrorInvalidValue when no kernel matches the geometry.
hipError_t launch_rowsum_u8(const RowSumU8Params &p, const Geometry &g, bool flush, dim3 grid,
                            hipStream_t stream);
hipError_t launch_rowsum_f32(const RowSumF32Params &p, const Geometry &g, dim3 grid,
                             hipStream_t stream);
bool rowsum_u8_supported(uint32_t items_d0, uint32_t items_d1);
bool rowsum_f32_supported(uint32_t items_d0, uint32_t items_d1);
// Single step over channel-block patterns (kernels_u8_factored.hip); g.staged == 2.
hipError_t launch_rowsum_u8_factored(const RowSumU8Params &p, const Geometry &g, hipStream_t stream);
// Pipelined kernels (kernels_u8_staged.hip / kernels_f32_staged.hip).
hipError_t launch_rowsum_u8_staged(const RowSumU8Params &p, const Geometry &g, bool flush, hipStream_t stream);
hipError_t launch_rowsum_f32_staged(const RowSumF32Params &p, const Geometry &g, hipStream_t stream);
struct TileDesc {
  uint32_t job, tile;
};

// Work queues of a persistent launch.  Tile descriptors (job, logical tile)
// are split into kQueues contiguous ranges, one per XCD; queue q owns
// tiles[begin[q], begin[q + 1]).  heads[q * kHeadStride] counts the claims
// made on queue q and is zero when the launch starts.  A workgroup drains
// the queue of its own XCD first, then the other queues in turn, and leaves
// when a claim on the last of them fails: every workgroup makes at most one
// failed claim per queue, so the launch ends once all tiles are claimed.
constexpr uint32_t kQueues = 8;
constexpr uint32_t kHeadStride = 64;  // claim counters 256 bytes apart
constexpr uint32_t kNoTile = 0xFFFFFFFFu;

struct TileQueues {
  const TileDesc *tiles;   // (job index, logical tile of that job)
  const uint32_t *begin;   // [kQueues + 1]
  uint32_t *heads;         // [kQueues * kHeadStride]
};

// Persistent batch kernels (kernels_u8_staged.hip / kernels_f32_staged.hip):
// `workgroups` resident workgroups drain the queues; job j of d_jobs is a
// staged launch description (n_tiles_x/y set) of one beam or plan.
hipError_t launch_rowsum_u8_batch(const RowSumU8Params *d_jobs, const TileQueues &q, const Geometry &g, bool flush,
                                  uint32_t workgroups, hipStream_t stream);
hipError_t launch_rowsum_f32_batch(const RowSumF32Params *d_jobs, const TileQueues &q, const Geometry &g,
                                   uint32_t workgroups, hipStream_t stream);
hipError_t rowsum_u8_batch_occupancy(const Geometry &g, bool flush, int *per_cu);
// One workgroup per listed descriptor (no work queue): grid = n.
hipError_t launch_rowsum_u8_list(const RowSumU8Params *d_jobs, const TileDesc *d_tiles, uint32_t n, const Geometry &g,
                                 bool flush, hipStream_t stream);
hipError_t launch_rowsum_f32_list(const RowSumF32Params *d_jobs, const TileDesc *d_tiles, uint32_t n,
                                  const Geometry &g, hipStream_t stream);
hipError_t launch_rowsum_u16_list(const RowSumF32Params *d_jobs, const TileDesc *d_tiles, uint32_t n,
                                  const Geometry &g, hipStream_t stream);
hipError_t rowsum_f32_batch_occupancy(const Geometry &g, int *per_cu);
hipError_t launch_clear_words(uint32_t *d_words, uint32_t n, hipStream_t stream);
// Step two over uint16 step-one sums (kernels_u16_staged.hip): the staged
// pipeline with a window of g.win uint16 elements per row, two copies per
// row (shifted by one element) so that every read is dword aligned.
hipError_t launch_rowsum_u16_staged(const RowSumF32Params &p, const Geometry &g, hipStream_t stream);
hipError_t launch_rowsum_u16_batch(const RowSumF32Params *d_jobs, const TileQueues &q, const Geometry &g,
                                   uint32_t workgroups, hipStream_t stream);
hipError_t rowsum_u16_batch_occupancy(const Geometry &g, int *per_cu);
// The same step two with subbands added in pairs first (kernels_u16_pairs.hip;
// RowSumF32Params::pair_slots set by prepare_step_two_u16).
hipError_t launch_rowsum_u16_pairs_batch(const RowSumF32Params *d_jobs, const TileQueues &q, const Geometry &g,
                                         uint32_t workgroups, hipStream_t stream);
hipError_t launch_rowsum_u16_pairs_list(const RowSumF32Params *d_jobs, const TileDesc *d_tiles, uint32_t n,
                                        const Geometry &g, hipStream_t stream);
hipError_t rowsum_u16_pairs_batch_occupancy(const Geometry &g, int *per_cu);
inline uint32_t u16_staged_chunks(uint32_t run) { return run / 8 + 1; }  // 16-byte loads per lane per row
constexpr uint32_t kStagedSlotsPerWave = 2;  // rows staged per wavefront per group
constexpr uint32_t kU8StagedChunks = 2;      // 16-byte staging loads per lane per row (window <= 2 KiB)
inline uint32_t f32_staged_chunks(uint32_t run) { return run / 4 + 1; }

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
// Step one's delay table: each channel's delay minus that of its subband's
// top channel, or (absolute) the channel's own delay.
hipError_t launch_relative_shifts(const float *d_table, uint32_t channels, uint32_t per_subband, bool absolute,
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
