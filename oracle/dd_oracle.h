/*
 * dd_oracle.h -- CPU oracle for the dedispersion hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference's CPU algorithm (TRASAL/Dedispersion) against which the HIP
 * kernels are checked.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it.  The product library never links or
 * calls it.
 *
 * Every function cites the reference file:line it restates.  Floating-point
 * expressions keep the reference's operand order and are compiled with
 * -ffp-contract=off so that no multiply-add is fused.
 *
 * Pinning status (how each function is checked, see tests/test_oracle_*.py):
 *   orc_pad, observation getters ...... closed forms
 *   orc_get_shifts .................... float64 closed form (<=2 ulp), exact
 *                                       numpy float32 restatement, invariants
 *   orc_get_shifts_step_two ........... same as orc_get_shifts
 *   orc_dispersed_samples ............. closed form + containment invariant
 *   orc_dedispersion .................. constant input, impulse response,
 *                                       DM-0 channel sum, linearity, zapped
 *                                       invariance, brute force (numpy) on
 *                                       tiny shapes, committed fixtures
 *   orc_step_one ...................... constant input, impulse, DM-0 per
 *                                       subband sum, brute force, fixtures
 *   orc_step_two ...................... constant input, impulse, first-DM
 *                                       subband sum, float sequential order,
 *                                       brute force, fixtures
 * The AstroData::Observation semantics (maximum frequency, subband
 * frequencies, padding, beam mapping) come from an external library that is
 * absent; they are defined here as documented below and are parity-unpinned
 * against that library.
 */
#ifndef DD_ORACLE_H
#define DD_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Observation metadata used by the path (reference: AstroData::Observation
 * getters listed at include/Dedispersion.hpp:72-171, src/Shifts.cpp:21-40). */
typedef struct orc_obs {
  uint32_t nr_beams;
  uint32_t nr_synthesized_beams;
  uint32_t nr_channels;
  uint32_t nr_subbands;
  float min_freq;
  float channel_bandwidth;
  float sampling_time;
  uint32_t downsampling;
  uint32_t nr_samples_per_batch;                /* S       */
  uint32_t nr_samples_per_dispersed_batch;      /* S_disp  */
  uint32_t nr_samples_per_batch_sub;            /* S1      */
  uint32_t nr_samples_per_dispersed_batch_sub;  /* S1_disp */
  uint32_t nr_dms;                              /* D or Df */
  float first_dm;
  float dm_step;
  uint32_t nr_dms_sub;                          /* Dc      */
  float first_dm_sub;
  float dm_step_sub;
  /* The reference's inputBits argument (include/Dedispersion.hpp:59-60):
   * 1, 2 or 4 = packed samples, 8 (or 0) = one byte per sample, 32 = float
   * samples (the templates with I = float). */
  uint32_t input_bits;
  /* Step one's channel delay: 0 = relative to the subband's top channel
   * (the CPU function, :124), 1 = absolute (the OpenCL kernel, :595-597). */
  uint32_t step_one_absolute;
} orc_obs;

/* isa::utils::pad: round v up to a multiple of k; k == 0 means no padding. */
uint32_t orc_pad(uint32_t v, uint32_t k);

/* Derived frequencies (AstroData semantics as defined here). */
float orc_max_freq(const orc_obs *o);
float orc_subband_bandwidth(const orc_obs *o);
float orc_subband_min_freq(const orc_obs *o);
float orc_subband_max_freq(const orc_obs *o);
uint32_t orc_channels_per_subband(const orc_obs *o);

/* src/Shifts.cpp:20-30.  out has pad(C, padding/4) entries. */
void orc_get_shifts(const orc_obs *o, uint32_t padding, float *out);
/* src/Shifts.cpp:32-42.  out has pad(Nsb, padding/4) entries. */
void orc_get_shifts_step_two(const orc_obs *o, uint32_t padding, float *out);

/* src/DedispersionTest.cpp:129,145,146: ceil(S + shift0 * (first + (D-1)*step)) */
uint32_t orc_dispersed_samples(uint32_t samples, float shift0, float first_dm,
                               float dm_step, uint32_t nr_dms);

/* The per-(channel, DM) integer shift of include/Dedispersion.hpp:81,124,168. */
uint32_t orc_shift(float first_dm, float dm_step, uint32_t dm, float table);

/* Input elements per channel row: pad(S_disp / (8 / inputBits), padding)
 * bytes as include/Dedispersion.hpp:90,97,133,140 pad them, with the
 * division rounded up so that a row holds all S_disp samples; for float
 * input (inputBits 32) pad(S_disp, padding / 4) floats.  padding in bytes. */
uint64_t orc_input_pitch(uint32_t dispersed_samples, uint32_t input_bits, uint32_t padding);

/* include/Dedispersion.hpp:70-109.
 * Rows [row_begin, row_end) of the flattened (sBeam, dm) output are
 * computed; pass 0 and sBeams*D for the whole output. */
void orc_dedispersion(const orc_obs *o, const uint32_t *zapped,
                      const uint32_t *beam_mapping, const void *input,
                      float *output, const float *shifts, uint32_t padding,
                      uint64_t row_begin, uint64_t row_end);

/* include/Dedispersion.hpp:111-153 (the delay per o->step_one_absolute).
 * Rows are the flattened (beam, coarse dm) pairs. */
void orc_step_one(const orc_obs *o, const uint32_t *zapped,
                  const void *input, float *output, const float *shifts,
                  uint32_t padding, uint64_t row_begin, uint64_t row_end);

/* include/Dedispersion.hpp:155-176.  Rows are the flattened
 * (sBeam, coarse dm, fine dm) triples. */
void orc_step_two(const orc_obs *o, const uint32_t *beam_mapping,
                  const float *input, float *output, const float *shifts,
                  uint32_t padding, uint64_t row_begin, uint64_t row_end);

/* Same results as the three functions above, rows split over nthreads
 * POSIX threads (used only to time the CPU baseline). */
void orc_dedispersion_mt(const orc_obs *o, const uint32_t *zapped,
                         const uint32_t *beam_mapping, const void *input,
                         float *output, const float *shifts, uint32_t padding,
                         uint64_t row_begin, uint64_t row_end, int nthreads);
void orc_step_one_mt(const orc_obs *o, const uint32_t *zapped,
                     const void *input, float *output, const float *shifts,
                     uint32_t padding, uint64_t row_begin, uint64_t row_end,
                     int nthreads);
void orc_step_two_mt(const orc_obs *o, const uint32_t *beam_mapping,
                     const float *input, float *output, const float *shifts,
                     uint32_t padding, uint64_t row_begin, uint64_t row_end,
                     int nthreads);

/* The step after the path (no counterpart in the reference): for each of
 * `rows` integer-valued rows (values truncated to integers), the boxcar of
 * width 2^j (j < nr_widths) and start t (t + w <= samples) with the highest
 * (B - w * mean) / (sigma * sqrt(w)), mean and sigma (population) from the
 * exact integer sums; ties to the smaller width, then the smaller start.
 * out_* have `rows` entries. */
void orc_detect_peaks(const float *rows_in, uint64_t rows, uint32_t samples, uint64_t pitch,
                      uint32_t nr_widths, float *out_snr, float *out_mean, float *out_sigma,
                      uint32_t *out_sample, uint32_t *out_width);

#ifdef __cplusplus
}
#endif
#endif
