/*
 * dd_oracle.c -- CPU oracle for the dedispersion hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see dd_oracle.h).  A sequential restatement of
 * the reference's CPU functions; compile with -ffp-contract=off.
 *
 * Reference files (TRASAL/Dedispersion):
 *   src/Shifts.cpp:20-42          shift tables
 *   include/Dedispersion.hpp:70-176  the three CPU loops
 *   src/DedispersionTest.cpp:129,145,146  dispersed-batch sizing
 */
#include "dd_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

uint32_t orc_pad(uint32_t v, uint32_t k) {
  if (k == 0) return v;
  return ((v + k - 1) / k) * k;
}

/* Build-defined AstroData::Observation semantics (parity-unpinned):
 *   max frequency      = min + (C - 1) * bandwidth
 *   subband bandwidth  = channelsPerSubband * bandwidth
 *   subband frequency  = frequency of the subband's top channel, so
 *   subband min freq   = min + (channelsPerSubband - 1) * bandwidth
 *   subband max freq   = subband min freq + (Nsb - 1) * subband bandwidth */
uint32_t orc_channels_per_subband(const orc_obs *o) {
  return o->nr_channels / o->nr_subbands;
}

float orc_max_freq(const orc_obs *o) {
  float span = (float)(o->nr_channels - 1) * o->channel_bandwidth;
  return o->min_freq + span;
}

float orc_subband_bandwidth(const orc_obs *o) {
  return (float)orc_channels_per_subband(o) * o->channel_bandwidth;
}

float orc_subband_min_freq(const orc_obs *o) {
  float span = (float)(orc_channels_per_subband(o) - 1) * o->channel_bandwidth;
  return o->min_freq + span;
}

float orc_subband_max_freq(const orc_obs *o) {
  float span = (float)(o->nr_subbands - 1) * orc_subband_bandwidth(o);
  return orc_subband_min_freq(o) + span;
}

/* src/Shifts.cpp:20-30.  std::pow(x, 2.0f) is taken as x * x (the value
 * g++ produces at -O1 and above). */
void orc_get_shifts(const orc_obs *o, uint32_t padding, float *out) {
  const float max_freq = orc_max_freq(o);
  const float inverse_high_freq = 1.0f / (max_freq * max_freq);      /* :21 */
  const float time_unit = (float)o->nr_samples_per_batch * o->sampling_time; /* :22 */
  const uint32_t n = orc_pad(o->nr_channels, padding / sizeof(float));       /* :23 */
  const float samples = (float)(o->nr_samples_per_batch / o->downsampling);
  memset(out, 0, n * sizeof(float));
  for (uint32_t channel = 0; channel + 1 < o->nr_channels; channel++) {   /* :24 */
    float freq = o->min_freq + ((float)channel * o->channel_bandwidth);   /* :25 */
    float inverse_freq = 1.0f / (freq * freq);
    float delay = 4148.808f * (inverse_freq - inverse_high_freq);        /* :26 */
    delay = delay * samples;
    out[channel] = delay / time_unit;
  }
  out[o->nr_channels - 1] = 0.0f;                                       /* :28 */
}

/* src/Shifts.cpp:32-42 */
void orc_get_shifts_step_two(const orc_obs *o, uint32_t padding, float *out) {
  const float max_freq = orc_subband_max_freq(o);
  const float inverse_high_freq = 1.0f / (max_freq * max_freq);      /* :33 */
  const float time_unit = (float)o->nr_samples_per_batch * o->sampling_time; /* :34 */
  const uint32_t n = orc_pad(o->nr_subbands, padding / sizeof(float));       /* :35 */
  const float samples = (float)(o->nr_samples_per_batch / o->downsampling);
  const float sb_min = orc_subband_min_freq(o);
  const float sb_bw = orc_subband_bandwidth(o);
  memset(out, 0, n * sizeof(float));
  for (uint32_t subband = 0; subband + 1 < o->nr_subbands; subband++) {   /* :36 */
    float freq = sb_min + ((float)subband * sb_bw);                       /* :37 */
    float inverse_freq = 1.0f / (freq * freq);
    float delay = 4148.808f * (inverse_freq - inverse_high_freq);        /* :38 */
    delay = delay * samples;
    out[subband] = delay / time_unit;
  }
  out[o->nr_subbands - 1] = 0.0f;                                       /* :40 */
}

/* src/DedispersionTest.cpp:129 (and :145, :146 with the step-two and
 * step-one tables). */
uint32_t orc_dispersed_samples(uint32_t samples, float shift0, float first_dm,
                               float dm_step, uint32_t nr_dms) {
  float last_dm = first_dm + ((float)(nr_dms - 1) * dm_step);
  float total = (float)samples + (shift0 * last_dm);
  return (uint32_t)ceilf(total);
}

/* include/Dedispersion.hpp:81 / :124 / :168 */
uint32_t orc_shift(float first_dm, float dm_step, uint32_t dm, float table) {
  float dm_value = first_dm + ((float)dm * dm_step);
  float delay = dm_value * table;
  return (uint32_t)delay;
}

/* include/Dedispersion.hpp:94-101: the inputBits-bit field of sample x of a
 * packed row, copied bit by bit, least significant bit first, into a value
 * whose higher bits are zero.  Byte x / (8 / inputBits) holds the sample,
 * starting at bit (x % (8 / inputBits)) * inputBits. */
static uint32_t orc_packed_sample(const uint8_t *row, uint64_t x, uint32_t bits) {
  const uint32_t per_byte = 8 / bits;
  const uint8_t buffer = row[x / per_byte];
  const uint32_t first_bit = (uint32_t)(x % per_byte) * bits;
  uint32_t value = 0;
  for (uint32_t bit = 0; bit < bits; bit++) value |= ((uint32_t)(buffer >> (first_bit + bit)) & 1u) << bit;
  return value;
}

/* Bits per input sample: 1, 2, 4 (packed bytes), 8 (bytes) or 32 (floats,
 * the reference's templates with I = float). */
static uint32_t orc_bits(const orc_obs *o) {
  return o->input_bits == 0 ? 8u : o->input_bits;
}

/* Elements of the input type per row: pad(S_disp, padding / sizeof(I)) for
 * whole elements (include/Dedispersion.hpp:90, :133), packed bytes rounded up
 * for sub-byte samples (:97, :140 round down).  padding in bytes. */
uint64_t orc_input_pitch(uint32_t dispersed_samples, uint32_t input_bits, uint32_t padding) {
  if (input_bits == 32) return orc_pad(dispersed_samples, padding / sizeof(float));
  if (input_bits == 0 || input_bits >= 8) return orc_pad(dispersed_samples, padding);
  const uint32_t per_byte = 8 / input_bits;
  return orc_pad((dispersed_samples + per_byte - 1) / per_byte, padding);
}

/* Sample x of an input row as the accumulator type (:90-91 for whole
 * elements of uint8 or float input, :94-101 for packed fields). */
static float orc_sample(const void *input, uint64_t row_offset, uint64_t x, uint32_t bits) {
  if (bits == 32) return ((const float *)input)[row_offset + x];
  const uint8_t *row = (const uint8_t *)input + row_offset;
  if (bits == 8) return (float)row[x];
  return (float)orc_packed_sample(row, x, bits);
}


/* include/Dedispersion.hpp:70-109: the inputBits >= 8 branch (:88-91) and the
 * sub-byte branch (:92-103).  The sub-byte branch of the reference indexes
 * the input with the mapped beam number alone (:97, no beam stride), which
 * only reads the right row for beam 0; here both branches use the beam
 * stride of :90. */
void orc_dedispersion(const orc_obs *o, const uint32_t *zapped,
                      const uint32_t *beam_mapping, const void *input,
                      float *output, const float *shifts, uint32_t padding,
                      uint64_t row_begin, uint64_t row_end) {
  const uint32_t C = o->nr_channels;
  const uint32_t D = o->nr_dms;
  const uint32_t samples = o->nr_samples_per_batch / o->downsampling;      /* :76 */
  const uint32_t bits = orc_bits(o);
  const uint64_t in_pitch = orc_input_pitch(o->nr_samples_per_dispersed_batch, bits, padding);
  const uint64_t out_pitch = orc_pad(samples, padding / sizeof(float));
  const uint64_t map_pitch = orc_pad(C, padding / sizeof(uint32_t));
  for (uint64_t row = row_begin; row < row_end; row++) {
    const uint64_t sbeam = row / D;                                      /* :72 */
    const uint32_t dm = (uint32_t)(row % D);                             /* :74 */
    for (uint32_t sample = 0; sample < samples; sample++) {
      float acc = 0.0f;                                                  /* :78 */
      for (uint32_t channel = 0; channel < C; channel++) {               /* :79 */
        uint32_t shift = orc_shift(o->first_dm, o->dm_step, dm, shifts[channel]); /* :81 */
        if (zapped[channel] != 0) continue;                              /* :83 */
        uint64_t beam = beam_mapping[sbeam * map_pitch + channel];
        const uint64_t row = beam * C * in_pitch + (uint64_t)channel * in_pitch; /* :90 */
        acc += orc_sample(input, row, (uint64_t)sample + shift, bits);   /* :90, :94-101 */
      }
      output[sbeam * D * out_pitch + (uint64_t)dm * out_pitch + sample] = acc; /* :105 */
    }
  }
}

/* include/Dedispersion.hpp:111-153: the inputBits >= 8 branch (:131-134)
 * and the sub-byte branch (:135-146).  The output rows are padded as the
 * 8-bit case pads them; the reference pads S1 / (8 / inputBits) there (:148),
 * which makes consecutive sub-byte rows overlap.  A channel's delay is its
 * table entry minus that of its subband's top channel (:124); with
 * step_one_absolute it is the table entry itself, the delay the reference's
 * OpenCL step-one kernel applies (:595-597, :731, :733). */
void orc_step_one(const orc_obs *o, const uint32_t *zapped,
                  const void *input, float *output, const float *shifts,
                  uint32_t padding, uint64_t row_begin, uint64_t row_end) {
  const uint32_t C = o->nr_channels;
  const uint32_t Nsb = o->nr_subbands;
  const uint32_t cps = orc_channels_per_subband(o);
  const uint32_t Dc = o->nr_dms_sub;
  const uint32_t samples = o->nr_samples_per_batch_sub / o->downsampling; /* :119 */
  const uint32_t bits = orc_bits(o);
  const uint64_t in_pitch = orc_input_pitch(o->nr_samples_per_dispersed_batch_sub, bits, padding);
  const uint64_t out_pitch = orc_pad(o->nr_samples_per_batch_sub, padding / sizeof(float));
  for (uint64_t row = row_begin; row < row_end; row++) {
    const uint64_t beam = row / Dc;                                      /* :113 */
    const uint32_t dm = (uint32_t)(row % Dc);                            /* :115 */
    for (uint32_t subband = 0; subband < Nsb; subband++) {               /* :117 */
      const float top = shifts[((subband + 1) * cps) - 1];
      for (uint32_t sample = 0; sample < samples; sample++) {
        float acc = 0.0f;
        for (uint32_t channel = subband * cps; channel < (subband + 1) * cps; channel++) { /* :122 */
          float delay = o->step_one_absolute ? shifts[channel]             /* :595-597 */
                                             : shifts[channel] - top;    /* :124 */
          uint32_t shift = orc_shift(o->first_dm_sub, o->dm_step_sub, dm, delay);
          if (zapped[channel] != 0) continue;                            /* :126 */
          const uint64_t row = beam * C * in_pitch + (uint64_t)channel * in_pitch; /* :133, :140 */
          acc += orc_sample(input, row, (uint64_t)sample + shift, bits); /* :133, :137-144 */
        }
        output[beam * Dc * Nsb * out_pitch + (uint64_t)dm * Nsb * out_pitch +
               (uint64_t)subband * out_pitch + sample] = acc;           /* :148 */
      }
    }
  }
}

/* include/Dedispersion.hpp:155-176 */
void orc_step_two(const orc_obs *o, const uint32_t *beam_mapping,
                  const float *input, float *output, const float *shifts,
                  uint32_t padding, uint64_t row_begin, uint64_t row_end) {
  const uint32_t Nsb = o->nr_subbands;
  const uint32_t Dc = o->nr_dms_sub;
  const uint32_t D = o->nr_dms;
  const uint32_t samples = o->nr_samples_per_batch / o->downsampling;      /* :163 */
  const uint64_t in_pitch = orc_pad(o->nr_samples_per_batch_sub / o->downsampling, padding / sizeof(float));
  const uint64_t out_pitch = orc_pad(samples, padding / sizeof(float));
  const uint64_t map_pitch = orc_pad(Nsb, padding / sizeof(uint32_t));
  for (uint64_t row = row_begin; row < row_end; row++) {
    const uint64_t sbeam = row / ((uint64_t)Dc * D);                      /* :157 */
    const uint32_t coarse = (uint32_t)((row / D) % Dc);                  /* :159 */
    const uint32_t dm = (uint32_t)(row % D);                             /* :161 */
    for (uint32_t sample = 0; sample < samples; sample++) {
      float acc = 0.0f;
      for (uint32_t subband = 0; subband < Nsb; subband++) {             /* :166 */
        uint32_t shift = orc_shift(o->first_dm, o->dm_step, dm, shifts[subband]); /* :168 */
        uint64_t beam = beam_mapping[sbeam * map_pitch + subband];
        uint64_t idx = beam * Dc * Nsb * in_pitch + (uint64_t)coarse * Nsb * in_pitch +
                       (uint64_t)subband * in_pitch + sample + shift;  /* :169 */
        acc += input[idx];
      }
      output[sbeam * ((uint64_t)Dc * D) * out_pitch + ((uint64_t)coarse * D + dm) * out_pitch + sample] = acc; /* :171 */
    }
  }
}

/* ---- row-parallel drivers (bit-identical: each row is computed by exactly
 * the same sequential code, rows are independent) ---- */

typedef struct mt_job {
  int kind;
  const orc_obs *o;
  const uint32_t *zapped;
  const uint32_t *beam_mapping;
  const void *input;
  float *output;
  const float *shifts;
  uint32_t padding;
  uint64_t begin, end;
} mt_job;

static void *mt_run(void *arg) {
  mt_job *j = (mt_job *)arg;
  if (j->kind == 0)
    orc_dedispersion(j->o, j->zapped, j->beam_mapping, j->input,
                     j->output, j->shifts, j->padding, j->begin, j->end);
  else if (j->kind == 1)
    orc_step_one(j->o, j->zapped, j->input, j->output,
                 j->shifts, j->padding, j->begin, j->end);
  else
    orc_step_two(j->o, j->beam_mapping, (const float *)j->input, j->output,
                 j->shifts, j->padding, j->begin, j->end);
  return NULL;
}

static void mt_dispatch(mt_job proto, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  uint64_t rows = proto.end - proto.begin;
  if ((uint64_t)nthreads > rows) nthreads = rows ? (int)rows : 1;
  pthread_t *tids = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
  mt_job *jobs = (mt_job *)calloc((size_t)nthreads, sizeof(mt_job));
  for (int t = 0; t < nthreads; t++) {
    jobs[t] = proto;
    jobs[t].begin = proto.begin + rows * (uint64_t)t / (uint64_t)nthreads;
    jobs[t].end = proto.begin + rows * (uint64_t)(t + 1) / (uint64_t)nthreads;
    pthread_create(&tids[t], NULL, mt_run, &jobs[t]);
  }
  for (int t = 0; t < nthreads; t++) pthread_join(tids[t], NULL);
  free(tids);
  free(jobs);
}

void orc_dedispersion_mt(const orc_obs *o, const uint32_t *zapped,
                         const uint32_t *beam_mapping, const void *input,
                         float *output, const float *shifts, uint32_t padding,
                         uint64_t row_begin, uint64_t row_end, int nthreads) {
  mt_job j = {0, o, zapped, beam_mapping, input, output, shifts, padding, row_begin, row_end};
  mt_dispatch(j, nthreads);
}

void orc_step_one_mt(const orc_obs *o, const uint32_t *zapped,
                     const void *input, float *output, const float *shifts,
                     uint32_t padding, uint64_t row_begin, uint64_t row_end,
                     int nthreads) {
  mt_job j = {1, o, zapped, NULL, input, output, shifts, padding, row_begin, row_end};
  mt_dispatch(j, nthreads);
}

void orc_step_two_mt(const orc_obs *o, const uint32_t *beam_mapping,
                     const float *input, float *output, const float *shifts,
                     uint32_t padding, uint64_t row_begin, uint64_t row_end,
                     int nthreads) {
  mt_job j = {2, o, NULL, beam_mapping, input, output, shifts, padding, row_begin, row_end};
  mt_dispatch(j, nthreads);
}

/* Boxcar peak search, straight loops: every window of every width is summed
 * from the samples (no prefix sums), so a mistake in a running sum of the
 * device kernel cannot be shared here. */
void orc_detect_peaks(const float *rows_in, uint64_t rows, uint32_t samples, uint64_t pitch,
                      uint32_t nr_widths, float *out_snr, float *out_mean, float *out_sigma,
                      uint32_t *out_sample, uint32_t *out_width) {
  for (uint64_t r = 0; r < rows; r++) {
    const float *y = rows_in + r * pitch;
    int64_t sum = 0, sq = 0;
    for (uint32_t t = 0; t < samples; t++) {
      const int64_t v = (int64_t)(int32_t)y[t];
      sum += v;
      sq += v * v;
    }
    const double dS = (double)samples;
    const double mean = (double)sum / dS;
    double var = (double)sq / dS - mean * mean;
    if (!(var > 0.0)) var = 0.0;
    const double sigma = sqrt(var);
    double best = 0.0;
    uint32_t best_t = 0, best_w = 0;
    int found = 0;
    for (uint32_t j = 0; j < nr_widths; j++) {
      const uint32_t w = 1u << j;
      if (w > samples) break;
      int64_t bmax = -1;
      uint32_t bt = 0;
      for (uint32_t t = 0; t + w <= samples; t++) {
        int64_t b = 0;
        for (uint32_t i = 0; i < w; i++) b += (int64_t)(int32_t)y[t + i];
        if (b > bmax) {
          bmax = b;
          bt = t;
        }
      }
      const double width = (double)w;
      const double snr = sigma > 0.0 ? ((double)bmax - width * mean) / (sigma * sqrt(width)) : 0.0;
      if (!found || snr > best) {
        found = 1;
        best = snr;
        best_t = bt;
        best_w = w;
      }
    }
    out_snr[r] = (float)best;
    out_mean[r] = (float)mean;
    out_sigma[r] = (float)sigma;
    out_sample[r] = best_t;
    out_width[r] = best_w;
  }
}
