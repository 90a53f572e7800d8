// plan.cpp -- the plan object behind the C ABI (dedispersion_amd.h).
//
// A plan fixes the observation and the padding, owns the device-resident
// tables and turns each reference launch (src/DedispersionTest.cpp:367-399)
// into one HIP kernel launch with validated geometry.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "plan_internal.hpp"

namespace ddamd {

static thread_local std::string g_last_error;

void set_error(const std::string &message) { g_last_error = message; }

dd_status fail(dd_status s, const std::string &msg) {
  set_error(msg);
  return s;
}

dd_status hip_fail(hipError_t e, const char *what) {
  return fail(DD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

namespace {

uint32_t cps(const dd_plan *p) { return p->obs.nr_channels / p->obs.nr_subbands; }

// Observation getters with the semantics stated in include/Observation.hpp.
float max_freq(const dd_observation &o) {
  float span = (float)(o.nr_channels - 1) * o.channel_bandwidth;
  return o.min_freq + span;
}
float subband_bandwidth(const dd_observation &o) {
  return (float)(o.nr_channels / o.nr_subbands) * o.channel_bandwidth;
}
float subband_min_freq(const dd_observation &o) {
  float span = (float)(o.nr_channels / o.nr_subbands - 1) * o.channel_bandwidth;
  return o.min_freq + span;
}
float subband_max_freq(const dd_observation &o) {
  float span = (float)(o.nr_subbands - 1) * subband_bandwidth(o);
  return subband_min_freq(o) + span;
}

// log2 of the samples per input byte: 0 for bytes (and float input), 1, 2, 3
// for 4-, 2- and 1-bit samples.
uint32_t input_log2f(const dd_plan *p) {
  return p->input_bits == 1 ? 3u : p->input_bits == 2 ? 2u : p->input_bits == 4 ? 1u : 0u;
}

// Input row pitch and sizes, in elements of the input type.  A row is
// pad(S_disp / (8 / inputBits), padding) bytes (include/Dedispersion.hpp:90,
// :97, :133, :140), the division rounded up so that the row holds every
// sample when 8 / inputBits does not divide S_disp; float rows are
// pad(S_disp, padding / 4) floats.
void size_inputs(dd_plan *p) {
  dd_sizes &z = p->sizes;
  const uint32_t per = 1u << input_log2f(p);
  const uint32_t pad_elems = p->input_bits == 32 ? p->padding / 4 : p->padding;
  const uint64_t C = p->obs.nr_channels;
  z.input_bits = p->input_bits;
  z.in_pitch = padto((z.nr_dispersed_samples + per - 1) / per, pad_elems);
  z.input_elements = (uint64_t)p->obs.nr_beams * C * z.in_pitch;
  if (p->two_step) {
    z.sub_in_pitch = padto((z.nr_dispersed_samples_subbanding + per - 1) / per, pad_elems);
    z.sub_input_elements = (uint64_t)p->obs.nr_beams * C * z.sub_in_pitch;
  }
}

// src/DedispersionTest.cpp:129 / :145 / :146
uint32_t dispersed(uint32_t samples, float shift0, float first, float step, uint32_t dms) {
  float last = first + ((float)(uint32_t)(dms - 1) * step);
  float total = (float)samples + (shift0 * last);
  return (uint32_t)std::ceil(total);
}

dd_status build_integer(dd_plan *p, ShiftTable &t, const float *d_table, uint32_t rows, uint32_t dms,
                        float first, float step) {
  t.rows = rows;
  t.dms = dms;
  t.spread_cache.clear();
  t.pair_cache.clear();
  t.pitch = roundup(std::max(dms, 1u), 64) + 64;
  DD_HIP(t.dev.allocate((size_t)rows * t.pitch));
  if (dms == 0 || rows == 0) {
    t.host.assign((size_t)rows * t.pitch, 0);
    return DD_OK;
  }
  DD_HIP(launch_integer_shifts(d_table, rows, dms, t.pitch, first, step, t.dev.ptr, nullptr));
  DD_HIP(hipDeviceSynchronize());
  DD_HIP(t.dev.download(t.host));
  (void)p;
  return DD_OK;
}

dd_status rebuild_integer_tables(dd_plan *p) {
  const dd_observation &o = p->obs;
  dd_status s = build_integer(p, p->ish_single, p->shifts.ptr, o.nr_channels, o.nr_dms, o.first_dm, o.dm_step);
  if (s != DD_OK) return s;
  if (p->two_step) {
    DD_HIP(p->relative.allocate(o.nr_channels));
    DD_HIP(launch_relative_shifts(p->shifts.ptr, o.nr_channels, cps(p), p->step_one_absolute, p->relative.ptr,
                                  nullptr));
    s = build_integer(p, p->ish_one, p->relative.ptr, o.nr_channels, o.nr_dms_subbanding, o.first_dm_subbanding,
                      o.dm_step_subbanding);
    if (s != DD_OK) return s;
    s = build_integer(p, p->ish_two, p->shifts2.ptr, o.nr_subbands, o.nr_dms, o.first_dm, o.dm_step);
    if (s != DD_OK) return s;
  }
  p->fact_valid = false;
  p->sizes.max_shift = p->ish_single.max_shift();
  p->sizes.max_shift_step_one = p->two_step ? p->ish_one.max_shift() : 0;
  p->sizes.max_shift_step_two = p->two_step ? p->ish_two.max_shift() : 0;
  return DD_OK;
}

dd_status rebuild_live(dd_plan *p) {
  p->fact_valid = false;
  const uint32_t C = p->obs.nr_channels;
  std::vector<uint32_t> single, count1(1, 0);
  for (uint32_t ch = 0; ch < C; ++ch)
    if (p->h_zapped[ch] == 0) single.push_back(ch);
  count1[0] = (uint32_t)single.size();
  if (single.empty()) single.push_back(0);
  DD_HIP(p->live_single.upload(single));
  DD_HIP(p->live_single_count.upload(count1));
  if (p->two_step) {
    const uint32_t n = cps(p), nsb = p->obs.nr_subbands;
    std::vector<uint32_t> sub((size_t)nsb * n, 0), counts(nsb, 0);
    for (uint32_t sb = 0; sb < nsb; ++sb)
      for (uint32_t ch = sb * n; ch < (sb + 1) * n; ++ch)
        if (p->h_zapped[ch] == 0) sub[(size_t)sb * n + counts[sb]++] = ch;
    DD_HIP(p->live_sub.upload(sub));
    DD_HIP(p->live_sub_count.upload(counts));
    p->max_live_sub = *std::max_element(counts.begin(), counts.end());
  }
  return DD_OK;
}

// Defaults: the fastest configurations of exhaustive sweeps on MI355X of the
// ARTS layouts -- the single step (1536 channels, 2048 DMs, 24576 samples)
// and the two-step path (32 subbands, 64 coarse x 32 fine DMs).
dd_conf default_conf(dd_mode mode) {
  dd_conf c{};
  c.threads_d0 = 64;
  c.split_batches = 0;
  c.local = 1;
  switch (mode) {
  case DD_SINGLE_STEP:
    c.threads_d1 = 8;
    c.items_d0 = 16;
    c.items_d1 = 8;
    c.unroll = 4;
    break;
  case DD_STEP_ONE:
    c.threads_d1 = 4;
    c.items_d0 = 16;
    c.items_d1 = 8;
    c.unroll = 4;
    break;
  default:
    c.threads_d1 = 8;
    c.items_d0 = 8;
    c.items_d1 = 4;
    c.unroll = 8;
    c.local = 2;
    break;
  }
  return c;
}

// Pattern records of the single step for DM tiles of kFactoredDMs DMs and
// blocks of four live channels (layout in internal.hpp).  fact_ok is false
// when a record would need more than kFactoredMaxPatterns patterns or an
// offset above 255.
dd_status build_factored(dd_plan *p) {
  if (p->fact_valid) return DD_OK;
  p->fact_valid = true;
  p->fact_ok = false;
  const ShiftTable &t = p->ish_single;
  std::vector<uint32_t> live;
  for (uint32_t ch = 0; ch < p->obs.nr_channels; ++ch)
    if (p->h_zapped[ch] == 0) live.push_back(ch);
  if (live.empty()) return DD_OK;
  const uint32_t D = p->obs.nr_dms, TD = kFactoredDMs, REC = kFactoredRecordDwords;
  const uint32_t ny = (D + TD - 1) / TD, nb = ((uint32_t)live.size() + 3) / 4;
  std::vector<uint32_t> rec((size_t)ny * nb * REC, 0);
  uint32_t maxp = 0, base_spread = 0, delta_spread = 0;
  for (uint32_t y = 0; y < ny; ++y) {
    const uint32_t dfirst = y * TD;
    for (uint32_t b = 0; b < nb; ++b) {
      const uint32_t nrows = std::min<uint32_t>(4, (uint32_t)live.size() - 4 * b);
      const uint32_t base = live[4 * b + nrows - 1];
      uint32_t *r = rec.data() + ((size_t)y * nb + b) * REC;
      const uint32_t s0 = t.at(base, dfirst);
      std::vector<uint32_t> pats;
      uint32_t mins = 0xFFFFFFFFu, maxs = 0;
      for (uint32_t j = 0; j < TD; ++j) {
        const uint32_t d = std::min(dfirst + j, D - 1);
        const uint32_t sb = t.at(base, d);
        if (sb < s0 || sb - s0 > 255) return DD_OK;
        uint32_t pat = 0;
        for (uint32_t i = 0; i < nrows; ++i) {
          const uint32_t si = t.at(live[4 * b + i], d);
          if (si < sb || si - sb > 255) return DD_OK;
          pat |= (si - sb) << (8 * i);
        }
        uint32_t idx = 0;
        while (idx < pats.size() && pats[idx] != pat) ++idx;
        if (idx == pats.size()) {
          if (pats.size() == kFactoredMaxPatterns) return DD_OK;
          pats.push_back(pat);
        }
        r[kFactoredDMWord + j / 2] |= ((sb - s0) | (idx << 8)) << (16 * (j & 1));
        base_spread = std::max(base_spread, sb - s0);
      }
      uint32_t lo = 0, hi = 0;
      for (uint32_t i = 0; i < 4; ++i) {
        uint32_t mn = 255, mx = 0;
        for (uint32_t q : pats) {
          const uint32_t v = i < nrows ? (q >> (8 * i)) & 0xFFu : 0;
          mn = std::min(mn, v);
          mx = std::max(mx, v);
        }
        lo |= mn << (8 * i);
        hi |= mx << (8 * i);
        delta_spread = std::max(delta_spread, mx - mn);
      }
      (void)mins;
      (void)maxs;
      r[0] = (uint32_t)pats.size() | (nrows << 8);
      r[1] = s0;
      r[2] = lo;
      r[3] = hi;
      for (size_t q = 0; q < pats.size(); ++q) r[kFactoredPatternWord + q] = pats[q];
      maxp = std::max<uint32_t>(maxp, (uint32_t)pats.size());
    }
  }
  DD_HIP(p->fact_records.upload(rec));
  p->fact_max_patterns = maxp;
  p->fact_base_spread = base_spread;
  p->fact_delta_spread = delta_spread;
  p->fact_ok = true;
  return DD_OK;
}

} // namespace

dd_status geometry_of(const dd_plan *p, dd_mode mode, const dd_conf &c, Geometry &g) {
  if (c.split_batches) return fail(DD_ERR_UNSUPPORTED, "split batches mode is not implemented");
  if (c.threads_d0 != 64) return fail(DD_ERR_CONFIGURATION, "threads_d0 must be 64 (one wavefront along samples)");
  if (c.threads_d1 < 1 || c.threads_d1 > 8) return fail(DD_ERR_CONFIGURATION, "threads_d1 must be in [1, 8]");
  if (c.unroll < 1 || c.unroll > 16) return fail(DD_ERR_CONFIGURATION, "unroll must be in [1, 16]");
  const uint32_t tile = c.threads_d1 * c.items_d1;
  if (tile > 64) return fail(DD_ERR_CONFIGURATION, "threads_d1 * items_d1 must not exceed 64 DMs");
  g.threads_d1 = c.threads_d1;
  g.items_d0 = c.items_d0;
  g.items_d1 = c.items_d1;
  g.packed = 0;
  g.pairs = 0;
  g.unroll = c.unroll;
  // rows of floats: step two, and the single step and step one over float input
  const bool float_rows = mode == DD_STEP_TWO || p->input_bits == 32;
  if (float_rows && mode != DD_STEP_TWO && c.local == 3) {
    // the channel-block pattern kernel sums bytes: float rows run the
    // staged pipeline of the same shape
    dd_conf staged = c;
    staged.local = 1;
    staged.unroll = std::min(c.unroll, kStagedSlotsPerWave * c.threads_d1);
    return geometry_of(p, mode, staged, g);
  }
  if (mode != DD_STEP_TWO && p->input_bits < 8 &&
      (c.local != 1 || c.unroll > kStagedSlotsPerWave * c.threads_d1)) {
    // packed rows are unpacked as they are staged: only the pipelined kernel
    // reads them, in the same shape
    dd_conf staged = c;
    staged.local = 1;
    staged.unroll = std::min(c.unroll, kStagedSlotsPerWave * c.threads_d1);
    return geometry_of(p, mode, staged, g);
  }
  if (c.local == 3) {
    // channel-block patterns: single step only, 32-DM tiles
    if (mode != DD_SINGLE_STEP) return fail(DD_ERR_CONFIGURATION, "local 3 (channel-block patterns) is single step only");
    if ((c.items_d0 != 8 && c.items_d0 != 16) || (c.items_d1 != 4 && c.items_d1 != 8) ||
        c.threads_d1 * c.items_d1 != kFactoredDMs)
      return fail(DD_ERR_CONFIGURATION,
                  "local 3 needs items_d0 8 or 16 and threads_d1 x items_d1 = 32 with items_d1 4 or 8");
    if (c.unroll > 8) return fail(DD_ERR_CONFIGURATION, "local 3 takes at most 8 channel blocks per phase (unroll)");
    dd_status s = build_factored(const_cast<dd_plan *>(p));
    if (s != DD_OK) return s;
    if (!p->fact_ok) {
      // too many patterns or too large offsets: the staged pipeline, same shape
      dd_conf staged = c;
      staged.local = 1;
      staged.unroll = std::min(c.unroll, kStagedSlotsPerWave * c.threads_d1);
      return geometry_of(p, mode, staged, g);
    }
    const uint32_t ts = 64 * c.items_d0;
    const uint32_t qw = roundup(ts + p->fact_base_spread + 2, 16);
    const uint32_t rw = roundup(qw + p->fact_delta_spread + 3 + 8, 16);
    const uint32_t slots = std::max<uint32_t>(p->fact_max_patterns, 4 * c.unroll);
    g.threads_d1 = c.threads_d1;
    g.items_d0 = c.items_d0;
    g.items_d1 = c.items_d1;
    g.local = 1;
    g.staged = 2;
    g.win = rw;
    g.lds_bytes = 4 * (4 * c.unroll * (rw / 4) + slots * qw + c.unroll * kFactoredRecordDwords + 4 * c.unroll);
    if (rw / 16 > 64 * 8 || g.lds_bytes > p->max_lds) {
      // the phase does not fit in LDS: the staged pipeline, same shape
      dd_conf staged = c;
      staged.local = 1;
      staged.unroll = std::min(c.unroll, kStagedSlotsPerWave * c.threads_d1);
      return geometry_of(p, mode, staged, g);
    }
    return DD_OK;
  }
  if (c.local > 2) return fail(DD_ERR_CONFIGURATION, "local must be 0, 1, 2 or 3");
  g.local = c.local ? 1 : 0;
  g.win = 0;
  g.lds_bytes = 0;
  g.staged = 0;
  const uint32_t ts = 64 * c.items_d0;
  const bool stageable = c.local == 1 && c.unroll <= kStagedSlotsPerWave * c.threads_d1;
  if (float_rows) {
    if (!rowsum_f32_supported(c.items_d0, c.items_d1))
      return fail(DD_ERR_CONFIGURATION, "items_d0 must be 4, 8 or 16 and items_d1 1, 2, 4 or 8");
    if (g.local) {
      // a lane reads floats [spread + lane + 64k] for k < items_d0
      const ShiftTable &t = mode == DD_SINGLE_STEP ? p->ish_single : mode == DD_STEP_ONE ? p->ish_one : p->ish_two;
      g.win = roundup(ts + t.max_spread(tile), 4);
      g.lds_bytes = c.unroll * g.win * 4;
      if (stageable && g.win / 4 <= 64 * f32_staged_chunks(c.items_d0)) {
        // rows (row | beam, delay) of a slice | [2 buffers][unroll rows][win] | DM offsets;
        // rows staged per barrier are capped at what fits in LDS
        const uint32_t slots = mode == DD_STEP_TWO ? p->obs.nr_subbands
                               : mode == DD_SINGLE_STEP ? p->obs.nr_channels : cps(p);
        auto lds_of = [&](uint32_t u) { return 4 * (roundup(2 * slots, 4) + 2 * u * g.win + 2 * u * tile); };
        uint32_t unroll = c.unroll;
        while (unroll > 1 && lds_of(unroll) > p->max_lds) --unroll;
        if (lds_of(unroll) <= p->max_lds) {
          g.staged = 1;
          g.unroll = unroll;
          g.lds_bytes = lds_of(unroll);
        }
      }
    }
  } else {
    if (!rowsum_u8_supported(c.items_d0, c.items_d1))
      return fail(DD_ERR_CONFIGURATION, "items_d0 must be 4, 8 or 16 and items_d1 1, 2, 4 or 8");
    if (g.local) {
      // window start misalignment (< 4 bytes, each of `per` samples) + spread
      // + the dword after each group; whole 16-byte chunks of packed input
      const ShiftTable &t = mode == DD_SINGLE_STEP ? p->ish_single : p->ish_one;
      const uint32_t per = 1u << input_log2f(p);
      g.win = roundup(ts + t.max_spread(tile) + 4 * per + 4, 16 * per);
      g.lds_bytes = c.unroll * g.win;
      if (stageable && g.win <= 16 * 64 * kU8StagedChunks) {
        const uint32_t live_pitch = mode == DD_SINGLE_STEP ? p->obs.nr_channels : cps(p);
        // rows | [2 buffers][unroll rows][4 shifted copies][win / 4 dwords] |
        // DM words (packed rows are unpacked before the copies are built);
        // rows staged per barrier are capped at what fits in LDS
        auto lds_of = [&](uint32_t u) { return 4 * (roundup(2 * live_pitch, 4) + 2 * u * 4 * (g.win / 4) + 2 * u * tile); };
        uint32_t unroll = c.unroll;
        while (unroll > 1 && lds_of(unroll) > p->max_lds) --unroll;
        if (lds_of(unroll) > p->max_lds && per == 1) {
          // not even one row per barrier fits beside the row table: the
          // one-group staging (or the direct kernel) below
        } else {
          g.staged = 1;
          g.packed = per > 1;
          g.unroll = unroll;
          g.lds_bytes = lds_of(unroll);
        }
      } else if (per > 1) {
        return fail(DD_ERR_CONFIGURATION, "packed input needs the pipelined kernel, whose row window is at most " +
                                              std::to_string(16 * 64 * kU8StagedChunks) + " bytes; this one needs " +
                                              std::to_string(g.win) + " (fewer samples per tile or DMs per tile)");
      }
    }
  }
  if (g.local && !g.staged && g.lds_bytes > p->max_lds) {
    // the rows' windows (tile plus the spread of the tile's shifts) do not
    // fit in LDS: the direct kernel, same shape, reads the rows through the
    // caches (the same sums)
    g.local = 0;
    g.win = 0;
    g.lds_bytes = 0;
  }
  if (g.lds_bytes > p->max_lds)
    return fail(DD_ERR_CONFIGURATION, "LDS demand " + std::to_string(g.lds_bytes) + " bytes exceeds the device limit");
  return DD_OK;
}

namespace {
bool aligned16(const void *ptr) { return (reinterpret_cast<uintptr_t>(ptr) & 15) == 0; }
} // namespace

dd_status prepare_single(dd_plan *plan, const dd_conf &c, const uint8_t *d_input, float *d_output,
                         uint32_t first_sbeam, uint32_t nr_sbeams, U8Launch &L) {
  if (!plan || !d_input || !d_output) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  if (plan->input_bits == 32) return fail(DD_ERR_INVALID_ARGUMENT, "the plan takes float input (dd_dedisperse_f32)");
  const dd_observation &o = plan->obs;
  const dd_sizes &z = plan->sizes;
  if (nr_sbeams == 0) nr_sbeams = o.nr_synthesized_beams - std::min(first_sbeam, o.nr_synthesized_beams);
  if (nr_sbeams == 0 || first_sbeam + nr_sbeams > o.nr_synthesized_beams)
    return fail(DD_ERR_INVALID_ARGUMENT, "synthesized beam range out of bounds");
  if ((uint64_t)z.nr_samples + z.max_shift > z.nr_dispersed_samples)
    return fail(DD_ERR_INVALID_ARGUMENT, "dispersed batch shorter than samples + largest shift");
  dd_status s = geometry_of(plan, DD_SINGLE_STEP, c, L.g);
  if (s != DD_OK) return s;
  RowSumU8Params &k = L.k;
  k = RowSumU8Params{};
  k.in = d_input;
  k.in_bytes = z.input_elements;
  k.out = d_output;
  k.ishift = plan->ish_single.dev.ptr;
  k.dm_pitch = plan->ish_single.pitch;
  k.live = plan->live_single.ptr;
  k.live_count = plan->live_single_count.ptr;
  k.live_pitch = o.nr_channels;
  k.map = plan->map_single.ptr;
  k.map_pitch = padto(o.nr_channels, plan->padding / 4);
  k.first_sbeam = first_sbeam;
  k.in_row_pitch = z.in_pitch;
  k.in_beam_pitch = (uint64_t)o.nr_channels * z.in_pitch;
  k.in_log2f = input_log2f(plan);
  k.out_dm_pitch = z.out_pitch;
  k.out_beam_pitch = (uint64_t)o.nr_dms * z.out_pitch;
  k.out_group_pitch = 0;
  k.n_samples = z.nr_samples;
  k.n_dms = o.nr_dms;
  k.n_groups = 1;
  k.unroll = L.g.unroll;
  k.win_bytes = L.g.win;
  k.nontemporal = 1;
  k.vec_store = (z.out_pitch % 4 == 0) && aligned16(d_output);
  k.table = plan->shifts.ptr;
  k.first_dm = o.first_dm;
  k.dm_step = o.dm_step;
  const uint32_t ts = 64 * c.items_d0, td = c.threads_d1 * c.items_d1;
  L.grid = dim3((z.nr_samples + ts - 1) / ts, (o.nr_dms + td - 1) / td, nr_sbeams);
  L.flush = o.nr_channels > 257;
  if (L.g.staged == 2) {
    k.fact_records = plan->fact_records.ptr;
    k.fact_slots = std::max<uint32_t>(plan->fact_max_patterns, 4 * L.g.unroll);
    k.fact_qw = roundup(ts + plan->fact_base_spread + 2, 16);
    k.vec_store = (z.out_pitch % 2 == 0) && (reinterpret_cast<uintptr_t>(d_output) & 7) == 0;
  }
  k.n_tiles_x = L.grid.x;
  k.n_tiles_y = L.grid.y;
  k.n_tiles = L.grid.x * L.grid.y * L.grid.z;
  return DD_OK;
}

dd_status prepare_step_one(dd_plan *plan, const dd_conf &c, const uint8_t *d_input, float *d_intermediate,
                           U8Launch &L) {
  if (!plan || !d_input || !d_intermediate) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  if (!plan->two_step) return fail(DD_ERR_INVALID_ARGUMENT, "plan has no subband path");
  if (plan->input_bits == 32)
    return fail(DD_ERR_INVALID_ARGUMENT, "the plan takes float input (dd_subband_step_one_f32)");
  const dd_observation &o = plan->obs;
  const dd_sizes &z = plan->sizes;
  if ((uint64_t)z.nr_samples_step_one + z.max_shift_step_one > z.nr_dispersed_samples_subbanding)
    return fail(DD_ERR_INVALID_ARGUMENT, "dispersed batch shorter than samples + largest shift");
  dd_status s = geometry_of(plan, DD_STEP_ONE, c, L.g);
  if (s != DD_OK) return s;
  RowSumU8Params &k = L.k;
  k = RowSumU8Params{};
  k.in = d_input;
  k.in_bytes = z.sub_input_elements;
  k.out = d_intermediate;
  k.ishift = plan->ish_one.dev.ptr;
  k.dm_pitch = plan->ish_one.pitch;
  k.live = plan->live_sub.ptr;
  k.live_count = plan->live_sub_count.ptr;
  k.live_pitch = cps(plan);
  k.map = nullptr;
  k.map_pitch = 0;
  k.first_sbeam = 0;
  k.in_row_pitch = z.sub_in_pitch;
  k.in_beam_pitch = (uint64_t)o.nr_channels * z.sub_in_pitch;
  k.in_log2f = input_log2f(plan);
  k.out_dm_pitch = (uint64_t)o.nr_subbands * z.sub_out_pitch;
  k.out_beam_pitch = (uint64_t)o.nr_dms_subbanding * o.nr_subbands * z.sub_out_pitch;
  k.out_group_pitch = z.sub_out_pitch;
  k.n_samples = z.nr_samples_step_one;
  k.n_dms = o.nr_dms_subbanding;
  k.n_groups = o.nr_subbands;
  k.unroll = L.g.unroll;
  k.win_bytes = L.g.win;
  k.nontemporal = 0;  // step two re-reads the intermediate
  k.vec_store = (z.sub_out_pitch % 4 == 0) && aligned16(d_intermediate);
  k.table = plan->relative.ptr;
  k.first_dm = o.first_dm_subbanding;
  k.dm_step = o.dm_step_subbanding;
  const uint32_t ts = 64 * c.items_d0, td = c.threads_d1 * c.items_d1;
  L.grid = dim3((z.nr_samples_step_one + ts - 1) / ts, (o.nr_dms_subbanding + td - 1) / td,
                o.nr_beams * o.nr_subbands);
  L.flush = cps(plan) > 257;
  k.n_tiles_x = L.grid.x;
  k.n_tiles_y = L.grid.y;
  k.n_tiles = L.grid.x * L.grid.y * L.grid.z;
  return DD_OK;
}

dd_status prepare_step_two(dd_plan *plan, const dd_conf &c, const float *d_intermediate, float *d_output,
                           uint32_t first_sbeam, uint32_t nr_sbeams, F32Launch &L) {
  if (!plan || !d_intermediate || !d_output) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  if (!plan->two_step) return fail(DD_ERR_INVALID_ARGUMENT, "plan has no subband path");
  const dd_observation &o = plan->obs;
  const dd_sizes &z = plan->sizes;
  if (nr_sbeams == 0) nr_sbeams = o.nr_synthesized_beams - std::min(first_sbeam, o.nr_synthesized_beams);
  if (nr_sbeams == 0 || first_sbeam + nr_sbeams > o.nr_synthesized_beams)
    return fail(DD_ERR_INVALID_ARGUMENT, "synthesized beam range out of bounds");
  if ((uint64_t)z.nr_samples + z.max_shift_step_two > z.nr_samples_step_one)
    return fail(DD_ERR_INVALID_ARGUMENT, "subbanded batch shorter than samples + largest shift");
  dd_status s = geometry_of(plan, DD_STEP_TWO, c, L.g);
  if (s != DD_OK) return s;
  RowSumF32Params &k = L.k;
  k = RowSumF32Params{};
  k.in = d_intermediate;
  k.in_floats = z.step_two_input_elements;
  k.out = d_output;
  k.ishift = plan->ish_two.dev.ptr;
  k.dm_pitch = plan->ish_two.pitch;
  k.map = plan->map_two.ptr;
  k.map_pitch = padto(o.nr_subbands, plan->padding / 4);
  k.first_sbeam = first_sbeam;
  k.n_rows = o.nr_subbands;
  k.n_coarse = o.nr_dms_subbanding;
  k.in_row_pitch = z.step_two_in_pitch;
  k.in_coarse_pitch = (uint64_t)o.nr_subbands * z.step_two_in_pitch;
  k.in_beam_pitch = (uint64_t)o.nr_dms_subbanding * o.nr_subbands * z.step_two_in_pitch;
  k.out_dm_pitch = z.out_pitch;
  k.out_beam_pitch = (uint64_t)o.nr_dms_subbanding * o.nr_dms * z.out_pitch;
  k.out_inner_pitch = (uint64_t)o.nr_dms * z.out_pitch;
  k.n_samples = z.nr_samples;
  k.n_dms = o.nr_dms;
  k.unroll = L.g.unroll;
  k.win_floats = L.g.win;
  k.nontemporal = 1;
  k.vec_store = (z.out_pitch % 4 == 0) && aligned16(d_output);
  k.table = plan->shifts2.ptr;
  k.first_dm = o.first_dm;
  k.dm_step = o.dm_step;
  const uint32_t ts = 64 * c.items_d0, td = c.threads_d1 * c.items_d1;
  L.grid = dim3((z.nr_samples + ts - 1) / ts, (o.nr_dms + td - 1) / td, nr_sbeams * o.nr_dms_subbanding);
  k.n_tiles_x = L.grid.x;
  k.n_tiles_y = L.grid.y;
  k.n_tiles = L.grid.x * L.grid.y * L.grid.z;
  return DD_OK;
}

dd_status prepare_step_two_u16(dd_plan *plan, const dd_conf &c, const void *d_intermediate, float *d_output,
                               uint32_t first_sbeam, uint32_t nr_sbeams, F32Launch &L, bool pairs) {
  dd_status s = prepare_step_two(plan, c, reinterpret_cast<const float *>(d_intermediate), d_output, first_sbeam,
                                 nr_sbeams, L);
  if (s != DD_OK) return s;
  const uint32_t bound = 255u * std::max(plan->max_live_sub, 1u);
  if (bound > 65535u) return fail(DD_ERR_CONFIGURATION, "subband sums do not fit 16 bits");
  if (c.local != 1 || c.unroll > kStagedSlotsPerWave * c.threads_d1)
    return fail(DD_ERR_CONFIGURATION, "the uint16 step two needs the staged pipeline (local = 1, unroll <= 2 threads_d1)");
  const uint32_t ts = 64 * c.items_d0, tile = c.threads_d1 * c.items_d1;
  const uint32_t win = roundup(ts + plan->ish_two.max_spread(tile) + 8, 8);
  if (win / 8 > 64 * u16_staged_chunks(c.items_d0))
    return fail(DD_ERR_CONFIGURATION, "step-two shift spread too large for the uint16 window");
  L.g.win = win;
  L.g.staged = 1;
  // rows | [2 buffers][unroll rows][4 copies][win / 2 dwords] | DM addresses;
  // rows staged per barrier are capped at what fits in LDS
  auto lds_of = [&](uint32_t u) { return 4 * (roundup(2 * plan->obs.nr_subbands, 4) + 4 * u * win + 2 * u * tile); };
  uint32_t unroll = c.unroll;
  while (unroll > 1 && lds_of(unroll) > plan->max_lds) --unroll;
  L.g.lds_bytes = lds_of(unroll);
  if (L.g.lds_bytes > plan->max_lds)
    return fail(DD_ERR_CONFIGURATION, "LDS demand " + std::to_string(L.g.lds_bytes) + " bytes exceeds the device limit");
  L.g.unroll = unroll;
  L.k.unroll = unroll;
  L.k.win_floats = win;
  L.k.combine_rows = 65535u / bound >= 4 ? 4 : 65535u / bound >= 2 ? 2 : 1;
  L.k.pair_slots = 0;
  L.g.pairs = 0;
  if (!pairs || 2 * bound > 65535u) return DD_OK;
  // the paired kernel: windows [2 buffers][pairs][slots][2 copies][win / 2
  // dwords]; pairs per barrier at most the wavefronts (one staging wavefront
  // or more per pair), half the configuration's rows, capped by LDS
  const uint32_t slots = plan->ish_two.pair_range(tile);
  const uint64_t aligned = L.k.in_row_pitch | L.k.in_coarse_pitch | L.k.in_beam_pitch;
  if (slots > kPairSlotsMax || (aligned & 7) != 0) return DD_OK;
  auto lds_pairs = [&](uint32_t gp) {
    return 4 * (roundup(2 * plan->obs.nr_subbands, 4) + 2 * gp * slots * win + 2 * gp * tile);
  };
  uint32_t gp = std::max(1u, std::min(c.unroll / 2, c.threads_d1));
  while (gp > 1 && lds_pairs(gp) > plan->max_lds) --gp;
  if (lds_pairs(gp) > plan->max_lds) return DD_OK;
  L.g.pairs = 1;
  L.g.lds_bytes = lds_pairs(gp);
  L.g.unroll = gp;
  L.k.unroll = gp;
  L.k.pair_slots = slots;
  L.k.combine_rows = 65535u / (2 * bound) >= 2 ? 2 : 1;
  return DD_OK;
}

dd_status prepare_single_f32(dd_plan *plan, const dd_conf &c, const float *d_input, float *d_output,
                             uint32_t first_sbeam, uint32_t nr_sbeams, F32Launch &L) {
  if (!plan || !d_input || !d_output) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  if (plan->input_bits != 32) return fail(DD_ERR_INVALID_ARGUMENT, "float input needs a plan with input_bits 32");
  const dd_observation &o = plan->obs;
  const dd_sizes &z = plan->sizes;
  if (nr_sbeams == 0) nr_sbeams = o.nr_synthesized_beams - std::min(first_sbeam, o.nr_synthesized_beams);
  if (nr_sbeams == 0 || first_sbeam + nr_sbeams > o.nr_synthesized_beams)
    return fail(DD_ERR_INVALID_ARGUMENT, "synthesized beam range out of bounds");
  if ((uint64_t)z.nr_samples + z.max_shift > z.nr_dispersed_samples)
    return fail(DD_ERR_INVALID_ARGUMENT, "dispersed batch shorter than samples + largest shift");
  dd_status s = geometry_of(plan, DD_SINGLE_STEP, c, L.g);
  if (s != DD_OK) return s;
  RowSumF32Params &k = L.k;
  k = RowSumF32Params{};
  k.in = d_input;
  k.in_floats = z.input_elements;
  k.out = d_output;
  k.ishift = plan->ish_single.dev.ptr;
  k.dm_pitch = plan->ish_single.pitch;
  k.map = plan->map_single.ptr;
  k.map_pitch = padto(o.nr_channels, plan->padding / 4);
  k.first_sbeam = first_sbeam;
  k.n_rows = o.nr_channels;
  k.n_coarse = 1;
  k.rows = plan->live_single.ptr;
  k.rows_count = plan->live_single_count.ptr;
  k.rows_pitch = o.nr_channels;
  k.in_row_pitch = z.in_pitch;
  k.in_coarse_pitch = 0;
  k.in_beam_pitch = (uint64_t)o.nr_channels * z.in_pitch;
  k.out_dm_pitch = z.out_pitch;
  k.out_beam_pitch = (uint64_t)o.nr_dms * z.out_pitch;
  k.out_inner_pitch = 0;
  k.n_samples = z.nr_samples;
  k.n_dms = o.nr_dms;
  k.unroll = L.g.unroll;
  k.win_floats = L.g.win;
  k.nontemporal = 1;
  k.table = plan->shifts.ptr;
  k.first_dm = o.first_dm;
  k.dm_step = o.dm_step;
  const uint32_t ts = 64 * c.items_d0, td = c.threads_d1 * c.items_d1;
  L.grid = dim3((z.nr_samples + ts - 1) / ts, (o.nr_dms + td - 1) / td, nr_sbeams);
  k.n_tiles_x = L.grid.x;
  k.n_tiles_y = L.grid.y;
  k.n_tiles = L.grid.x * L.grid.y * L.grid.z;
  return DD_OK;
}

dd_status prepare_step_one_f32(dd_plan *plan, const dd_conf &c, const float *d_input, float *d_intermediate,
                               F32Launch &L) {
  if (!plan || !d_input || !d_intermediate) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  if (!plan->two_step) return fail(DD_ERR_INVALID_ARGUMENT, "plan has no subband path");
  if (plan->input_bits != 32) return fail(DD_ERR_INVALID_ARGUMENT, "float input needs a plan with input_bits 32");
  const dd_observation &o = plan->obs;
  const dd_sizes &z = plan->sizes;
  if ((uint64_t)z.nr_samples_step_one + z.max_shift_step_one > z.nr_dispersed_samples_subbanding)
    return fail(DD_ERR_INVALID_ARGUMENT, "dispersed batch shorter than samples + largest shift");
  dd_status s = geometry_of(plan, DD_STEP_ONE, c, L.g);
  if (s != DD_OK) return s;
  RowSumF32Params &k = L.k;
  k = RowSumF32Params{};
  k.in = d_input;
  k.in_floats = z.sub_input_elements;
  k.out = d_intermediate;
  k.ishift = plan->ish_one.dev.ptr;
  k.dm_pitch = plan->ish_one.pitch;
  k.map = nullptr;  // the input beam is the beam itself
  k.n_rows = cps(plan);
  k.n_coarse = o.nr_subbands;
  k.rows = plan->live_sub.ptr;
  k.rows_count = plan->live_sub_count.ptr;
  k.rows_pitch = cps(plan);
  k.in_row_pitch = z.sub_in_pitch;
  k.in_coarse_pitch = 0;
  k.in_beam_pitch = (uint64_t)o.nr_channels * z.sub_in_pitch;
  k.out_dm_pitch = (uint64_t)o.nr_subbands * z.sub_out_pitch;
  k.out_beam_pitch = (uint64_t)o.nr_dms_subbanding * o.nr_subbands * z.sub_out_pitch;
  k.out_inner_pitch = z.sub_out_pitch;
  k.n_samples = z.nr_samples_step_one;
  k.n_dms = o.nr_dms_subbanding;
  k.unroll = L.g.unroll;
  k.win_floats = L.g.win;
  k.nontemporal = 0;  // step two re-reads the intermediate
  k.table = plan->relative.ptr;
  k.first_dm = o.first_dm_subbanding;
  k.dm_step = o.dm_step_subbanding;
  const uint32_t ts = 64 * c.items_d0, td = c.threads_d1 * c.items_d1;
  L.grid = dim3((z.nr_samples_step_one + ts - 1) / ts, (o.nr_dms_subbanding + td - 1) / td,
                o.nr_beams * o.nr_subbands);
  k.n_tiles_x = L.grid.x;
  k.n_tiles_y = L.grid.y;
  k.n_tiles = L.grid.x * L.grid.y * L.grid.z;
  return DD_OK;
}

dd_status run(const U8Launch &L, hipStream_t stream) {
  if (L.g.staged == 2) {
    DD_HIP(launch_rowsum_u8_factored(L.k, L.g, stream));
    return DD_OK;
  }
  if (L.g.staged) {
    DD_HIP(launch_rowsum_u8_staged(L.k, L.g, L.flush, stream));
    return DD_OK;
  }
  if (L.grid.y > 65535 || L.grid.z > 65535) return fail(DD_ERR_UNSUPPORTED, "grid too large");
  DD_HIP(launch_rowsum_u8(L.k, L.g, L.flush, L.grid, stream));
  return DD_OK;
}

dd_status run(const F32Launch &L, hipStream_t stream) {
  if (L.g.staged) {
    DD_HIP(launch_rowsum_f32_staged(L.k, L.g, stream));
    return DD_OK;
  }
  if (L.grid.y > 65535 || L.grid.z > 65535) return fail(DD_ERR_UNSUPPORTED, "grid too large");
  DD_HIP(launch_rowsum_f32(L.k, L.g, L.grid, stream));
  return DD_OK;
}

} // namespace ddamd

using namespace ddamd;

extern "C" {

const char *dd_last_error(void) { return g_last_error.c_str(); }

const char *dd_status_string(dd_status status) {
  switch (status) {
  case DD_OK: return "ok";
  case DD_ERR_INVALID_ARGUMENT: return "invalid argument";
  case DD_ERR_UNSUPPORTED: return "unsupported";
  case DD_ERR_HIP: return "HIP runtime error";
  case DD_ERR_OUT_OF_MEMORY: return "out of memory";
  case DD_ERR_CONFIGURATION: return "invalid kernel configuration";
  case DD_ERR_NOT_READY: return "not ready";
  case DD_ERR_IO: return "I/O error";
  }
  return "unknown status";
}

uint32_t dd_abi_version(void) { return DD_ABI_VERSION; }

dd_status dd_plan_create(dd_plan **out, const dd_observation *obs_in, uint32_t padding) {
  if (!out || !obs_in) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  const dd_observation &o = *obs_in;
  if (o.nr_beams == 0 || o.nr_synthesized_beams == 0) return fail(DD_ERR_INVALID_ARGUMENT, "beam counts must be positive");
  if (o.nr_channels < 1) return fail(DD_ERR_INVALID_ARGUMENT, "nr_channels must be positive");
  if (o.nr_subbands < 1 || o.nr_channels % o.nr_subbands != 0)
    return fail(DD_ERR_INVALID_ARGUMENT, "nr_subbands must divide nr_channels");
  if (o.nr_samples_per_batch == 0 || o.downsampling == 0 || o.nr_samples_per_batch / o.downsampling == 0)
    return fail(DD_ERR_INVALID_ARGUMENT, "nr_samples_per_batch / downsampling must be positive");
  if (!(o.sampling_time > 0.0f)) return fail(DD_ERR_INVALID_ARGUMENT, "sampling_time must be positive");
  if (o.nr_dms == 0) return fail(DD_ERR_INVALID_ARGUMENT, "nr_dms must be positive");
  if (!(o.first_dm >= 0.0f) || !(o.dm_step >= 0.0f) || !(o.first_dm_subbanding >= 0.0f) ||
      !(o.dm_step_subbanding >= 0.0f))
    return fail(DD_ERR_INVALID_ARGUMENT, "DM ranges must be non-negative");
  if (padding % 4 != 0) return fail(DD_ERR_INVALID_ARGUMENT, "padding must be a multiple of 4 bytes");
  if (o.nr_channels > 65535 || o.nr_beams > 65535 || o.nr_dms > (1u << 22) || o.nr_dms_subbanding > (1u << 22))
    return fail(DD_ERR_UNSUPPORTED, "too many channels, beams or DMs");

  dd_plan *p = new dd_plan();
  p->obs = o;
  p->padding = padding;
  p->two_step = o.nr_dms_subbanding > 0;
  (void)hipGetDevice(&p->device);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, p->device) == hipSuccess) p->max_lds = prop.sharedMemPerBlock;

  auto cleanup = [&](dd_status s) {
    delete p;
    return s;
  };

  // Float tables on the device (src/Shifts.cpp:20-42), copied back for sizing.
  ShiftFormula f{};
  f.first_freq = o.min_freq;
  f.freq_step = o.channel_bandwidth;
  f.max_freq = max_freq(o);
  f.samples_per_batch = o.nr_samples_per_batch;
  f.downsampling = o.downsampling;
  f.sampling_time = o.sampling_time;
  f.rows = o.nr_channels;
  f.padded_rows = padto(o.nr_channels, padding / 4);
  hipError_t e = p->shifts.allocate(f.padded_rows);
  if (e == hipSuccess) e = launch_build_shifts(f, p->shifts.ptr, nullptr);
  ShiftFormula f2 = f;
  f2.first_freq = subband_min_freq(o);
  f2.freq_step = subband_bandwidth(o);
  f2.max_freq = subband_max_freq(o);
  f2.rows = o.nr_subbands;
  f2.padded_rows = padto(o.nr_subbands, padding / 4);
  if (e == hipSuccess) e = p->shifts2.allocate(f2.padded_rows);
  if (e == hipSuccess) e = launch_build_shifts(f2, p->shifts2.ptr, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = p->shifts.download(p->h_shifts);
  if (e == hipSuccess) e = p->shifts2.download(p->h_shifts2);
  if (e != hipSuccess) return cleanup(hip_fail(e, "building shift tables"));

  dd_status s = rebuild_integer_tables(p);
  if (s != DD_OK) return cleanup(s);

  // Sizes (src/DedispersionTest.cpp:129, 145, 146).
  dd_sizes &z = p->sizes;
  z.nr_samples = o.nr_samples_per_batch / o.downsampling;
  z.nr_dispersed_samples = o.nr_samples_per_dispersed_batch
                               ? o.nr_samples_per_dispersed_batch
                               : dispersed(o.nr_samples_per_batch, p->h_shifts[0], o.first_dm, o.dm_step, o.nr_dms);
  if (p->two_step) {
    z.nr_samples_subbanding = o.nr_samples_per_batch_subbanding
                                  ? o.nr_samples_per_batch_subbanding
                                  : dispersed(o.nr_samples_per_batch, p->h_shifts2[0], o.first_dm, o.dm_step, o.nr_dms);
    z.nr_dispersed_samples_subbanding =
        o.nr_samples_per_dispersed_batch_subbanding
            ? o.nr_samples_per_dispersed_batch_subbanding
            : dispersed(z.nr_samples_subbanding, p->h_shifts[0], o.first_dm_subbanding, o.dm_step_subbanding,
                        o.nr_dms_subbanding);
  }
  const uint32_t C = o.nr_channels;
  size_inputs(p);
  z.out_pitch = padto(z.nr_samples, padding / 4);
  z.output_elements = (uint64_t)o.nr_synthesized_beams * o.nr_dms * z.out_pitch;
  if (p->two_step) {
    // The intermediate rows are padded as for 8-bit input whatever inputBits
    // is; include/Dedispersion.hpp:148 pads S1 / (8 / inputBits), which makes
    // the rows of packed input overlap.  Step one writes S1 / downsampling
    // samples per row (:119) into rows of pad(S1) (:148); step two reads rows
    // of pad(S1 / downsampling) (:163, :169).  The two pitches differ when
    // downsampling > 1, as they do in the reference.
    z.sub_out_pitch = padto(z.nr_samples_subbanding, padding / 4);
    z.intermediate_elements = (uint64_t)o.nr_beams * o.nr_dms_subbanding * o.nr_subbands * z.sub_out_pitch;
    z.sub_output_elements = (uint64_t)o.nr_synthesized_beams * o.nr_dms_subbanding * o.nr_dms * z.out_pitch;
    z.nr_samples_step_one = z.nr_samples_subbanding / o.downsampling;
    z.step_two_in_pitch = padto(z.nr_samples_step_one, padding / 4);
    z.step_two_input_elements = (uint64_t)o.nr_beams * o.nr_dms_subbanding * o.nr_subbands * z.step_two_in_pitch;
  }

  // Default tables: no zapped channel, sBeam -> beam sBeam % nrBeams.
  p->h_zapped.assign(padto(C, padding / 4), 0);
  s = rebuild_live(p);
  if (s != DD_OK) return cleanup(s);
  {
    const uint32_t mp = padto(C, padding / 4);
    std::vector<uint32_t> m((size_t)o.nr_synthesized_beams * mp, 0);
    for (uint32_t b = 0; b < o.nr_synthesized_beams; ++b)
      for (uint32_t ch = 0; ch < mp; ++ch) m[(size_t)b * mp + ch] = b % o.nr_beams;
    e = p->map_single.upload(m);
    if (e != hipSuccess) return cleanup(hip_fail(e, "uploading beam mapping"));
    const uint32_t mp2 = padto(o.nr_subbands, padding / 4);
    std::vector<uint32_t> m2((size_t)o.nr_synthesized_beams * mp2, 0);
    for (uint32_t b = 0; b < o.nr_synthesized_beams; ++b)
      for (uint32_t sb = 0; sb < mp2; ++sb) m2[(size_t)b * mp2 + sb] = b % o.nr_beams;
    e = p->map_two.upload(m2);
    if (e != hipSuccess) return cleanup(hip_fail(e, "uploading beam mapping"));
  }
  for (int m = 0; m < 3; ++m) p->conf[m] = default_conf((dd_mode)m);
  *out = p;
  return DD_OK;
}

dd_status dd_plan_set_input_bits(dd_plan *plan, uint32_t input_bits) {
  if (!plan) return fail(DD_ERR_INVALID_ARGUMENT, "null plan");
  if (input_bits != 1 && input_bits != 2 && input_bits != 4 && input_bits != 8 && input_bits != 32)
    return fail(DD_ERR_INVALID_ARGUMENT, "input_bits must be 1, 2, 4, 8 or 32 (float)");
  plan->input_bits = input_bits;
  size_inputs(plan);
  return DD_OK;
}

dd_status dd_plan_set_step_one_shifts(dd_plan *plan, dd_step_one_shifts delays) {
  if (!plan) return fail(DD_ERR_INVALID_ARGUMENT, "null plan");
  if (delays != DD_STEP_ONE_RELATIVE && delays != DD_STEP_ONE_ABSOLUTE)
    return fail(DD_ERR_INVALID_ARGUMENT, "step-one delays must be DD_STEP_ONE_RELATIVE or DD_STEP_ONE_ABSOLUTE");
  if (!plan->two_step) return fail(DD_ERR_INVALID_ARGUMENT, "plan has no subband path");
  plan->step_one_absolute = delays == DD_STEP_ONE_ABSOLUTE;
  return rebuild_integer_tables(plan);
}

dd_status dd_plan_destroy(dd_plan *plan) {
  delete plan;
  return DD_OK;
}

dd_status dd_plan_get_sizes(const dd_plan *plan, dd_sizes *sizes) {
  if (!plan || !sizes) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  *sizes = plan->sizes;
  return DD_OK;
}

dd_status dd_plan_set_conf(dd_plan *plan, dd_mode mode, const dd_conf *conf) {
  if (!plan || !conf || (int)mode < 0 || (int)mode > 2) return fail(DD_ERR_INVALID_ARGUMENT, "invalid argument");
  if (mode != DD_SINGLE_STEP && !plan->two_step) return fail(DD_ERR_INVALID_ARGUMENT, "plan has no subband path");
  Geometry g;
  dd_status s = geometry_of(plan, mode, *conf, g);
  if (s != DD_OK) return s;
  plan->conf[mode] = *conf;
  return DD_OK;
}

dd_status dd_plan_get_conf(const dd_plan *plan, dd_mode mode, dd_conf *conf) {
  if (!plan || !conf || (int)mode < 0 || (int)mode > 2) return fail(DD_ERR_INVALID_ARGUMENT, "invalid argument");
  *conf = plan->conf[mode];
  return DD_OK;
}

dd_status dd_plan_set_zapped_channels(dd_plan *plan, const uint32_t *zapped) {
  if (!plan || !zapped) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  plan->h_zapped.assign(zapped, zapped + plan->h_zapped.size());
  return rebuild_live(plan);
}

dd_status dd_plan_set_beam_mapping(dd_plan *plan, const uint32_t *single_step, const uint32_t *step_two) {
  if (!plan) return fail(DD_ERR_INVALID_ARGUMENT, "null plan");
  const dd_observation &o = plan->obs;
  if (single_step) {
    const size_t n = (size_t)o.nr_synthesized_beams * padto(o.nr_channels, plan->padding / 4);
    std::vector<uint32_t> m(single_step, single_step + n);
    for (uint32_t b = 0; b < o.nr_synthesized_beams; ++b)
      for (uint32_t ch = 0; ch < o.nr_channels; ++ch)
        if (m[(size_t)b * padto(o.nr_channels, plan->padding / 4) + ch] >= o.nr_beams)
          return fail(DD_ERR_INVALID_ARGUMENT, "beam mapping entry out of range");
    DD_HIP(plan->map_single.upload(m));
  }
  if (step_two) {
    const uint32_t mp = padto(o.nr_subbands, plan->padding / 4);
    const size_t n = (size_t)o.nr_synthesized_beams * mp;
    std::vector<uint32_t> m(step_two, step_two + n);
    for (uint32_t b = 0; b < o.nr_synthesized_beams; ++b)
      for (uint32_t sb = 0; sb < o.nr_subbands; ++sb)
        if (m[(size_t)b * mp + sb] >= o.nr_beams) return fail(DD_ERR_INVALID_ARGUMENT, "beam mapping entry out of range");
    DD_HIP(plan->map_two.upload(m));
  }
  return DD_OK;
}

dd_status dd_plan_set_shifts(dd_plan *plan, const float *shifts, const float *shifts_step_two) {
  if (!plan) return fail(DD_ERR_INVALID_ARGUMENT, "null plan");
  auto check = [](const float *t, size_t n) {
    for (size_t i = 0; i < n; ++i)
      if (!(t[i] >= 0.0f) || !std::isfinite(t[i])) return false;
    return true;
  };
  if (shifts) {
    if (!check(shifts, plan->obs.nr_channels)) return fail(DD_ERR_INVALID_ARGUMENT, "shift table entries must be finite and non-negative");
    plan->h_shifts.assign(shifts, shifts + plan->h_shifts.size());
    DD_HIP(plan->shifts.upload(plan->h_shifts));
  }
  if (shifts_step_two) {
    if (!check(shifts_step_two, plan->obs.nr_subbands)) return fail(DD_ERR_INVALID_ARGUMENT, "shift table entries must be finite and non-negative");
    plan->h_shifts2.assign(shifts_step_two, shifts_step_two + plan->h_shifts2.size());
    DD_HIP(plan->shifts2.upload(plan->h_shifts2));
  }
  return rebuild_integer_tables(plan);
}

dd_status dd_plan_get_shifts(const dd_plan *plan, float *shifts, float *shifts_step_two) {
  if (!plan) return fail(DD_ERR_INVALID_ARGUMENT, "null plan");
  if (shifts) std::memcpy(shifts, plan->h_shifts.data(), plan->h_shifts.size() * sizeof(float));
  if (shifts_step_two) std::memcpy(shifts_step_two, plan->h_shifts2.data(), plan->h_shifts2.size() * sizeof(float));
  return DD_OK;
}

dd_status dd_plan_get_integer_shifts(const dd_plan *plan, dd_mode mode, uint32_t *table) {
  if (!plan || !table) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  const ShiftTable *t = mode == DD_SINGLE_STEP ? &plan->ish_single : mode == DD_STEP_ONE ? &plan->ish_one : &plan->ish_two;
  if (mode != DD_SINGLE_STEP && !plan->two_step) return fail(DD_ERR_INVALID_ARGUMENT, "plan has no subband path");
  for (uint32_t r = 0; r < t->rows; ++r)
    for (uint32_t d = 0; d < t->dms; ++d) table[(size_t)r * t->dms + d] = t->at(r, d);
  return DD_OK;
}

dd_status dd_dedisperse(dd_plan *plan, const uint8_t *d_input, float *d_output, uint32_t first_sbeam,
                        uint32_t nr_sbeams, void *stream) {
  if (!plan) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  U8Launch L;
  dd_status s = prepare_single(plan, plan->conf[DD_SINGLE_STEP], d_input, d_output, first_sbeam, nr_sbeams, L);
  return s != DD_OK ? s : run(L, (hipStream_t)stream);
}

dd_status dd_subband_step_one(dd_plan *plan, const uint8_t *d_input, float *d_intermediate, void *stream) {
  if (!plan) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  U8Launch L;
  dd_status s = prepare_step_one(plan, plan->conf[DD_STEP_ONE], d_input, d_intermediate, L);
  return s != DD_OK ? s : run(L, (hipStream_t)stream);
}

dd_status dd_dedisperse_f32(dd_plan *plan, const float *d_input, float *d_output, uint32_t first_sbeam,
                            uint32_t nr_sbeams, void *stream) {
  if (!plan) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  F32Launch L;
  dd_status s = prepare_single_f32(plan, plan->conf[DD_SINGLE_STEP], d_input, d_output, first_sbeam, nr_sbeams, L);
  return s != DD_OK ? s : run(L, (hipStream_t)stream);
}

dd_status dd_subband_step_one_f32(dd_plan *plan, const float *d_input, float *d_intermediate, void *stream) {
  if (!plan) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  F32Launch L;
  dd_status s = prepare_step_one_f32(plan, plan->conf[DD_STEP_ONE], d_input, d_intermediate, L);
  return s != DD_OK ? s : run(L, (hipStream_t)stream);
}

dd_status dd_subband_step_two(dd_plan *plan, const float *d_intermediate, float *d_output, uint32_t first_sbeam,
                              uint32_t nr_sbeams, void *stream) {
  if (!plan) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  F32Launch L;
  dd_status s = prepare_step_two(plan, plan->conf[DD_STEP_TWO], d_intermediate, d_output, first_sbeam, nr_sbeams, L);
  return s != DD_OK ? s : run(L, (hipStream_t)stream);
}

dd_status dd_tune_time(dd_plan *plan, dd_mode mode, const dd_conf *conf, const void *d_input, void *d_output,
                       uint32_t warmup, uint32_t iterations, void *stream, float *mean_ms, float *std_ms) {
  if (!plan || !conf || !mean_ms || !std_ms || iterations == 0 || (int)mode < 0 || (int)mode > 2)
    return fail(DD_ERR_INVALID_ARGUMENT, "invalid argument");
  if (mode != DD_SINGLE_STEP && !plan->two_step) return fail(DD_ERR_INVALID_ARGUMENT, "plan has no subband path");
  Geometry g;
  dd_status s = geometry_of(plan, mode, *conf, g);
  if (s != DD_OK) return s;
  const dd_conf saved = plan->conf[mode];
  plan->conf[mode] = *conf;
  const bool float_input = plan->input_bits == 32;
  auto launch = [&]() -> dd_status {
    switch (mode) {
    case DD_SINGLE_STEP:
      if (float_input) return dd_dedisperse_f32(plan, (const float *)d_input, (float *)d_output, 0, 0, stream);
      return dd_dedisperse(plan, (const uint8_t *)d_input, (float *)d_output, 0, 0, stream);
    case DD_STEP_ONE:
      if (float_input) return dd_subband_step_one_f32(plan, (const float *)d_input, (float *)d_output, stream);
      return dd_subband_step_one(plan, (const uint8_t *)d_input, (float *)d_output, stream);
    default:
      return dd_subband_step_two(plan, (const float *)d_input, (float *)d_output, 0, 0, stream);
    }
  };
  std::vector<hipEvent_t> ev(2 * iterations, nullptr);
  std::vector<float> ms(iterations, 0.0f);
  hipError_t e = hipSuccess;
  for (auto &x : ev)
    if (e == hipSuccess) e = hipEventCreate(&x);
  for (uint32_t i = 0; i < warmup && e == hipSuccess && s == DD_OK; ++i) s = launch();
  for (uint32_t i = 0; i < iterations && e == hipSuccess && s == DD_OK; ++i) {
    e = hipEventRecord(ev[2 * i], (hipStream_t)stream);
    if (e == hipSuccess) s = launch();
    if (e == hipSuccess && s == DD_OK) e = hipEventRecord(ev[2 * i + 1], (hipStream_t)stream);
  }
  if (e == hipSuccess && s == DD_OK) e = hipStreamSynchronize((hipStream_t)stream);
  for (uint32_t i = 0; i < iterations && e == hipSuccess && s == DD_OK; ++i)
    e = hipEventElapsedTime(&ms[i], ev[2 * i], ev[2 * i + 1]);
  for (auto &x : ev)
    if (x) (void)hipEventDestroy(x);
  plan->conf[mode] = saved;
  if (s != DD_OK) return s;
  if (e != hipSuccess) return hip_fail(e, "timing launches");
  double sum = 0, sq = 0;
  for (float v : ms) sum += v;
  const double mean = sum / iterations;
  for (float v : ms) sq += (v - mean) * (v - mean);
  *mean_ms = (float)mean;
  *std_ms = (float)std::sqrt(sq / iterations);
  return DD_OK;
}

dd_status dd_synth_gaussian_u8(uint8_t *d_data, uint64_t rows, uint64_t cols, uint64_t pitch, float mean,
                               float std_dev, uint64_t seed, void *stream) {
  if (!d_data || cols > pitch) return fail(DD_ERR_INVALID_ARGUMENT, "invalid buffer");
  DD_HIP(launch_synth_gaussian(d_data, rows, cols, pitch, mean, std_dev, seed, false, (hipStream_t)stream));
  return DD_OK;
}

dd_status dd_synth_uniform_u8(uint8_t *d_data, uint64_t rows, uint64_t cols, uint64_t pitch, uint64_t seed,
                              void *stream) {
  if (!d_data || cols > pitch) return fail(DD_ERR_INVALID_ARGUMENT, "invalid buffer");
  DD_HIP(launch_synth_gaussian(d_data, rows, cols, pitch, 0.0f, 0.0f, seed, true, (hipStream_t)stream));
  return DD_OK;
}

dd_status dd_synth_inject_pulse(dd_plan *plan, uint8_t *d_beam_input, uint64_t pitch, uint32_t row_len,
                                float dm_value, uint32_t t0, uint32_t width, uint32_t amplitude, void *stream) {
  if (!plan || !d_beam_input || row_len > pitch) return fail(DD_ERR_INVALID_ARGUMENT, "invalid argument");
  if (!(dm_value >= 0.0f)) return fail(DD_ERR_INVALID_ARGUMENT, "dm_value must be non-negative");
  DD_HIP(launch_inject_pulse(d_beam_input, pitch, plan->shifts.ptr, plan->obs.nr_channels, nullptr, dm_value, t0,
                             width, amplitude, row_len, (hipStream_t)stream));
  return DD_OK;
}

} // extern "C"
