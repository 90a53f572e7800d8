// kernels_u8_staged.hip -- pipelined shifted row sums over uint8 rows.
//
// Same operation and arithmetic as kernels_u8.hip (single-step
// dedispersion, include/Dedispersion.hpp:70-109, and step one, :111-153), with
// the per-row overheads taken off the critical path:
//
//  * Rows are staged in LDS in groups of `unroll`, double buffered: the
//    global loads of group k+1 are issued before group k is summed and written
//    to the other buffer after it, so one barrier per group remains and the
//    load latency hides behind the additions.
//  * A row's integer shifts for the tile's DMs are computed once, by the lanes
//    of the wavefront that stages the row, from the float delay table with the
//    reference's unfused arithmetic, and published through LDS; the summing
//    wavefronts read them from LDS instead of chasing dependent scalar loads.
//  * The live rows (and their input beams) of the workgroup are copied to LDS
//    once at the start.
//  * Workgroups are ordered XCD-aware (staged_common.hpp).
//
// Rows are staged in four copies, copy c shifted down by c bytes (built in
// registers with v_alignbyte), so that every four-sample group of a DM is one
// aligned ds_read_b32.  Packed rows (inputBits 1, 2 or 4,
// include/Dedispersion.hpp:92-103) are loaded as packed bytes and unpacked
// to one byte per sample in registers before the copies are built.
//
// The summation keeps two 32-bit accumulators per four-sample group: `all`,
// the plain 32-bit sum of the groups' dwords (bytes carry into their
// neighbours), and `odd`, the odd bytes extracted by v_perm_b32 into 16-bit
// pairs.  all - (odd << 8) is then the even bytes' 16-bit pairs, exactly, as
// long as every pair sum stays below 2^16; partial sums are exact integers.  Each half-wavefront sums one DM at
// a time (lane l: DM 2s + l / 32 of the wavefront's DMs, over the groups
// (l % 32) + 32q), so every lane reads its own DM's LDS address.
#include "staged_common.hpp"

namespace ddamd {
namespace {

using namespace staged;

// The 16 samples of window group m (one byte each, 4 dwords u) and the 4
// after them (`next`) from the packed dwords d[0..2], which start at the
// dword holding packed byte 16m >> LF; LF = log2 of samples per byte (1..3).
// Samples are the fields of each byte from the least significant bit up.
template <int LF>
__device__ __forceinline__ void unpack_group(const u32x4 &d, uint32_t m, uint32_t (&u)[4], uint32_t &next) {
  if constexpr (LF == 1) {
    // 4-bit fields, 64 bits from d[0] and d[1] (the next 4 samples in the low
    // half of d[2]): even fields to the bytes of `lo`, odd fields to those
    // of `hi`, interleaved
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      const uint32_t lo = d[e] & 0x0F0F0F0Fu, hi = (d[e] >> 4) & 0x0F0F0F0Fu;
      const uint32_t a = __builtin_amdgcn_perm(hi, lo, 0x05010400u);
      if (e < 2) {
        u[2 * e] = a;
        u[2 * e + 1] = __builtin_amdgcn_perm(hi, lo, 0x07030602u);
      } else {
        next = a;
      }
    }
  } else if constexpr (LF == 2) {
    // 2-bit fields, 32 bits from d[0] (the next 4 samples in the low byte of
    // d[1]): m_j byte k = sample 4k + j; two rounds of byte interleave
    const uint32_t x = d[0];
    const uint32_t m0 = x & 0x03030303u, m1 = (x >> 2) & 0x03030303u;
    const uint32_t m2 = (x >> 4) & 0x03030303u, m3 = (x >> 6) & 0x03030303u;
    const uint32_t a01 = __builtin_amdgcn_perm(m1, m0, 0x05010400u);  // s0 s1 s4 s5
    const uint32_t a23 = __builtin_amdgcn_perm(m3, m2, 0x05010400u);  // s2 s3 s6 s7
    const uint32_t b01 = __builtin_amdgcn_perm(m1, m0, 0x07030602u);  // s8 s9 s12 s13
    const uint32_t b23 = __builtin_amdgcn_perm(m3, m2, 0x07030602u);  // s10 s11 s14 s15
    u[0] = __builtin_amdgcn_perm(a23, a01, 0x05040100u);
    u[1] = __builtin_amdgcn_perm(a23, a01, 0x07060302u);
    u[2] = __builtin_amdgcn_perm(b23, b01, 0x05040100u);
    u[3] = __builtin_amdgcn_perm(b23, b01, 0x07060302u);
    const uint32_t y = d[1];
    const uint32_t c01 = __builtin_amdgcn_perm((y >> 2) & 0x03030303u, y & 0x03030303u, 0x05010400u);
    const uint32_t c23 = __builtin_amdgcn_perm((y >> 6) & 0x03030303u, (y >> 4) & 0x03030303u, 0x05010400u);
    next = __builtin_amdgcn_perm(c23, c01, 0x05040100u);
  } else {
    // 1-bit fields, 16 bits from bit 16 (m & 1) of d[0]: a nibble n spread to
    // bytes by n * (1 + 2^7 + 2^14 + 2^21), whose four shifted copies do not
    // overlap, then bit 0 of each byte kept
    const uint32_t x = __builtin_amdgcn_alignbit(d[1], d[0], 16 * (m & 1));
#pragma unroll
    for (int q = 0; q < 4; ++q) u[q] = (((x >> (4 * q)) & 0xFu) * 0x204081u) & 0x01010101u;
    next = (((x >> 16) & 0xFu) * 0x204081u) & 0x01010101u;
  }
}

// Writes the byte run u (n dwords, followed by the dword `next`) to the
// four copies of a row window at dword position `at`: copy c holds the run
// shifted down by c bytes.
template <int N>
__device__ __forceinline__ void store_copies(uint32_t *w, uint32_t WD, uint32_t at, const uint32_t (&u)[N],
                                             uint32_t next) {
  static_assert(N % 4 == 0, "whole 16-byte groups");
#pragma unroll
  for (int i = 0; i < N; i += 4) {
    *reinterpret_cast<u32x4 *>(w + at + i) = u32x4{u[i], u[i + 1], u[i + 2], u[i + 3]};
#pragma unroll
    for (int c = 1; c < 4; ++c) {
      u32x4 r;
#pragma unroll
      for (int e = 0; e < 4; ++e)
        r[e] = __builtin_amdgcn_alignbyte(i + e + 1 < N ? u[i + e + 1] : next, u[i + e], c);
      *reinterpret_cast<u32x4 *>(w + c * WD + at + i) = r;
    }
  }
}

// One (beam x group, DM tile, sample tile) block: logical tile index
// `logical` of the launch described by p (DM tile fastest).
template <int RUN, int DPW, bool FLUSH, bool PACKED>
__device__ __forceinline__ void u8_tile(const RowSumU8Params &p, uint32_t logical, uint32_t *lds) {
  constexpr int NG = RUN / 4;   // four-sample groups per lane
  constexpr int TS = 64 * RUN;  // samples per tile
  constexpr int SPW = kStagedSlotsPerWave;
  constexpr int CPL = kU8StagedChunks;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wave = uniform(threadIdx.x >> 6);
  const uint32_t nwaves = blockDim.x >> 6;
  const uint32_t TD = nwaves * DPW;
  const Tile tile = tile_split(logical, p.n_tiles_x, p.n_tiles_y);
  const uint32_t t0 = tile.x * TS;
  const uint32_t d0 = tile.y * TD;
  const uint32_t dw = wave * DPW;  // first DM of this wavefront, relative to d0
  // a wavefront whose DMs all lie past the last DM (a partial DM tile) only stages rows
  const bool busy = d0 + dw < p.n_dms;
  const uint32_t group = tile.z % p.n_groups;
  const uint32_t beam = tile.z / p.n_groups;
  const uint32_t nlive = uniform(p.live_count[group]);
  const uint32_t *live = p.live + (uint64_t)group * p.live_pitch;
  const uint32_t G = uniform(p.unroll);
  const uint32_t WD = p.win_bytes / 4;  // dwords per window
  const uint32_t NCH = WD / 4;          // 16-byte chunks per window
  const uint32_t lf = PACKED ? p.in_log2f : 0u;  // log2 of samples per input byte
  const uint32_t dmax = p.n_dms - 1;

  // LDS: rows (packed row | beam << 16, float delay) | windows [2][G][4][WD] | DM words [2][G][TD]
  // Rows are kept in four copies, copy c shifted down by c bytes, so a DM's
  // four-sample group at any offset is one aligned dword of copy
  // (offset & 3); packed rows are unpacked to one byte per sample first.
  const uint32_t win_off = (2 * p.live_pitch + 3) & ~3u;
  uint32_t *lrow = lds;
  float *ltab = reinterpret_cast<float *>(lds + p.live_pitch);
  uint32_t *windows = lds + win_off;
  uint32_t *meta = windows + 2 * G * 4 * WD;

  for (uint32_t k = threadIdx.x; k < nlive; k += blockDim.x) {
    const uint32_t row = live[k];
    const uint32_t in_beam = p.map ? p.map[(uint64_t)(p.first_sbeam + beam) * p.map_pitch + row] : beam;
    lrow[k] = row | (in_beam << 16);
    ltab[k] = p.table[row];
  }
  __syncthreads();

  // ---- staging: wavefront w stages rows w and w + nwaves of a group ----
  u32x4 sv[SPW][CPL];
  uint32_t sx[SPW][CPL];  // the dword after each 16-byte chunk (byte rows)
  uint32_t so[SPW];
  auto stage_load = [&](uint32_t k0, uint32_t nn) {
#pragma unroll
    for (int s = 0; s < SPW; ++s) {
      const uint32_t g = wave + s * nwaves;
      if (g < nn) {
        const uint32_t packed = uniform(lrow[k0 + g]);
        const float tab = uniformf(ltab[k0 + g]);
        const uint32_t row = packed & 0xFFFFu, in_beam = packed >> 16;
        const uint32_t sh = shift_of(p.first_dm, p.dm_step, min(d0 + lane, dmax), tab);
        const uint32_t shmin = uniform(sh);
        // first sample of the window and the byte that holds it (packed rows
        // hold 1 << lf samples per byte)
        const uint32_t x = t0 + shmin;
        const uint64_t off = uniform64((uint64_t)in_beam * p.in_beam_pitch + (uint64_t)row * p.in_row_pitch + (x >> lf));
        const uint64_t a0 = off & ~uint64_t(3);
        // offset, in unpacked samples, of each DM's samples in the window
        // (the window starts at the first sample of byte a0)
        so[s] = sh - shmin + ((uint32_t)(off & 3) << lf) + (x & ((1u << lf) - 1));
        const uint32_t recs = records_from(p.in_bytes, a0);
        const __amdgpu_buffer_rsrc_t rsrc = make_rsrc(p.in + a0, recs);
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          const uint32_t m = lane + 64 * c;
          if (m >= NCH) continue;
          if (!PACKED || lf == 0) {
            // 16 bytes of the window and the dword after them
            if (16 * m + 16 <= recs) {
              sv[s][c] = __builtin_amdgcn_raw_buffer_load_b128(rsrc, 16 * m, 0, 0);
            } else {
#pragma unroll
              for (int e = 0; e < 4; ++e) sv[s][c][e] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, 16 * m + 4 * e, 0, 0);
            }
            sx[s][c] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, 16 * m + 16, 0, 0);
          } else {
            // the packed bytes of the window's 16 samples from 16m and of the
            // 4 after them: three dwords from the dword that holds packed
            // byte 16m >> lf
            const uint32_t pd = (16 * m >> lf) & ~3u;
            if (pd + 12 <= recs) {
              const u32x3 t = __builtin_amdgcn_raw_buffer_load_b96(rsrc, pd, 0, 0);
              sv[s][c][0] = t[0];
              sv[s][c][1] = t[1];
              sv[s][c][2] = t[2];
            } else {
#pragma unroll
              for (int e = 0; e < 3; ++e) sv[s][c][e] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, pd + 4 * e, 0, 0);
            }
          }
        }
      }
    }
  };
  auto stage_store = [&](uint32_t buf, uint32_t nn) {
#pragma unroll
    for (int s = 0; s < SPW; ++s) {
      const uint32_t g = wave + s * nwaves;
      if (g < nn) {
        uint32_t *w = windows + (buf * G + g) * 4 * WD;
        auto unpacked = [&](auto lfc) {
          constexpr int LF = lfc.value;
#pragma unroll
          for (int c = 0; c < CPL; ++c) {
            const uint32_t m = lane + 64 * c;
            if (m < NCH) {
              uint32_t u[4];
              uint32_t next;
              unpack_group<LF>(sv[s][c], m, u, next);
              store_copies(w, WD, 4 * m, u, next);
            }
          }
        };
        if (!PACKED || lf == 0) {
#pragma unroll
          for (int c = 0; c < CPL; ++c) {
            const uint32_t m = lane + 64 * c;
            if (m < NCH) {
              const uint32_t u[4] = {sv[s][c][0], sv[s][c][1], sv[s][c][2], sv[s][c][3]};
              store_copies(w, WD, 4 * m, u, sx[s][c]);
            }
          }
        } else if constexpr (PACKED) {
          if (lf == 1) unpacked(std::integral_constant<int, 1>{});
          else if (lf == 2) unpacked(std::integral_constant<int, 2>{});
          else unpacked(std::integral_constant<int, 3>{});
        }
        // the DM's LDS byte address: copy (offset & 3), dword offset >> 2
        const uint32_t o = so[s];
        if (lane < TD) meta[(buf * G + g) * TD + lane] = 4 * (win_off + ((buf * G + g) * 4 + (o & 3)) * WD + (o >> 2));
      }
    }
  };

  // ---- accumulation ----
  // all[j][k]: 32-bit sum of the group dwords, B0 + 2^8 B1 + 2^16 B2 + 2^24 B3
  // (mod 2^32) for byte sums B0..B3; odd[j][k] = B1 + 2^16 B3.  Then
  // all - (odd << 8) = B0 + 2^16 B2 exactly while every B stays below 2^16.
  uint32_t all[DPW][NG], odd[DPW][NG];
  uint32_t wide[DPW][RUN];
#pragma unroll
  for (int j = 0; j < DPW; ++j)
#pragma unroll
    for (int k = 0; k < NG; ++k) all[j][k] = odd[j][k] = 0;
  if constexpr (FLUSH) {
#pragma unroll
    for (int j = 0; j < DPW; ++j)
#pragma unroll
      for (int k = 0; k < RUN; ++k) wide[j][k] = 0;
  }
  auto even_of = [&](int j, int k) -> uint32_t { return all[j][k] - (odd[j][k] << 8); };
  auto flush = [&]() {
    if constexpr (FLUSH) {
#pragma unroll
      for (int j = 0; j < DPW; ++j)
#pragma unroll
        for (int k = 0; k < NG; ++k) {
          const uint32_t even = even_of(j, k);
          wide[j][4 * k + 0] += even & 0xFFFFu;
          wide[j][4 * k + 2] += even >> 16;
          wide[j][4 * k + 1] += odd[j][k] & 0xFFFFu;
          wide[j][4 * k + 3] += odd[j][k] >> 16;
          all[j][k] = odd[j][k] = 0;
        }
    }
  };

  // Rows are taken two at a time so that the two extracted pairs of a DM
  // and sample group go into its accumulator with one three-input add.
  //
  // Each half-wavefront takes one DM at a time.  With LPD = 32
  // lanes per DM (64 when the wavefront has one DM), lane l sums, for DM
  // slot s < SL, DM dw + s CPS + l / LPD over the four-sample groups
  // (l % LPD) + LPD q, q < NGS: the 32 lanes of a half read 32 consecutive
  // dwords (no bank conflicts), and each lane reads its DM's word itself
  // (no v_readlane).  Chunk c (NG groups) of the lane is accumulator [c][k],
  // slot c / CPS, groups (c % CPS) NG + k.  A group is one aligned dword of
  // copy (offset & 3); it is added whole to `all`, and its odd samples,
  // extracted to 16-bit pairs with a constant v_perm selector, to `odd`.
  constexpr uint32_t SEL_ODD = 0x0c030c01u;
  constexpr int LPD = DPW == 1 ? 64 : 32;  // lanes per DM
  constexpr int CPS = 64 / LPD;            // DMs per slot = chunks per slot
  constexpr int SL = DPW / CPS;            // DM slots per lane
  const uint32_t lane_dm = lane / LPD;
  const char *split_base = reinterpret_cast<const char *>(lds) + 4 * (lane % LPD);
  const uint32_t *lane_meta = meta + dw + lane_dm;
  auto word_of = [&](uint32_t buf, uint32_t g, int sl) { return lane_meta[uniform((buf * G + g) * TD) + sl * CPS]; };
  auto compute_bytes = [&](uint32_t buf, uint32_t n) {
    uint32_t g = 0;
    for (; g + 1 < n; g += 2) {
      const uint32_t *src0[SL], *src1[SL];
#pragma unroll
      for (int sl = 0; sl < SL; ++sl) {
        src0[sl] = reinterpret_cast<const uint32_t *>(split_base + word_of(buf, g, sl));
        src1[sl] = reinterpret_cast<const uint32_t *>(split_base + word_of(buf, g + 1, sl));
      }
      // the groups of chunk c + 1 are loaded before chunk c is summed
      uint32_t x[2][2][NG];  // [chunk parity][row][group]
      auto load = [&](int c, uint32_t (&x0)[NG], uint32_t (&x1)[NG]) {
        const int sl = c / CPS, q0 = (c % CPS) * NG;
#pragma unroll
        for (int k = 0; k < NG; ++k) {
          x0[k] = src0[sl][LPD * (q0 + k)];
          x1[k] = src1[sl][LPD * (q0 + k)];
        }
      };
      load(0, x[0][0], x[0][1]);
#pragma unroll
      for (int c = 0; c < DPW; ++c) {
        if (c + 1 < DPW) load(c + 1, x[(c + 1) & 1][0], x[(c + 1) & 1][1]);
        const uint32_t(&x0)[NG] = x[c & 1][0];
        const uint32_t(&x1)[NG] = x[c & 1][1];
#pragma unroll
        for (int k = 0; k < NG; ++k) {
          all[c][k] += x0[k] + x1[k];
          odd[c][k] += __builtin_amdgcn_perm(x0[k], x0[k], SEL_ODD) + __builtin_amdgcn_perm(x1[k], x1[k], SEL_ODD);
        }
      }
    }
    if (g < n) {
#pragma unroll
      for (int sl = 0; sl < SL; ++sl) {
        const uint32_t *src = reinterpret_cast<const uint32_t *>(split_base + word_of(buf, g, sl));
#pragma unroll
        for (int cc = 0; cc < CPS; ++cc) {
          const int c = sl * CPS + cc;
#pragma unroll
          for (int k = 0; k < NG; ++k) {
            const uint32_t v = src[LPD * (cc * NG + k)];
            all[c][k] += v;
            odd[c][k] += __builtin_amdgcn_perm(v, v, SEL_ODD);
          }
        }
      }
    }
  };
  auto compute = [&](uint32_t buf, uint32_t n) { compute_bytes(buf, n); };

  // ---- pipeline over groups of rows ----
  uint32_t cur = 0;
  if (nlive > 0) {
    const uint32_t n0 = min(G, nlive);
    stage_load(0, n0);
    stage_store(0, n0);
  }
  __syncthreads();
  uint32_t since_flush = 0;
  for (uint32_t k0 = 0; k0 < nlive; k0 += G) {
    const uint32_t n = min(G, nlive - k0);
    const uint32_t kn = k0 + G;
    const uint32_t nn = kn < nlive ? min(G, nlive - kn) : 0;
    if (nn) stage_load(kn, nn);
    if constexpr (FLUSH) {
      if (since_flush + n > 256) {
        flush();
        since_flush = 0;
      }
      since_flush += n;
    }
    if (busy) compute(cur, n);
    if (nn) stage_store(cur ^ 1, nn);
    __syncthreads();
    cur ^= 1;
  }
  flush();

  // the DM and first sample of accumulator [j][k]
  auto dm_of = [&](int j) -> uint32_t { return d0 + dw + (j / CPS) * CPS + lane_dm; };
  auto t_of = [&](int j, int k) -> uint32_t { return t0 + 4 * (lane % LPD) + 4 * LPD * ((j % CPS) * NG + k); };
  if (p.out_u16) {
    // compact output: the four sums of a group as uint16 (each below 2^16)
#pragma unroll
    for (int j = 0; j < DPW; ++j) {
      const uint32_t dm = dm_of(j);
      if (dm >= p.n_dms) continue;
      uint16_t *orow = reinterpret_cast<uint16_t *>(p.out) + (uint64_t)beam * p.out_beam_pitch +
                       (uint64_t)group * p.out_group_pitch + (uint64_t)dm * p.out_dm_pitch;
#pragma unroll
      for (int k = 0; k < NG; ++k) {
        uint32_t w0, w1;  // samples (0, 1) and (2, 3)
        if constexpr (FLUSH) {
          w0 = (wide[j][4 * k] & 0xFFFFu) | (wide[j][4 * k + 1] << 16);
          w1 = (wide[j][4 * k + 2] & 0xFFFFu) | (wide[j][4 * k + 3] << 16);
        } else {
          const uint32_t even = even_of(j, k);
          w0 = __builtin_amdgcn_perm(odd[j][k], even, 0x05040100u);
          w1 = __builtin_amdgcn_perm(odd[j][k], even, 0x07060302u);
        }
        const uint32_t t = t_of(j, k);
        if (p.vec_store && t + 3 < p.n_samples) {
          u32x2 v = {w0, w1};
          *reinterpret_cast<u32x2 *>(orow + t) = v;
        } else {
          const uint32_t e[4] = {w0 & 0xFFFFu, w0 >> 16, w1 & 0xFFFFu, w1 >> 16};
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (t + q < p.n_samples) orow[t + q] = (uint16_t)e[q];
        }
      }
    }
    return;
  }

#pragma unroll
  for (int j = 0; j < DPW; ++j) {
    const uint32_t dm = dm_of(j);
    if (dm >= p.n_dms) continue;
    float *orow = p.out + (uint64_t)beam * p.out_beam_pitch + (uint64_t)group * p.out_group_pitch +
                  (uint64_t)dm * p.out_dm_pitch;
#pragma unroll
    for (int k = 0; k < NG; ++k) {
      float v[4];
      if constexpr (FLUSH) {
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = (float)wide[j][4 * k + e];
      } else {
        const uint32_t even = even_of(j, k);
        v[0] = (float)(even & 0xFFFFu);
        v[1] = (float)(odd[j][k] & 0xFFFFu);
        v[2] = (float)(even >> 16);
        v[3] = (float)(odd[j][k] >> 16);
      }
      const uint32_t t = t_of(j, k);
      if (p.vec_store && t + 3 < p.n_samples) {
        f32x4 f = {v[0], v[1], v[2], v[3]};
        if (p.nontemporal) __builtin_nontemporal_store(f, reinterpret_cast<f32x4 *>(orow + t));
        else *reinterpret_cast<f32x4 *>(orow + t) = f;
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (t + e < p.n_samples) orow[t + e] = v[e];
      }
    }
  }
}

// Shapes whose sums fit in 64 VGPRs (RUN x DPW <= 128, no 32-bit flush
// sums) are held to four waves per SIMD: the loads issued a chunk ahead and
// the queue loop of the persistent form can otherwise push the allocation
// just past 128 and cost a wave.  Kernels that also read packed rows need
// registers for the unpacking and are held to four waves only up to
// RUN x DPW = 64.
template <int RUN, int DPW, bool FLUSH, bool PACKED>
constexpr int u8_min_waves() {
  return (!FLUSH && RUN * DPW <= (PACKED ? 64 : 128)) ? 4 : 1;
}

template <int RUN, int DPW, bool FLUSH, bool PACKED>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(u8_min_waves<RUN, DPW, FLUSH, PACKED>())))
rowsum_u8_staged(const RowSumU8Params p) {
  extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
  u8_tile<RUN, DPW, FLUSH, PACKED>(p, xcd_logical(blockIdx.x, p.n_tiles), lds);
}

// Persistent form: workgroups take (beam, tile) descriptors from XCD-local
// work queues until every queue is empty (see staged_common.hpp).
template <int RUN, int DPW, bool FLUSH, bool PACKED>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(u8_min_waves<RUN, DPW, FLUSH, PACKED>())))
rowsum_u8_batch(const RowSumU8Params *__restrict__ beams, const TileQueues q) {
  extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
  __shared__ uint32_t claimed;
  QueueCursor cur;
  for (;;) {
    const uint32_t t = cur.next(q, &claimed);
    if (t == kNoTile) break;
    const TileDesc d = q.tiles[t];
    u8_tile<RUN, DPW, FLUSH, PACKED>(beams[d.job], d.tile, lds);
  }
}

// One workgroup per listed (job, tile) descriptor, in XCD-aware order.
template <int RUN, int DPW, bool FLUSH, bool PACKED>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(u8_min_waves<RUN, DPW, FLUSH, PACKED>())))
rowsum_u8_list(const RowSumU8Params *__restrict__ jobs, const TileDesc *__restrict__ tiles, uint32_t n) {
  extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
  const TileDesc d = tiles[xcd_logical(blockIdx.x, n)];
  u8_tile<RUN, DPW, FLUSH, PACKED>(jobs[d.job], d.tile, lds);
}

// Calls f(RUN, DPW, FLUSH, PACKED) with the compile-time constants of
// geometry g; PACKED kernels also read packed rows (g.packed: some job of the
// launch has them), the others bytes only.
template <typename F> hipError_t dispatch(const Geometry &g, bool flush, F &&f) {
  auto by_dpw = [&](auto run, auto fl, auto pk) -> hipError_t {
    switch (g.items_d1) {
    case 1: return f(run, std::integral_constant<int, 1>{}, fl, pk);
    case 2: return f(run, std::integral_constant<int, 2>{}, fl, pk);
    case 4: return f(run, std::integral_constant<int, 4>{}, fl, pk);
    case 8: return f(run, std::integral_constant<int, 8>{}, fl, pk);
    default: return hipErrorInvalidValue;
    }
  };
  auto by_run = [&](auto fl, auto pk) -> hipError_t {
    switch (g.items_d0) {
    case 4: return by_dpw(std::integral_constant<int, 4>{}, fl, pk);
    case 8: return by_dpw(std::integral_constant<int, 8>{}, fl, pk);
    case 16: return by_dpw(std::integral_constant<int, 16>{}, fl, pk);
    default: return hipErrorInvalidValue;
    }
  };
  auto by_flush = [&](auto pk) -> hipError_t {
    return flush ? by_run(std::true_type{}, pk) : by_run(std::false_type{}, pk);
  };
  return g.packed ? by_flush(std::true_type{}) : by_flush(std::false_type{});
}

} // namespace

hipError_t launch_rowsum_u8_staged(const RowSumU8Params &p, const Geometry &g, bool flush, hipStream_t stream) {
  return dispatch(g, flush, [&](auto run, auto dpw, auto fl, auto pk) {
    hipLaunchKernelGGL((rowsum_u8_staged<run.value, dpw.value, fl.value, pk.value>), dim3(p.n_tiles), dim3(64 * g.threads_d1),
                       g.lds_bytes, stream, p);
    return hipGetLastError();
  });
}

hipError_t launch_rowsum_u8_batch(const RowSumU8Params *d_jobs, const TileQueues &q, const Geometry &g, bool flush,
                                  uint32_t workgroups, hipStream_t stream) {
  return dispatch(g, flush, [&](auto run, auto dpw, auto fl, auto pk) {
    hipLaunchKernelGGL((rowsum_u8_batch<run.value, dpw.value, fl.value, pk.value>), dim3(workgroups), dim3(64 * g.threads_d1),
                       g.lds_bytes, stream, d_jobs, q);
    return hipGetLastError();
  });
}

hipError_t launch_rowsum_u8_list(const RowSumU8Params *d_jobs, const TileDesc *d_tiles, uint32_t n, const Geometry &g,
                                 bool flush, hipStream_t stream) {
  return dispatch(g, flush, [&](auto run, auto dpw, auto fl, auto pk) {
    hipLaunchKernelGGL((rowsum_u8_list<run.value, dpw.value, fl.value, pk.value>), dim3(n), dim3(64 * g.threads_d1),
                       g.lds_bytes, stream, d_jobs, d_tiles, n);
    return hipGetLastError();
  });
}

hipError_t rowsum_u8_batch_occupancy(const Geometry &g, bool flush, int *per_cu) {
  return dispatch(g, flush, [&](auto run, auto dpw, auto fl, auto pk) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, rowsum_u8_batch<run.value, dpw.value, fl.value, pk.value>,
                                                        (int)(64 * g.threads_d1), (size_t)g.lds_bytes);
  });
}

} // namespace ddamd
