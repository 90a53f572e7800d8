// kernels_u8_factored.hip -- single-step dedispersion with shared channel
// patterns (the operation of include/Dedispersion.hpp:70-109).
//
// The live channels are taken four at a time (a block).  Over a tile of 32
// consecutive DMs, the shifts of a block's channels relative to its last
// channel -- the block's pattern -- take only a few distinct values (3.7 on
// average for 1536 ARTS channels at a DM step of 0.2).  For every distinct
// pattern q the block sum
//     Q_q[u] = sum_i in[c_i][u + D_i(q)]        (D_i = s_i - s_base)
// is built once per tile, and a DM d with pattern q(d) adds Q_q(d)[t + s_base(d)]
// to its output: one addition per block instead of four.  All sums are
// integers below 2^24, so any order is exact and the result equals the
// reference's float sum bit for bit.
//
// Host-built records (one per DM tile and block, from the same integer shift
// table as the other kernels) give the patterns, each DM's pattern and base
// offset, and the base shift of the tile's first DM.  A workgroup owns a
// (beam, DM tile, sample tile) block and walks the channel blocks in phases:
//   A. stage the rows of up to `unroll` blocks into LDS byte windows;
//   B. build the patterns' Q rows as uint16 pairs, each with a copy shifted
//      by one element so that every DM read is dword aligned (16-sample tasks
//      spread over all threads);
//   D. every wavefront adds, for each of its DMs and each block, one dword
//      (two samples) per pair of output samples into 16-bit pair sums, which
//      are widened into 32-bit sums every 64 blocks (64 x 1020 < 2^16).
#include "staged_common.hpp"

namespace ddamd {
namespace {

using namespace staged;

__device__ __forceinline__ uint32_t sel_even_of(uint32_t r) { return r | 0x0c00u | ((r + 2) << 16) | 0x0c000000u; }
__device__ __forceinline__ uint32_t sel_odd_of(uint32_t r) { return (r + 1) | 0x0c00u | ((r + 3) << 16) | 0x0c000000u; }

template <int RUN, int DPW>
__global__ void __launch_bounds__(512) rowsum_u8_factored(const RowSumU8Params p) {
  constexpr int NP = RUN / 2;           // sample pairs per lane
  constexpr int TS = 64 * RUN;          // samples per tile
  constexpr uint32_t TD = kFactoredDMs; // DMs per tile
  constexpr uint32_t REC = kFactoredRecordDwords;
  extern __shared__ __attribute__((aligned(16))) uint32_t lds[];

  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wave = uniform(threadIdx.x >> 6);
  const uint32_t nwaves = blockDim.x >> 6;
  const Tile tile = tile_split(xcd_logical(blockIdx.x, p.n_tiles), p.n_tiles_x, p.n_tiles_y);
  const uint32_t t0 = tile.x * TS;
  const uint32_t d0 = tile.y * TD;
  const uint32_t dw = wave * DPW;
  const uint32_t sbeam = tile.z;
  const uint32_t nlive = p.live_count[0];
  const uint32_t nblocks = (nlive + 3) / 4;
  const uint32_t KB = p.unroll;           // blocks per phase (at most)
  const uint32_t S = p.fact_slots;        // Q rows per phase
  const uint32_t QW = p.fact_qw;          // uint16 elements per Q row (multiple of 16)
  const uint32_t QD = QW / 2;             // dwords per Q copy
  const uint32_t RW = p.win_bytes;        // bytes per staged row window (multiple of 16)
  const uint32_t RWD = RW / 4;
  const uint32_t *recs = p.fact_records + (uint64_t)tile.y * nblocks * REC;

  // LDS: rows [4 KB][RWD] | Q slots [S][2][QD] | block records [KB][REC] | row misalignment [4 KB]
  uint32_t *rows = lds;
  uint32_t *slots = rows + 4 * KB * RWD;
  uint32_t *brec = slots + S * 2 * QD;
  uint32_t *rmis = brec + KB * REC;

  uint32_t acc[DPW][NP], lo[DPW][NP], hi[DPW][NP];
#pragma unroll
  for (int j = 0; j < DPW; ++j)
#pragma unroll
    for (int k = 0; k < NP; ++k) acc[j][k] = lo[j][k] = hi[j][k] = 0;
  uint32_t since = 0;  // blocks added since the last widening

  for (uint32_t kb = 0; kb < nblocks;) {
    // blocks of this phase: as many as fit the Q slots (every record has P <= S)
    uint32_t nb = 0, used = 0;
    while (kb + nb < nblocks && nb < KB) {
      const uint32_t P = recs[(uint64_t)(kb + nb) * REC] & 0xFFu;
      if (used + P > S) break;
      used += P;
      ++nb;
    }
    nb = uniform(nb);
    if (nb == 0) break;  // cannot happen: the host sizes the slots for the largest record

    // ---- A: records to LDS, rows staged ----
    for (uint32_t q = threadIdx.x; q < nb * REC; q += blockDim.x) brec[q] = recs[(uint64_t)kb * REC + q];
    for (uint32_t r = wave; r < 4 * nb; r += nwaves) {
      const uint32_t b = r / 4, i = r % 4;
      const uint32_t *rec = recs + (uint64_t)(kb + b) * REC;
      const uint32_t nrows = (rec[0] >> 8) & 0xFFu;
      if (i >= nrows) continue;
      const uint32_t row = p.live[(kb + b) * 4 + i];
      const uint32_t in_beam = p.map ? p.map[(uint64_t)(p.first_sbeam + sbeam) * p.map_pitch + row] : sbeam;
      const uint32_t min_d = (rec[2] >> (8 * i)) & 0xFFu;
      const uint64_t start = (uint64_t)in_beam * p.in_beam_pitch + (uint64_t)row * p.in_row_pitch + t0 + rec[1] + min_d;
      const uint64_t a0 = uniform64(start & ~uint64_t(3));
      if (lane == 0) rmis[r] = (uint32_t)(start - a0);
      const uint32_t recsb = records_from(p.in_bytes, a0);
      const __amdgpu_buffer_rsrc_t rsrc = make_rsrc(p.in + a0, recsb);
      uint32_t *w = rows + r * RWD;
      for (uint32_t m = lane; m < RW / 16; m += 64) {
        u32x4 v;
        if (16 * m + 16 <= recsb) {
          v = __builtin_amdgcn_raw_buffer_load_b128(rsrc, 16 * m, 0, 0);
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e) v[e] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, 16 * m + 4 * e, 0, 0);
        }
        *reinterpret_cast<u32x4 *>(w + 4 * m) = v;
      }
    }
    __syncthreads();

    // ---- B: Q rows and their shifted copies.  Tasks of 16 samples (four
    // groups of four) of one (block, pattern) are spread over all threads;
    // a group's two dwords per row also hold the fifth sample, so the copy
    // shifted by one element is written in the same pass. ----
    {
      const uint32_t G16 = QW / 16;  // tasks per Q row
      for (uint32_t task = threadIdx.x; task < used * G16; task += blockDim.x) {
        const uint32_t slot = task / G16, g0 = 4 * (task % G16);
        // the (block, pattern) of this slot
        uint32_t b = 0, first = 0;
        for (;;) {
          const uint32_t P = brec[b * REC] & 0xFFu;
          if (slot < first + P || b + 1 == nb) break;
          first += P;
          ++b;
        }
        const uint32_t *rec = brec + b * REC;
        const uint32_t nrows = (rec[0] >> 8) & 0xFFu;
        const uint32_t pat = rec[kFactoredPatternWord + (slot - first)], mins = rec[2];
        uint32_t *q0 = slots + slot * 2 * QD, *q1 = q0 + QD;
        uint32_t off[4], se[4], sd[4], s5[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const uint32_t o = ((pat >> (8 * i)) & 0xFFu) - ((mins >> (8 * i)) & 0xFFu) + rmis[4 * b + i];
          off[i] = o >> 2;
          se[i] = sel_even_of(o & 3);
          sd[i] = sel_odd_of(o & 3);
          s5[i] = ((o & 3) + 4) | 0x0c0c0c00u;  // the fifth byte, zero-extended
        }
#pragma unroll
        for (int gg = 0; gg < 4; ++gg) {
          const uint32_t g = g0 + gg;
          uint32_t ev = 0, od = 0, fifth = 0;
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            if ((uint32_t)i < nrows) {
              const uint32_t *src = rows + (4 * b + i) * RWD + off[i] + g;
              const uint32_t x0 = src[0], x1 = src[1];
              ev += __builtin_amdgcn_perm(x1, x0, se[i]);
              od += __builtin_amdgcn_perm(x1, x0, sd[i]);
              fifth += __builtin_amdgcn_perm(x1, x0, s5[i]);
            }
          }
          // copy 0: (Q0, Q1), (Q2, Q3); copy 1: (Q1, Q2), (Q3, Q4)
          u32x2 v0 = {__builtin_amdgcn_perm(od, ev, 0x05040100u), __builtin_amdgcn_perm(od, ev, 0x07060302u)};
          u32x2 v1 = {__builtin_amdgcn_perm(ev, od, 0x07060100u), __builtin_amdgcn_perm(fifth, od, 0x05040302u)};
          *reinterpret_cast<u32x2 *>(q0 + 2 * g) = v0;
          *reinterpret_cast<u32x2 *>(q1 + 2 * g) = v1;
        }
      }
    }
    __syncthreads();

    // ---- D: DM sums ----
    {
      uint32_t slot_base = 0;
      for (uint32_t b = 0; b < nb; ++b) {
        const uint32_t *rec = brec + b * REC;
        const uint32_t *pr = slots + slot_base * 2 * QD + lane;
#pragma unroll
        for (int j = 0; j < DPW; ++j) {
          const uint32_t dm_word = rec[kFactoredDMWord + (dw + j) / 2];  // broadcast read
          const uint32_t e = ((dw + j) & 1) ? dm_word >> 16 : dm_word & 0xFFFFu;
          const uint32_t o = e & 0xFFu, pidx = e >> 8;
          const uint32_t *src = pr + pidx * 2 * QD + (o & 1) * QD + (o >> 1);
#pragma unroll
          for (int k = 0; k < NP; ++k) acc[j][k] += src[64 * k];
        }
        slot_base += uniform(rec[0]) & 0xFFu;
        if (++since == 64) {
#pragma unroll
          for (int j = 0; j < DPW; ++j)
#pragma unroll
            for (int k = 0; k < NP; ++k) {
              lo[j][k] += acc[j][k] & 0xFFFFu;
              hi[j][k] += acc[j][k] >> 16;
              acc[j][k] = 0;
            }
          since = 0;
        }
      }
    }
    __syncthreads();
    kb += nb;
  }

#pragma unroll
  for (int j = 0; j < DPW; ++j)
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      lo[j][k] += acc[j][k] & 0xFFFFu;
      hi[j][k] += acc[j][k] >> 16;
    }

  typedef float f32x2 __attribute__((ext_vector_type(2)));
#pragma unroll
  for (int j = 0; j < DPW; ++j) {
    const uint32_t dm = d0 + dw + j;
    if (dm >= p.n_dms) continue;
    float *orow = p.out + (uint64_t)sbeam * p.out_beam_pitch + (uint64_t)dm * p.out_dm_pitch;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const uint32_t t = t0 + 2 * lane + 128 * k;
      const float a = (float)lo[j][k], b = (float)hi[j][k];
      if (p.vec_store && t + 1 < p.n_samples) {
        f32x2 v = {a, b};
        if (p.nontemporal) __builtin_nontemporal_store(v, reinterpret_cast<f32x2 *>(orow + t));
        else *reinterpret_cast<f32x2 *>(orow + t) = v;
      } else {
        if (t < p.n_samples) orow[t] = a;
        if (t + 1 < p.n_samples) orow[t + 1] = b;
      }
    }
  }
}

template <int RUN>
hipError_t by_dpw(const RowSumU8Params &p, const Geometry &g, hipStream_t s) {
  switch (g.items_d1) {
  case 4:
    hipLaunchKernelGGL((rowsum_u8_factored<RUN, 4>), dim3(p.n_tiles), dim3(64 * g.threads_d1), g.lds_bytes, s, p);
    return hipGetLastError();
  case 8:
    hipLaunchKernelGGL((rowsum_u8_factored<RUN, 8>), dim3(p.n_tiles), dim3(64 * g.threads_d1), g.lds_bytes, s, p);
    return hipGetLastError();
  default: return hipErrorInvalidValue;
  }
}

} // namespace

hipError_t launch_rowsum_u8_factored(const RowSumU8Params &p, const Geometry &g, hipStream_t stream) {
  if (g.threads_d1 * g.items_d1 != kFactoredDMs || p.in_log2f != 0) return hipErrorInvalidValue;
  switch (g.items_d0) {
  case 8: return by_dpw<8>(p, g, stream);
  case 16: return by_dpw<16>(p, g, stream);
  default: return hipErrorInvalidValue;
  }
}

} // namespace ddamd
