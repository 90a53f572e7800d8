// kernels_u16_pairs.hip -- step two over uint16 step-one sums, subbands added
// in pairs before the DMs read them.
//
// Step two (include/Dedispersion.hpp:155-176) sums, for fine DM d and sample
// t, the step-one rows of every subband at t + shift(d, subband).  For two
// subbands r0 and r1 (a pair) and a DM d, the two terms are
// in[r0][t + s0(d)] + in[r1][t + s1(d)] = Q_delta[t + s0(d)], with
// delta = s1(d) - s0(d) and Q_delta[u] = in[r0][u] + in[r1][u + delta].  Q
// depends on d only through delta, and neighbouring subbands have nearly equal
// delays, so the DMs of a tile take only a few values of delta per pair (two
// to four for the 32 fine DMs of a 1536-channel, 32-subband ARTS band).  Each
// workgroup therefore stages, per pair, one window of Q for every delta its
// DMs take, and every DM then reads one row per pair instead of one per
// subband: half the LDS reads and about half the additions of
// kernels_u16_staged.hip, for a few more windows staged.
//
// All sums are integers: a pair sum is at most twice a subband sum (the host
// takes this kernel only when that fits 16 bits), so Q is added lane-wise
// (v_pk_add_u16) and every DM's total is exact in any order; its float
// conversion equals the reference's sequential float sum of the same
// integer-valued floats.  Positions past a row's last valid sample may hold
// anything; their sums only reach outputs past the last sample, which are not
// stored, and a 16-bit lane never carries into the lane above it except from
// such a position.
//
// Layout and lanes:
//  * Windows are kept in two copies, copy 1 shifted down by one element, so a
//    DM's run starts dword-aligned in copy (offset & 1).  Lane l owns the
//    dwords l + 64 j (samples t0 + 2 l + 128 j and the next) of its DMs'
//    runs: every read is a conflict-free ds_read2st64_b32 pair.
//  * The staging lanes compute the shifts of the tile's DMs for the pair's two
//    subbands (the reference's unfused arithmetic), reduce delta to its range
//    over the tile, and publish for every DM the LDS byte address of its Q
//    window, copy and position.
//  * Double-buffered staging, one barrier per group of pairs, XCD-aware tile
//    order (staged_common.hpp), as in the other staged kernels.
#include "staged_common.hpp"

namespace ddamd {
namespace {

using namespace staged;

// Lane-wise sum of two pairs of uint16 (v_pk_add_u16; each half wraps on its own).
__device__ __forceinline__ uint32_t add_u16x2(uint32_t a, uint32_t b) {
  typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));
  return __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2, a) + __builtin_bit_cast(u16x2, b));
}

// Wave-wide minimum and maximum with DPP moves (VALU only, no LDS round
// trips): row_shr 1, 2, 4, 8 leave each row's extreme in its lane 15,
// row_bcast 15 carries lane 15 of rows 0 and 2 into rows 1 and 3, row_bcast
// 31 carries lane 31 into rows 2 and 3, and lane 63 then holds the extreme of
// all 64 lanes.  Lanes without a source keep their own value (bound_ctrl
// off, old = v), which min and max leave unchanged.
template <bool MAX> __device__ __forceinline__ int wave_extreme(int v) {
  auto pick = [](int a, int b) { return MAX ? max(a, b) : min(a, b); };
  v = pick(v, __builtin_amdgcn_update_dpp(v, v, 0x111, 0xf, 0xf, false));  // row_shr:1
  v = pick(v, __builtin_amdgcn_update_dpp(v, v, 0x112, 0xf, 0xf, false));  // row_shr:2
  v = pick(v, __builtin_amdgcn_update_dpp(v, v, 0x114, 0xf, 0xf, false));  // row_shr:4
  v = pick(v, __builtin_amdgcn_update_dpp(v, v, 0x118, 0xf, 0xf, false));  // row_shr:8
  v = pick(v, __builtin_amdgcn_update_dpp(v, v, 0x142, 0xa, 0xf, false));  // row_bcast:15
  v = pick(v, __builtin_amdgcn_update_dpp(v, v, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return __builtin_amdgcn_readlane(v, 63);
}

template <int RUN, int DPW>
__device__ __forceinline__ void pair_tile(const RowSumF32Params &p, uint32_t logical, uint32_t *lds) {
  constexpr int ND = RUN / 2;       // dwords (sample pairs) per lane per DM
  constexpr int TS = 64 * RUN;      // samples per tile
  constexpr int CPL = RUN / 8 + 1;  // 8-element chunks per lane per window
  constexpr int KMAX = kPairSlotsMax;

  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wave = uniform(threadIdx.x >> 6);
  const uint32_t nwaves = blockDim.x >> 6;
  const uint32_t TD = nwaves * DPW;
  const Tile tile = tile_split(logical, p.n_tiles_x, p.n_tiles_y);
  const uint32_t t0 = tile.x * TS;
  const uint32_t d0 = tile.y * TD;
  const uint32_t dw = wave * DPW;
  // a wavefront whose DMs all lie past the last DM (a partial DM tile) only stages
  const bool busy = d0 + dw < p.n_dms;
  const uint32_t coarse = tile.z % p.n_coarse;
  const uint32_t sbeam = tile.z / p.n_coarse;
  const uint32_t KS = p.pair_slots;      // Q windows per pair (1..KMAX)
  const uint32_t GP = p.unroll;          // pairs staged per barrier (<= waves)
  const uint32_t NP = (p.n_rows + 1) / 2;
  const uint32_t WE = p.win_floats;      // uint16 elements per window (multiple of 8)
  const uint32_t WD = WE / 2;            // dwords per window copy
  const uint32_t NCH = WE / 8;           // 8-element chunks per window
  const uint32_t dmax = p.n_dms - 1;
  const uint32_t rows_pad = (2 * p.n_rows + 3) & ~3u;
  const uint16_t *in16 = reinterpret_cast<const uint16_t *>(p.in);
  // staging: WPP waves per pair, each taking the chunks c with c % WPP == part
  const uint32_t WPP = max(1u, nwaves / GP);
  const uint32_t stage_pair = wave / WPP, part = wave % WPP;

  // LDS: rows (input beam, float delay) | windows [2][GP][KS][2 copies][WD] | DM addresses [2][GP][TD]
  uint32_t *lbeam = lds;
  float *ltab = reinterpret_cast<float *>(lds + p.n_rows);
  uint32_t *windows = lds + rows_pad;
  uint32_t *meta = windows + 2 * GP * KS * 2 * WD;

  for (uint32_t sb = threadIdx.x; sb < p.n_rows; sb += blockDim.x) {
    lbeam[sb] = p.map[(uint64_t)(p.first_sbeam + sbeam) * p.map_pitch + sb];
    ltab[sb] = p.table[sb];
  }
  __syncthreads();

  // the DM of this lane for the staging arithmetic (lanes past the tile
  // repeat its last DM)
  const uint32_t lane_dm = min(d0 + min(lane, TD - 1), dmax);

  // staging registers of the pair this wave stages, per chunk: r0 (4 dwords
  // and the next) and r1 (7 dwords from the dword holding element dmin)
  u32x4 a0v[CPL];
  uint32_t a0x[CPL];
  u32x4 b0v[CPL];
  u32x3 b1v[CPL];
  uint32_t s_k = 0, s_par = 0, s_word = 0;  // windows of the pair, r1 parity, the lane DM's LDS word
  bool s_has1 = false;
  auto stage_load = [&](uint32_t p0, uint32_t nn) {
    const uint32_t g = stage_pair;
    if (g >= nn) return;
    const uint32_t r0 = 2 * (p0 + g), r1 = r0 + 1;
    const bool has1 = r1 < p.n_rows;
    const uint32_t beam0 = uniform(lbeam[r0]);
    const float tab0 = uniformf(ltab[r0]);
    const uint32_t sh0 = shift_of(p.first_dm, p.dm_step, lane_dm, tab0);
    int delta = 0;
    uint32_t beam1 = 0;
    if (has1) {
      beam1 = uniform(lbeam[r1]);
      const float tab1 = uniformf(ltab[r1]);
      delta = (int)shift_of(p.first_dm, p.dm_step, lane_dm, tab1) - (int)sh0;
    }
    const int dlo = has1 ? wave_extreme<false>(delta) : 0;
    const int dhi = has1 ? wave_extreme<true>(delta) : 0;
    const uint32_t k = min((uint32_t)(dhi - dlo) + 1u, KS);
    const uint32_t shmin = uniform(sh0);  // lane 0: the tile's first DM, the smallest shift
    const uint64_t R0 = (uint64_t)beam0 * p.in_beam_pitch + (uint64_t)coarse * p.in_coarse_pitch +
                        (uint64_t)r0 * p.in_row_pitch;
    const uint64_t off = uniform64(R0 + t0 + shmin);
    const uint64_t a0 = off & ~uint64_t(7);  // 16-byte aligned window start
    const uint32_t o = sh0 - shmin + (uint32_t)(off - a0);  // element offset of the DM's run
    const uint32_t slot = min((uint32_t)(delta - dlo), KS - 1);
    s_word = 4 * (slot * 2 * WD + (o & 1) * WD + (o >> 1));
    s_k = uniform(k);
    s_has1 = has1;
    const uint32_t recs0 = records_from(2 * p.in_floats, 2 * a0);
    const __amdgpu_buffer_rsrc_t rsrc0 = make_rsrc(in16 + a0, recs0);
    // r1 element of window element e at delta: R1 + (a0 - R0) + e + delta
    // (a0 >= R0 and R1 >= one row pitch > |delta|: the host takes this kernel
    // only for rows 8-element aligned)
    uint64_t b1 = 0;
    uint32_t recs1 = 0;
    if (has1) {
      const uint64_t R1 = (uint64_t)beam1 * p.in_beam_pitch + (uint64_t)coarse * p.in_coarse_pitch +
                          (uint64_t)r1 * p.in_row_pitch;
      const uint64_t first = uniform64((uint64_t)((int64_t)(R1 + (a0 - R0)) + dlo));
      b1 = first & ~uint64_t(1);
      s_par = (uint32_t)(first & 1);
      recs1 = records_from(2 * p.in_floats, 2 * b1);
    }
    const __amdgpu_buffer_rsrc_t rsrc1 = make_rsrc(in16 + b1, recs1);
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      const uint32_t m = lane + 64 * c;
      if (c % WPP != part || m >= NCH) continue;
      if (16 * m + 20 <= recs0) {
        a0v[c] = __builtin_amdgcn_raw_buffer_load_b128(rsrc0, 16 * m, 0, 0);
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) a0v[c][e] = __builtin_amdgcn_raw_buffer_load_b32(rsrc0, 16 * m + 4 * e, 0, 0);
      }
      a0x[c] = __builtin_amdgcn_raw_buffer_load_b32(rsrc0, 16 * m + 16, 0, 0);
      if (has1) {
        if (16 * m + 28 <= recs1) {
          b0v[c] = __builtin_amdgcn_raw_buffer_load_b128(rsrc1, 16 * m, 0, 0);
          b1v[c] = __builtin_amdgcn_raw_buffer_load_b96(rsrc1, 16 * m + 16, 0, 0);
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e) b0v[c][e] = __builtin_amdgcn_raw_buffer_load_b32(rsrc1, 16 * m + 4 * e, 0, 0);
#pragma unroll
          for (int e = 0; e < 3; ++e) b1v[c][e] = __builtin_amdgcn_raw_buffer_load_b32(rsrc1, 16 * m + 16 + 4 * e, 0, 0);
        }
      }
    }
  };
  // Window kk of the pair holds Q at delta = dmin + kk; copy 1 is copy 0
  // shifted down by one element.
  auto stage_store = [&](uint32_t buf, uint32_t nn) {
    const uint32_t g = stage_pair;
    if (g >= nn) return;
    uint32_t *wp = windows + (buf * GP + g) * KS * 2 * WD;
    const uint32_t sh_par = 16 * s_par;
#pragma unroll
    for (int c = 0; c < CPL; ++c) {
      const uint32_t m = lane + 64 * c;
      if (c % WPP != part || m >= NCH) continue;
      const uint32_t r0d[5] = {a0v[c][0], a0v[c][1], a0v[c][2], a0v[c][3], a0x[c]};
      if (!s_has1) {
        // a lone last subband: Q is the row itself
        u32x4 q1;
#pragma unroll
        for (int i = 0; i < 4; ++i) q1[i] = __builtin_amdgcn_alignbit(r0d[i + 1], r0d[i], 16);
        *reinterpret_cast<u32x4 *>(wp + 4 * m) = a0v[c];
        *reinterpret_cast<u32x4 *>(wp + WD + 4 * m) = q1;
        continue;
      }
      // r1 from the dword holding element dmin of chunk m: 14 elements, of
      // which slot kk reads par + kk .. par + kk + 9 (kk < 4)
      const uint32_t r1d[8] = {b0v[c][0], b0v[c][1], b0v[c][2], b0v[c][3], b1v[c][0], b1v[c][1], b1v[c][2], 0u};
      // al[i] holds elements 2i, 2i + 1 from element dmin on (al[6]'s upper
      // element, past what any slot reads, is left zero)
      uint32_t al[7];
#pragma unroll
      for (int i = 0; i < 7; ++i) al[i] = __builtin_amdgcn_alignbit(r1d[i + 1], r1d[i], sh_par);
#pragma unroll
      for (int kk = 0; kk < KMAX; ++kk) {
        if ((uint32_t)kk >= s_k) break;
        uint32_t q[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) {
          const uint32_t r1k = (kk & 1) ? __builtin_amdgcn_alignbit(al[i + kk / 2 + 1], al[i + kk / 2], 16)
                                        : al[i + kk / 2];
          q[i] = add_u16x2(r0d[i], r1k);
        }
        u32x4 q0, q1;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          q0[i] = q[i];
          q1[i] = __builtin_amdgcn_alignbit(q[i + 1], q[i], 16);
        }
        uint32_t *w = wp + kk * 2 * WD;
        *reinterpret_cast<u32x4 *>(w + 4 * m) = q0;
        *reinterpret_cast<u32x4 *>(w + WD + 4 * m) = q1;
      }
    }
    // the DM's LDS byte address for the summing lanes
    if (part == 0 && lane < TD) meta[(buf * GP + g) * TD + lane] = 4 * (rows_pad + (buf * GP + g) * KS * 2 * WD) + s_word;
  };

  uint32_t lo[DPW][ND], hi[DPW][ND];
#pragma unroll
  for (int j = 0; j < DPW; ++j)
#pragma unroll
    for (int k = 0; k < ND; ++k) lo[j][k] = hi[j][k] = 0;

  // R pairs are first added as 16-bit lanes (R x the largest pair sum stays
  // below 2^16: p.combine_rows >= R), then each half goes to its 32-bit sum.
  // The DM words of the R pairs reach the wavefront with one LDS read (lane
  // l < R DPW reads the word of pair l / DPW, DM l % DPW) and are taken into
  // scalar registers (v_readlane); a DM's address is its word plus 4 x lane.
  const char *lane_base = reinterpret_cast<const char *>(lds) + 4 * lane;
  auto pairs = [&](auto R, uint32_t buf, uint32_t g) {
    constexpr int NR = R.value;
    const uint32_t words = meta[(buf * GP + g + (lane / DPW) % NR) * TD + dw + lane % DPW];
    auto src_of = [&](int r, int j) {
      return reinterpret_cast<const uint32_t *>(lane_base + (uint32_t)__builtin_amdgcn_readlane((int)words, r * DPW + j));
    };
    // pair 0 of DM j + 1 is loaded before DM j is summed, pair 1 at DM j's own step
    uint32_t xa[2][ND], xb[ND];  // [DM parity][dword], [dword]
    auto load = [&](int r, int j, uint32_t (&d)[ND]) {
      const uint32_t *src = src_of(r, j);
#pragma unroll
      for (int k = 0; k < ND; ++k) d[k] = src[64 * k];
    };
    load(0, 0, xa[0]);
#pragma unroll
    for (int j = 0; j < DPW; ++j) {
      if (NR > 1) load(1, j, xb);
      if (j + 1 < DPW) load(0, j + 1, xa[(j + 1) & 1]);
#pragma unroll
      for (int k = 0; k < ND; ++k) {
        uint32_t v = xa[j & 1][k];
        if (NR > 1) v += xb[k];
        lo[j][k] += v & 0xFFFFu;
        hi[j][k] += v >> 16;
      }
    }
  };
  const uint32_t combine = p.combine_rows;
  auto compute = [&](uint32_t buf, uint32_t n) {
    uint32_t g = 0;
    if (combine >= 2)
      for (; g + 2 <= n; g += 2) pairs(std::integral_constant<int, 2>{}, buf, g);
    for (; g < n; ++g) pairs(std::integral_constant<int, 1>{}, buf, g);
  };

  uint32_t cur = 0;
  {
    const uint32_t n0 = min(GP, NP);
    stage_load(0, n0);
    stage_store(0, n0);
  }
  __syncthreads();
  for (uint32_t p0 = 0; p0 < NP; p0 += GP) {
    const uint32_t n = min(GP, NP - p0);
    const uint32_t pn = p0 + GP;
    const uint32_t nn = pn < NP ? min(GP, NP - pn) : 0;
    if (nn) stage_load(pn, nn);
    if (busy) compute(cur, n);
    if (nn) stage_store(cur ^ 1, nn);
    __syncthreads();
    cur ^= 1;
  }

#pragma unroll
  for (int j = 0; j < DPW; ++j) {
    const uint32_t dm = d0 + dw + j;
    if (dm >= p.n_dms) continue;
    float *orow = p.out + (uint64_t)sbeam * p.out_beam_pitch + ((uint64_t)coarse * p.n_dms + dm) * p.out_dm_pitch;
#pragma unroll
    for (int k = 0; k < ND; ++k) {
      const uint32_t t = t0 + 2 * lane + 128 * k;
      const float v0 = (float)lo[j][k], v1 = (float)hi[j][k];
      if (p.vec_store && t + 1 < p.n_samples) {
        typedef float f32x2 __attribute__((ext_vector_type(2)));
        const f32x2 f = {v0, v1};
        if (p.nontemporal) __builtin_nontemporal_store(f, reinterpret_cast<f32x2 *>(orow + t));
        else *reinterpret_cast<f32x2 *>(orow + t) = f;
      } else {
        if (t < p.n_samples) orow[t] = v0;
        if (t + 1 < p.n_samples) orow[t + 1] = v1;
      }
    }
  }
}

template <int RUN, int DPW>
__global__ void __launch_bounds__(512)
rowsum_u16_pairs_batch(const RowSumF32Params *__restrict__ beams, const TileQueues q) {
  extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
  __shared__ uint32_t claimed;
  QueueCursor cur;
  for (;;) {
    const uint32_t t = cur.next(q, &claimed);
    if (t == kNoTile) break;
    const TileDesc d = q.tiles[t];
    pair_tile<RUN, DPW>(beams[d.job], d.tile, lds);
  }
}

// One workgroup per listed (job, tile) descriptor, in XCD-aware order.
template <int RUN, int DPW>
__global__ void __launch_bounds__(512)
rowsum_u16_pairs_list(const RowSumF32Params *__restrict__ jobs, const TileDesc *__restrict__ tiles, uint32_t n) {
  extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
  const TileDesc d = tiles[xcd_logical(blockIdx.x, n)];
  pair_tile<RUN, DPW>(jobs[d.job], d.tile, lds);
}

template <typename F> hipError_t dispatch(const Geometry &g, F &&f) {
  auto by_dpw = [&](auto run) -> hipError_t {
    switch (g.items_d1) {
    case 1: return f(run, std::integral_constant<int, 1>{});
    case 2: return f(run, std::integral_constant<int, 2>{});
    case 4: return f(run, std::integral_constant<int, 4>{});
    case 8: return f(run, std::integral_constant<int, 8>{});
    default: return hipErrorInvalidValue;
    }
  };
  switch (g.items_d0) {
  case 4: return by_dpw(std::integral_constant<int, 4>{});
  case 8: return by_dpw(std::integral_constant<int, 8>{});
  case 16: return by_dpw(std::integral_constant<int, 16>{});
  default: return hipErrorInvalidValue;
  }
}

} // namespace

hipError_t launch_rowsum_u16_pairs_batch(const RowSumF32Params *d_jobs, const TileQueues &q, const Geometry &g,
                                         uint32_t workgroups, hipStream_t stream) {
  return dispatch(g, [&](auto run, auto dpw) {
    hipLaunchKernelGGL((rowsum_u16_pairs_batch<run.value, dpw.value>), dim3(workgroups), dim3(64 * g.threads_d1),
                       g.lds_bytes, stream, d_jobs, q);
    return hipGetLastError();
  });
}

hipError_t launch_rowsum_u16_pairs_list(const RowSumF32Params *d_jobs, const TileDesc *d_tiles, uint32_t n,
                                        const Geometry &g, hipStream_t stream) {
  return dispatch(g, [&](auto run, auto dpw) {
    hipLaunchKernelGGL((rowsum_u16_pairs_list<run.value, dpw.value>), dim3(n), dim3(64 * g.threads_d1), g.lds_bytes,
                       stream, d_jobs, d_tiles, n);
    return hipGetLastError();
  });
}

hipError_t rowsum_u16_pairs_batch_occupancy(const Geometry &g, int *per_cu) {
  return dispatch(g, [&](auto run, auto dpw) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, rowsum_u16_pairs_batch<run.value, dpw.value>,
                                                        (int)(64 * g.threads_d1), (size_t)g.lds_bytes);
  });
}

} // namespace ddamd
