// kernels_u16_staged.hip -- step two over uint16 step-one sums.
//
// In a two-step pipeline over uint8 input, every step-one output is an
// integer sum of at most (channels per subband) uint8 values, so a batch that
// owns its intermediate can keep it as uint16 (exact when 255 x live
// channels of a subband <= 65535).  Step two (include/Dedispersion.hpp:155-176)
// then adds integers: the sum over subbands is exact in any order and its
// float conversion equals the reference's sequential float sum of the same
// integer-valued floats, bit for bit.
//
// The pipeline is that of kernels_f32_staged.hip (double-buffered subband
// staging, shifts from the float delay table published through LDS,
// XCD-aware order).  What changes:
//  * Lane l owns the sample quads t0 + 4l + 256k; a quad is one ds_read_b64,
//    so a DM's run of 64 quads is one conflict-free 8-byte read per lane and
//    the LDS reads per sample are a quarter of those of float rows.
//  * A DM's offset into the window is any element; every staged row is kept
//    in four copies, copy c shifted down by c elements, so all reads are
//    8-byte aligned (misaligned LDS reads are correct on gfx950 but several
//    times slower).  The shifted copies are built in registers with
//    v_alignbit from the 16-byte chunk plus the next two dwords.
//  * Up to four rows are added as 16-bit pairs (exact while four subband
//    sums stay below 2^16), then each half of the result is added to its own
//    32-bit sum (SDWA adds, v_add_u32 with src0_sel WORD_0 / WORD_1).
//  * The staging lanes store each DM's LDS byte address (row window, copy
//    and position) in the row's metadata; the summing wavefront reads the
//    words of a step's rows with one lane-distinct read, takes each into a
//    scalar register (v_readlane) and adds its lane offset, one add per row
//    and DM.
#include "staged_common.hpp"

namespace ddamd {
namespace {

using namespace staged;

template <int RUN, int DPW>
__device__ __forceinline__ void u16_tile(const RowSumF32Params &p, uint32_t logical, uint32_t *lds) {
  constexpr int NQ = RUN / 4;   // sample quads per lane
  constexpr int TS = 64 * RUN;  // samples per tile
  constexpr int SPW = kStagedSlotsPerWave;
  constexpr int CPL = RUN / 8 + 1;  // 16-byte staging loads per lane per row
  typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wave = uniform(threadIdx.x >> 6);
  const uint32_t nwaves = blockDim.x >> 6;
  const uint32_t TD = nwaves * DPW;
  const Tile tile = tile_split(logical, p.n_tiles_x, p.n_tiles_y);
  const uint32_t t0 = tile.x * TS;
  const uint32_t d0 = tile.y * TD;
  const uint32_t dw = wave * DPW;
  // a wavefront whose DMs all lie past the last DM (a partial DM tile) only stages rows
  const bool busy = d0 + dw < p.n_dms;
  const uint32_t coarse = tile.z % p.n_coarse;
  const uint32_t sbeam = tile.z / p.n_coarse;
  const uint32_t G = p.unroll;
  const uint32_t WE = p.win_floats;  // uint16 elements per window (multiple of 8)
  const uint32_t WD = WE / 2;        // dwords per window copy
  const uint32_t NCH = WE / 8;       // 16-byte chunks per window
  const uint32_t dmax = p.n_dms - 1;
  const uint32_t rows_pad = (2 * p.n_rows + 3) & ~3u;
  const uint16_t *in16 = reinterpret_cast<const uint16_t *>(p.in);

  // LDS: rows (input beam, float delay) | windows [2][G][4 copies][WD] | DM addresses [2][G][TD]
  uint32_t *lbeam = lds;
  float *ltab = reinterpret_cast<float *>(lds + p.n_rows);
  uint32_t *windows = lds + rows_pad;
  uint32_t *meta = windows + 2 * G * 4 * WD;

  for (uint32_t sb = threadIdx.x; sb < p.n_rows; sb += blockDim.x) {
    lbeam[sb] = p.map[(uint64_t)(p.first_sbeam + sbeam) * p.map_pitch + sb];
    ltab[sb] = p.table[sb];
  }
  __syncthreads();

  u32x4 sv[SPW][CPL];
  u32x2 sx[SPW][CPL];
  uint32_t so[SPW];
  auto stage_load = [&](uint32_t s0, uint32_t nn) {
#pragma unroll
    for (int s = 0; s < SPW; ++s) {
      const uint32_t g = wave + s * nwaves;
      if (g < nn) {
        const uint32_t sb = s0 + g;
        const uint32_t in_beam = uniform(lbeam[sb]);
        const float tab = uniformf(ltab[sb]);
        const uint32_t sh = shift_of(p.first_dm, p.dm_step, min(d0 + lane, dmax), tab);
        const uint32_t shmin = uniform(sh);
        const uint64_t off = uniform64((uint64_t)in_beam * p.in_beam_pitch + (uint64_t)coarse * p.in_coarse_pitch +
                                       (uint64_t)sb * p.in_row_pitch + t0 + shmin);
        const uint64_t a0 = off & ~uint64_t(7);  // 16-byte aligned window start
        const uint32_t o = sh - shmin + (uint32_t)(off - a0);  // element offset in the window
        // byte offset in the row's four copies: copy (o & 3), quad o >> 2
        so[s] = (o & 3) * 4 * WD + 8 * (o >> 2);
        const uint32_t recs = records_from(2 * p.in_floats, 2 * a0);
        const __amdgpu_buffer_rsrc_t rsrc = make_rsrc(in16 + a0, recs);
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          const uint32_t m = lane + 64 * c;
          if (m < NCH) {
            if (16 * m + 24 <= recs) {
              sv[s][c] = __builtin_amdgcn_raw_buffer_load_b128(rsrc, 16 * m, 0, 0);
              sx[s][c] = __builtin_amdgcn_raw_buffer_load_b64(rsrc, 16 * m + 16, 0, 0);
            } else {
#pragma unroll
              for (int e = 0; e < 4; ++e) sv[s][c][e] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, 16 * m + 4 * e, 0, 0);
              sx[s][c][0] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, 16 * m + 16, 0, 0);
              sx[s][c][1] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, 16 * m + 20, 0, 0);
            }
          }
        }
      }
    }
  };
  // Copy c of a row holds the window shifted down by c elements, so the quad
  // at any element offset o starts 8-byte aligned in copy o & 3.
  auto stage_store = [&](uint32_t buf, uint32_t nn) {
#pragma unroll
    for (int s = 0; s < SPW; ++s) {
      const uint32_t g = wave + s * nwaves;
      if (g < nn) {
        uint32_t *w0 = windows + ((buf * G + g) * 4) * WD;
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          const uint32_t m = lane + 64 * c;
          if (m < NCH) {
            const u32x4 v = sv[s][c];
            const uint32_t x0 = sx[s][c][0], x1 = sx[s][c][1];
            u32x4 r1, r2, r3;
            r1[0] = __builtin_amdgcn_alignbit(v[1], v[0], 16);
            r1[1] = __builtin_amdgcn_alignbit(v[2], v[1], 16);
            r1[2] = __builtin_amdgcn_alignbit(v[3], v[2], 16);
            r1[3] = __builtin_amdgcn_alignbit(x0, v[3], 16);
            r2[0] = v[1];
            r2[1] = v[2];
            r2[2] = v[3];
            r2[3] = x0;
            r3[0] = r1[1];
            r3[1] = r1[2];
            r3[2] = r1[3];
            r3[3] = __builtin_amdgcn_alignbit(x1, x0, 16);
            *reinterpret_cast<u32x4 *>(w0 + 4 * m) = v;
            *reinterpret_cast<u32x4 *>(w0 + WD + 4 * m) = r1;
            *reinterpret_cast<u32x4 *>(w0 + 2 * WD + 4 * m) = r2;
            *reinterpret_cast<u32x4 *>(w0 + 3 * WD + 4 * m) = r3;
          }
        }
        // the DM's LDS byte address for the summing lanes
        if (lane < TD) meta[(buf * G + g) * TD + lane] = 4 * (rows_pad + ((buf * G + g) * 4) * WD) + so[s];
      }
    }
  };

  uint32_t lo[DPW][2 * NQ], hi[DPW][2 * NQ];
#pragma unroll
  for (int j = 0; j < DPW; ++j)
#pragma unroll
    for (int k = 0; k < 2 * NQ; ++k) lo[j][k] = hi[j][k] = 0;

  // R rows are first added as 16-bit pairs (R x the largest subband sum
  // stays below 2^16: p.combine_rows >= R), then each half goes to its
  // 32-bit sum with one SDWA add.  The DM words of the R rows reach the
  // wavefront with one LDS read (lane l < R DPW reads the word of row
  // l / DPW, DM l % DPW) and are taken from their lanes into scalar
  // registers (v_readlane); a DM's address is its word plus the lane
  // offset, and one ds_read_b64 fetches a quad.
  const char *lane_base = reinterpret_cast<const char *>(lds) + 8 * lane;
  auto rows = [&](auto R, uint32_t buf, uint32_t g) {
    constexpr int NR = R.value;
    const uint32_t words = meta[(buf * G + g + (lane / DPW) % NR) * TD + dw + lane % DPW];
    // rows [0, PF) of DM j + 1 are loaded before DM j is summed, the others
    // (R = 4) at DM j's own step
    constexpr int PF = NR > 2 ? 2 : NR;
    auto src_of = [&](int r, int j) {
      return reinterpret_cast<const u32x2 *>(lane_base + (uint32_t)__builtin_amdgcn_readlane((int)words, r * DPW + j));
    };
    u32x2 xa[2][PF][NQ];  // [DM parity][row][quad]
    auto load = [&](int j, u32x2 (&d)[PF][NQ]) {
#pragma unroll
      for (int r = 0; r < PF; ++r) {
        const u32x2 *src = src_of(r, j);
#pragma unroll
        for (int k = 0; k < NQ; ++k) d[r][k] = src[64 * k];
      }
    };
    load(0, xa[0]);
#pragma unroll
    for (int j = 0; j < DPW; ++j) {
      u32x2 xb[NR - PF + 1][NQ];
#pragma unroll
      for (int r = PF; r < NR; ++r) {
        const u32x2 *src = src_of(r, j);
#pragma unroll
        for (int k = 0; k < NQ; ++k) xb[r - PF][k] = src[64 * k];
      }
      if (j + 1 < DPW) load(j + 1, xa[(j + 1) & 1]);
#pragma unroll
      for (int k = 0; k < NQ; ++k) {
        u32x2 v = xa[j & 1][0][k];
#pragma unroll
        for (int r = 1; r < PF; ++r) v += xa[j & 1][r][k];
#pragma unroll
        for (int r = PF; r < NR; ++r) v += xb[r - PF][k];
        lo[j][2 * k] += v[0] & 0xFFFFu;
        hi[j][2 * k] += v[0] >> 16;
        lo[j][2 * k + 1] += v[1] & 0xFFFFu;
        hi[j][2 * k + 1] += v[1] >> 16;
      }
    }
  };
  const uint32_t combine = p.combine_rows;
  auto compute = [&](uint32_t buf, uint32_t n) {
    uint32_t g = 0;
    if (combine >= 4)
      for (; g + 4 <= n; g += 4) rows(std::integral_constant<int, 4>{}, buf, g);
    if (combine >= 2)
      for (; g + 2 <= n; g += 2) rows(std::integral_constant<int, 2>{}, buf, g);
    for (; g < n; ++g) rows(std::integral_constant<int, 1>{}, buf, g);
  };

  uint32_t cur = 0;
  {
    const uint32_t n0 = min(G, p.n_rows);
    stage_load(0, n0);
    stage_store(0, n0);
  }
  __syncthreads();
  for (uint32_t s0 = 0; s0 < p.n_rows; s0 += G) {
    const uint32_t n = min(G, p.n_rows - s0);
    const uint32_t sn = s0 + G;
    const uint32_t nn = sn < p.n_rows ? min(G, p.n_rows - sn) : 0;
    if (nn) stage_load(sn, nn);
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
    for (int k = 0; k < NQ; ++k) {
      const uint32_t t = t0 + 4 * lane + 256 * k;
      const float v[4] = {(float)lo[j][2 * k], (float)hi[j][2 * k], (float)lo[j][2 * k + 1], (float)hi[j][2 * k + 1]};
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

template <int RUN, int DPW>
__global__ void __launch_bounds__(512)
rowsum_u16_staged(const RowSumF32Params p) {
  extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
  u16_tile<RUN, DPW>(p, xcd_logical(blockIdx.x, p.n_tiles), lds);
}

template <int RUN, int DPW>
__global__ void __launch_bounds__(512)
rowsum_u16_batch(const RowSumF32Params *__restrict__ beams, const TileQueues q) {
  extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
  __shared__ uint32_t claimed;
  QueueCursor cur;
  for (;;) {
    const uint32_t t = cur.next(q, &claimed);
    if (t == kNoTile) break;
    const TileDesc d = q.tiles[t];
    u16_tile<RUN, DPW>(beams[d.job], d.tile, lds);
  }
}

// One workgroup per listed (job, tile) descriptor, in XCD-aware order.
template <int RUN, int DPW>
__global__ void __launch_bounds__(512)
rowsum_u16_list(const RowSumF32Params *__restrict__ jobs, const TileDesc *__restrict__ tiles, uint32_t n) {
  extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
  const TileDesc d = tiles[xcd_logical(blockIdx.x, n)];
  u16_tile<RUN, DPW>(jobs[d.job], d.tile, lds);
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

hipError_t launch_rowsum_u16_staged(const RowSumF32Params &p, const Geometry &g, hipStream_t stream) {
  return dispatch(g, [&](auto run, auto dpw) {
    hipLaunchKernelGGL((rowsum_u16_staged<run.value, dpw.value>), dim3(p.n_tiles), dim3(64 * g.threads_d1),
                       g.lds_bytes, stream, p);
    return hipGetLastError();
  });
}

hipError_t launch_rowsum_u16_batch(const RowSumF32Params *d_jobs, const TileQueues &q, const Geometry &g,
                                   uint32_t workgroups, hipStream_t stream) {
  return dispatch(g, [&](auto run, auto dpw) {
    hipLaunchKernelGGL((rowsum_u16_batch<run.value, dpw.value>), dim3(workgroups), dim3(64 * g.threads_d1),
                       g.lds_bytes, stream, d_jobs, q);
    return hipGetLastError();
  });
}

hipError_t launch_rowsum_u16_list(const RowSumF32Params *d_jobs, const TileDesc *d_tiles, uint32_t n,
                                  const Geometry &g, hipStream_t stream) {
  return dispatch(g, [&](auto run, auto dpw) {
    hipLaunchKernelGGL((rowsum_u16_list<run.value, dpw.value>), dim3(n), dim3(64 * g.threads_d1), g.lds_bytes, stream,
                       d_jobs, d_tiles, n);
    return hipGetLastError();
  });
}

hipError_t rowsum_u16_batch_occupancy(const Geometry &g, int *per_cu) {
  return dispatch(g, [&](auto run, auto dpw) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, rowsum_u16_batch<run.value, dpw.value>,
                                                        (int)(64 * g.threads_d1), (size_t)g.lds_bytes);
  });
}

} // namespace ddamd
