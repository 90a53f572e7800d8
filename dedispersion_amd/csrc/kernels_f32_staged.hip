// kernels_f32_staged.hip -- pipelined shifted row sums over float rows.
//
// Same operation and arithmetic as kernels_f32.hip: step two
// (include/Dedispersion.hpp:155-176), and the single step and step one over
// float input (:70-153 with I = float).  A slice's rows (subbands, or live
// channels) are added in order into float accumulators that start at 0,
// bit-identical to the reference's sequential float sum for any input.  The
// pipeline is that of kernels_u8_staged.hip: double-buffered row staging in
// LDS with one barrier per group, integer shifts computed once per staged row
// from the float delay table and published through LDS, the slice's rows and
// input beams copied to LDS, XCD-aware workgroup order.  Lane l owns samples
// t0 + l + 64k and reads a DM's run with ds_read2st64_b32 pairs from one
// base address.
#include "staged_common.hpp"

namespace ddamd {
namespace {

using namespace staged;

// One (slice, DM tile, sample tile) block: logical tile index `logical` of
// the launch described by p (DM tile fastest).
template <int RUN, int DPW>
__device__ __forceinline__ void f32_tile(const RowSumF32Params &p, uint32_t logical, float *ldsf) {
  constexpr int TS = 64 * RUN;
  constexpr int SPW = kStagedSlotsPerWave;
  constexpr int CPL = RUN / 4 + 1;  // 16-byte staging loads per lane per row

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
  const uint32_t inner = tile.z % p.n_coarse;
  const uint32_t outer = tile.z / p.n_coarse;
  const uint32_t G = p.unroll;
  const uint32_t WF = p.win_floats;  // floats per window (multiple of 4)
  const uint32_t NCH = WF / 4;
  const uint32_t dmax = p.n_dms - 1;
  const uint32_t *list = p.rows ? p.rows + (uint64_t)inner * p.rows_pitch : nullptr;
  const uint32_t nrows = list ? uniform(p.rows_count[inner]) : p.n_rows;
  const uint32_t slots = list ? p.rows_pitch : p.n_rows;  // row entries in LDS
  const uint32_t rows_pad = (2 * slots + 3) & ~3u;

  // LDS: rows (row | input beam << 16, float delay) | windows [2][G][WF] | DM offsets [2][G][TD]
  uint32_t *lrow = reinterpret_cast<uint32_t *>(ldsf);
  float *ltab = ldsf + slots;
  float *windows = ldsf + rows_pad;
  uint32_t *meta = reinterpret_cast<uint32_t *>(windows + 2 * G * WF);

  for (uint32_t k = threadIdx.x; k < nrows; k += blockDim.x) {
    const uint32_t row = list ? list[k] : k;
    const uint32_t in_beam = p.map ? p.map[(uint64_t)(p.first_sbeam + outer) * p.map_pitch + row] : outer;
    lrow[k] = row | (in_beam << 16);
    ltab[k] = p.table[row];
  }
  __syncthreads();

  u32x4 sv[SPW][CPL];
  uint32_t so[SPW];
  auto stage_load = [&](uint32_t s0, uint32_t nn) {
#pragma unroll
    for (int s = 0; s < SPW; ++s) {
      const uint32_t g = wave + s * nwaves;
      if (g < nn) {
        const uint32_t packed = uniform(lrow[s0 + g]);
        const float tab = uniformf(ltab[s0 + g]);
        const uint32_t row = packed & 0xFFFFu, in_beam = packed >> 16;
        const uint32_t sh = shift_of(p.first_dm, p.dm_step, min(d0 + lane, dmax), tab);
        const uint32_t shmin = uniform(sh);
        so[s] = sh - shmin;
        const uint64_t off = uniform64((uint64_t)in_beam * p.in_beam_pitch + (uint64_t)inner * p.in_coarse_pitch +
                                       (uint64_t)row * p.in_row_pitch + t0 + shmin);
        const uint32_t recs = records_from_floats(p.in_floats, off);
        const __amdgpu_buffer_rsrc_t rsrc = make_rsrc(p.in + off, recs);
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          const uint32_t m = lane + 64 * c;
          if (m < NCH) {
            if (16 * m + 16 <= recs) {
              sv[s][c] = __builtin_amdgcn_raw_buffer_load_b128(rsrc, 16 * m, 0, 0);
            } else {
#pragma unroll
              for (int e = 0; e < 4; ++e) sv[s][c][e] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, 16 * m + 4 * e, 0, 0);
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
        float *w = windows + (buf * G + g) * WF;
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          const uint32_t m = lane + 64 * c;
          if (m < NCH) *reinterpret_cast<u32x4 *>(w + 4 * m) = sv[s][c];
        }
        if (lane < TD) meta[(buf * G + g) * TD + lane] = so[s];
      }
    }
  };

  float acc[DPW][RUN];
#pragma unroll
  for (int j = 0; j < DPW; ++j)
#pragma unroll
    for (int k = 0; k < RUN; ++k) acc[j][k] = 0.0f;

  auto compute = [&](uint32_t buf, uint32_t n) {
    for (uint32_t g = 0; g < n; ++g) {
      const uint32_t *mrow = meta + (buf * G + g) * TD + dw;  // broadcast reads
      const float *w = windows + (buf * G + g) * WF + lane;
#pragma unroll
      for (int j = 0; j < DPW; ++j) {
        const float *src = w + mrow[j];
#pragma unroll
        for (int k = 0; k < RUN; ++k) acc[j][k] += src[64 * k];
      }
    }
  };

  uint32_t cur = 0;
  if (nrows > 0) {
    const uint32_t n0 = min(G, nrows);
    stage_load(0, n0);
    stage_store(0, n0);
  }
  __syncthreads();
  for (uint32_t s0 = 0; s0 < nrows; s0 += G) {
    const uint32_t n = min(G, nrows - s0);
    const uint32_t sn = s0 + G;
    const uint32_t nn = sn < nrows ? min(G, nrows - sn) : 0;
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
    float *orow = p.out + (uint64_t)outer * p.out_beam_pitch + (uint64_t)inner * p.out_inner_pitch +
                  (uint64_t)dm * p.out_dm_pitch;
#pragma unroll
    for (int k = 0; k < RUN; ++k) {
      const uint32_t t = t0 + lane + 64 * k;
      if (t < p.n_samples) {
        if (p.nontemporal) __builtin_nontemporal_store(acc[j][k], orow + t);
        else orow[t] = acc[j][k];
      }
    }
  }
}

template <int RUN, int DPW>
__global__ void __launch_bounds__(512) rowsum_f32_staged(const RowSumF32Params p) {
  extern __shared__ __attribute__((aligned(16))) float ldsf[];
  f32_tile<RUN, DPW>(p, xcd_logical(blockIdx.x, p.n_tiles), ldsf);
}

// Persistent form over (beam, tile) descriptors (see rowsum_u8_batch).
template <int RUN, int DPW>
__global__ void __launch_bounds__(512) rowsum_f32_batch(const RowSumF32Params *__restrict__ beams,
                                                          const TileQueues q) {
  extern __shared__ __attribute__((aligned(16))) float ldsf[];
  __shared__ uint32_t claimed;
  QueueCursor cur;
  for (;;) {
    const uint32_t t = cur.next(q, &claimed);
    if (t == kNoTile) break;
    const TileDesc d = q.tiles[t];
    f32_tile<RUN, DPW>(beams[d.job], d.tile, ldsf);
  }
}

// One workgroup per listed (job, tile) descriptor, in XCD-aware order.
template <int RUN, int DPW>
__global__ void __launch_bounds__(512) rowsum_f32_list(const RowSumF32Params *__restrict__ jobs, const TileDesc *__restrict__ tiles,
                                                      uint32_t n) {
  extern __shared__ __attribute__((aligned(16))) float ldsf[];
  const TileDesc d = tiles[xcd_logical(blockIdx.x, n)];
  f32_tile<RUN, DPW>(jobs[d.job], d.tile, ldsf);
}

// Calls f(RUN, DPW) with the compile-time constants of geometry g.
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

hipError_t launch_rowsum_f32_staged(const RowSumF32Params &p, const Geometry &g, hipStream_t stream) {
  return dispatch(g, [&](auto run, auto dpw) {
    hipLaunchKernelGGL((rowsum_f32_staged<run.value, dpw.value>), dim3(p.n_tiles), dim3(64 * g.threads_d1),
                       g.lds_bytes, stream, p);
    return hipGetLastError();
  });
}

hipError_t launch_rowsum_f32_batch(const RowSumF32Params *d_jobs, const TileQueues &q, const Geometry &g,
                                   uint32_t workgroups, hipStream_t stream) {
  return dispatch(g, [&](auto run, auto dpw) {
    hipLaunchKernelGGL((rowsum_f32_batch<run.value, dpw.value>), dim3(workgroups), dim3(64 * g.threads_d1),
                       g.lds_bytes, stream, d_jobs, q);
    return hipGetLastError();
  });
}

hipError_t launch_rowsum_f32_list(const RowSumF32Params *d_jobs, const TileDesc *d_tiles, uint32_t n, const Geometry &g,
                                  hipStream_t stream) {
  return dispatch(g, [&](auto run, auto dpw) {
    hipLaunchKernelGGL((rowsum_f32_list<run.value, dpw.value>), dim3(n), dim3(64 * g.threads_d1), g.lds_bytes, stream,
                       d_jobs, d_tiles, n);
    return hipGetLastError();
  });
}

hipError_t rowsum_f32_batch_occupancy(const Geometry &g, int *per_cu) {
  return dispatch(g, [&](auto run, auto dpw) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, rowsum_f32_batch<run.value, dpw.value>,
                                                        (int)(64 * g.threads_d1), (size_t)g.lds_bytes);
  });
}

// Zeroes the claim counters of the work queues (a kernel rather than a
// memset so that the reset can sit in a captured graph).
__global__ void clear_words(uint32_t *words, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) words[i] = 0;
}

hipError_t launch_clear_words(uint32_t *d_words, uint32_t n, hipStream_t stream) {
  hipLaunchKernelGGL(clear_words, dim3(1), dim3(256), 0, stream, d_words, n);
  return hipGetLastError();
}

} // namespace ddamd
