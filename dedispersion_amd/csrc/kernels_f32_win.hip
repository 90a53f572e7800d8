// kernels_f32_win.hip -- register-window step two over float rows.
//
// Same operation as kernels_f32.hip (include/Dedispersion.hpp:155-176): the
// subbands are added in order 0..Nsb-1 into float accumulators, so results are
// bit-identical to the reference's sequential float sum for any input.
//
// Organisation (see kernels_u8_win.hip): wavefront w owns DPW consecutive fine
// DMs, lane l the RUN consecutive samples t0 + RUN*l + [0, RUN); subband rows
// are staged in LDS double buffered; per row a lane reads one float window
// covering the runs of all its wavefront's DMs with 16-byte LDS reads and each
// DM takes its run from registers at a wave-uniform offset chosen by a scalar
// branch tree.  With RUN = 16 the lane stride is 64 bytes, so the LDS window
// is stored skewed (4 floats of padding after every 16) to keep the 16-byte
// reads of eight consecutive lanes on distinct banks.
#include "window_common.hpp"

namespace ddamd {
namespace {

using namespace win;

template <int RUN, int DPW, int SPAN>
__global__ void __launch_bounds__(512) rowsum_f32_win(const RowSumF32Params p) {
  constexpr bool SKEW = RUN == 16;
  constexpr int TS = 64 * RUN;
  constexpr int NRF = ((3 + SPAN + RUN + 3) / 4) * 4;  // floats per lane window
  constexpr int NR = NRF / 4;                          // 16-byte reads per lane window
  constexpr int EMAX = 3 + SPAN;                       // largest offset of a DM run in a lane window
  constexpr int SPW = kWinSlotsPerWave;
  constexpr int CPL = RUN / 4 + 1;
  extern __shared__ float ldsf[];

  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wave = uniform(threadIdx.x >> 6);
  const uint32_t nwaves = blockDim.x >> 6;
  const uint32_t TD = nwaves * DPW;
  const Tile tile = tile_of(blockIdx.x, p.n_tiles_x, p.n_tiles_y, p.n_tiles);
  const uint32_t t0 = tile.x * TS;
  const uint32_t d0 = tile.y * TD;
  const uint32_t dw = wave * DPW;
  const uint32_t coarse = tile.z % p.n_coarse;
  const uint32_t sbeam = tile.z / p.n_coarse;
  const uint32_t G = p.unroll;
  const uint32_t WF = p.win_floats;                   // logical floats per window (multiple of 16)
  const uint32_t WP = SKEW ? WF + WF / 4 : WF;        // floats per window in LDS
  const uint32_t NCH = WF / 4;
  const uint32_t dmax = p.n_dms - 1;
  const uint32_t rows_pad = (2 * p.n_rows + 3) & ~3u;

  uint32_t *lbeam = reinterpret_cast<uint32_t *>(ldsf);
  float *ltab = ldsf + p.n_rows;
  float *windows = ldsf + rows_pad;
  uint32_t *meta = reinterpret_cast<uint32_t *>(windows + 2 * G * WP);

  for (uint32_t sb = threadIdx.x; sb < p.n_rows; sb += blockDim.x) {
    lbeam[sb] = p.map[(uint64_t)(p.first_sbeam + sbeam) * p.map_pitch + sb];
    ltab[sb] = p.table[sb];
  }
  __syncthreads();

  u32x4 sv[SPW][CPL];
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
        so[s] = sh - shmin;
        const uint64_t off = uniform64((uint64_t)in_beam * p.in_beam_pitch + (uint64_t)coarse * p.in_coarse_pitch +
                                       (uint64_t)sb * p.in_row_pitch + t0 + shmin);
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
        float *w = windows + (buf * G + g) * WP;
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          const uint32_t m = lane + 64 * c;
          if (m < NCH) {
            const uint32_t at = SKEW ? 4 * m + 4 * (m / 4) : 4 * m;
            *reinterpret_cast<u32x4 *>(w + at) = sv[s][c];
          }
        }
        if (lane < TD) meta[(buf * G + g) * TD + lane] = so[s];
      }
    }
  };

  float acc[DPW][RUN];
#pragma unroll
  for (int j = 0; j < DPW; ++j)
#pragma unroll
    for (int i = 0; i < RUN; ++i) acc[j][i] = 0.0f;

  auto compute = [&](uint32_t buf, uint32_t n) {
    for (uint32_t g = 0; g < n; ++g) {
      const uint32_t *mrow = meta + (buf * G + g) * TD + dw;
      uint32_t o[DPW];
#pragma unroll
      for (int j = 0; j < DPW; ++j) o[j] = uniform(mrow[j]);
      const uint32_t base = o[0] & ~3u;
      const float *w = windows + (buf * G + g) * WP;
      float R[NRF];
#pragma unroll
      for (int c = 0; c < NR; ++c) {
        // window float RUN*lane + base + 4c
        uint32_t at;
        if constexpr (SKEW) {
          const uint32_t q = base + 4 * c;  // multiple of 4, wave-uniform
          at = 20 * lane + q + 4 * (q >> 4);
        } else {
          at = RUN * lane + base + 4 * c;
        }
        const f32x4 v = *reinterpret_cast<const f32x4 *>(w + at);
        R[4 * c + 0] = v.x;
        R[4 * c + 1] = v.y;
        R[4 * c + 2] = v.z;
        R[4 * c + 3] = v.w;
      }
      uint32_t e[DPW];
#pragma unroll
      for (int j = 0; j < DPW; ++j) e[j] = o[j] - base;
      // Every DM takes its run at its own wave-uniform offset; walking the
      // offsets at compile time keeps each update a plain conditional add.
#pragma unroll
      for (int x = 0; x <= EMAX; ++x) {
#pragma unroll
        for (int j = 0; j < DPW; ++j) {
          if (e[j] == (uint32_t)x) {
            asm volatile("");
#pragma unroll
            for (int i = 0; i < RUN; ++i) acc[j][i] += R[x + i];
          }
        }
      }
    }
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
    compute(cur, n);
    if (nn) stage_store(cur ^ 1, nn);
    __syncthreads();
    cur ^= 1;
  }

#pragma unroll
  for (int j = 0; j < DPW; ++j) {
    const uint32_t dm = d0 + dw + j;
    if (dm >= p.n_dms) break;
    float *orow = p.out + (uint64_t)sbeam * p.out_beam_pitch + ((uint64_t)coarse * p.n_dms + dm) * p.out_dm_pitch;
#pragma unroll
    for (int k = 0; k < RUN / 4; ++k) {
      const uint32_t t = t0 + RUN * lane + 4 * k;
      if (p.vec_store && t + 3 < p.n_samples) {
        f32x4 f = {acc[j][4 * k], acc[j][4 * k + 1], acc[j][4 * k + 2], acc[j][4 * k + 3]};
        if (p.nontemporal) __builtin_nontemporal_store(f, reinterpret_cast<f32x4 *>(orow + t));
        else *reinterpret_cast<f32x4 *>(orow + t) = f;
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (t + e < p.n_samples) orow[t + e] = acc[j][4 * k + e];
      }
    }
  }
}

template <int RUN, int DPW, int SPAN>
hipError_t launch_one(const RowSumF32Params &p, const Geometry &g, hipStream_t s) {
  hipLaunchKernelGGL((rowsum_f32_win<RUN, DPW, SPAN>), dim3(p.n_tiles), dim3(64 * g.threads_d1), g.lds_bytes, s, p);
  return hipGetLastError();
}

template <int RUN, int DPW>
hipError_t by_span(const RowSumF32Params &p, const Geometry &g, hipStream_t s) {
  switch (g.span) {
  case 8: return launch_one<RUN, DPW, 8>(p, g, s);
  case 16: return launch_one<RUN, DPW, 16>(p, g, s);
  case 32: return launch_one<RUN, DPW, 32>(p, g, s);
  default: return hipErrorInvalidValue;
  }
}

template <int RUN>
hipError_t by_dpw(const RowSumF32Params &p, const Geometry &g, hipStream_t s) {
  switch (g.items_d1) {
  case 1: return by_span<RUN, 1>(p, g, s);
  case 2: return by_span<RUN, 2>(p, g, s);
  case 4: return by_span<RUN, 4>(p, g, s);
  case 8: return by_span<RUN, 8>(p, g, s);
  default: return hipErrorInvalidValue;
  }
}

} // namespace

uint32_t f32_win_span(uint32_t need) {
  for (uint32_t s : {8u, 16u, 32u})
    if (need <= s) return s;
  return 0;
}

hipError_t launch_rowsum_f32_win(const RowSumF32Params &p, const Geometry &g, hipStream_t stream) {
  switch (g.items_d0) {
  case 4: return by_dpw<4>(p, g, stream);
  case 16: return by_dpw<16>(p, g, stream);
  default: return hipErrorInvalidValue;
  }
}

} // namespace ddamd
