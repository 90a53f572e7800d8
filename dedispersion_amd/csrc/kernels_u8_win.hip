// kernels_u8_win.hip -- register-window shifted row sums over uint8 rows.
//
// Same operation as kernels_u8.hip (single-step dedispersion,
// include/Dedispersion.hpp:70-109, and step one, :111-153), organised so that
// VALU work is the only limit:
//
//  * A workgroup owns a (DM tile x sample tile) block; wavefront w owns DPW
//    consecutive DMs, lane l the RUN consecutive samples t0 + RUN*l + [0, RUN).
//  * Rows are staged in LDS in groups of `unroll`, double buffered: the
//    global loads of group k+1 are issued before group k is summed and
//    written to the other buffer after it, so one barrier per group remains
//    and the load latency hides behind the additions.
//  * Per row, a lane reads ONE window of bytes covering the runs of all DPW
//    DMs of its wavefront with RUN-byte LDS reads at a RUN-byte lane stride
//    (conflict-free), i.e. about 0.75 LDS bytes per addition.  Each DM then
//    takes its run out of the registers: the dword offset is wave-uniform
//    and selected by a scalar branch tree, the byte offset by v_perm_b32
//    selectors held in SGPRs.
//  * Integer shifts are computed on the fly from the float delay table with
//    the reference's unfused arithmetic, once per staged row, and published
//    to the wavefronts through LDS.
//  * Workgroups are ordered XCD-aware (window_common.hpp).
//
// Sums use 16-bit-pair accumulators exactly as kernels_u8.hip; all partial
// sums are integers below 2^24, so the float output is bit-exact.
#include "window_common.hpp"

namespace ddamd {
namespace {

using namespace win;

__device__ __forceinline__ uint32_t sel_even_of(uint32_t r) { return r | 0x0c00u | ((r + 2) << 16) | 0x0c000000u; }
__device__ __forceinline__ uint32_t sel_odd_of(uint32_t r) { return (r + 1) | 0x0c00u | ((r + 3) << 16) | 0x0c000000u; }

template <int RUN, int DPW, int SPAN, bool FLUSH>
__global__ void __launch_bounds__(512) rowsum_u8_win(const RowSumU8Params p) {
  constexpr int NG = RUN / 4;                   // four-sample groups per lane run
  constexpr int TS = 64 * RUN;                  // samples per tile
  constexpr int EMAX = RUN - 1 + SPAN;          // largest byte offset of a DM run in a lane window
  constexpr int NWD = EMAX / 4 + NG + 1;        // dwords a lane window must cover
  constexpr int NR = (NWD * 4 + RUN - 1) / RUN; // RUN-byte LDS reads per lane window
  constexpr int NRD = NR * RUN / 4;             // dwords read per lane window
  constexpr int SPW = kWinSlotsPerWave;
  constexpr int CPL = kU8WinChunks;
  extern __shared__ uint32_t lds[];

  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wave = uniform(threadIdx.x >> 6);
  const uint32_t nwaves = blockDim.x >> 6;
  const uint32_t TD = nwaves * DPW;
  const Tile tile = tile_of(blockIdx.x, p.n_tiles_x, p.n_tiles_y, p.n_tiles);
  const uint32_t t0 = tile.x * TS;
  const uint32_t d0 = tile.y * TD;
  const uint32_t dw = wave * DPW;  // first DM of this wavefront, relative to d0
  const uint32_t group = tile.z % p.n_groups;
  const uint32_t beam = tile.z / p.n_groups;
  const uint32_t nlive = p.live_count[group];
  const uint32_t *live = p.live + (uint64_t)group * p.live_pitch;
  const uint32_t G = p.unroll;
  const uint32_t WD = p.win_bytes / 4;  // dwords per window
  const uint32_t NCH = WD / 4;          // 16-byte chunks per window
  const uint32_t dmax = p.n_dms - 1;

  // LDS: rows (packed row | beam << 16, float delay) | windows [2][G][WD] | meta [2][G][TD]
  uint32_t *lrow = lds;
  float *ltab = reinterpret_cast<float *>(lds + p.live_pitch);
  uint32_t *windows = lds + ((2 * p.live_pitch + 3) & ~3u);
  uint32_t *meta = windows + 2 * G * WD;

  for (uint32_t k = threadIdx.x; k < nlive; k += blockDim.x) {
    const uint32_t row = live[k];
    const uint32_t in_beam = p.map ? p.map[(uint64_t)(p.first_sbeam + beam) * p.map_pitch + row] : beam;
    lrow[k] = row | (in_beam << 16);
    ltab[k] = p.table[row];
  }
  __syncthreads();

  // ---- staging: wavefront w stages rows w and w + nwaves of a group ----
  u32x4 sv[SPW][CPL];
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
        const uint64_t off = uniform64((uint64_t)in_beam * p.in_beam_pitch + (uint64_t)row * p.in_row_pitch + t0 + shmin);
        const uint64_t a0 = off & ~uint64_t(3);
        so[s] = sh - shmin + (uint32_t)(off & 3);
        const uint32_t recs = records_from(p.in_bytes, a0);
        const __amdgpu_buffer_rsrc_t rsrc = make_rsrc(p.in + a0, recs);
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
        uint32_t *w = windows + (buf * G + g) * WD;
#pragma unroll
        for (int c = 0; c < CPL; ++c) {
          const uint32_t m = lane + 64 * c;
          if (m < NCH) *reinterpret_cast<u32x4 *>(w + 4 * m) = sv[s][c];
        }
        if (lane < TD) meta[(buf * G + g) * TD + lane] = so[s];
      }
    }
  };

  // ---- accumulation ----
  uint32_t even[DPW][NG], odd[DPW][NG];
  uint32_t wide[DPW][RUN];
#pragma unroll
  for (int j = 0; j < DPW; ++j)
#pragma unroll
    for (int k = 0; k < NG; ++k) even[j][k] = odd[j][k] = 0;
  if constexpr (FLUSH) {
#pragma unroll
    for (int j = 0; j < DPW; ++j)
#pragma unroll
      for (int k = 0; k < RUN; ++k) wide[j][k] = 0;
  }
  auto flush = [&]() {
    if constexpr (FLUSH) {
#pragma unroll
      for (int j = 0; j < DPW; ++j)
#pragma unroll
        for (int k = 0; k < NG; ++k) {
          wide[j][4 * k + 0] += even[j][k] & 0xFFFFu;
          wide[j][4 * k + 2] += even[j][k] >> 16;
          wide[j][4 * k + 1] += odd[j][k] & 0xFFFFu;
          wide[j][4 * k + 3] += odd[j][k] >> 16;
          even[j][k] = odd[j][k] = 0;
        }
    }
  };

  auto compute = [&](uint32_t buf, uint32_t n) {
    for (uint32_t g = 0; g < n; ++g) {
      const uint32_t *mrow = meta + (buf * G + g) * TD + dw;
      uint32_t o[DPW];
#pragma unroll
      for (int j = 0; j < DPW; ++j) o[j] = uniform(mrow[j]);
      const uint32_t base = o[0] & ~uint32_t(RUN - 1);
      const uint32_t *src = windows + (buf * G + g) * WD + (RUN * lane + base) / 4;
      uint32_t R[NRD];
#pragma unroll
      for (int c = 0; c < NR; ++c) {
        if constexpr (RUN == 16) {
          const u32x4 v = *reinterpret_cast<const u32x4 *>(src + 4 * c);
          R[4 * c + 0] = v.x;
          R[4 * c + 1] = v.y;
          R[4 * c + 2] = v.z;
          R[4 * c + 3] = v.w;
        } else if constexpr (RUN == 8) {
          const u32x2 v = *reinterpret_cast<const u32x2 *>(src + 2 * c);
          R[2 * c + 0] = v.x;
          R[2 * c + 1] = v.y;
        } else {
          R[c] = src[c];
        }
      }
      uint32_t q[DPW], se[DPW], sd[DPW];
#pragma unroll
      for (int j = 0; j < DPW; ++j) {
        const uint32_t e = o[j] - base;
        q[j] = e >> 2;
        se[j] = sel_even_of(e & 3);
        sd[j] = sel_odd_of(e & 3);
      }
      // Every DM takes its run at its own wave-uniform dword offset; walking
      // the offsets at compile time keeps each update a plain conditional add.
#pragma unroll
      for (int x = 0; x <= EMAX / 4; ++x) {
#pragma unroll
        for (int j = 0; j < DPW; ++j) {
          if (q[j] == (uint32_t)x) {
            asm volatile("");
#pragma unroll
            for (int k = 0; k < NG; ++k) {
              even[j][k] += __builtin_amdgcn_perm(R[x + k + 1], R[x + k], se[j]);
              odd[j][k] += __builtin_amdgcn_perm(R[x + k + 1], R[x + k], sd[j]);
            }
          }
        }
      }
    }
  };

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
    compute(cur, n);
    if (nn) stage_store(cur ^ 1, nn);
    __syncthreads();
    cur ^= 1;
  }
  flush();

#pragma unroll
  for (int j = 0; j < DPW; ++j) {
    const uint32_t dm = d0 + dw + j;
    if (dm >= p.n_dms) break;
    float *orow = p.out + (uint64_t)beam * p.out_beam_pitch + (uint64_t)group * p.out_group_pitch +
                  (uint64_t)dm * p.out_dm_pitch;
#pragma unroll
    for (int k = 0; k < NG; ++k) {
      float v[4];
      if constexpr (FLUSH) {
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = (float)wide[j][4 * k + e];
      } else {
        v[0] = (float)(even[j][k] & 0xFFFFu);
        v[1] = (float)(odd[j][k] & 0xFFFFu);
        v[2] = (float)(even[j][k] >> 16);
        v[3] = (float)(odd[j][k] >> 16);
      }
      const uint32_t t = t0 + RUN * lane + 4 * k;
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

template <int RUN, int DPW, int SPAN, bool FLUSH>
hipError_t launch_one(const RowSumU8Params &p, const Geometry &g, hipStream_t s) {
  hipLaunchKernelGGL((rowsum_u8_win<RUN, DPW, SPAN, FLUSH>), dim3(p.n_tiles), dim3(64 * g.threads_d1), g.lds_bytes,
                     s, p);
  return hipGetLastError();
}

template <int RUN, int DPW, bool FLUSH>
hipError_t by_span(const RowSumU8Params &p, const Geometry &g, hipStream_t s) {
  switch (g.span) {
  case 8: return launch_one<RUN, DPW, 8, FLUSH>(p, g, s);
  case 16: return launch_one<RUN, DPW, 16, FLUSH>(p, g, s);
  case 32: return launch_one<RUN, DPW, 32, FLUSH>(p, g, s);
  default: return hipErrorInvalidValue;
  }
}

template <int RUN, bool FLUSH>
hipError_t by_dpw(const RowSumU8Params &p, const Geometry &g, hipStream_t s) {
  switch (g.items_d1) {
  case 1: return by_span<RUN, 1, FLUSH>(p, g, s);
  case 2: return by_span<RUN, 2, FLUSH>(p, g, s);
  case 4: return by_span<RUN, 4, FLUSH>(p, g, s);
  case 8: return by_span<RUN, 8, FLUSH>(p, g, s);
  default: return hipErrorInvalidValue;
  }
}

template <bool FLUSH>
hipError_t by_run(const RowSumU8Params &p, const Geometry &g, hipStream_t s) {
  switch (g.items_d0) {
  case 4: return by_dpw<4, FLUSH>(p, g, s);
  case 8: return by_dpw<8, FLUSH>(p, g, s);
  case 16: return by_dpw<16, FLUSH>(p, g, s);
  default: return hipErrorInvalidValue;
  }
}

} // namespace

uint32_t u8_win_span(uint32_t need) {
  for (uint32_t s : {8u, 16u, 32u})
    if (need <= s) return s;
  return 0;
}

hipError_t launch_rowsum_u8_win(const RowSumU8Params &p, const Geometry &g, bool flush, hipStream_t stream) {
  return flush ? by_run<true>(p, g, stream) : by_run<false>(p, g, stream);
}

} // namespace ddamd
