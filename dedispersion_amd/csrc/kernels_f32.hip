// kernels_f32.hip -- shifted row sums over float rows on CDNA4 (gfx950):
// step two of the subband path, and the single step and step one over float
// input.
//
// Reference: CPU loops include/Dedispersion.hpp:155-176 (and :70-153 with
// I = float), OpenCL generator :884-1122.  A workgroup owns a (DM tile x
// sample tile) block of one slice -- (synthesized beam, coarse DM) for step
// two -- and walks the slice's rows in order (subbands 0..Nsb-1, or the live
// channels), adding into float accumulators.  The order matches the
// reference's sequential float sum, so the result is bit-identical for any
// float input, not only for the integer-valued output of step one.
//
// Each wavefront owns `DPW` consecutive fine DMs (wave-uniform shifts);
// lane l owns samples t0 + l + 64k, k < RUN, so LDS reads and global stores
// are consecutive across the wavefront.  Per subband the workgroup stages the
// float window [t0 + shift(dmFirst), t0 + TS + shift(dmLast)) in LDS; the
// `local = 0` variant reads the same elements through the vector caches.
#include "internal.hpp"

namespace ddamd {
namespace {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ uint32_t records_from(uint64_t total_floats, uint64_t offset_floats) {
  if (offset_floats >= total_floats) return 0;
  uint64_t n = (total_floats - offset_floats) * 4;
  return n > 0xFFFFFFF0ull ? 0xFFFFFFF0u : (uint32_t)n;
}

template <int RUN, int DPW, bool LDS>
__global__ void __launch_bounds__(512) rowsum_f32(const RowSumF32Params p) {
  constexpr int TS = 64 * RUN;
  extern __shared__ __attribute__((aligned(16))) float ldsf[];

  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wave = uniform(threadIdx.x >> 6);
  const uint32_t nwaves = blockDim.x >> 6;
  const uint32_t t0 = blockIdx.x * TS;
  const uint32_t d0 = blockIdx.y * nwaves * DPW;
  const uint32_t dw = d0 + wave * DPW;
  const uint32_t inner = blockIdx.z % p.n_coarse;
  const uint32_t outer = blockIdx.z / p.n_coarse;
  const uint32_t W = p.win_floats;
  const uint32_t *list = p.rows ? p.rows + (uint64_t)inner * p.rows_pitch : nullptr;
  const uint32_t nrows = list ? uniform(p.rows_count[inner]) : p.n_rows;

  float acc[DPW][RUN];
#pragma unroll
  for (int j = 0; j < DPW; ++j)
#pragma unroll
    for (int k = 0; k < RUN; ++k) acc[j][k] = 0.0f;

  // the row of list entry k, and the input offset of a row's window start
  auto row_of = [&](uint32_t k) -> uint32_t { return list ? uniform(list[k]) : k; };
  auto row_offset = [&](uint32_t sb) -> uint64_t {
    const uint32_t in_beam = p.map ? p.map[(uint64_t)(p.first_sbeam + outer) * p.map_pitch + sb] : outer;
    return (uint64_t)in_beam * p.in_beam_pitch + (uint64_t)inner * p.in_coarse_pitch +
           (uint64_t)sb * p.in_row_pitch + t0;
  };

  const uint32_t G = p.unroll;
  for (uint32_t s0 = 0; s0 < nrows; s0 += G) {
    const uint32_t n = min(G, nrows - s0);
    if constexpr (LDS) {
      for (uint32_t g = wave; g < n; g += nwaves) {
        const uint32_t sb = row_of(s0 + g);
        const uint32_t shmin = p.ishift[(uint64_t)sb * p.dm_pitch + d0];
        const uint64_t off = row_offset(sb) + shmin;
        const uint32_t recs = records_from(p.in_floats, off);
        const __amdgpu_buffer_rsrc_t rsrc = make_rsrc(p.in + off, recs);
        float *win = ldsf + g * W;
        for (uint32_t m = lane; m < W / 4; m += 64) {
          u32x4 v;
          if (16 * m + 16 <= recs) {
            v = __builtin_amdgcn_raw_buffer_load_b128(rsrc, 16 * m, 0, 0);
          } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) v[e] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, 16 * m + 4 * e, 0, 0);
          }
          *reinterpret_cast<u32x4 *>(win + 4 * m) = v;
        }
      }
      __syncthreads();
    }
    for (uint32_t g = 0; g < n; ++g) {
      const uint32_t sb = row_of(s0 + g);
      const uint32_t *shrow = p.ishift + (uint64_t)sb * p.dm_pitch;
      if constexpr (LDS) {
        const uint32_t shmin = shrow[d0];
        const float *win = ldsf + g * W;
#pragma unroll
        for (int j = 0; j < DPW; ++j) {
          const float *src = win + (shrow[dw + j] - shmin) + lane;
#pragma unroll
          for (int k = 0; k < RUN; ++k) acc[j][k] += src[64 * k];
        }
      } else {
        const uint64_t off = row_offset(sb);
#pragma unroll
        for (int j = 0; j < DPW; ++j) {
          const uint64_t start = off + shrow[dw + j];
          const __amdgpu_buffer_rsrc_t rsrc = make_rsrc(p.in + start, records_from(p.in_floats, start));
#pragma unroll
          for (int k = 0; k < RUN; ++k)
            acc[j][k] += __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc, 4 * (lane + 64 * k), 0, 0));
        }
      }
    }
    if constexpr (LDS) __syncthreads();
  }

#pragma unroll
  for (int j = 0; j < DPW; ++j) {
    const uint32_t dm = dw + j;
    if (dm >= p.n_dms) break;
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
hipError_t launch_one(const RowSumF32Params &p, const Geometry &g, dim3 grid, hipStream_t s) {
  const uint32_t threads = 64 * g.threads_d1;
  if (g.local)
    hipLaunchKernelGGL((rowsum_f32<RUN, DPW, true>), grid, dim3(threads), g.lds_bytes, s, p);
  else
    hipLaunchKernelGGL((rowsum_f32<RUN, DPW, false>), grid, dim3(threads), 0, s, p);
  return hipGetLastError();
}

template <int RUN>
hipError_t dispatch_dpw(const RowSumF32Params &p, const Geometry &g, dim3 grid, hipStream_t s) {
  switch (g.items_d1) {
  case 1: return launch_one<RUN, 1>(p, g, grid, s);
  case 2: return launch_one<RUN, 2>(p, g, grid, s);
  case 4: return launch_one<RUN, 4>(p, g, grid, s);
  case 8: return launch_one<RUN, 8>(p, g, grid, s);
  default: return hipErrorInvalidValue;
  }
}

} // namespace

bool rowsum_f32_supported(uint32_t items_d0, uint32_t items_d1) {
  bool run_ok = items_d0 == 4 || items_d0 == 8 || items_d0 == 16;
  bool dpw_ok = items_d1 == 1 || items_d1 == 2 || items_d1 == 4 || items_d1 == 8;
  return run_ok && dpw_ok;
}

hipError_t launch_rowsum_f32(const RowSumF32Params &p, const Geometry &g, dim3 grid, hipStream_t stream) {
  switch (g.items_d0) {
  case 4: return dispatch_dpw<4>(p, g, grid, stream);
  case 8: return dispatch_dpw<8>(p, g, grid, stream);
  case 16: return dispatch_dpw<16>(p, g, grid, stream);
  default: return hipErrorInvalidValue;
  }
}

} // namespace ddamd
