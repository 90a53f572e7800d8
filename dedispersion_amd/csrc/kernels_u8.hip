// kernels_u8.hip -- shifted row sums over uint8 rows on CDNA4 (gfx950).
//
// Serves single-step dedispersion (reference CPU loop include/Dedispersion.hpp:70-109,
// OpenCL generator :203-540) and step one of the subband path (CPU loop
// :111-153, generator :543-882).
//
// Work decomposition.  A workgroup owns a (DM tile x sample tile) block of one
// output beam (and one subband for step one) and walks the live rows
// (channels).  Each wavefront owns `DPW` consecutive DMs, so every integer
// shift it applies is wave-uniform and lives in SGPRs.  Lane l owns the
// RUN/4 groups of four consecutive samples t0 + 4l + 256k + [0, 4).  Per row
// the workgroup stages the byte window [t0 + shift(dmFirst), t0 + TS +
// shift(dmLast)) in LDS; for one DM a lane then reads the two dwords covering
// each of its groups at dword (l + 64k + q), so consecutive lanes read
// consecutive dwords (no bank conflicts) and every read of a DM is one base
// VGPR plus an immediate offset.
//
// Arithmetic.  Input values are bytes; sums are integers.  A lane extracts the
// four bytes of a group at byte offset r = (window offset) % 4 from its two
// dwords with one v_perm_b32 per pair of samples, zero-extending them into
// 16-bit fields (samples 4k and 4k+2 in one register, 4k+1 and 4k+3 in
// another), and adds that register to a 16-bit-pair accumulator with one
// v_add_u32: one VALU instruction per sample addition.  A 16-bit field holds
// 257 bytes of 255, so accumulators of rows longer than 257 are flushed into
// 32-bit sums every 256 rows.  All sums are exact integers below 2^24, so
// the float output equals the reference's sequential float sum bit for bit.
#include "internal.hpp"

namespace ddamd {
namespace {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Bytes addressable from `offset` in a buffer of `total` bytes, rounded up to a
// whole dword (the dword holding the last byte is safe to read), clamped to the
// 32-bit range of a buffer resource.
__device__ __forceinline__ uint32_t records_from(uint64_t total, uint64_t offset) {
  uint64_t end = (total + 3) & ~uint64_t(3);
  if (offset >= end) return 0;
  uint64_t n = end - offset;
  return n > 0xFFFFFFF0ull ? 0xFFFFFFF0u : (uint32_t)n;
}

// Byte offset (from p.in) of sample t0 + shift(first DM of the tile) in a row,
// and that shift; both wave-uniform.
struct RowInfo {
  uint64_t off;
  uint32_t shmin;
};

__device__ __forceinline__ RowInfo row_info(const RowSumU8Params &p, uint32_t row, uint32_t beam, uint32_t t0,
                                            uint32_t d0) {
  RowInfo ri;
  uint32_t in_beam = beam;
  if (p.map) in_beam = p.map[(uint64_t)(p.first_sbeam + beam) * p.map_pitch + row];
  ri.shmin = p.ishift[(uint64_t)row * p.dm_pitch + d0];
  ri.off = (uint64_t)in_beam * p.in_beam_pitch + (uint64_t)row * p.in_row_pitch + t0 + ri.shmin;
  return ri;
}

// v_perm_b32 selectors picking bytes r, r+2 (even) and r+1, r+3 (odd) of the
// 8-byte pair {hi, lo} into the low byte of each 16-bit half (0x0c = zero).
__device__ __forceinline__ uint32_t sel_even_of(uint32_t r) { return r | 0x0c00u | ((r + 2) << 16) | 0x0c000000u; }
__device__ __forceinline__ uint32_t sel_odd_of(uint32_t r) { return (r + 1) | 0x0c00u | ((r + 3) << 16) | 0x0c000000u; }

template <int NG>
__device__ __forceinline__ void accumulate(const uint32_t (&lo)[NG], const uint32_t (&hi)[NG], uint32_t r,
                                           uint32_t (&even)[NG], uint32_t (&odd)[NG]) {
  const uint32_t se = sel_even_of(r), so = sel_odd_of(r);
#pragma unroll
  for (int k = 0; k < NG; ++k) {
    even[k] += __builtin_amdgcn_perm(hi[k], lo[k], se);
    odd[k] += __builtin_amdgcn_perm(hi[k], lo[k], so);
  }
}

template <int RUN, int DPW, bool LDS, bool FLUSH>
__global__ void __launch_bounds__(512) rowsum_u8(const RowSumU8Params p) {
  constexpr int NG = RUN / 4;   // four-sample groups per lane
  constexpr int TS = 64 * RUN;  // samples per tile
  extern __shared__ __attribute__((aligned(16))) uint32_t lds[];

  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wave = uniform(threadIdx.x >> 6);
  const uint32_t nwaves = blockDim.x >> 6;
  const uint32_t t0 = blockIdx.x * TS;
  const uint32_t d0 = blockIdx.y * nwaves * DPW;
  const uint32_t dw = d0 + wave * DPW;  // first DM of this wavefront
  const uint32_t group = blockIdx.z % p.n_groups;
  const uint32_t beam = blockIdx.z / p.n_groups;
  const uint32_t nlive = p.live_count[group];
  const uint32_t *live = p.live + (uint64_t)group * p.live_pitch;
  const uint32_t WD = p.win_bytes / 4;  // dwords per staged window

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

  const uint32_t G = p.unroll;
  uint32_t since_flush = 0;
  for (uint32_t k0 = 0; k0 < nlive; k0 += G) {
    const uint32_t n = min(G, nlive - k0);
    if constexpr (FLUSH) {
      if (since_flush + n > 256) {
        flush();
        since_flush = 0;
      }
      since_flush += n;
    }
    if constexpr (LDS) {
      // Stage the n row windows: wavefront w stages rows w, w + nwaves, ...
      for (uint32_t g = wave; g < n; g += nwaves) {
        const RowInfo ri = row_info(p, live[k0 + g], beam, t0, d0);
        const uint64_t a0 = ri.off & ~uint64_t(3);
        const uint32_t recs = records_from(p.in_bytes, a0);
        const __amdgpu_buffer_rsrc_t rsrc = make_rsrc(p.in + a0, recs);
        uint32_t *win = lds + g * WD;
        for (uint32_t m = lane; m < WD / 4; m += 64) {
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
      const uint32_t row = live[k0 + g];
      const RowInfo ri = row_info(p, row, beam, t0, d0);
      const uint32_t *shrow = p.ishift + (uint64_t)row * p.dm_pitch + dw;
      if constexpr (LDS) {
        const uint32_t a = (uint32_t)(ri.off & 3);
        const uint32_t *win = lds + g * WD + lane;
#pragma unroll
        for (int j = 0; j < DPW; ++j) {
          const uint32_t o = shrow[j] - ri.shmin + a;  // byte offset of the DM's samples in the window
          const uint32_t *src = win + (o >> 2);
          uint32_t lo[NG], hi[NG];
#pragma unroll
          for (int k = 0; k < NG; ++k) {
            lo[k] = src[64 * k];
            hi[k] = src[64 * k + 1];
          }
          accumulate<NG>(lo, hi, o & 3, even[j], odd[j]);
        }
      } else {
        const uint64_t base = ri.off - ri.shmin;  // byte offset of sample t0
#pragma unroll
        for (int j = 0; j < DPW; ++j) {
          const uint64_t start = base + shrow[j];
          const uint64_t a0 = start & ~uint64_t(3);
          const __amdgpu_buffer_rsrc_t rsrc = make_rsrc(p.in + a0, records_from(p.in_bytes, a0));
          uint32_t lo[NG], hi[NG];
#pragma unroll
          for (int k = 0; k < NG; ++k) {
            lo[k] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, 4 * (lane + 64 * k), 0, 0);
            hi[k] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, 4 * (lane + 64 * k) + 4, 0, 0);
          }
          accumulate<NG>(lo, hi, (uint32_t)(start & 3), even[j], odd[j]);
        }
      }
    }
    if constexpr (LDS) __syncthreads();
  }
  flush();

#pragma unroll
  for (int j = 0; j < DPW; ++j) {
    const uint32_t dm = dw + j;
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
      const uint32_t t = t0 + 4 * lane + 256 * k;
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

template <int RUN, int DPW, bool FLUSH>
hipError_t launch_one(const RowSumU8Params &p, const Geometry &g, dim3 grid, hipStream_t s) {
  const uint32_t threads = 64 * g.threads_d1;
  if (g.local)
    hipLaunchKernelGGL((rowsum_u8<RUN, DPW, true, FLUSH>), grid, dim3(threads), g.lds_bytes, s, p);
  else
    hipLaunchKernelGGL((rowsum_u8<RUN, DPW, false, FLUSH>), grid, dim3(threads), 0, s, p);
  return hipGetLastError();
}

template <int RUN, bool FLUSH>
hipError_t dispatch_dpw(const RowSumU8Params &p, const Geometry &g, dim3 grid, hipStream_t s) {
  switch (g.items_d1) {
  case 1: return launch_one<RUN, 1, FLUSH>(p, g, grid, s);
  case 2: return launch_one<RUN, 2, FLUSH>(p, g, grid, s);
  case 4: return launch_one<RUN, 4, FLUSH>(p, g, grid, s);
  case 8: return launch_one<RUN, 8, FLUSH>(p, g, grid, s);
  default: return hipErrorInvalidValue;
  }
}

template <bool FLUSH>
hipError_t dispatch_run(const RowSumU8Params &p, const Geometry &g, dim3 grid, hipStream_t s) {
  switch (g.items_d0) {
  case 4: return dispatch_dpw<4, FLUSH>(p, g, grid, s);
  case 8: return dispatch_dpw<8, FLUSH>(p, g, grid, s);
  case 16: return dispatch_dpw<16, FLUSH>(p, g, grid, s);
  default: return hipErrorInvalidValue;
  }
}

} // namespace

bool rowsum_u8_supported(uint32_t items_d0, uint32_t items_d1) {
  bool run_ok = items_d0 == 4 || items_d0 == 8 || items_d0 == 16;
  bool dpw_ok = items_d1 == 1 || items_d1 == 2 || items_d1 == 4 || items_d1 == 8;
  return run_ok && dpw_ok;
}

hipError_t launch_rowsum_u8(const RowSumU8Params &p, const Geometry &g, bool flush, dim3 grid,
                            hipStream_t stream) {
  if (p.in_log2f != 0) return hipErrorInvalidValue;  // packed rows: pipelined kernels only
  return flush ? dispatch_run<true>(p, g, grid, stream) : dispatch_run<false>(p, g, grid, stream);
}

} // namespace ddamd
