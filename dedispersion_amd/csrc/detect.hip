// detect.hip -- the step after dedispersion: a boxcar peak search over every
// dedispersed row (one DM trial of one beam).
//
// The reference stops at the dedispersed array; the search that follows it
// in a pipeline is not part of its repository.  This kernel reads each row
// once and reports, for boxcar widths 1, 2, 4, ... 2^(n-1), the window with
// the highest signal-to-noise ratio
//     snr = (B_w(t) - w * mean) / (sigma * sqrt(w)),   B_w(t) = sum_{i<w} y(t + i),
// over starts t with t + w <= S, where mean and sigma (population) are those
// of the whole row.
//
// The rows are integer-valued (sums of uint8 or packed samples, each below
// 2^24; a fractional part is truncated): the row sum, the sum of squares
// (below 2^63 while S x max^2 is) and every boxcar are exact integers, so the
// maximum of each width and its first start are exact, whatever the order of
// the work.  Only the final ratios are floating point (double, in a fixed
// order).
//
// Layout: a workgroup of 256 lanes owns a row and walks it in segments of
// 8192 starts, staging the segment plus the 64-sample reach of the widest
// boxcar in LDS as integers.  An in-place prefix scan (each lane a run of 33
// samples, then the 256 run totals across the workgroup) turns the segment
// into inclusive prefix sums Q, kept modulo 2^32: a boxcar
// B_w(t) = Q[t + w - 1] - Q[t - 1] is exact because every boxcar is below
// 2^31.  Lane l then takes the starts l + 256m, so the 64 lanes of a
// wavefront read consecutive words (no bank conflicts), and keeps, per width,
// the largest boxcar and its first start.  No array is indexed at run time.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <string>

#include "dedispersion_amd.h"
#include "internal.hpp"

namespace {

using ddamd::set_error;

constexpr int kThreads = 256;
constexpr int kSeg = 8192;               // starts per segment
constexpr int kWidths = 7;               // widths 1 .. 64
constexpr int kMaxWidth = 1 << (kWidths - 1);
constexpr int kStage = kSeg + kMaxWidth; // samples staged per segment
constexpr int kChunk = (kStage + kThreads - 1) / kThreads;  // scan run per lane (33)

struct DetectParams {
  const float *rows;
  uint64_t pitch;       // floats between rows
  uint64_t n_rows;
  uint32_t n_samples;
  uint32_t n_widths;    // widths 1 .. 2^(n_widths - 1)
  uint32_t vec;         // rows 16-byte aligned with a pitch of whole float4s
  dd_peak *out;
};

// (max, first start) of two candidates of one width
__device__ __forceinline__ void better(int32_t &b, uint32_t &t, int32_t ob, uint32_t ot) {
  if (ob > b || (ob == b && ot < t)) {
    b = ob;
    t = ot;
  }
}

__global__ void __launch_bounds__(kThreads) detect_peaks(const DetectParams p) {
  __shared__ __attribute__((aligned(16))) uint32_t q[kChunk * kThreads];
  __shared__ uint32_t wave_tot[kThreads / 64];
  __shared__ int64_t red_sum[kThreads / 64], red_sq[kThreads / 64];
  __shared__ int32_t red_b[kThreads / 64][kWidths];
  __shared__ uint32_t red_t[kThreads / 64][kWidths];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t S = p.n_samples;

  for (uint64_t row = blockIdx.x; row < p.n_rows; row += gridDim.x) {
    const float *y = p.rows + row * p.pitch;
    int64_t sum = 0, sq64 = 0;
    int32_t bmax[kWidths];
    uint32_t bt[kWidths];
#pragma unroll
    for (int w = 0; w < kWidths; ++w) {
      bmax[w] = INT_MIN;
      bt[w] = UINT_MAX;
    }

    for (uint32_t s0 = 0; s0 < S; s0 += kSeg) {
      const uint32_t own = min((uint32_t)kSeg, S - s0);     // starts of this segment
      const uint32_t nload = min((uint32_t)kStage, S - s0);  // samples staged
      __syncthreads();
      // ---- stage the samples as integers (zeros past the row) ----
      if (p.vec) {
        for (uint32_t i = 4 * threadIdx.x; i < (uint32_t)(kChunk * kThreads); i += 4 * kThreads) {
          uint32_t v0 = 0, v1 = 0, v2 = 0, v3 = 0;
          if (i + 3 < nload) {
            const float4 f = *reinterpret_cast<const float4 *>(y + s0 + i);
            v0 = (uint32_t)(int32_t)f.x;
            v1 = (uint32_t)(int32_t)f.y;
            v2 = (uint32_t)(int32_t)f.z;
            v3 = (uint32_t)(int32_t)f.w;
          } else {
            if (i < nload) v0 = (uint32_t)(int32_t)y[s0 + i];
            if (i + 1 < nload) v1 = (uint32_t)(int32_t)y[s0 + i + 1];
            if (i + 2 < nload) v2 = (uint32_t)(int32_t)y[s0 + i + 2];
          }
          q[i] = v0;
          q[i + 1] = v1;
          q[i + 2] = v2;
          q[i + 3] = v3;
        }
      } else {
        for (uint32_t i = threadIdx.x; i < (uint32_t)(kChunk * kThreads); i += kThreads)
          q[i] = i < nload ? (uint32_t)(int32_t)y[s0 + i] : 0u;
      }
      __syncthreads();

      // ---- row statistics over this segment's own samples, and the scan ----
      // lane l: run [33 l, 33 l + 33); inclusive prefix within the run
      const uint32_t base = kChunk * threadIdx.x;
      uint32_t run = 0;
      if (base + kChunk <= own) {
        uint32_t s32 = 0;  // 33 samples below 2^24: no overflow
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
          const uint32_t v = q[base + j];
          s32 += v;
          sq64 += (int64_t)(int32_t)v * (int32_t)v;
          run += v;
          q[base + j] = run;
        }
        sum += s32;
      } else {
        for (int j = 0; j < kChunk; ++j) {
          const uint32_t v = q[base + j];
          if (base + j < own) {
            sum += (int32_t)v;
            sq64 += (int64_t)(int32_t)v * (int32_t)v;
          }
          run += v;
          q[base + j] = run;
        }
      }
      // exclusive prefix of the run totals: wavefront scan, then the wavefronts
      uint32_t incl = run;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t n = __shfl_up(incl, off);
        if ((int)lane >= off) incl += n;
      }
      if (lane == 63) wave_tot[wave] = incl;
      __syncthreads();
      uint32_t excl = incl - run;
      for (uint32_t v = 0; v < wave; ++v) excl += wave_tot[v];
      for (int j = 0; j < kChunk; ++j) q[base + j] += excl;
      __syncthreads();

      // ---- boxcars: starts t = l + 256 m of this segment ----
      const bool all_widths = p.n_widths == (uint32_t)kWidths;
      for (uint32_t t = threadIdx.x; t < own; t += kThreads) {
        const uint32_t before = t > 0 ? q[t - 1] : 0u;
        const uint32_t left = S - s0 - t;  // samples from t to the row's end
        if (all_widths && left >= (uint32_t)kMaxWidth) {
          // every width fits: no checks
#pragma unroll
          for (int w = 0; w < kWidths; ++w) {
            const int32_t b = (int32_t)(q[t + (1u << w) - 1] - before);
            const bool up = b > bmax[w];
            bmax[w] = up ? b : bmax[w];
            bt[w] = up ? s0 + t : bt[w];
          }
        } else {
#pragma unroll
          for (int w = 0; w < kWidths; ++w) {
            if ((uint32_t)w < p.n_widths && (1u << w) <= left) {
              const int32_t b = (int32_t)(q[t + (1u << w) - 1] - before);
              if (b > bmax[w]) {
                bmax[w] = b;
                bt[w] = s0 + t;
              }
            }
          }
        }
      }
    }

    // ---- reductions: wavefront shuffles, then the four wavefronts in LDS ----
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      sum += __shfl_xor(sum, off);
      sq64 += __shfl_xor(sq64, off);
#pragma unroll
      for (int w = 0; w < kWidths; ++w) {
        const int32_t ob = __shfl_xor(bmax[w], off);
        const uint32_t ot = __shfl_xor(bt[w], off);
        better(bmax[w], bt[w], ob, ot);
      }
    }
    if (lane == 0) {
      red_sum[wave] = sum;
      red_sq[wave] = sq64;
#pragma unroll
      for (int w = 0; w < kWidths; ++w) {
        red_b[wave][w] = bmax[w];
        red_t[wave][w] = bt[w];
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int64_t tsum = 0, tsq = 0;
#pragma unroll
      for (int v = 0; v < kThreads / 64; ++v) {
        tsum += red_sum[v];
        tsq += red_sq[v];
      }
      const double dS = (double)S;
      const double mean = (double)tsum / dS;
      double var = (double)tsq / dS - mean * mean;
      if (!(var > 0.0)) var = 0.0;
      const double sigma = sqrt(var);
      double best = 0.0;
      uint32_t bw = 0, bts = 0;
      bool found = false;
#pragma unroll
      for (int w = 0; w < kWidths; ++w) {
        int32_t b = red_b[0][w];
        uint32_t t = red_t[0][w];
#pragma unroll
        for (int v = 1; v < kThreads / 64; ++v) better(b, t, red_b[v][w], red_t[v][w]);
        if ((uint32_t)w < p.n_widths && t != UINT_MAX) {
          const double width = (double)(1u << w);
          const double snr = sigma > 0.0 ? ((double)b - width * mean) / (sigma * sqrt(width)) : 0.0;
          if (!found || snr > best) {
            found = true;
            best = snr;
            bw = 1u << w;
            bts = t;
          }
        }
      }
      dd_peak o;
      o.snr = (float)best;
      o.mean = (float)mean;
      o.sigma = (float)sigma;
      o.sample = found ? bts : 0;
      o.width = found ? bw : 0;
      p.out[row] = o;
    }
  }
}

dd_status fail_detect(dd_status s, const std::string &msg) {
  set_error(msg);
  return s;
}

} // namespace

extern "C" {

dd_status dd_detect_peaks(const float *d_rows, uint64_t nr_rows, uint32_t nr_samples, uint64_t pitch,
                          uint32_t nr_widths, dd_peak *d_peaks, void *stream) {
  if (!d_rows || !d_peaks) return fail_detect(DD_ERR_INVALID_ARGUMENT, "null argument");
  if (nr_rows == 0) return DD_OK;
  if (nr_samples == 0 || pitch < nr_samples) return fail_detect(DD_ERR_INVALID_ARGUMENT, "need 0 < nr_samples <= pitch");
  if (nr_widths < 1 || nr_widths > (uint32_t)kWidths)
    return fail_detect(DD_ERR_INVALID_ARGUMENT, "nr_widths must be in [1, 7] (widths 1 .. 64)");
  DetectParams p{};
  p.rows = d_rows;
  p.pitch = pitch;
  p.n_rows = nr_rows;
  p.n_samples = nr_samples;
  p.n_widths = nr_widths;
  p.vec = (pitch % 4 == 0) && ((reinterpret_cast<uintptr_t>(d_rows) & 15) == 0);
  p.out = d_peaks;
  int device = 0, cus = 256;
  (void)hipGetDevice(&device);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  const uint64_t grid = std::min<uint64_t>(nr_rows, (uint64_t)cus * 8);
  hipLaunchKernelGGL(detect_peaks, dim3((uint32_t)grid), dim3(kThreads), 0, static_cast<hipStream_t>(stream), p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_detect(DD_ERR_HIP, std::string("detect_peaks launch: ") + hipGetErrorString(e));
  return DD_OK;
}

} // extern "C"
