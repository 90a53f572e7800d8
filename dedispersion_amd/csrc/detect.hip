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
// The rows are non-negative integer-valued floats (sums of uint8 or packed
// samples, each below 2^24; a fractional part is truncated): the row sum,
// the sum of squares and every boxcar are exact integers, so the maximum of
// each width and its first start are exact, whatever the order of the work.
// Only the final ratios are floating point (double, in a fixed order).
//
// Layout: one wavefront walks whole rows, a chunk of 64 runs of 68 samples
// at a time.  The chunk is read with coalesced 16-byte loads (the next
// chunk's loads are issued before the current one is searched), converted to
// integers and passed through LDS so that lane l holds run l in registers.
// Each lane turns its run into prefix sums Q (modulo 2^32; every boxcar is
// below 2^31, so differences are exact) and takes, through a DPP shift, the
// prefix sums of the next lane's run: the window of any start in the run,
// up to 64 wide, then ends in one of the two, and B_w(i) = Q(i + w - 1) -
// Q(i - 1) is one subtraction.  Lane 63 only supplies the last lane's
// neighbour, so a chunk advances 63 runs (4284 starts).
//
// Samples past the row's end are read as zeros (buffer loads bounded by the
// row).  Because the samples are non-negative, a window that runs past the
// end sums to at most the last window that fits, B_w(S - w), whose start is
// earlier: the maximum of each width and its first start do not change, so
// no start needs a bounds check.
//
// Per lane and width the search keeps the largest boxcar and the start of the
// first run (68 starts) where it occurs; at the row's end the wavefront
// reduces them, picks the width of the highest ratio, and finds the first
// start inside that width's first run by rescanning those 131 samples.
//
// One launch takes up to 64 sets of rows (the outputs of as many beams,
// lengths and pitches free to differ) from its kernel arguments and deals
// all their rows to 2048 wavefronts round robin, so a batch of beams pays
// one launch's ramp and drain instead of one per beam.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <string>

#include "dedispersion_amd.h"
#include "internal.hpp"
#include "staged_common.hpp"

namespace {

using ddamd::set_error;
using ddamd::staged::make_rsrc;
using ddamd::staged::records_from_floats;
using ddamd::staged::u32x4;

constexpr int kWidths = 7;                    // widths 1 .. 64
constexpr int kMaxWidth = 1 << (kWidths - 1);
constexpr int kRun = 68;                      // starts per lane per chunk
constexpr int kVec = kRun / 4;                // 16-byte pieces per lane per chunk
constexpr int kHalo = kMaxWidth - 1;          // samples past a run that a window reaches
constexpr uint32_t kStage = 64 * kRun;        // samples staged per chunk (4352)
constexpr uint32_t kStep = 63 * kRun;         // starts searched per chunk (4284)
constexpr uint32_t kRescan = kRun + kHalo;    // samples of one run's windows (131)

static_assert(kRun % 4 == 0 && kRun >= kHalo, "a run covers the reach of the widest window");

// One set of rows: nr_rows rows of n_samples floats, pitch floats apart.
struct DetectJob {
  const float *rows;
  uint64_t pitch;
  dd_peak *out;
  uint32_t n_samples;
  uint32_t chunks;      // chunks per row, ceil(n_samples / kStep)
};

constexpr uint32_t kMaxJobs = 64;  // sets of rows per launch (kernel arguments)

// The rows of all jobs in a launch, numbered job after job: job j holds rows
// first_row[j] .. first_row[j + 1] - 1.
struct DetectParams {
  DetectJob jobs[kMaxJobs];
  uint64_t first_row[kMaxJobs + 1];
  uint32_t n_jobs;
  uint32_t n_widths;    // widths 1 .. 2^(n_widths - 1)
};

// Reads the kStage samples of chunk c of a row, piece j of lane l being
// samples 4 (64 j + l) .. +3 of the chunk: 16-byte loads when the rows are
// 16-byte aligned (VEC), else single floats.  Zeros past the row's end.
template <bool VEC>
__device__ __forceinline__ void load_chunk(const DetectJob &job, uint64_t row, uint32_t c, uint32_t lane,
                                           u32x4 (&v)[kVec]) {
  const uint32_t s0 = c * kStep;
  const uint32_t left = min(job.n_samples - s0, kStage);  // samples of the row in this chunk
  const float *base = job.rows + row * job.pitch + s0;
  const auto exact = make_rsrc(base, 4 * left);
  if (VEC) {
    // whole pieces inside the row; the piece holding the row's last samples,
    // if the row ends inside one, is read a float at a time
    const auto whole = make_rsrc(base, 16 * (left / 4));
#pragma unroll
    for (int j = 0; j < kVec; ++j) v[j] = __builtin_amdgcn_raw_buffer_load_b128(whole, 16 * (64 * j + lane), 0, 0);
    const uint32_t q = left / 4;  // the piece ending the row, if partial
    if ((left & 3) != 0) {
#pragma unroll
      for (int j = 0; j < kVec; ++j) {
        if (q == 64u * j + lane) {
#pragma unroll
          for (int e = 0; e < 4; ++e) v[j][e] = __builtin_amdgcn_raw_buffer_load_b32(exact, 16 * q + 4 * e, 0, 0);
        }
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < kVec; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        v[j][e] = __builtin_amdgcn_raw_buffer_load_b32(exact, 16 * (64 * j + lane) + 4 * e, 0, 0);
  }
}

// lane l <- lane l + 1 (DPP wave_shl:1); lane 63 reads 0
__device__ __forceinline__ uint32_t from_next_lane(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x130, 0xf, 0xf, true);
}

// max(m, a, b) as one v_max3_u32 (the compiler otherwise pairs the two new
// values first, three instructions for every four values)
__device__ __forceinline__ uint32_t max3(uint32_t m, uint32_t a, uint32_t b) {
  uint32_t r;
  asm("v_max3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(m), "v"(a), "v"(b));
  return r;
}

// Searches lane `lane`'s run of the staged chunk: m[w] = the largest boxcar of
// width 2^w over the run's 68 starts, T = the run's sum, sq = its sum of
// squares.  Starts are taken in order, all widths at each, and prefix sum
// 68 + j of the run's frame (this run's total plus the next lane's prefix
// sum j) is formed just before the first start whose window reaches it, so
// that few prefix sums are live at once.
__device__ __forceinline__ void search_run(const uint32_t *stage, uint32_t lane, uint32_t (&m)[kWidths], uint32_t &T,
                                           uint64_t &sq) {
  uint32_t Q[kRun];
#pragma unroll
  for (int k = 0; k < kVec; ++k) {
    const u32x4 v = *reinterpret_cast<const u32x4 *>(stage + kRun * lane + 4 * k);
    Q[4 * k] = v[0];
    Q[4 * k + 1] = v[1];
    Q[4 * k + 2] = v[2];
    Q[4 * k + 3] = v[3];
  }
  uint64_t s2 = 0;
  uint32_t m0 = 0;
#pragma unroll
  for (int i = 0; i < kRun; i += 2) {
    const uint32_t x0 = Q[i], x1 = Q[i + 1];
    s2 += (uint64_t)x0 * x0;  // each below 2^48
    asm volatile("" : "+v"(s2));  // one multiply-add per sample
    s2 += (uint64_t)x1 * x1;
    m0 = max3(m0, x0, x1);
    Q[i] = (i > 0 ? Q[i - 1] : 0u) + x0;
    Q[i + 1] = Q[i] + x1;
    // opaque to the optimizer, which would otherwise rewrite short boxcars
    // Q[e] - Q[i - 1] as sums of the samples and keep every sample live
    asm("" : "+v"(Q[i]), "+v"(Q[i + 1]));
    // the squares here, not sunk to the end with every sample held for them
    asm volatile("" : "+v"(s2));
    __builtin_amdgcn_sched_barrier(0);
  }
  T = Q[kRun - 1];
  m[0] = m0;
#pragma unroll
  for (int w = 1; w < kWidths; ++w) m[w] = 0;

  uint32_t N[kHalo];
#pragma unroll
  for (int i = 0; i < kRun; i += 2) {
    // the widest window of start i + 1 ends at i + 64
#pragma unroll
    for (int j = i + kMaxWidth - kRun - 1; j <= i + kMaxWidth - kRun; ++j)
      if (j >= 0 && j < kHalo) N[j] = from_next_lane(Q[j]) + T;
    const uint32_t lo0 = i > 0 ? Q[i - 1] : 0u, lo1 = Q[i];
#pragma unroll
    for (int w = 1; w < kWidths; ++w) {
      const int e0 = i + (1 << w) - 1, e1 = e0 + 1;
      const uint32_t h0 = e0 < kRun ? Q[e0] : N[e0 - kRun];
      const uint32_t h1 = e1 < kRun ? Q[e1] : N[e1 - kRun];
      m[w] = max3(m[w], h0 - lo0, h1 - lo1);
    }
    // keep this order: the scheduler would otherwise form every N early
    __builtin_amdgcn_sched_barrier(0);
  }
  sq = s2;
}

// Wavefront maximum / minimum, the result read from lane 63.
template <bool MAX> __device__ __forceinline__ uint32_t wave_extreme(uint32_t v) {
  auto pick = [](uint32_t a, uint32_t b) { return MAX ? max(a, b) : min(a, b); };
  const int keep = MAX ? 0 : -1;
  v = pick(v, (uint32_t)__builtin_amdgcn_update_dpp(keep, (int)v, 0x111, 0xf, 0xf, false));  // row_shr:1
  v = pick(v, (uint32_t)__builtin_amdgcn_update_dpp(keep, (int)v, 0x112, 0xf, 0xf, false));  // row_shr:2
  v = pick(v, (uint32_t)__builtin_amdgcn_update_dpp(keep, (int)v, 0x114, 0xf, 0xf, false));  // row_shr:4
  v = pick(v, (uint32_t)__builtin_amdgcn_update_dpp(keep, (int)v, 0x118, 0xf, 0xf, false));  // row_shr:8
  v = pick(v, (uint32_t)__builtin_amdgcn_update_dpp(keep, (int)v, 0x142, 0xa, 0xf, false));  // row_bcast:15
  v = pick(v, (uint32_t)__builtin_amdgcn_update_dpp(keep, (int)v, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// inclusive prefix sum over the wavefront's lanes
__device__ __forceinline__ uint32_t wave_scan(uint32_t v, uint32_t lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t n = __shfl_up(v, off);
    if (lane >= (uint32_t)off) v += n;
  }
  return v;
}

// The row's statistics and ratios, in double with no contraction, in the
// order of the straight-loop definition.
struct Pick {
  double snr, mean, sigma;
  uint32_t w;  // width index
};

__device__ Pick pick_width(uint64_t tsum, uint64_t tsq, const uint32_t (&M)[kWidths], uint32_t S, uint32_t nw) {
#pragma clang fp contract(off)
  const double dS = (double)S;
  const double mean = (double)tsum / dS;
  double var = (double)tsq / dS - mean * mean;
  if (!(var > 0.0)) var = 0.0;
  const double sigma = sqrt(var);
  Pick r{0.0, mean, sigma, 0};
  bool found = false;
#pragma unroll
  for (int w = 0; w < kWidths; ++w) {
    if ((uint32_t)w < nw && (1u << w) <= S) {
      const double width = (double)(1u << w);
      const double snr = sigma > 0.0 ? ((double)M[w] - width * mean) / (sigma * sqrt(width)) : 0.0;
      if (!found || snr > r.snr) {
        found = true;
        r.snr = snr;
        r.w = w;
      }
    }
  }
  return r;
}

// First start t in [P, P + 68) with B_W(t) == M (zeros past the row's end).
__device__ uint32_t first_start(const float *y, uint32_t S, uint32_t P, uint32_t W, uint32_t M, uint32_t lane) {
  const auto rsrc = make_rsrc(y + P, records_from_floats(S, P));
  auto rd = [&](uint32_t i) { return (uint32_t)__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc, 4 * i, 0, 0)); };
  const uint32_t a = rd(lane), b = rd(64 + lane);
  const uint32_t c = lane < kRescan - 128 ? rd(128 + lane) : 0u;
  const uint32_t sa = wave_scan(a, lane);
  const uint32_t sb = wave_scan(b, lane) + (uint32_t)__builtin_amdgcn_readlane((int)sa, 63);
  const uint32_t sc = wave_scan(c, lane) + (uint32_t)__builtin_amdgcn_readlane((int)sb, 63);
  // Q(n) = sum of samples P .. P + n, n < 131
  auto Q = [&](uint32_t n) {
    const int src = (int)(n & 63);
    const uint32_t qa = (uint32_t)__shfl((int)sa, src), qb = (uint32_t)__shfl((int)sb, src),
                   qc = (uint32_t)__shfl((int)sc, src);
    return n < 64 ? qa : (n < 128 ? qb : qc);
  };
  // starts k0 = lane and k1 = 64 + lane (lanes 0-3); every lane shuffles
  const uint32_t k0 = lane, k1 = min(64 + lane, (uint32_t)kRun - 1);
  const uint32_t before0 = Q(k0 > 0 ? k0 - 1 : 0u);
  const uint32_t b0 = Q(k0 + W - 1) - (k0 > 0 ? before0 : 0u);
  const uint32_t b1 = Q(k1 + W - 1) - Q(k1 - 1);
  const uint64_t hit0 = __ballot(b0 == M);
  const uint64_t hit1 = __ballot(b1 == M && 64 + lane < (uint32_t)kRun);
  if (hit0) return P + (uint32_t)__builtin_ctzll(hit0);
  if (hit1) return P + 64 + (uint32_t)__builtin_ctzll(hit1);
  return P;  // not reached: M occurs in this run
}

// One wavefront per workgroup.  Wavefront b takes rows b, b + gridDim.x, ...
// of the launch's numbering, each whole, chunk after chunk; the next chunk
// (of this row or the next) is loaded while this one is searched.
template <bool VEC> __global__ void __launch_bounds__(64) detect_peaks(const DetectParams p) {
  __shared__ __attribute__((aligned(16))) uint32_t stage[kStage];
  const uint32_t lane = threadIdx.x;
  const uint64_t G = gridDim.x, n_rows = p.first_row[p.n_jobs];
  uint64_t g = blockIdx.x;  // row of the launch
  if (g >= n_rows) return;
  uint32_t j = 0;           // its job
  while (g >= p.first_row[j + 1]) ++j;
  uint32_t c = 0;           // its chunk

  // aligned rows: the next chunk's loads are in flight while this one is
  // searched; single-float reads (unaligned rows) load each chunk in turn
  u32x4 pre[kVec];
  if (VEC) load_chunk<VEC>(p.jobs[j], g - p.first_row[j], 0, lane, pre);

  uint32_t best[kWidths], pos[kWidths];
  uint64_t sum = 0, sq = 0;
#pragma unroll
  for (int w = 0; w < kWidths; ++w) {
    best[w] = 0;
    pos[w] = kRun * lane;
  }

  for (;;) {
    const DetectJob &job = p.jobs[j];
    const uint64_t row = g - p.first_row[j];
    if (!VEC) load_chunk<VEC>(job, row, c, lane, pre);
    // ---- stage this chunk as integers, then issue the next chunk's loads ----
#pragma unroll
    for (int k = 0; k < kVec; ++k) {
      u32x4 v;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        // through a scalar: __builtin_bit_cast of a vector element reads
        // element 0 whatever e is (hipcc 7.2)
        const uint32_t bits = pre[k][e];
        v[e] = (uint32_t)__builtin_bit_cast(float, bits);
      }
      *reinterpret_cast<u32x4 *>(stage + 4 * (64 * k + lane)) = v;
    }
    const bool row_end = c + 1 == job.chunks;
    uint64_t ng = g;
    uint32_t nj = j, nc = c + 1;
    if (row_end) {
      ng += G;
      nc = 0;
      while (nj < p.n_jobs && ng >= p.first_row[nj + 1]) ++nj;
    }
    const bool more = ng < n_rows;
    if (VEC && more) load_chunk<VEC>(p.jobs[nj], ng - p.first_row[nj], nc, lane, pre);
    __syncthreads();

    // ---- search this lane's run ----
    uint32_t m[kWidths], T;
    uint64_t s2;
    search_run(stage, lane, m, T, s2);
    const uint32_t start = c * kStep + kRun * lane;
#pragma unroll
    for (int w = 0; w < kWidths; ++w) {
      const bool up = m[w] > best[w];
      best[w] = up ? m[w] : best[w];
      pos[w] = up ? start : pos[w];
    }
    if (lane < 63) {  // lane 63's run is searched again as the next chunk's first
      sum += T;
      sq += s2;
    }
    __syncthreads();

    // ---- the row's end: reduce, pick the width, find its first start ----
    if (row_end) {
      const uint32_t S = job.n_samples;
      const uint64_t tsum = wave_sum64(sum), tsq = wave_sum64(sq);
      uint32_t M[kWidths], P[kWidths];
#pragma unroll
      for (int w = 0; w < kWidths; ++w) {
        M[w] = wave_extreme<true>(best[w]);
        P[w] = wave_extreme<false>(best[w] == M[w] ? pos[w] : UINT_MAX);
      }
      const Pick k = pick_width(tsum, tsq, M, S, p.n_widths);
      uint32_t Pk = P[0], Mk = M[0];
#pragma unroll
      for (int w = 1; w < kWidths; ++w) {
        Pk = k.w == (uint32_t)w ? P[w] : Pk;
        Mk = k.w == (uint32_t)w ? M[w] : Mk;
      }
      const uint32_t t = first_start(job.rows + row * job.pitch, S, Pk, 1u << k.w, Mk, lane);
      if (lane == 0) {
        dd_peak o;
        o.snr = (float)k.snr;
        o.mean = (float)k.mean;
        o.sigma = (float)k.sigma;
        o.sample = t;
        o.width = 1u << k.w;
        job.out[row] = o;
      }
      sum = sq = 0;
#pragma unroll
      for (int w = 0; w < kWidths; ++w) {
        best[w] = 0;
        pos[w] = kRun * lane;
      }
    }
    if (!more) break;
    g = ng;
    j = nj;
    c = nc;
  }
}

dd_status fail_detect(dd_status s, const std::string &msg) {
  set_error(msg);
  return s;
}

dd_status check_job(const dd_detect_job &j) {
  if (j.nr_rows == 0) return DD_OK;
  if (!j.d_rows || !j.d_peaks) return fail_detect(DD_ERR_INVALID_ARGUMENT, "null argument");
  if (j.nr_samples == 0 || j.pitch < j.nr_samples)
    return fail_detect(DD_ERR_INVALID_ARGUMENT, "need 0 < nr_samples <= pitch");
  if (j.nr_samples > 0xFFFFFFFFu - kStage)
    return fail_detect(DD_ERR_INVALID_ARGUMENT, "nr_samples too large for the search's 32-bit starts");
  return DD_OK;
}

// 16-byte loads need the rows 16-byte aligned and a pitch of whole float4s.
bool aligned(const dd_detect_job &j) {
  return j.pitch % 4 == 0 && (reinterpret_cast<uintptr_t>(j.d_rows) & 15) == 0;
}

// One launch over up to kMaxJobs jobs (all aligned, or all not).
dd_status launch(const dd_detect_job *const *jobs, uint32_t n, uint32_t nr_widths, bool vec, hipStream_t stream) {
  DetectParams p{};
  p.n_jobs = n;
  p.n_widths = nr_widths;
  for (uint32_t i = 0; i < n; ++i) {
    const dd_detect_job &j = *jobs[i];
    p.jobs[i].rows = j.d_rows;
    p.jobs[i].pitch = j.pitch;
    p.jobs[i].out = j.d_peaks;
    p.jobs[i].n_samples = j.nr_samples;
    p.jobs[i].chunks = (j.nr_samples + kStep - 1) / kStep;
    p.first_row[i + 1] = p.first_row[i] + j.nr_rows;
  }
  const uint64_t rows = p.first_row[n];
  if (rows == 0) return DD_OK;
  int device = 0, cus = 256;
  (void)hipGetDevice(&device);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  // eight wavefronts per CU are resident (two per SIMD, by registers)
  const uint64_t grid = std::min<uint64_t>(rows, (uint64_t)cus * 8);
  auto kernel = vec ? detect_peaks<true> : detect_peaks<false>;
  hipLaunchKernelGGL(kernel, dim3((uint32_t)grid), dim3(64), 0, stream, p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_detect(DD_ERR_HIP, std::string("detect_peaks launch: ") + hipGetErrorString(e));
  return DD_OK;
}

} // namespace

extern "C" {

dd_status dd_detect_peaks_batch(const dd_detect_job *jobs, uint32_t nr_jobs, uint32_t nr_widths, void *stream) {
  if (nr_jobs == 0) return DD_OK;
  if (!jobs) return fail_detect(DD_ERR_INVALID_ARGUMENT, "null argument");
  if (nr_widths < 1 || nr_widths > (uint32_t)kWidths)
    return fail_detect(DD_ERR_INVALID_ARGUMENT, "nr_widths must be in [1, 7] (widths 1 .. 64)");
  for (uint32_t i = 0; i < nr_jobs; ++i) {
    const dd_status s = check_job(jobs[i]);
    if (s != DD_OK) return fail_detect(s, "job " + std::to_string(i) + ": " + dd_last_error());
  }
  const hipStream_t st = static_cast<hipStream_t>(stream);
  for (int vec = 1; vec >= 0; --vec) {
    const dd_detect_job *group[kMaxJobs];
    uint32_t n = 0;
    for (uint32_t i = 0; i < nr_jobs; ++i) {
      if (jobs[i].nr_rows == 0 || aligned(jobs[i]) != (vec == 1)) continue;
      group[n++] = &jobs[i];
      if (n == kMaxJobs) {
        const dd_status s = launch(group, n, nr_widths, vec == 1, st);
        if (s != DD_OK) return s;
        n = 0;
      }
    }
    if (n) {
      const dd_status s = launch(group, n, nr_widths, vec == 1, st);
      if (s != DD_OK) return s;
    }
  }
  return DD_OK;
}

dd_status dd_detect_peaks(const float *d_rows, uint64_t nr_rows, uint32_t nr_samples, uint64_t pitch,
                          uint32_t nr_widths, dd_peak *d_peaks, void *stream) {
  if (!d_rows || !d_peaks) return fail_detect(DD_ERR_INVALID_ARGUMENT, "null argument");
  if (nr_rows == 0) return DD_OK;
  if (nr_widths < 1 || nr_widths > (uint32_t)kWidths)
    return fail_detect(DD_ERR_INVALID_ARGUMENT, "nr_widths must be in [1, 7] (widths 1 .. 64)");
  dd_detect_job j{};
  j.d_rows = d_rows;
  j.nr_rows = nr_rows;
  j.pitch = pitch;
  j.nr_samples = nr_samples;
  j.d_peaks = d_peaks;
  const dd_status s = check_job(j);
  if (s != DD_OK) return s;
  const dd_detect_job *one = &j;
  return launch(&one, 1, nr_widths, aligned(j), static_cast<hipStream_t>(stream));
}

} // extern "C"
