// batch.cpp -- many plans in one persistent launch per kernel kind.
//
// The reference enqueues one kernel per beam and step
// (src/DedispersionTest.cpp:399-416).  A batch instead lists the tiles of
// every job of one kernel kind -- (job, logical tile) descriptors, each job
// a prepared launch of the single-plan entry points -- and runs them with a
// persistent kernel whose workgroups claim tiles from XCD-local queues
// (TileQueues).  Each tile is computed by the same device function, with
// the same parameters, as in the job's own launch, so results are
// bit-identical to the per-job launches.
//
// Launch order on the stream: clear the claim counters, then single-step
// jobs, step one of the two-step jobs, step two of the two-step jobs.  The
// kernel boundary between step one and step two orders every intermediate
// write before its reads.
//
// Pipelined mode (dd_batch_set_pipeline, chunks >= 2): the two-step jobs are
// split, in order, into chunks of near-equal work; step one of chunk c runs
// on one batch stream and step two of chunk c on another after an event, so
// step two of chunk c overlaps step one of chunk c + 1 on the device.  Each
// launch is a plain grid over the chunk's (job, tile) list, which lets the
// dispatcher interleave the workgroups of both kernels on the compute units.
// The fork and join are events, so the whole pipeline is captured in the
// graph.
//
// Compact intermediate (the default when every two-step job allows it): a
// batch owns its jobs' intermediate buffers, so step one writes its sums --
// integers below 2^16 when 255 x the live channels of any subband fit --
// as uint16 into the first half of each buffer, and step two reads them with
// the uint16 kernel (kernels_u16_staged.hip).  Final outputs are
// bit-identical; the intermediate buffers then hold uint16 sums.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "batch_internal.hpp"

using namespace ddamd;

namespace {

enum JobKind { kSingleStep = 0, kTwoStep = 1 };

struct Job {
  dd_plan *plan;
  JobKind kind;
  const uint8_t *in;
  float *mid;
  float *out;
  uint32_t first_sbeam, nr_sbeams;
};

// The launches of one kernel kind over all its jobs.
struct KindQueue {
  std::vector<U8Launch> u8;    // single step or step one, one per job of the kind
  std::vector<F32Launch> f32;  // step two
  Geometry g{};
  bool flush = false;
  uint32_t n_tiles = 0, workgroups = 0;
  DeviceBuffer<RowSumU8Params> d_u8;
  DeviceBuffer<RowSumF32Params> d_f32;
  DeviceBuffer<TileDesc> d_tiles;
  DeviceBuffer<uint32_t> d_begin;
};

} // namespace

struct dd_batch {
  std::vector<Job> jobs;
  dd_conf conf[3]{};
  bool compact_requested = true;
  bool compact = false;  // in effect after build()
  bool paired_requested = false;  // the compact step two adds subbands in pairs when every job allows it
  bool built = false;
  // one ordinary launch per job (float intermediate): dd_batch_run_per_job
  std::vector<U8Launch> pj_single, pj_one;
  std::vector<F32Launch> pj_two;
  // jobs whose row windows do not fit the batch kernels under the batch's
  // configuration: launched beside them with their plan's own configuration
  // and the float intermediate
  std::vector<U8Launch> solo_single, solo_one;
  std::vector<F32Launch> solo_two;
  // per job, in job order: runs beside the batch kernels; its intermediate holds uint16 sums
  std::vector<bool> beside, compact_job;
  KindQueue q[3];
  DeviceBuffer<uint32_t> heads;  // claim counters, kQueues * kHeadStride per kind
  uint64_t units = 0;
  int cus = 0;
  hipStream_t capture = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // pipelined mode
  uint32_t chunks = 3;                         // <= 1: one persistent launch per kind
  uint32_t lds_reserve[3] = {0, 0, 0};         // least LDS bytes per workgroup, per kind
  std::vector<uint32_t> off1, n1, off2, n2;    // per chunk: tile list offsets and counts
  DeviceBuffer<TileDesc> d_list1, d_list2;     // step one / step two tiles, chunk after chunk
  hipStream_t side[2] = {nullptr, nullptr};    // step one, step two
  std::vector<hipEvent_t> events;              // fork, join x2, one per chunk
  // the allocation-free run: the host image of its last device area
  ddamd::StaticBatchImage static_image;

  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  void invalidate() {
    built = false;
    drop_graph();
  }
  ~dd_batch() {
    drop_graph();
    if (capture) (void)hipStreamDestroy(capture);
    for (auto st : side)
      if (st) (void)hipStreamDestroy(st);
    for (auto e : events)
      if (e) (void)hipEventDestroy(e);
  }
};

namespace {

dd_conf batch_default_conf(dd_mode mode) {
  // The fastest staged shape of an exhaustive sweep of both kernel kinds over
  // 40 uneven 1536-channel beams on MI355X.
  dd_conf c{};
  c.threads_d0 = 64;
  c.split_batches = 0;
  c.local = 1;
  c.threads_d1 = 4;
  c.items_d0 = 16;
  c.items_d1 = 8;
  c.unroll = 8;
  return c;
}

// Splits n tiles into kQueues contiguous ranges of near-equal length.
std::vector<uint32_t> queue_bounds(uint32_t n) {
  std::vector<uint32_t> b(kQueues + 1);
  for (uint32_t q = 0; q <= kQueues; ++q) b[q] = (uint32_t)(((uint64_t)n * q) / kQueues);
  return b;
}

// Whether a job's launch under the batch configuration is the staged
// pipeline (its row windows fit), which the batch kernels run.
template <typename L> bool staged(const L &launch) { return launch.g.staged == 1; }

template <typename L> void merge_geometry(KindQueue &k, const L &launch, bool first) {
  if (first) {
    k.g = launch.g;
  } else {
    k.g.lds_bytes = std::max(k.g.lds_bytes, launch.g.lds_bytes);
    k.g.win = std::max(k.g.win, launch.g.win);
    k.g.packed |= launch.g.packed;
  }
}

// Every job's launches under the batch configuration (and its plan's own),
// sorted into the kind queues or beside them; `pairs`: prepare the compact
// step two's paired kernel where a job allows it.
dd_status plan_jobs(dd_batch *b, bool pairs) {
  for (auto &k : b->q) {
    k.u8.clear();
    k.f32.clear();
    k.n_tiles = 0;
    k.workgroups = 0;
    k.flush = false;
  }
  b->units = 0;
  b->pj_single.clear();
  b->pj_one.clear();
  b->pj_two.clear();
  b->solo_single.clear();
  b->solo_one.clear();
  b->solo_two.clear();
  b->beside.clear();
  b->compact_job.clear();
  for (size_t j = 0; j < b->jobs.size(); ++j) {
    const Job &job = b->jobs[j];
    const dd_observation &o = job.plan->obs;
    const dd_sizes &z = job.plan->sizes;
    const uint64_t sbeams = job.nr_sbeams ? job.nr_sbeams
                                          : o.nr_synthesized_beams - std::min(job.first_sbeam, o.nr_synthesized_beams);
    dd_status s;
    if (job.kind == kSingleStep) {
      U8Launch L, P;
      s = prepare_single(job.plan, job.plan->conf[DD_SINGLE_STEP], job.in, job.out, job.first_sbeam, job.nr_sbeams,
                         P);
      if (s != DD_OK) return s;
      s = prepare_single(job.plan, b->conf[DD_SINGLE_STEP], job.in, job.out, job.first_sbeam, job.nr_sbeams, L);
      if (s != DD_OK && s != DD_ERR_CONFIGURATION) return s;
      b->pj_single.push_back(P);
      b->units += sbeams * o.nr_dms * z.nr_samples;
      b->beside.push_back(s != DD_OK || !staged(L));
      b->compact_job.push_back(false);
      if (b->beside.back()) {
        b->solo_single.push_back(P);
        continue;
      }
      KindQueue &k = b->q[DD_SINGLE_STEP];
      merge_geometry(k, L, k.u8.empty());
      k.flush = k.flush || L.flush;
      k.u8.push_back(L);
    } else {
      // the per-job launches use each plan's own configuration
      U8Launch P1;
      F32Launch P2;
      s = prepare_step_one(job.plan, job.plan->conf[DD_STEP_ONE], job.in, job.mid, P1);
      if (s == DD_OK)
        s = prepare_step_two(job.plan, job.plan->conf[DD_STEP_TWO], job.mid, job.out, job.first_sbeam,
                             job.nr_sbeams, P2);
      if (s != DD_OK) return s;
      b->pj_one.push_back(P1);
      b->pj_two.push_back(P2);
      b->units += sbeams * o.nr_dms_subbanding * o.nr_dms * z.nr_samples;
      U8Launch L1;
      F32Launch L2;
      s = prepare_step_one(job.plan, b->conf[DD_STEP_ONE], job.in, job.mid, L1);
      if (s == DD_OK)
        s = prepare_step_two(job.plan, b->conf[DD_STEP_TWO], job.mid, job.out, job.first_sbeam, job.nr_sbeams, L2);
      // the plan's own launches prepared above; a batch configuration that
      // cannot run this job (its LDS, its windows) leaves it beside the batch
      if (s != DD_OK && s != DD_ERR_CONFIGURATION) return s;
      bool fits = s == DD_OK && staged(L1);
      if (fits && b->compact) {
        L1.k.out_u16 = 1;  // vector stores of four uint16 sums need 8-byte aligned rows
        L1.k.vec_store = (z.sub_out_pitch % 4 == 0) && (reinterpret_cast<uintptr_t>(job.mid) & 7) == 0;
        // the configuration is staged and its float form prepared: what fails
        // here is the uint16 row window or its LDS
        s = prepare_step_two_u16(job.plan, b->conf[DD_STEP_TWO], job.mid, job.out, job.first_sbeam, job.nr_sbeams,
                                 L2, pairs);
        if (s == DD_ERR_CONFIGURATION) fits = false;
        else if (s != DD_OK) return s;
      } else if (fits) {
        fits = staged(L2);
      }
      b->beside.push_back(!fits);
      b->compact_job.push_back(fits && b->compact);
      if (!fits) {
        b->solo_one.push_back(P1);
        b->solo_two.push_back(P2);
        continue;
      }
      KindQueue &k1 = b->q[DD_STEP_ONE];
      merge_geometry(k1, L1, k1.u8.empty());
      k1.flush = k1.flush || L1.flush;
      k1.u8.push_back(L1);
      KindQueue &k2 = b->q[DD_STEP_TWO];
      merge_geometry(k2, L2, k2.f32.empty());
      k2.f32.push_back(L2);
    }
  }
  return DD_OK;
}

// The host part of build(): every job's launches, the per-kind geometry,
// occupancy and tile lists (tiles[m] for kind m); no device memory is
// touched.
dd_status plan_batch(dd_batch *b, std::vector<TileDesc> (&tiles_of)[3]) {
  if (b->cus == 0) {
    int dev = 0;
    DD_HIP(hipGetDevice(&dev));
    DD_HIP(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev));
  }
  // the batch kernels are the staged ones: a configuration that can never
  // select them is an error; a job whose windows do not fit runs on its own
  for (const Job &job : b->jobs) {
    const int m0 = job.kind == kSingleStep ? DD_SINGLE_STEP : DD_STEP_ONE;
    const int m1 = job.kind == kSingleStep ? DD_SINGLE_STEP : DD_STEP_TWO;
    for (int m = m0; m <= m1; ++m) {
      const dd_conf &c = b->conf[m];
      if (c.local != 1 || c.unroll > kStagedSlotsPerWave * c.threads_d1)
        return fail(DD_ERR_CONFIGURATION, "batch configuration of mode " + std::to_string(m) +
                                              " does not select the staged pipeline (needs local = 1 and unroll <= 2 "
                                              "* threads_d1)");
    }
  }
  // compact intermediate: every two-step job's subband sums fit 16 bits
  b->compact = b->compact_requested;
  for (const Job &job : b->jobs)
    if (job.kind == kTwoStep && 255u * std::max(job.plan->max_live_sub, 1u) > 65535u) b->compact = false;
  // one step-two kernel for the whole batch: the paired one only when every
  // job in the batch kernels took it
  bool pairs = b->compact && b->paired_requested;
  dd_status st = plan_jobs(b, pairs);
  if (st != DD_OK) return st;
  if (pairs) {
    bool all = true;
    for (const F32Launch &L : b->q[DD_STEP_TWO].f32) all = all && L.g.pairs;
    if (!all) {
      st = plan_jobs(b, false);
      if (st != DD_OK) return st;
    }
  }
  for (int m = 0; m < 3; ++m) {
    KindQueue &k = b->q[m];
    std::vector<TileDesc> &tiles = tiles_of[m];
    tiles.clear();
    const size_t njobs = m == DD_STEP_TWO ? k.f32.size() : k.u8.size();
    for (size_t j = 0; j < njobs; ++j) {
      const uint32_t n = m == DD_STEP_TWO ? k.f32[j].k.n_tiles : k.u8[j].k.n_tiles;
      for (uint32_t t = 0; t < n; ++t) tiles.push_back(TileDesc{(uint32_t)j, t});
    }
    if (tiles.size() > 0xFFFFFFF0u) return fail(DD_ERR_UNSUPPORTED, "too many tiles in one batch");
    k.n_tiles = (uint32_t)tiles.size();
    if (k.n_tiles == 0) continue;
    // the persistent kernels also hold 16 bytes of static LDS (the claim slot, aligned)
    if (k.g.lds_bytes + 16 > b->jobs.front().plan->max_lds || b->lds_reserve[m] > b->jobs.front().plan->max_lds)
      return fail(DD_ERR_CONFIGURATION, "batch LDS demand exceeds the device limit");
    int per_cu = 0;
    if (m == DD_STEP_TWO) {
      if (b->compact && k.g.pairs) DD_HIP(rowsum_u16_pairs_batch_occupancy(k.g, &per_cu));
      else if (b->compact) DD_HIP(rowsum_u16_batch_occupancy(k.g, &per_cu));
      else DD_HIP(rowsum_f32_batch_occupancy(k.g, &per_cu));
    } else {
      DD_HIP(rowsum_u8_batch_occupancy(k.g, k.flush, &per_cu));
    }
    if (per_cu < 1) return fail(DD_ERR_CONFIGURATION, "batch kernel cannot be resident on a compute unit");
    const uint32_t resident = (uint32_t)per_cu * (uint32_t)b->cus;
    k.workgroups = std::min(resident, roundup(k.n_tiles, kQueues));
  }
  return DD_OK;
}

dd_status build(dd_batch *b) {
  if (b->built) return DD_OK;
  std::vector<TileDesc> tiles_of[3];
  dd_status planned = plan_batch(b, tiles_of);
  if (planned != DD_OK) return planned;
  for (int m = 0; m < 3; ++m) {
    KindQueue &k = b->q[m];
    if (k.n_tiles == 0) continue;
    if (m == DD_STEP_TWO) {
      std::vector<RowSumF32Params> params;
      for (auto &L : k.f32) params.push_back(L.k);
      DD_HIP(k.d_f32.upload(params));
    } else {
      std::vector<RowSumU8Params> params;
      for (auto &L : k.u8) params.push_back(L.k);
      DD_HIP(k.d_u8.upload(params));
    }
    DD_HIP(k.d_tiles.upload(tiles_of[m]));
    DD_HIP(k.d_begin.upload(queue_bounds(k.n_tiles)));
  }
  if (b->heads.count == 0) DD_HIP(b->heads.allocate(3 * kQueues * kHeadStride));
  b->off1.clear();
  b->n1.clear();
  b->off2.clear();
  b->n2.clear();
  const size_t n_two = b->q[DD_STEP_ONE].u8.size();
  if (b->chunks >= 2 && n_two > 0) {
    // chunks of near-equal step-two tiles, jobs in order
    uint64_t total = 0;
    for (auto &L : b->q[DD_STEP_TWO].f32) total += L.k.n_tiles;
    std::vector<TileDesc> t1, t2;
    size_t j = 0;
    uint64_t done = 0;
    for (uint32_t c = 0; c < b->chunks && j < n_two; ++c) {
      const uint64_t goal = total * (c + 1) / b->chunks;
      b->off1.push_back((uint32_t)t1.size());
      b->off2.push_back((uint32_t)t2.size());
      do {  // at least one job per chunk; the last chunk takes the rest
        for (uint32_t t = 0; t < b->q[DD_STEP_ONE].u8[j].k.n_tiles; ++t) t1.push_back(TileDesc{(uint32_t)j, t});
        for (uint32_t t = 0; t < b->q[DD_STEP_TWO].f32[j].k.n_tiles; ++t) t2.push_back(TileDesc{(uint32_t)j, t});
        done += b->q[DD_STEP_TWO].f32[j].k.n_tiles;
        ++j;
      } while (j < n_two && (done < goal || c + 1 == b->chunks));
      b->n1.push_back((uint32_t)t1.size() - b->off1.back());
      b->n2.push_back((uint32_t)t2.size() - b->off2.back());
    }
    DD_HIP(b->d_list1.upload(t1));
    DD_HIP(b->d_list2.upload(t2));
    for (auto &st : b->side)
      if (!st) DD_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    const size_t need = 3 + b->n1.size();
    while (b->events.size() < need) {
      hipEvent_t e = nullptr;
      DD_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      b->events.push_back(e);
    }
  }
  b->built = true;
  return DD_OK;
}

// The compact step two's persistent and tile-list kernels: the paired form
// when the geometry says so.
hipError_t launch_u16_batch(const RowSumF32Params *d_jobs, const TileQueues &tq, const Geometry &g,
                            uint32_t workgroups, hipStream_t s) {
  return g.pairs ? launch_rowsum_u16_pairs_batch(d_jobs, tq, g, workgroups, s)
                 : launch_rowsum_u16_batch(d_jobs, tq, g, workgroups, s);
}
hipError_t launch_u16_list(const RowSumF32Params *d_jobs, const TileDesc *d_tiles, uint32_t n, const Geometry &g,
                           hipStream_t s) {
  return g.pairs ? launch_rowsum_u16_pairs_list(d_jobs, d_tiles, n, g, s)
                 : launch_rowsum_u16_list(d_jobs, d_tiles, n, g, s);
}

// The ordinary launches of kind m of the jobs that run beside the batch.
dd_status enqueue_solo(dd_batch *b, int m, hipStream_t s) {
  dd_status st = DD_OK;
  if (m == DD_SINGLE_STEP)
    for (size_t i = 0; i < b->solo_single.size() && st == DD_OK; ++i) st = run(b->solo_single[i], s);
  else if (m == DD_STEP_ONE)
    for (size_t i = 0; i < b->solo_one.size() && st == DD_OK; ++i) st = run(b->solo_one[i], s);
  else
    for (size_t i = 0; i < b->solo_two.size() && st == DD_OK; ++i) st = run(b->solo_two[i], s);
  return st;
}

// Launches kind m of the jobs beside the batch, then clears the claim
// counters of kind m and launches its persistent kernel.
dd_status enqueue_kind(dd_batch *b, int m, hipStream_t s) {
  dd_status st = enqueue_solo(b, m, s);
  if (st != DD_OK) return st;
  KindQueue &k = b->q[m];
  if (k.n_tiles == 0) return DD_OK;
  uint32_t *heads = b->heads.ptr + m * kQueues * kHeadStride;
  DD_HIP(launch_clear_words(heads, kQueues * kHeadStride, s));
  TileQueues tq{k.d_tiles.ptr, k.d_begin.ptr, heads};
  if (m == DD_STEP_TWO && b->compact) DD_HIP(launch_u16_batch(k.d_f32.ptr, tq, k.g, k.workgroups, s));
  else if (m == DD_STEP_TWO) DD_HIP(launch_rowsum_f32_batch(k.d_f32.ptr, tq, k.g, k.workgroups, s));
  else DD_HIP(launch_rowsum_u8_batch(k.d_u8.ptr, tq, k.g, k.flush, k.workgroups, s));
  return DD_OK;
}

// Step one of chunk c on side[0], step two of chunk c on side[1] after it;
// forked from and joined back into s with events.
dd_status enqueue_pipelined(dd_batch *b, hipStream_t s) {
  hipStream_t a = b->side[0], c2 = b->side[1];
  hipEvent_t fork = b->events[0], join_a = b->events[1], join_b = b->events[2];
  DD_HIP(hipEventRecord(fork, s));
  DD_HIP(hipStreamWaitEvent(a, fork, 0));
  DD_HIP(hipStreamWaitEvent(c2, fork, 0));
  dd_status st = enqueue_kind(b, DD_SINGLE_STEP, a);
  // step one of the jobs beside the batch goes first on the step-one stream,
  // so the step-two stream's first wait covers it; their step two goes last
  if (st == DD_OK) st = enqueue_solo(b, DD_STEP_ONE, a);
  if (st != DD_OK) return st;
  const KindQueue &k1 = b->q[DD_STEP_ONE], &k2 = b->q[DD_STEP_TWO];
  // the pipelined launches request the LDS reserves (room for the other kind)
  Geometry g1 = k1.g, g2 = k2.g;
  g1.lds_bytes = std::max(g1.lds_bytes, b->lds_reserve[DD_STEP_ONE]);
  g2.lds_bytes = std::max(g2.lds_bytes, b->lds_reserve[DD_STEP_TWO]);
  for (size_t c = 0; c < b->n1.size(); ++c) {
    if (b->n1[c])
      DD_HIP(launch_rowsum_u8_list(k1.d_u8.ptr, b->d_list1.ptr + b->off1[c], b->n1[c], g1, k1.flush, a));
    DD_HIP(hipEventRecord(b->events[3 + c], a));
    DD_HIP(hipStreamWaitEvent(c2, b->events[3 + c], 0));
    if (b->n2[c] && b->compact)
      DD_HIP(launch_u16_list(k2.d_f32.ptr, b->d_list2.ptr + b->off2[c], b->n2[c], g2, c2));
    else if (b->n2[c])
      DD_HIP(launch_rowsum_f32_list(k2.d_f32.ptr, b->d_list2.ptr + b->off2[c], b->n2[c], g2, c2));
  }
  st = enqueue_solo(b, DD_STEP_TWO, c2);
  if (st != DD_OK) return st;
  DD_HIP(hipEventRecord(join_a, a));
  DD_HIP(hipEventRecord(join_b, c2));
  DD_HIP(hipStreamWaitEvent(s, join_a, 0));
  DD_HIP(hipStreamWaitEvent(s, join_b, 0));
  return DD_OK;
}

dd_status enqueue(dd_batch *b, hipStream_t s) {
  if (!b->n1.empty()) return enqueue_pipelined(b, s);
  for (int m = 0; m < 3; ++m) {
    dd_status st = enqueue_kind(b, m, s);
    if (st != DD_OK) return st;
  }
  return DD_OK;
}

dd_status add_job(dd_batch *b, const Job &job) {
  if (job.plan->input_bits == 32)
    return fail(DD_ERR_UNSUPPORTED, "batches take 8-bit and packed input; launch float-input plans on their own");
  if (job.kind == kTwoStep && job.plan->obs.downsampling != 1)
    return fail(DD_ERR_UNSUPPORTED, "a batch chains step one into step two, whose intermediate pitches differ at "
                                    "downsampling > 1; launch the two steps of such a plan on their own");
  b->jobs.push_back(job);
  b->invalidate();
  return DD_OK;
}

} // namespace

namespace ddamd {

size_t batch_job_count(const dd_batch *b) { return b->jobs.size(); }

StaticBatchImage &batch_image(dd_batch *b) { return b->static_image; }

// Lays out the persistent launches of a batch -- claim counters, then per
// kernel kind its parameter blocks, tile descriptors and queue bounds, each
// at a 256-byte aligned offset -- as one host image for a device area of
// `capacity` bytes.  Jobs that would run beside the batch kernels are
// refused: their plans' own launches are not part of the image.
dd_status batch_static_image(dd_batch *b, uint64_t capacity, StaticBatchImage &img) {
  std::vector<TileDesc> tiles_of[3];
  dd_status st = plan_batch(b, tiles_of);
  if (st != DD_OK) return st;
  if (!b->solo_single.empty() || !b->solo_one.empty())
    return fail(DD_ERR_UNSUPPORTED, "a job's row windows do not fit the batch kernels; the allocation-free run "
                                    "takes only jobs that the batch kernels run");
  auto align = [](uint64_t v) { return (v + 255) & ~uint64_t(255); };
  uint64_t at = align(3ull * kQueues * kHeadStride * sizeof(uint32_t));
  img.heads = 0;
  for (int m = 0; m < 3; ++m) {
    const KindQueue &k = b->q[m];
    img.params[m] = img.tiles[m] = img.begin[m] = 0;
    if (k.n_tiles == 0) continue;
    const uint64_t pbytes = m == DD_STEP_TWO ? k.f32.size() * sizeof(RowSumF32Params)
                                             : k.u8.size() * sizeof(RowSumU8Params);
    img.params[m] = at;
    at = align(at + pbytes);
    img.tiles[m] = at;
    at = align(at + (uint64_t)k.n_tiles * sizeof(TileDesc));
    img.begin[m] = at;
    at = align(at + (kQueues + 1) * sizeof(uint32_t));
  }
  img.bytes = at;
  if (img.bytes > capacity)
    return fail(DD_ERR_UNSUPPORTED, "the batch's work queues need " + std::to_string(img.bytes) +
                                        " bytes; the allocation-free run holds at most " + std::to_string(capacity));
  img.host.assign(img.bytes, 0);
  for (int m = 0; m < 3; ++m) {
    const KindQueue &k = b->q[m];
    if (k.n_tiles == 0) continue;
    uint8_t *base = img.host.data();
    if (m == DD_STEP_TWO)
      for (size_t j = 0; j < k.f32.size(); ++j)
        std::memcpy(base + img.params[m] + j * sizeof(RowSumF32Params), &k.f32[j].k, sizeof(RowSumF32Params));
    else
      for (size_t j = 0; j < k.u8.size(); ++j)
        std::memcpy(base + img.params[m] + j * sizeof(RowSumU8Params), &k.u8[j].k, sizeof(RowSumU8Params));
    std::memcpy(base + img.tiles[m], tiles_of[m].data(), tiles_of[m].size() * sizeof(TileDesc));
    const std::vector<uint32_t> bounds = queue_bounds(k.n_tiles);
    std::memcpy(base + img.begin[m], bounds.data(), bounds.size() * sizeof(uint32_t));
  }
  return DD_OK;
}

// The persistent launches of the image at device address `dev` (copied
// there beforehand): per kernel kind, its claim counters cleared by a kernel,
// then its persistent kernel, on stream s.
dd_status batch_enqueue_static(dd_batch *b, uint8_t *dev, const StaticBatchImage &img, hipStream_t s) {
  for (int m = 0; m < 3; ++m) {
    const KindQueue &k = b->q[m];
    if (k.n_tiles == 0) continue;
    uint32_t *heads = reinterpret_cast<uint32_t *>(dev + img.heads) + m * kQueues * kHeadStride;
    DD_HIP(launch_clear_words(heads, kQueues * kHeadStride, s));
    TileQueues tq{reinterpret_cast<const TileDesc *>(dev + img.tiles[m]),
                  reinterpret_cast<const uint32_t *>(dev + img.begin[m]), heads};
    if (m == DD_STEP_TWO) {
      const auto *params = reinterpret_cast<const RowSumF32Params *>(dev + img.params[m]);
      if (b->compact) DD_HIP(launch_u16_batch(params, tq, k.g, k.workgroups, s));
      else DD_HIP(launch_rowsum_f32_batch(params, tq, k.g, k.workgroups, s));
    } else {
      const auto *params = reinterpret_cast<const RowSumU8Params *>(dev + img.params[m]);
      DD_HIP(launch_rowsum_u8_batch(params, tq, k.g, k.flush, k.workgroups, s));
    }
  }
  return DD_OK;
}

} // namespace ddamd

extern "C" {

dd_status dd_batch_create(dd_batch **batch) {
  if (!batch) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  dd_batch *b = new dd_batch();
  for (int m = 0; m < 3; ++m) b->conf[m] = batch_default_conf((dd_mode)m);
  *batch = b;
  return DD_OK;
}

dd_status dd_batch_destroy(dd_batch *batch) {
  delete batch;
  return DD_OK;
}

dd_status dd_batch_add_single_step(dd_batch *batch, dd_plan *plan, const uint8_t *d_input, float *d_output,
                                   uint32_t first_sbeam, uint32_t nr_sbeams) {
  if (!batch || !plan || !d_input || !d_output) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  return add_job(batch, Job{plan, kSingleStep, d_input, nullptr, d_output, first_sbeam, nr_sbeams});
}

dd_status dd_batch_add_two_step(dd_batch *batch, dd_plan *plan, const uint8_t *d_input, float *d_intermediate,
                                float *d_output, uint32_t first_sbeam, uint32_t nr_sbeams) {
  if (!batch || !plan || !d_input || !d_intermediate || !d_output)
    return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  if (!plan->two_step) return fail(DD_ERR_INVALID_ARGUMENT, "plan has no subband path");
  return add_job(batch, Job{plan, kTwoStep, d_input, d_intermediate, d_output, first_sbeam, nr_sbeams});
}

dd_status dd_batch_set_conf(dd_batch *batch, dd_mode mode, const dd_conf *conf) {
  if (!batch || !conf || (int)mode < 0 || (int)mode > 2) return fail(DD_ERR_INVALID_ARGUMENT, "invalid argument");
  batch->conf[mode] = *conf;
  batch->invalidate();
  return DD_OK;
}

dd_status dd_batch_get_conf(const dd_batch *batch, dd_mode mode, dd_conf *conf) {
  if (!batch || !conf || (int)mode < 0 || (int)mode > 2) return fail(DD_ERR_INVALID_ARGUMENT, "invalid argument");
  *conf = batch->conf[mode];
  return DD_OK;
}

dd_status dd_batch_run(dd_batch *batch, int use_graph, void *stream) {
  if (!batch) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  if (batch->jobs.empty()) return DD_OK;
  dd_status s = build(batch);
  if (s != DD_OK) return s;
  if (!use_graph) return enqueue(batch, (hipStream_t)stream);
  if (!batch->exec) {
    if (!batch->capture) DD_HIP(hipStreamCreateWithFlags(&batch->capture, hipStreamNonBlocking));
    DD_HIP(hipStreamBeginCapture(batch->capture, hipStreamCaptureModeThreadLocal));
    s = enqueue(batch, batch->capture);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(batch->capture, &g);
    if (s != DD_OK) {
      if (g) (void)hipGraphDestroy(g);
      return s;
    }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    batch->graph = g;
    e = hipGraphInstantiate(&batch->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      batch->drop_graph();
      return hip_fail(e, "hipGraphInstantiate");
    }
  }
  DD_HIP(hipGraphLaunch(batch->exec, (hipStream_t)stream));
  return DD_OK;
}

dd_status dd_batch_run_mode(dd_batch *batch, dd_mode mode, void *stream) {
  if (!batch || (int)mode < 0 || (int)mode > 2) return fail(DD_ERR_INVALID_ARGUMENT, "invalid argument");
  if (batch->jobs.empty()) return DD_OK;
  dd_status s = build(batch);
  return s != DD_OK ? s : enqueue_kind(batch, (int)mode, (hipStream_t)stream);
}

dd_status dd_batch_run_per_job(dd_batch *batch, void *stream) {
  if (!batch) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  dd_status s = build(batch);
  if (s != DD_OK) return s;
  size_t i_single = 0, i_two = 0;
  for (const Job &job : batch->jobs) {
    if (job.kind == kSingleStep) {
      s = run(batch->pj_single[i_single++], (hipStream_t)stream);
    } else {
      s = run(batch->pj_one[i_two], (hipStream_t)stream);
      if (s == DD_OK) s = run(batch->pj_two[i_two], (hipStream_t)stream);
      ++i_two;
    }
    if (s != DD_OK) return s;
  }
  return DD_OK;
}

dd_status dd_batch_set_pipeline(dd_batch *batch, uint32_t chunks) {
  if (!batch) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  if (chunks > 64) return fail(DD_ERR_INVALID_ARGUMENT, "at most 64 pipeline chunks");
  batch->chunks = chunks;
  batch->invalidate();
  return DD_OK;
}

dd_status dd_batch_set_lds_reserve(dd_batch *batch, dd_mode mode, uint32_t bytes) {
  if (!batch || (int)mode < 0 || (int)mode > 2) return fail(DD_ERR_INVALID_ARGUMENT, "invalid argument");
  batch->lds_reserve[mode] = bytes;
  batch->invalidate();
  return DD_OK;
}

dd_status dd_batch_set_compact_intermediate(dd_batch *batch, int enable) {
  if (!batch) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  batch->compact_requested = enable != 0;
  batch->invalidate();
  return DD_OK;
}

dd_status dd_batch_set_paired_step_two(dd_batch *batch, int enable) {
  if (!batch) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  batch->paired_requested = enable != 0;
  batch->invalidate();
  return DD_OK;
}

dd_status dd_batch_get_job_placement(dd_batch *batch, uint32_t job, uint32_t *beside, uint32_t *compact) {
  if (!batch || !beside || !compact) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  if (job >= batch->jobs.size()) return fail(DD_ERR_INVALID_ARGUMENT, "job index out of range");
  dd_status s = build(batch);
  if (s != DD_OK) return s;
  *beside = batch->beside[job] ? 1u : 0u;
  *compact = batch->compact_job[job] ? 1u : 0u;
  return DD_OK;
}

dd_status dd_batch_get_info(dd_batch *batch, dd_batch_info *info) {
  if (!batch || !info) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  *info = dd_batch_info{};
  info->jobs = (uint32_t)batch->jobs.size();
  if (batch->jobs.empty()) return DD_OK;
  dd_status s = build(batch);
  if (s != DD_OK) return s;
  for (int m = 0; m < 3; ++m) {
    info->tiles[m] = batch->q[m].n_tiles;
    info->workgroups[m] = batch->q[m].workgroups;
    info->lds_bytes[m] = batch->q[m].n_tiles ? batch->q[m].g.lds_bytes : 0;
  }
  info->units = batch->units;
  info->compact_intermediate = batch->compact ? 1 : 0;
  info->pipeline_chunks = (uint32_t)batch->n1.size();
  info->solo_jobs = (uint32_t)(batch->solo_single.size() + batch->solo_one.size());
  info->paired_step_two = batch->compact && batch->q[DD_STEP_TWO].n_tiles > 0 && batch->q[DD_STEP_TWO].g.pairs;
  return DD_OK;
}

} // extern "C"
