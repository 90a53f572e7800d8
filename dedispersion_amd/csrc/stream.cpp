// stream.cpp -- continuous two-step dedispersion of chunked beams.
//
// Each beam delivers its samples in chunks of S = nr_samples_per_batch
// samples per channel ([C][S] uint8, host memory).  The two-step path needs
// H = S1_disp - S samples of look-back before each chunk (the dispersion
// delay plus the step-two margin), so the engine keeps, per beam, the last H
// samples of every channel on the device (the history).  The dispersed batch
// of chunk k of a beam is [history | chunk]; its output rows cover the beam's
// samples [k S - H, k S - H + S).  The first chunk of a beam sees a zero
// history.
//
// Requests are admitted into slots -- device input, intermediate and output
// buffers sized by the plan -- as long as the HBM budget allows; a request
// waits in its beam's FIFO until a slot is free and every earlier chunk of
// the beam has been admitted.  Each admitted request runs on one of several
// HIP streams: history -> slot, chunk -> slot (host to device), slot ->
// history, step one, step two, completion event.  Work of requests admitted
// while others are on the device overlaps with theirs; a beam's history
// copies are ordered across streams by an event per beam.
//
// The reference has no streaming mode (its test driver processes one batch,
// src/DedispersionTest.cpp:129-416); the kernels and layouts are those of
// dd_subband_step_one / dd_subband_step_two.
#include <hip/hip_runtime.h>

#include <chrono>
#include <deque>
#include <string>
#include <vector>

#include "plan_internal.hpp"

using namespace ddamd;

namespace {

struct Request {
  uint64_t tag;
  uint32_t beam;
  uint64_t chunk;  // index of the chunk within its beam
  const uint8_t *host;
  uint64_t host_pitch;
};

struct Slot {
  uint8_t *in = nullptr;
  float *mid = nullptr, *out = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;      // admitted, not yet released
  bool reported = false;  // completion returned to the caller
  Request req{};
  uint32_t stream = 0;
  double admitted_s = 0.0;
};

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

} // namespace

struct dd_stream {
  dd_plan *plan = nullptr;
  uint32_t nr_beams = 0;
  uint64_t chunk = 0, history = 0, batch = 0;  // samples: S, H, S + H
  uint64_t pitch = 0;                          // slot input row pitch (bytes)
  uint64_t hist_pitch = 0;
  uint64_t slot_bytes = 0;
  std::vector<Slot> slots;
  std::vector<hipStream_t> streams;
  uint32_t next_stream = 0;
  std::vector<uint8_t *> hist;        // per beam [C][hist_pitch]
  std::vector<hipEvent_t> hist_ready; // per beam, after its last history update
  std::vector<int> hist_stream;       // stream of that update, -1 = none
  std::vector<uint64_t> next_chunk;   // per beam, chunk index of the next submission
  std::deque<Request> queue;          // submitted, not admitted (FIFO)
  uint64_t submitted = 0, admitted = 0, completed = 0;

  ~dd_stream() {
    for (auto &s : slots) {
      (void)hipFree(s.in);
      (void)hipFree(s.mid);
      (void)hipFree(s.out);
      if (s.done) (void)hipEventDestroy(s.done);
    }
    for (auto *h : hist) (void)hipFree(h);
    for (auto e : hist_ready)
      if (e) (void)hipEventDestroy(e);
    for (auto s : streams)
      if (s) (void)hipStreamDestroy(s);
  }
};

namespace {

// Admits the oldest queued request that can run: a slot is free and no
// earlier chunk of its beam is still queued (the queue is FIFO, so the first
// queued request of each beam is its earliest).
dd_status admit(dd_stream *st, Request r, Slot &slot) {
  const dd_observation &o = st->plan->obs;
  const uint32_t C = o.nr_channels;
  const uint32_t si = st->next_stream;
  st->next_stream = (st->next_stream + 1) % (uint32_t)st->streams.size();
  hipStream_t s = st->streams[si];
  const uint32_t b = r.beam;
  if (st->hist_stream[b] >= 0 && (uint32_t)st->hist_stream[b] != si)
    DD_HIP(hipStreamWaitEvent(s, st->hist_ready[b], 0));
  // [history | chunk] -> slot input; the slot's last H samples -> history
  if (st->history > 0)
    DD_HIP(hipMemcpy2DAsync(slot.in, st->pitch, st->hist[b], st->hist_pitch, st->history, C, hipMemcpyDeviceToDevice,
                            s));
  DD_HIP(hipMemcpy2DAsync(slot.in + st->history, st->pitch, r.host, r.host_pitch, st->chunk, C,
                          hipMemcpyHostToDevice, s));
  if (st->history > 0)
    DD_HIP(hipMemcpy2DAsync(st->hist[b], st->hist_pitch, slot.in + st->chunk, st->pitch, st->history, C,
                            hipMemcpyDeviceToDevice, s));
  DD_HIP(hipEventRecord(st->hist_ready[b], s));
  st->hist_stream[b] = (int)si;
  dd_status ds = dd_subband_step_one(st->plan, slot.in, slot.mid, s);
  if (ds == DD_OK) ds = dd_subband_step_two(st->plan, slot.mid, slot.out, 0, 0, s);
  if (ds != DD_OK) return ds;
  DD_HIP(hipEventRecord(slot.done, s));
  slot.busy = true;
  slot.reported = false;
  slot.req = r;
  slot.stream = si;
  slot.admitted_s = now_s();
  st->admitted += 1;
  return DD_OK;
}

dd_status admit_ready(dd_stream *st) {
  while (!st->queue.empty()) {
    Slot *free_slot = nullptr;
    for (auto &s : st->slots)
      if (!s.busy) {
        free_slot = &s;
        break;
      }
    if (!free_slot) return DD_OK;
    const Request r = st->queue.front();
    st->queue.pop_front();
    dd_status ds = admit(st, r, *free_slot);
    if (ds != DD_OK) return ds;
  }
  return DD_OK;
}

} // namespace

extern "C" {

dd_status dd_stream_create(dd_stream **out, dd_plan *plan, uint32_t nr_beams, uint64_t hbm_budget_bytes,
                           uint32_t max_slots, uint32_t hip_streams) {
  if (!out || !plan) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (!plan->two_step) return fail(DD_ERR_INVALID_ARGUMENT, "streaming runs the subband path: plan has no coarse DMs");
  if (plan->input_bits != 8) return fail(DD_ERR_UNSUPPORTED, "streaming takes 8-bit input only");
  if (plan->obs.downsampling != 1)
    return fail(DD_ERR_UNSUPPORTED, "streaming chains step one into step two, whose intermediate pitches differ at "
                                    "downsampling > 1");
  const dd_observation &o = plan->obs;
  if (o.nr_beams != 1 || o.nr_synthesized_beams != 1)
    return fail(DD_ERR_INVALID_ARGUMENT, "a streaming plan describes one beam (nr_beams = nr_synthesized_beams = 1)");
  if (nr_beams == 0 || hip_streams == 0 || hip_streams > 16 || max_slots == 0)
    return fail(DD_ERR_INVALID_ARGUMENT, "nr_beams, max_slots and hip_streams (<= 16) must be positive");
  const dd_sizes &z = plan->sizes;
  dd_stream *st = new dd_stream();
  st->plan = plan;
  st->nr_beams = nr_beams;
  st->chunk = o.nr_samples_per_batch;
  st->batch = z.nr_dispersed_samples_subbanding;
  if (st->batch < st->chunk) {
    delete st;
    return fail(DD_ERR_INVALID_ARGUMENT, "dispersed batch shorter than the chunk");
  }
  st->history = st->batch - st->chunk;
  st->pitch = z.sub_in_pitch;
  st->hist_pitch = padto((uint32_t)std::max<uint64_t>(st->history, 1), 128);
  st->slot_bytes = z.sub_input_elements + 4 * (z.intermediate_elements + z.sub_output_elements);
  const uint64_t history_bytes = (uint64_t)nr_beams * o.nr_channels * st->hist_pitch;
  if (hbm_budget_bytes < history_bytes + st->slot_bytes) {
    delete st;
    return fail(DD_ERR_OUT_OF_MEMORY, "HBM budget smaller than the history plus one slot (" +
                                          std::to_string(history_bytes + st->slot_bytes) + " bytes)");
  }
  const uint64_t n_slots = std::min<uint64_t>(max_slots, (hbm_budget_bytes - history_bytes) / st->slot_bytes);
  auto cleanup = [&](dd_status s) {
    delete st;
    return s;
  };
  hipError_t e = hipSuccess;
  st->slots.resize(n_slots);
  for (auto &s : st->slots) {
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&s.in), z.sub_input_elements);
    if (e == hipSuccess) e = hipMemset(s.in, 0, z.sub_input_elements);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&s.mid), 4 * z.intermediate_elements);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&s.out), 4 * z.sub_output_elements);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
  }
  st->hist.assign(nr_beams, nullptr);
  st->hist_ready.assign(nr_beams, nullptr);
  st->hist_stream.assign(nr_beams, -1);
  st->next_chunk.assign(nr_beams, 0);
  for (uint32_t b = 0; b < nr_beams && e == hipSuccess; ++b) {
    e = hipMalloc(reinterpret_cast<void **>(&st->hist[b]), (size_t)o.nr_channels * st->hist_pitch);
    if (e == hipSuccess) e = hipMemset(st->hist[b], 0, (size_t)o.nr_channels * st->hist_pitch);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&st->hist_ready[b], hipEventDisableTiming);
  }
  st->streams.assign(hip_streams, nullptr);
  for (auto &s : st->streams)
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) return cleanup(hip_fail(e, "allocating streaming slots"));
  *out = st;
  return DD_OK;
}

dd_status dd_stream_destroy(dd_stream *st) {
  if (st) (void)hipDeviceSynchronize();
  delete st;
  return DD_OK;
}

dd_status dd_stream_get_info(const dd_stream *st, dd_stream_info *info) {
  if (!st || !info) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  info->slots = (uint32_t)st->slots.size();
  info->hip_streams = (uint32_t)st->streams.size();
  info->chunk_samples = st->chunk;
  info->history_samples = st->history;
  info->slot_bytes = st->slot_bytes;
  info->history_bytes = (uint64_t)st->nr_beams * st->plan->obs.nr_channels * st->hist_pitch;
  info->queued = (uint32_t)st->queue.size();
  uint32_t busy = 0;
  for (auto &s : st->slots) busy += s.busy ? 1 : 0;
  info->in_flight = busy;
  info->submitted = st->submitted;
  info->completed = st->completed;
  return DD_OK;
}

dd_status dd_stream_submit(dd_stream *st, uint32_t beam, const uint8_t *host_chunk, uint64_t host_pitch,
                           uint64_t tag) {
  if (!st || !host_chunk) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  if (beam >= st->nr_beams) return fail(DD_ERR_INVALID_ARGUMENT, "beam out of range");
  if (host_pitch < st->chunk) return fail(DD_ERR_INVALID_ARGUMENT, "host row pitch shorter than a chunk");
  st->queue.push_back(Request{tag, beam, st->next_chunk[beam]++, host_chunk, host_pitch});
  st->submitted += 1;
  return admit_ready(st);
}

dd_status dd_stream_progress(dd_stream *st, dd_stream_event *events, uint32_t max_events, uint32_t *n_events) {
  if (!st || !n_events || (max_events && !events)) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  *n_events = 0;
  dd_status ds = admit_ready(st);
  if (ds != DD_OK) return ds;
  for (uint32_t i = 0; i < st->slots.size() && *n_events < max_events; ++i) {
    Slot &s = st->slots[i];
    if (!s.busy || s.reported) continue;
    const hipError_t q = hipEventQuery(s.done);
    if (q == hipErrorNotReady) continue;
    if (q != hipSuccess) return hip_fail(q, "hipEventQuery");
    s.reported = true;
    st->completed += 1;
    dd_stream_event &ev = events[(*n_events)++];
    ev.tag = s.req.tag;
    ev.beam = s.req.beam;
    ev.chunk = s.req.chunk;
    ev.slot = i;
    ev.output = s.out;
    ev.intermediate = s.mid;
    ev.first_sample = (int64_t)(s.req.chunk * st->chunk) - (int64_t)st->history;
    ev.device_seconds = now_s() - s.admitted_s;
  }
  return DD_OK;
}

dd_status dd_stream_release(dd_stream *st, uint32_t slot) {
  if (!st || slot >= st->slots.size()) return fail(DD_ERR_INVALID_ARGUMENT, "slot out of range");
  Slot &s = st->slots[slot];
  if (!s.busy || !s.reported) return fail(DD_ERR_INVALID_ARGUMENT, "slot holds no completed request");
  s.busy = false;
  s.reported = false;
  return admit_ready(st);
}

dd_status dd_stream_drain(dd_stream *st) {
  if (!st) return fail(DD_ERR_INVALID_ARGUMENT, "null argument");
  for (auto s : st->streams) DD_HIP(hipStreamSynchronize(s));
  return DD_OK;
}

} // extern "C"
