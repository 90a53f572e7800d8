// batch_internal.hpp -- what the allocation-free batch run (batch_static.hip)
// needs from the batch implementation (batch.cpp).
#pragma once

#include <vector>

#include "plan_internal.hpp"

namespace ddamd {

// A batch's persistent launches laid out for one device area: byte offsets of
// the claim counters and, per kernel kind, of its parameter blocks, tile
// descriptors and queue bounds (0 for a kind without tiles), and the host
// copy of the whole area.
struct StaticBatchImage {
  uint64_t heads = 0;
  uint64_t params[3] = {0, 0, 0}, tiles[3] = {0, 0, 0}, begin[3] = {0, 0, 0};
  uint64_t bytes = 0;
  std::vector<uint8_t> host;
};

size_t batch_job_count(const dd_batch *b);
// The image a batch keeps of its last allocation-free run (the host copy
// stays valid until the next run of the batch or its destruction).
StaticBatchImage &batch_image(dd_batch *b);
// Builds the image (fails with DD_ERR_UNSUPPORTED beyond `capacity` bytes, or
// when a job would run beside the batch kernels); no device memory is used.
dd_status batch_static_image(dd_batch *b, uint64_t capacity, StaticBatchImage &img);
// Enqueues the persistent launches of an image already copied to `dev`.
dd_status batch_enqueue_static(dd_batch *b, uint8_t *dev, const StaticBatchImage &img, hipStream_t s);

} // namespace ddamd
