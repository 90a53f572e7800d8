// staged_common.hpp -- device helpers shared by the pipelined (staged) kernels.
#pragma once

#include "internal.hpp"

#include <type_traits>

namespace ddamd {
namespace staged {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x3 __attribute__((ext_vector_type(3)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Wave-uniform copies of a value (readfirstlane).  The builtin returns a
// signed int; results are taken as uint32_t before any widening.
__device__ __forceinline__ uint32_t uniform(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  return (uint64_t)uniform((uint32_t)v) | ((uint64_t)uniform((uint32_t)(v >> 32)) << 32);
}
__device__ __forceinline__ float uniformf(float v) {
  return __builtin_bit_cast(float, uniform(__builtin_bit_cast(uint32_t, v)));
}

// A raw buffer resource.  Both fields go through readfirstlane so that the
// resource is built in SGPRs even when its inputs were computed on the VALU
// (a resource in VGPRs makes the backend wrap every load in a waterfall loop).
// Each address half is taken as uint32_t before it is widened, or the low
// half would sign-extend into the high address bits.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *base, uint32_t bytes) {
  const uint64_t addr = reinterpret_cast<uintptr_t>(base);
  const uint64_t b = (uint64_t)uniform((uint32_t)addr) | ((uint64_t)uniform((uint32_t)(addr >> 32)) << 32);
  return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void *>(b), (short)0, (int)uniform(bytes), 0x00020000);
}

// Bytes addressable from `offset` in a buffer of `total` bytes, rounded up to a
// whole dword (the dword holding the last byte is safe to read), clamped to the
// 32-bit range of a buffer resource.
__device__ __forceinline__ uint32_t records_from(uint64_t total, uint64_t offset) {
  uint64_t end = (total + 3) & ~uint64_t(3);
  if (offset >= end) return 0;
  uint64_t n = end - offset;
  return n > 0xFFFFFFF0ull ? 0xFFFFFFF0u : (uint32_t)n;
}

// Bytes addressable from float offset `offset` in a buffer of `total` floats,
// clamped to the 32-bit range of a buffer resource (integer-only, so that a
// wave-uniform offset keeps the resource in SGPRs).
__device__ __forceinline__ uint32_t records_from_floats(uint64_t total, uint64_t offset) {
  if (offset >= total) return 0;
  const uint64_t n = total - offset;
  return n >= 0x3FFFFFFCull ? 0xFFFFFFF0u : (uint32_t)(n * 4);
}

// The integer shift of include/Dedispersion.hpp:81/124/168, separately rounded:
// (unsigned)((firstDM + dm * DMStep) * table).  Same arithmetic as tables.hip.
__device__ __forceinline__ uint32_t shift_of(float first, float step, uint32_t dm, float table) {
  return (uint32_t)__fmul_rn(__fadd_rn(first, __fmul_rn((float)dm, step)), table);
}

// Tiles.  A launch of n tiles has nx sample tiles and ny DM tiles per
// (beam x group) slice; the logical tile index runs DM tile fastest, then
// sample tile, then slice.
struct Tile {
  uint32_t x, y, z;
};

__device__ __forceinline__ Tile tile_split(uint32_t logical, uint32_t nx, uint32_t ny) {
  Tile t;
  t.y = uniform(logical % ny);
  t.x = uniform((logical / ny) % nx);
  t.z = uniform(logical / (ny * nx));
  return t;
}

// Logical tile of workgroup `id` in a 1-D grid of n tiles.  Workgroup ids are
// placed on the 8 XCDs round robin (XCD = id % 8); logical tiles are dealt to
// XCDs in contiguous chunks, so the DM tiles of one (beam, sample tile) --
// which read the same input rows -- run on one XCD and share its L2.
__device__ __forceinline__ uint32_t xcd_logical(uint32_t id, uint32_t n) {
  const uint32_t per = n / 8, rem = n % 8, xcd = id % 8, k = id / 8;
  return uniform(xcd * per + min(xcd, rem) + k);
}

// Claims on the work queues of a persistent launch (TileQueues).
struct QueueCursor {
  uint32_t probe = 0;  // queues found empty so far (thread 0 only)

  // The next tile of this workgroup, or kNoTile.  Called by every thread of
  // the workgroup; one barrier.  `slot` is a workgroup-shared word; the
  // caller's tile work must contain a barrier between two calls.
  __device__ __forceinline__ uint32_t next(const TileQueues &q, uint32_t *slot) {
    if (threadIdx.x == 0) {
      const uint32_t home = blockIdx.x % kQueues;
      uint32_t t = kNoTile;
      while (probe < kQueues) {
        const uint32_t qi = (home + probe) % kQueues;
        const uint32_t b = q.begin[qi], n = q.begin[qi + 1] - b;
        if (n != 0) {
          const uint32_t k = atomicAdd(q.heads + qi * kHeadStride, 1u);
          if (k < n) {
            t = b + k;
            break;
          }
        }
        ++probe;
      }
      *slot = t;
    }
    __syncthreads();
    return uniform(*slot);
  }
};

} // namespace staged
} // namespace ddamd
