// Host-side runtime helpers shared by the library's translation units:
// thread-local last-error text, HIP status checks, launch sizing.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "p2c_hip.h"

namespace p2c {

int set_error(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Copies a small host table (a batch's launch plan and scene descriptors) to
// device memory `dev` on stream s without waiting for the stream: the bytes
// pass through a pinned staging ring of the calling thread (one per device,
// held until the process ends), and a stretch of the ring is written again
// only after the copy that read it has completed.
int upload_table(const void* host, size_t bytes, void* dev, hipStream_t s);

// Grid for a grid-stride kernel over n items.
inline unsigned grid_for(int64_t n, int block, int64_t max_blocks = int64_t(1) << 20) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > max_blocks) g = max_blocks;
  return static_cast<unsigned>(g);
}

// Workgroups are dispatched to the 8 XCDs of an MI355X round-robin (block i
// to XCD i % 8).  This maps block i to a logical block so that each XCD gets
// one contiguous run of logical blocks: kernels whose neighbouring blocks
// touch neighbouring data (Morton-sorted queries) then keep each XCD's L2 on
// one region.  A bijection of [0, grid) for any grid size.
constexpr unsigned kXcds = 8;

__device__ inline unsigned xcd_block(unsigned i, unsigned grid) {
  const unsigned per = grid / kXcds, rem = grid % kXcds, x = i % kXcds;
  return x * per + (x < rem ? x : rem) + i / kXcds;
}

}  // namespace p2c

#define P2C_HIP_CHECK(expr)                         \
  do {                                              \
    hipError_t _e = (expr);                         \
    if (_e != hipSuccess) return p2c::hip_fail(_e, #expr); \
  } while (0)
