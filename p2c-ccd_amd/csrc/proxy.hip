// Proxy-scene raw candidates on gfx950 (replaces GenerateRawCandidatesCpu,
// src/cpp/rt_candidate/candidate_buffer.cpp:144-181, and the OptiX trace of
// optix_candidate_program.cu:96-130 with the CPU path's FP64 semantics):
// every pair i < j of proxy primitives with different objects, the same
// slab, strictly overlapping time intervals (:27-29) and inclusively
// overlapping AABBs (:18-25) is one raw hit, with pair_key (min << 32 | max)
// (:31-35), flags valid | aabb-overlap and the per-field maximum of the two
// motion bounds cast to float (:37-46).
//
// All pairs are tested (the reference's own complexity), one lane per i,
// with the j side staged through LDS in tiles of 256 proxies; count, then an
// exclusive scan on the host side, then fill at offsets[i] in ascending j,
// which reproduces the reference's (i, j) loop order exactly.
#include "p2c_hip.h"
#include "runtime.h"

namespace p2c {
namespace {

constexpr int kBlock = 256;

struct Tile {
  uint32_t object[kBlock];
  uint32_t slab[kBlock];
  double t0[kBlock];
  double t1[kBlock];
  double box[kBlock][6];
};

template <bool kFill>
__global__ __launch_bounds__(kBlock) void proxy_pairs_kernel(p2c_proxy_set s, uint32_t* __restrict__ counts,
                                                             const int64_t* __restrict__ offsets,
                                                             p2c_raw_hit* __restrict__ out, int64_t cap) {
  __shared__ Tile tile;
  const int64_t n = s.n;
  const int64_t first = int64_t(blockIdx.x) * kBlock;
  const int64_t i = first + threadIdx.x;
  const bool live = i < n;
  uint32_t oi = 0, si = 0;
  double t0i = 0, t1i = 0, bi[6] = {0, 0, 0, 0, 0, 0}, mi[4] = {0, 0, 0, 0};
  if (live) {
    oi = s.object_id[i];
    si = s.slab_id[i];
    t0i = s.t0[i];
    t1i = s.t1[i];
#pragma unroll
    for (int k = 0; k < 6; ++k) bi[k] = s.bounds[6 * i + k];
    if (kFill)
#pragma unroll
      for (int k = 0; k < 4; ++k) mi[k] = s.motion[4 * i + k];
  }
  uint32_t found = 0;
  const int64_t base = kFill && live ? offsets[i] : 0;
  for (int64_t t = first; t < n; t += kBlock) {
    __syncthreads();
    const int64_t j = t + threadIdx.x;
    if (j < n) {
      tile.object[threadIdx.x] = s.object_id[j];
      tile.slab[threadIdx.x] = s.slab_id[j];
      tile.t0[threadIdx.x] = s.t0[j];
      tile.t1[threadIdx.x] = s.t1[j];
#pragma unroll
      for (int k = 0; k < 6; ++k) tile.box[threadIdx.x][k] = s.bounds[6 * j + k];
    }
    __syncthreads();
    if (!live) continue;
    const int m = n - t < kBlock ? int(n - t) : kBlock;
    for (int q = 0; q < m; ++q) {
      const int64_t jj = t + q;
      if (jj <= i) continue;
      if (tile.object[q] == oi || tile.slab[q] != si) continue;
      if (!(t0i < tile.t1[q] && tile.t0[q] < t1i)) continue;
      const double* bj = tile.box[q];
      if (bi[3] < bj[0] || bj[3] < bi[0] || bi[4] < bj[1] || bj[4] < bi[1] || bi[5] < bj[2] || bj[5] < bi[2])
        continue;
      if (kFill) {
        const int64_t slot = base + found;
        if (slot < cap) {
          p2c_raw_hit h;
          h.query_id = s.query_id;
          h.pair_key = ((uint64_t)(uint32_t)i << 32) | (uint32_t)jj;
          h.proxy_a_index = (uint32_t)i;
          h.proxy_b_index = (uint32_t)jj;
          h.slab_id = si;
          h.rt_hit_count = 1;
          h.flags = P2C_RAW_VALID | P2C_RAW_AABB_OVERLAP;
          h.reserved0 = 0;
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const double mj = s.motion[4 * jj + k];
            h.motion_bound[k] = (float)(mi[k] < mj ? mj : mi[k]);
          }
          h.reserved1 = h.reserved2 = 0;
          out[slot] = h;
        }
      }
      ++found;
    }
  }
  if (!kFill && live) counts[i] = found;
}

int check_set(const p2c_proxy_set* s) {
  if (s == nullptr) return set_error(P2C_EINVAL, "proxy set is required");
  if (s->query_id == 0) return set_error(P2C_EINVAL, "query_id is required");
  if (s->n < 0 || s->n > 0xFFFFFFFFll) return set_error(P2C_EINVAL, "proxy count out of range");
  return P2C_OK;
}

}  // namespace
}  // namespace p2c

extern "C" {

int p2c_proxy_raw_count(const p2c_proxy_set* set, uint32_t* counts, void* stream) {
  if (int rc = p2c::check_set(set)) return rc;
  if (set->n == 0) return P2C_OK;
  hipLaunchKernelGGL(p2c::proxy_pairs_kernel<false>, dim3(p2c::grid_for(set->n, p2c::kBlock)), dim3(p2c::kBlock), 0,
                     p2c::as_stream(stream), *set, counts, nullptr, nullptr, 0);
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

int p2c_proxy_raw_fill(const p2c_proxy_set* set, const int64_t* offsets, p2c_raw_hit* out, int64_t cap,
                       void* stream) {
  if (int rc = p2c::check_set(set)) return rc;
  if (set->n == 0) return P2C_OK;
  hipLaunchKernelGGL(p2c::proxy_pairs_kernel<true>, dim3(p2c::grid_for(set->n, p2c::kBlock)), dim3(p2c::kBlock), 0,
                     p2c::as_stream(stream), *set, nullptr, offsets, out, cap);
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

}  // extern "C"
