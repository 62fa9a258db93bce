// External CCD query batches on gfx950: every query's swept boxes, their
// overlap test, the candidate record and its id, with the batched proxy
// scene of GenerateCandidatesForExternalBatch
// (src/cpp/rt_candidate/external_batch_candidate.cpp:231-339).
//
// The reference packs query q as two swept-AABB proxies (object 1: the vertex
// or first edge, object 2: the face or second edge) in a slab of its own,
// [2q, 2q + 1], shifted along x so the slabs do not touch, and runs the
// all-pairs proxy generator over the whole batch.  Proxies of different slabs
// never pair, so the result is at most one candidate per query, present
// exactly when that query's two shifted boxes overlap (inclusive).  Here one
// lane takes one query: O(n) work in place of the O(n^2) pair loop, with the
// same FP64 boxes (the host supplies each query's x shift, a sequential sum
// it computes in the reference's order) and so the same candidate set.
// Candidates come out in ascending runtime query id (the order the caller
// passes), numbered runtime_id * 1000003 + ordinal + 1.
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "p2c_hip.h"
#include "runtime.h"
#include "sort_scan.h"

namespace p2c {
namespace {

constexpr int kExtBlock = 256;
constexpr uint64_t kCandidateIdStride = 1000003ull;

struct ExtBoxes {
  double lo[2][3], hi[2][3];  // [0] first side, [1] second side, after the x shift
};

// AabbFromFeatureIndices (:77-93) for both sides, then TranslateAabbX.
__device__ inline ExtBoxes ext_boxes(const p2c_external_query& q) {
  ExtBoxes b;
  const bool vf = q.family == 0;
  for (int s = 0; s < 2; ++s) {
    for (int a = 0; a < 3; ++a) {
      b.lo[s][a] = __builtin_inf();
      b.hi[s][a] = -__builtin_inf();
    }
    // vertex-face: {0} and {1, 2, 3}; edge-edge: {0, 1} and {2, 3}
    const int first = s == 0 ? 0 : (vf ? 1 : 2);
    const int last = s == 0 ? (vf ? 0 : 1) : 3;
    for (int v = first; v <= last; ++v)
      for (int t = 0; t < 2; ++t) {
        const double* p = t == 0 ? q.vertices_t0[v] : q.vertices_t1[v];
        for (int a = 0; a < 3; ++a) {
          b.lo[s][a] = p[a] < b.lo[s][a] ? p[a] : b.lo[s][a];  // std::min(lo, p)
          b.hi[s][a] = b.hi[s][a] < p[a] ? p[a] : b.hi[s][a];  // std::max(hi, p)
        }
      }
    b.lo[s][0] += q.x_offset;
    b.hi[s][0] += q.x_offset;
  }
  return b;
}

// ValidateAabbBounds of ValidateProxyScene (proxy_scene.cpp:31-42).
__device__ inline bool ext_bounds_valid(const ExtBoxes& b) {
  for (int s = 0; s < 2; ++s)
    for (int a = 0; a < 3; ++a)
      if (!__builtin_isfinite(b.lo[s][a]) || !__builtin_isfinite(b.hi[s][a]) || b.lo[s][a] > b.hi[s][a]) return false;
  return true;
}

// AabbOverlap (candidate_buffer.cpp:18-25), inclusive.
__device__ inline bool ext_overlap(const ExtBoxes& b) {
  for (int a = 0; a < 3; ++a)
    if (b.hi[0][a] < b.lo[1][a] || b.hi[1][a] < b.lo[0][a]) return false;
  return true;
}

// MergeMotionBound of the two proxies' MotionBoundForAabb (:95-113): the
// only non-zero entry is max(conservative radius) = 0.5 * max extent.
__device__ inline float ext_radius(const ExtBoxes& b) {
  double r[2];
  for (int s = 0; s < 2; ++s) {
    double m = b.hi[s][0] - b.lo[s][0];
    const double ey = b.hi[s][1] - b.lo[s][1], ez = b.hi[s][2] - b.lo[s][2];
    m = m < ey ? ey : m;
    m = m < ez ? ez : m;
    r[s] = 0.5 * m;
  }
  return static_cast<float>(r[0] < r[1] ? r[1] : r[0]);
}

// status words: [0] candidate count, [1] first query whose shifted boxes are
// invalid, [2] first query whose merged motion bound is not a finite float,
// [3] first candidate position whose id wraps to 0; UINT64_MAX when none
__global__ void ext_init_kernel(unsigned long long* status) {
  if (threadIdx.x < 4) status[threadIdx.x] = threadIdx.x == 0 ? 0ull : ~0ull;
}

__global__ __launch_bounds__(kExtBlock) void ext_flag_kernel(const p2c_external_query* __restrict__ q,
                                                             const uint32_t* __restrict__ order, int64_t n,
                                                             uint32_t* __restrict__ flags,
                                                             unsigned long long* __restrict__ status) {
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += int64_t(gridDim.x) * blockDim.x) {
    const uint32_t i = order[k];
    const ExtBoxes b = ext_boxes(q[i]);
    const bool valid = ext_bounds_valid(b);
    const bool hit = valid && ext_overlap(b);
    if (!valid) atomicMin(&status[1], (unsigned long long)i);
    if (hit) {
      const float r = ext_radius(b);
      if (!__builtin_isfinite(r) || r < 0.0f) atomicMin(&status[2], (unsigned long long)i);
    }
    flags[k] = hit ? 1u : 0u;
  }
}

__global__ __launch_bounds__(kExtBlock) void ext_write_kernel(const p2c_external_query* __restrict__ q,
                                                              const uint32_t* __restrict__ order, int64_t n,
                                                              const uint32_t* __restrict__ flags,
                                                              const uint64_t* __restrict__ offsets,
                                                              p2c_candidate* __restrict__ out, int64_t cap,
                                                              unsigned long long* __restrict__ status) {
  for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += int64_t(gridDim.x) * blockDim.x) {
    if (k == 0) status[0] = offsets[n];
    if (!flags[k]) continue;
    const uint32_t i = order[k];
    const p2c_external_query& e = q[i];
    const uint64_t ord = offsets[k];
    p2c_candidate c;
    memset(&c, 0, sizeof(c));
    c.schema_version = 1;
    c.candidate_id = e.runtime_query_id * kCandidateIdStride + ord + 1ull;
    if (c.candidate_id == 0) atomicMin(&status[3], (unsigned long long)ord);
    c.query_id = e.runtime_query_id;
    c.slab_id = i;
    c.object_a_id = e.object_a_id;
    c.patch_a_id = e.family == 0 ? 0u : 1u;
    c.object_b_id = e.object_b_id;
    c.patch_b_id = e.source_query_index + 1u;
    c.proxy_type_a = 1;  // ProxyType::kSweptAabb
    c.proxy_type_b = 1;
    c.rt_hit_count = 1;
    c.motion_bound[3] = ext_radius(ext_boxes(e));
    c.flags = P2C_RAW_VALID | P2C_RAW_AABB_OVERLAP;
    if ((int64_t)ord < cap) out[ord] = c;
  }
}

struct ExtLayout {
  size_t flags, offsets, part, total;
};

ExtLayout ext_layout(int64_t n) {
  auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
  ExtLayout L;
  L.flags = 0;
  L.offsets = up(L.flags + 4 * size_t(n > 0 ? n : 1));
  L.part = up(L.offsets + 8 * size_t(n + 1));
  L.total = up(L.part + 8 * size_t(n / 1024 + 2));
  return L;
}

struct DevMem {
  void* p = nullptr;
  ~DevMem() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace
}  // namespace p2c

extern "C" {

size_t p2c_external_batch_workspace_bytes(int64_t n) { return p2c::ext_layout(n < 0 ? 0 : n).total; }

int p2c_external_batch_candidates(const p2c_external_query* queries, const uint32_t* order, int64_t n,
                                  void* workspace, size_t workspace_bytes, p2c_candidate* out, int64_t cap,
                                  uint64_t* status, void* stream) {
  using namespace p2c;
  if (n < 0 || n > 0xFFFFFFFFll) return set_error(P2C_EINVAL, "external query count out of range");
  if (status == nullptr) return set_error(P2C_EINVAL, "status output is required");
  const ExtLayout L = ext_layout(n);
  if (workspace == nullptr || workspace_bytes < L.total)
    return set_error(P2C_ECAPACITY, "external batch workspace is too small");
  if (n > 0 && (queries == nullptr || order == nullptr))
    return set_error(P2C_EINVAL, "query and order buffers are required");
  hipStream_t s = as_stream(stream);
  auto* st = reinterpret_cast<unsigned long long*>(status);
  hipLaunchKernelGGL(ext_init_kernel, dim3(1), dim3(64), 0, s, st);
  P2C_HIP_CHECK(hipGetLastError());
  if (n == 0) return P2C_OK;
  auto* flags = reinterpret_cast<uint32_t*>(static_cast<char*>(workspace) + L.flags);
  auto* offsets = reinterpret_cast<uint64_t*>(static_cast<char*>(workspace) + L.offsets);
  auto* part = reinterpret_cast<uint64_t*>(static_cast<char*>(workspace) + L.part);
  const unsigned grid = grid_for(n, kExtBlock, 4096);
  hipLaunchKernelGGL(ext_flag_kernel, dim3(grid), dim3(kExtBlock), 0, s, queries, order, n, flags, st);
  if (int rc = exclusive_scan_u32(flags, n, offsets, part, s)) return rc;
  hipLaunchKernelGGL(ext_write_kernel, dim3(grid), dim3(kExtBlock), 0, s, queries, order, n, flags, offsets, out,
                     cap, st);
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

int p2c_external_batch_host(const p2c_external_query* queries, const uint32_t* order, int64_t n,
                            p2c_candidate** cand, int64_t* n_cand, uint64_t* status, double* trace_ms) {
  using namespace p2c;
  if (cand == nullptr || n_cand == nullptr || status == nullptr)
    return set_error(P2C_EINVAL, "output pointers are required");
  *cand = nullptr;
  *n_cand = 0;
  if (n < 0 || n > 0xFFFFFFFFll) return set_error(P2C_EINVAL, "external query count out of range");
  DevMem dq, dorder, dws, dstatus, dout;
  const size_t ws_bytes = p2c_external_batch_workspace_bytes(n);
  const size_t qn = size_t(n > 0 ? n : 1);
  P2C_HIP_CHECK(hipMalloc(&dq.p, sizeof(p2c_external_query) * qn));
  P2C_HIP_CHECK(hipMalloc(&dorder.p, 4 * qn));
  P2C_HIP_CHECK(hipMalloc(&dws.p, ws_bytes));
  P2C_HIP_CHECK(hipMalloc(&dstatus.p, 32));
  P2C_HIP_CHECK(hipMalloc(&dout.p, sizeof(p2c_candidate) * qn));
  hipEvent_t ev[2];
  for (auto& e : ev) P2C_HIP_CHECK(hipEventCreate(&e));
  struct Events {
    hipEvent_t* e;
    ~Events() {
      (void)hipEventDestroy(e[0]);
      (void)hipEventDestroy(e[1]);
    }
  } guard{ev};
  if (n > 0) {
    P2C_HIP_CHECK(hipMemcpy(dq.p, queries, sizeof(p2c_external_query) * size_t(n), hipMemcpyHostToDevice));
    P2C_HIP_CHECK(hipMemcpy(dorder.p, order, 4 * size_t(n), hipMemcpyHostToDevice));
  }
  P2C_HIP_CHECK(hipEventRecord(ev[0], nullptr));
  if (int rc = p2c_external_batch_candidates(static_cast<const p2c_external_query*>(dq.p),
                                             static_cast<const uint32_t*>(dorder.p), n, dws.p, ws_bytes,
                                             static_cast<p2c_candidate*>(dout.p), n,
                                             static_cast<uint64_t*>(dstatus.p), nullptr))
    return rc;
  P2C_HIP_CHECK(hipEventRecord(ev[1], nullptr));
  P2C_HIP_CHECK(hipMemcpy(status, dstatus.p, 32, hipMemcpyDeviceToHost));
  if (trace_ms != nullptr) {
    float ms = 0.0f;
    P2C_HIP_CHECK(hipEventElapsedTime(&ms, ev[0], ev[1]));
    *trace_ms = ms;
  }
  const int64_t c = int64_t(status[0]);
  if (c > 0) {
    auto* host = static_cast<p2c_candidate*>(std::malloc(sizeof(p2c_candidate) * size_t(c)));
    if (host == nullptr) return set_error(P2C_EINVAL, "out of host memory for candidates");
    const hipError_t e = hipMemcpy(host, dout.p, sizeof(p2c_candidate) * size_t(c), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      std::free(host);
      return hip_fail(e, "hipMemcpy candidates D2H");
    }
    *cand = host;
  }
  *n_cand = c;
  return P2C_OK;
}

}  // extern "C"
