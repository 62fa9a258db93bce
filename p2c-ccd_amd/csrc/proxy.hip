// Proxy-scene raw candidates on gfx950 (replaces GenerateRawCandidatesCpu,
// src/cpp/rt_candidate/candidate_buffer.cpp:144-181, and the OptiX trace of
// optix_candidate_program.cu:96-130 with the CPU path's FP64 semantics):
// every pair i < j of proxy primitives with different objects, the same
// slab, strictly overlapping time intervals (:27-29) and inclusively
// overlapping AABBs (:18-25) is one raw hit, with pair_key (min << 32 | max)
// (:31-35), flags valid | aabb-overlap and the per-field maximum of the two
// motion bounds cast to float (:37-46).
//
// All pairs are tested (the reference's own complexity): one lane per row
// i, the wave's 64 rows against one segment of the columns j, each column
// read once per wave with a scalar load; count, an exclusive scan, then fill
// at offsets[i] in ascending j, which reproduces the reference's (i, j) loop
// order exactly.
//
// Compaction (CompactRawCandidates, candidate_buffer.cpp:183-228) on the
// device: each raw hit is validated (ValidateRawHit, :66-102) and keyed
// (MakeCandidateKey, :104-124: first = the smaller (object, patch, proxy
// id)); a stable LSD radix sort over the key fields, least significant
// first, puts the hits in std::map key order with equal keys in buffer
// order; segment starts are flagged and scanned into candidate ids; one lane
// per candidate then merges its hits in buffer order (rt_hit_count sum with
// the overflow check, std::max of the motion bounds, OR of the flags), so
// ties, signed zeros and the overflow point follow the reference exactly.
// The first failing hit in buffer order (validation or overflow) is
// reported through an atomic minimum over (hit index, code).
#include "p2c_hip.h"
#include "runtime.h"
#include "sort_scan.h"

#include <algorithm>
#include <cstdlib>

namespace p2c {
namespace {


constexpr int kRowWave = 64;            // rows i per wave, one per lane
constexpr int64_t kMinSegment = 1024;   // columns j per segment, at least
constexpr int kMaxSegments = 64;        // segments per row, at most

// Columns are cut into at most 64 segments so that a launch has about
// (n / 64) x (segments) / 2 live waves instead of n / 64 waves whose work
// shrinks along the triangle.
struct SegPlan {
  int64_t len;
  int count;
};

inline SegPlan seg_plan(int64_t n) {
  int64_t len = (n + kMaxSegments - 1) / kMaxSegments;
  if (len < kMinSegment) len = kMinSegment;
  int count = int((n + len - 1) / len);
  if (count < 1) count = 1;
  return {len, count};
}

// Raw hits of 64 rows i (one per lane) against the columns j > i of one
// segment.  Columns are uniform across the wave (scalar loads, one per column
// for all 64 rows), and each lane meets its row's columns in ascending j, so
// its hits come out in the reference's buffer order (candidate_buffer.cpp).
// The count pass stores each (segment, row) count in `seg`; after
// seg_prefix_kernel has turned them into per-row exclusive prefixes, the fill
// pass writes row i's hits of segment g from offsets[i] + seg[g][i],
// dropping slots at or beyond cap.
template <bool kFill>
__global__ __launch_bounds__(kRowWave) void proxy_rows_kernel(p2c_proxy_set s, int64_t seg_len,
                                                              uint32_t* __restrict__ seg,
                                                              const int64_t* __restrict__ offsets,
                                                              p2c_raw_hit* __restrict__ out, int64_t cap) {
  const int64_t n = s.n;
  const int64_t first = int64_t(blockIdx.x) * kRowWave;
  const int64_t i = first + threadIdx.x;
  const int64_t g = blockIdx.y;
  const bool row = i < n;
  const int64_t js = g * seg_len;
  const int64_t je = js + seg_len < n ? js + seg_len : n;
  uint32_t oi = 0, si = 0;
  double t0i = 0.0, t1i = 0.0, bi[6] = {0, 0, 0, 0, 0, 0};
  if (row) {
    oi = s.object_id[i];
    si = s.slab_id[i];
    t0i = s.t0[i];
    t1i = s.t1[i];
#pragma unroll
    for (int k = 0; k < 6; ++k) bi[k] = s.bounds[6 * i + k];
  }
  int64_t base = 0;
  if (kFill && row) base = offsets[i] + seg[g * n + i];
  uint32_t found = 0;
  const int64_t j_begin = js > first + 1 ? js : first + 1;
#pragma unroll 4
  for (int64_t j = j_begin; j < je; ++j) {
    const uint32_t oj = s.object_id[j];
    const uint32_t sj = s.slab_id[j];
    const double t0j = s.t0[j], t1j = s.t1[j];
    const double* bj = s.bounds + 6 * j;
    // all tests evaluated (bitwise, no short circuit): straight-line code
    const bool apart = (bi[3] < bj[0]) | (bj[3] < bi[0]) | (bi[4] < bj[1]) | (bj[4] < bi[1]) | (bi[5] < bj[2]) |
                       (bj[5] < bi[2]);
    const bool hit = row & (j > i) & (oj != oi) & (sj == si) & (t0i < t1j) & (t0j < t1i) & !apart;
    if (hit) {
      if (kFill) {
        const int64_t slot = base + found;
        if (slot < cap) {
          p2c_raw_hit h;
          h.query_id = s.query_id;
          h.pair_key = ((uint64_t)(uint32_t)i << 32) | (uint32_t)j;
          h.proxy_a_index = (uint32_t)i;
          h.proxy_b_index = (uint32_t)j;
          h.slab_id = si;
          h.rt_hit_count = 1;
          h.flags = P2C_RAW_VALID | P2C_RAW_AABB_OVERLAP;
          h.reserved0 = 0;
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const double mi = s.motion[4 * i + k];
            const double mj = s.motion[4 * j + k];
            h.motion_bound[k] = (float)(mi < mj ? mj : mi);
          }
          h.reserved1 = h.reserved2 = 0;
          out[slot] = h;
        }
      }
      ++found;
    }
  }
  if (!kFill && row) seg[g * n + i] = found;
}

// Per row: the segment counts become exclusive prefixes (in place) and their
// sum is the row's count.
__global__ __launch_bounds__(256) void seg_prefix_kernel(uint32_t* __restrict__ seg, int64_t n, int nseg,
                                                         uint32_t* __restrict__ counts) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    uint32_t run = 0;
    for (int g = 0; g < nseg; ++g) {
      const uint32_t c = seg[int64_t(g) * n + i];
      seg[int64_t(g) * n + i] = run;
      run += c;
    }
    counts[i] = run;
  }
}

// ---- compaction ----------------------------------------------------------
enum : uint32_t {
  kErrNotValid = 1,
  kErrQueryMissing,
  kErrQueryMismatch,
  kErrIndexRange,
  kErrSameProxy,
  kErrPairKey,
  kErrSameObject,
  kErrSlab,
  kErrMotion,
  kErrOverflow,
};

struct CompactLayout {
  size_t f_slab, f_oa, f_pa, f_ob, f_pb, f_ty, k0, k1, v0, v1, flag, segoff, hist, hist_off, part, err, total;
  int64_t tiles;
};

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

CompactLayout compact_layout(int64_t m) {
  CompactLayout L{};
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o = align256(o + (bytes ? bytes : 1));
    return at;
  };
  const size_t mm = size_t(m > 0 ? m : 1);
  L.tiles = (int64_t(mm) + kSortTile - 1) / kSortTile;
  L.f_slab = take(4 * mm);
  L.f_oa = take(4 * mm);
  L.f_pa = take(4 * mm);
  L.f_ob = take(4 * mm);
  L.f_pb = take(4 * mm);
  L.f_ty = take(4 * mm);
  L.k0 = take(4 * mm);
  L.k1 = take(4 * mm);
  L.v0 = take(4 * mm);
  L.v1 = take(4 * mm);
  L.flag = take(4 * mm);
  L.segoff = take(8 * (mm + 1));
  L.hist = take(4 * 256 * size_t(L.tiles));
  L.hist_off = take(8 * 256 * size_t(L.tiles));
  const size_t scan_items = std::max<size_t>(mm + 1, 256 * size_t(L.tiles));
  L.part = take(8 * (scan_items / 1024 + 2));
  L.err = take(16);
  L.total = o;
  return L;
}

template <typename T>
T* at(void* ws, size_t off) {
  return reinterpret_cast<T*>(static_cast<char*>(ws) + off);
}

__device__ inline void report(unsigned long long* err, int64_t h, uint32_t code) {
  atomicMin(err, ((unsigned long long)h << 4) | code);
}

// Validation in ValidateRawHit's order, then the key fields (zero for a
// failing hit, which never reaches the output: the error aborts the call).
__global__ __launch_bounds__(kScanBlock) void compact_key_kernel(p2c_proxy_set s, const p2c_raw_hit* __restrict__ raw,
                                                                 int64_t m, uint32_t* __restrict__ f_slab,
                                                                 uint32_t* __restrict__ f_oa, uint32_t* __restrict__ f_pa,
                                                                 uint32_t* __restrict__ f_ob, uint32_t* __restrict__ f_pb,
                                                                 uint32_t* __restrict__ f_ty,
                                                                 uint32_t* __restrict__ perm,
                                                                 unsigned long long* __restrict__ err) {
  for (int64_t h = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; h < m; h += int64_t(gridDim.x) * blockDim.x) {
    const p2c_raw_hit r = raw[h];
    perm[h] = uint32_t(h);
    uint32_t code = 0;
    const uint32_t a = r.proxy_a_index, b = r.proxy_b_index;
    const bool in_range = int64_t(a) < s.n && int64_t(b) < s.n;
    if ((r.flags & P2C_RAW_VALID) == 0u) code = kErrNotValid;
    else if (r.query_id == 0) code = kErrQueryMissing;
    else if (r.query_id != s.query_id) code = kErrQueryMismatch;
    else if (!in_range) code = kErrIndexRange;
    else if (a == b) code = kErrSameProxy;
    else if (r.pair_key != ((uint64_t(a < b ? a : b) << 32) | uint64_t(a < b ? b : a))) code = kErrPairKey;
    else if (s.object_id[a] == s.object_id[b]) code = kErrSameObject;
    else if (s.slab_id[a] != s.slab_id[b] || r.slab_id != s.slab_id[a]) code = kErrSlab;
    else {
      for (int k = 0; k < 4; ++k) {
        const float v = r.motion_bound[k];
        if (!__builtin_isfinite(v) || v < 0.0f) code = kErrMotion;
      }
    }
    uint32_t slab = 0, oa = 0, pa = 0, ob = 0, pb = 0, ty = 0;
    if (code != 0) {
      report(err, h, code);
    } else {
      const uint32_t o_a = s.object_id[a], o_b = s.object_id[b], p_a = s.patch_id[a], p_b = s.patch_id[b];
      const bool swap = o_a != o_b ? o_a > o_b : (p_a != p_b ? p_a > p_b : a > b);
      const uint32_t first = swap ? b : a, second = swap ? a : b;
      slab = r.slab_id;
      oa = s.object_id[first];
      pa = s.patch_id[first];
      ob = s.object_id[second];
      pb = s.patch_id[second];
      ty = ((s.proxy_type[first] & 0xffu) << 8) | (s.proxy_type[second] & 0xffu);
    }
    f_slab[h] = slab;
    f_oa[h] = oa;
    f_pa[h] = pa;
    f_ob[h] = ob;
    f_pb[h] = pb;
    f_ty[h] = ty;
  }
}

__global__ __launch_bounds__(kScanBlock) void gather_u32_kernel(const uint32_t* __restrict__ field,
                                                                const uint32_t* __restrict__ perm, int64_t m,
                                                                uint32_t* __restrict__ out) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += int64_t(gridDim.x) * blockDim.x)
    out[i] = field[perm[i]];
}

struct KeyFields {
  const uint32_t *slab, *oa, *pa, *ob, *pb, *ty;
};

__global__ __launch_bounds__(kScanBlock) void segment_flag_kernel(KeyFields f, const uint32_t* __restrict__ perm,
                                                                  int64_t m, uint32_t* __restrict__ flag) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += int64_t(gridDim.x) * blockDim.x) {
    uint32_t start = 1;
    if (i > 0) {
      const uint32_t x = perm[i], y = perm[i - 1];
      start = (f.slab[x] != f.slab[y] || f.oa[x] != f.oa[y] || f.pa[x] != f.pa[y] || f.ob[x] != f.ob[y] ||
               f.pb[x] != f.pb[y] || f.ty[x] != f.ty[y])
                  ? 1u
                  : 0u;
    }
    flag[i] = start;
  }
}

// One lane per candidate: merges its hits in buffer order.
__global__ __launch_bounds__(kScanBlock) void merge_kernel(KeyFields f, const p2c_raw_hit* __restrict__ raw,
                                                           const uint32_t* __restrict__ perm,
                                                           const uint32_t* __restrict__ flag,
                                                           const uint64_t* __restrict__ segoff, int64_t m,
                                                           uint64_t query_id, p2c_candidate* __restrict__ out,
                                                           int64_t cap, unsigned long long* __restrict__ err) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += int64_t(gridDim.x) * blockDim.x) {
    if (flag[i] == 0) continue;
    const uint64_t seg = segoff[i];
    const uint32_t h0 = perm[i];
    const p2c_raw_hit& first = raw[h0];
    uint64_t count = first.rt_hit_count;
    float mb[4] = {first.motion_bound[0], first.motion_bound[1], first.motion_bound[2], first.motion_bound[3]};
    uint32_t flags = first.flags;
    for (int64_t j = i + 1; j < m && flag[j] == 0; ++j) {
      const uint32_t h = perm[j];
      const p2c_raw_hit& r = raw[h];
      if (count > 0xFFFFFFFFull - r.rt_hit_count) {
        report(err, h, kErrOverflow);
        break;
      }
      count += r.rt_hit_count;
#pragma unroll
      for (int k = 0; k < 4; ++k) mb[k] = mb[k] < r.motion_bound[k] ? r.motion_bound[k] : mb[k];
      flags |= r.flags;
    }
    if (int64_t(seg) >= cap) continue;
    p2c_candidate c;
    c.schema_version = 1;
    c.reserved0 = 0;
    c.candidate_id = seg + 1;
    c.query_id = query_id;
    c.slab_id = f.slab[h0];
    c.object_a_id = f.oa[h0];
    c.patch_a_id = f.pa[h0];
    c.object_b_id = f.ob[h0];
    c.patch_b_id = f.pb[h0];
    c.proxy_type_a = uint8_t(f.ty[h0] >> 8);
    c.proxy_type_b = uint8_t(f.ty[h0] & 0xffu);
    c.reserved1 = 0;
    c.rt_hit_count = uint32_t(count);
#pragma unroll
    for (int k = 0; k < 4; ++k) c.motion_bound[k] = mb[k];
    c.proxy_features_offset = 0;
    c.flags = flags;
    c.reserved2 = 0;
    out[seg] = c;
  }
}

__global__ void compact_init_kernel(unsigned long long* err) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *err = ~0ull;
}

__global__ void compact_status_kernel(const uint64_t* __restrict__ segoff, int64_t m,
                                      const unsigned long long* __restrict__ err, uint64_t* __restrict__ status) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    status[0] = m > 0 ? segoff[m] : 0;
    status[1] = *err == ~0ull ? 0ull : uint64_t(*err) + 1ull;
  }
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

size_t p2c_proxy_raw_workspace_bytes(int64_t n) {
  if (n <= 0) return 256;
  const p2c::SegPlan plan = p2c::seg_plan(n);
  return size_t(plan.count) * size_t(n) * 4;
}

int p2c_proxy_raw_count(const p2c_proxy_set* set, uint32_t* counts, void* workspace, size_t workspace_bytes,
                        void* stream) {
  if (int rc = p2c::check_set(set)) return rc;
  const int64_t n = set->n;
  if (n == 0) return P2C_OK;
  if (counts == nullptr || workspace == nullptr) return p2c::set_error(P2C_EINVAL, "counts and workspace are required");
  if (workspace_bytes < p2c_proxy_raw_workspace_bytes(n))
    return p2c::set_error(P2C_EINVAL, "raw-hit workspace is smaller than p2c_proxy_raw_workspace_bytes(n)");
  const p2c::SegPlan plan = p2c::seg_plan(n);
  const unsigned rows = unsigned((n + p2c::kRowWave - 1) / p2c::kRowWave);
  uint32_t* seg = static_cast<uint32_t*>(workspace);
  hipStream_t s = p2c::as_stream(stream);
  hipLaunchKernelGGL(p2c::proxy_rows_kernel<false>, dim3(rows, unsigned(plan.count)), dim3(p2c::kRowWave), 0, s,
                     *set, plan.len, seg, nullptr, nullptr, 0);
  hipLaunchKernelGGL(p2c::seg_prefix_kernel, dim3(p2c::grid_for(n, 256)), dim3(256), 0, s, seg, n, plan.count,
                     counts);
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

int p2c_proxy_raw_fill(const p2c_proxy_set* set, const int64_t* offsets, const void* workspace,
                       size_t workspace_bytes, p2c_raw_hit* out, int64_t cap, void* stream) {
  if (int rc = p2c::check_set(set)) return rc;
  const int64_t n = set->n;
  if (n == 0) return P2C_OK;
  if (offsets == nullptr || workspace == nullptr) return p2c::set_error(P2C_EINVAL, "offsets and workspace are required");
  if (workspace_bytes < p2c_proxy_raw_workspace_bytes(n))
    return p2c::set_error(P2C_EINVAL, "raw-hit workspace is smaller than p2c_proxy_raw_workspace_bytes(n)");
  if (cap > 0 && out == nullptr) return p2c::set_error(P2C_EINVAL, "raw-hit output is required");
  const p2c::SegPlan plan = p2c::seg_plan(n);
  const unsigned rows = unsigned((n + p2c::kRowWave - 1) / p2c::kRowWave);
  hipLaunchKernelGGL(p2c::proxy_rows_kernel<true>, dim3(rows, unsigned(plan.count)), dim3(p2c::kRowWave), 0,
                     p2c::as_stream(stream), *set, plan.len,
                     const_cast<uint32_t*>(static_cast<const uint32_t*>(workspace)), offsets, out, cap);
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

size_t p2c_proxy_compact_workspace_bytes(int64_t n_raw) { return p2c::compact_layout(n_raw).total; }

int p2c_exclusive_scan_u32(const uint32_t* in, int64_t n, uint64_t* out, uint64_t* part, void* stream) {
  if (n < 0) return p2c::set_error(P2C_EINVAL, "scan length must be non-negative");
  if (n == 0) return P2C_OK;
  return p2c::exclusive_scan_u32(in, n, out, part, p2c::as_stream(stream));
}

int p2c_proxy_compact(const p2c_proxy_set* set, const p2c_raw_hit* raw, int64_t n_raw, void* workspace,
                      size_t workspace_bytes, p2c_candidate* out, int64_t cap, uint64_t* status, void* stream) {
  using namespace p2c;
  if (int rc = check_set(set)) return rc;
  if (n_raw < 0 || n_raw > 0xFFFFFFFFll) return set_error(P2C_EINVAL, "raw hit count out of range");
  const CompactLayout L = compact_layout(n_raw);
  if (workspace == nullptr || workspace_bytes < L.total)
    return set_error(P2C_ECAPACITY, "compaction workspace is too small");
  if (status == nullptr) return set_error(P2C_EINVAL, "status output is required");
  hipStream_t s = as_stream(stream);
  const int64_t m = n_raw;
  auto* err = at<unsigned long long>(workspace, L.err);
  auto* segoff = at<uint64_t>(workspace, L.segoff);
  hipLaunchKernelGGL(compact_init_kernel, dim3(1), dim3(64), 0, s, err);
  if (m > 0) {
    uint32_t* field[6] = {at<uint32_t>(workspace, L.f_slab), at<uint32_t>(workspace, L.f_oa),
                          at<uint32_t>(workspace, L.f_pa),   at<uint32_t>(workspace, L.f_ob),
                          at<uint32_t>(workspace, L.f_pb),   at<uint32_t>(workspace, L.f_ty)};
    uint32_t* k0 = at<uint32_t>(workspace, L.k0);
    uint32_t* k1 = at<uint32_t>(workspace, L.k1);
    uint32_t* v0 = at<uint32_t>(workspace, L.v0);
    uint32_t* v1 = at<uint32_t>(workspace, L.v1);
    const unsigned g = grid_for(m, kScanBlock, 1 << 16);
    hipLaunchKernelGGL(compact_key_kernel, dim3(g), dim3(kScanBlock), 0, s, *set, raw, m, field[0], field[1], field[2],
                       field[3], field[4], field[5], v0, err);
    // least significant field first: types (2 digits), then patch_b, object_b, patch_a, object_a, slab
    const int order[6] = {5, 4, 3, 2, 1, 0};
    for (int f : order) {
      const int passes = f == 5 ? 2 : 4;
      hipLaunchKernelGGL(gather_u32_kernel, dim3(g), dim3(kScanBlock), 0, s, field[f], v0, m, k0);
      if (int rc = radix_sort_passes(k0, v0, k1, v1, m, passes, at<uint32_t>(workspace, L.hist),
                                     at<uint64_t>(workspace, L.hist_off), at<uint64_t>(workspace, L.part), s))
        return rc;
      if (passes & 1) {
        std::swap(k0, k1);
        std::swap(v0, v1);
      }
    }
    const KeyFields kf{field[0], field[1], field[2], field[3], field[4], field[5]};
    uint32_t* flag = at<uint32_t>(workspace, L.flag);
    hipLaunchKernelGGL(segment_flag_kernel, dim3(g), dim3(kScanBlock), 0, s, kf, v0, m, flag);
    if (int rc = exclusive_scan_u32(flag, m, segoff, at<uint64_t>(workspace, L.part), s)) return rc;
    hipLaunchKernelGGL(merge_kernel, dim3(g), dim3(kScanBlock), 0, s, kf, raw, v0, flag, segoff, m, set->query_id,
                       out, cap, err);
  }
  hipLaunchKernelGGL(compact_status_kernel, dim3(1), dim3(64), 0, s, segoff, m, err, status);
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

// ---- host-owned forms (the reference's host-vector ownership model) -------
namespace {

// Device copy of a host proxy set; frees everything on destruction.
struct DeviceSet {
  p2c_proxy_set d{};
  void* block = nullptr;
  ~DeviceSet() {
    if (block) (void)hipFree(block);
  }
  int upload(const p2c_proxy_set& h) {
    const int64_t n = h.n;
    const size_t nn = size_t(n > 0 ? n : 1);
    const size_t sizes[8] = {4 * nn, 4 * nn, 4 * nn, 4 * nn, 8 * nn, 8 * nn, 48 * nn, 32 * nn};
    size_t bytes = 0;
    for (size_t b : sizes) bytes += (b + 255) & ~size_t(255);
    P2C_HIP_CHECK(hipMalloc(&block, bytes));
    char* p = static_cast<char*>(block);
    auto put = [&](const void* src, size_t b) -> void* {
      void* dst = p;
      p += (b + 255) & ~size_t(255);
      if (n > 0 && hipMemcpy(dst, src, b, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
      return dst;
    };
    d.object_id = static_cast<const uint32_t*>(put(h.object_id, 4 * size_t(n)));
    d.patch_id = static_cast<const uint32_t*>(put(h.patch_id, 4 * size_t(n)));
    d.slab_id = static_cast<const uint32_t*>(put(h.slab_id, 4 * size_t(n)));
    d.proxy_type = static_cast<const uint32_t*>(put(h.proxy_type, 4 * size_t(n)));
    d.t0 = static_cast<const double*>(put(h.t0, 8 * size_t(n)));
    d.t1 = static_cast<const double*>(put(h.t1, 8 * size_t(n)));
    d.bounds = static_cast<const double*>(put(h.bounds, 48 * size_t(n)));
    d.motion = static_cast<const double*>(put(h.motion, 32 * size_t(n)));
    if (!d.object_id || !d.patch_id || !d.slab_id || !d.proxy_type || !d.t0 || !d.t1 || !d.bounds || !d.motion)
      return p2c::set_error(P2C_EHIP, "hipMemcpy proxy set H2D");
    d.n = n;
    d.query_id = h.query_id;
    return P2C_OK;
  }
};

struct DeviceBlock {
  void* p = nullptr;
  ~DeviceBlock() {
    if (p) (void)hipFree(p);
  }
};

// Compacts device raw hits into a malloc'ed host candidate array.
int compact_to_host(const p2c_proxy_set& dset, const p2c_raw_hit* d_raw, int64_t n_raw, p2c_candidate** cand,
                    int64_t* n_cand) {
  const size_t ws = p2c_proxy_compact_workspace_bytes(n_raw);
  DeviceBlock w, out, st;
  P2C_HIP_CHECK(hipMalloc(&w.p, ws));
  P2C_HIP_CHECK(hipMalloc(&out.p, sizeof(p2c_candidate) * size_t(n_raw > 0 ? n_raw : 1)));
  P2C_HIP_CHECK(hipMalloc(&st.p, 16));
  if (int rc = p2c_proxy_compact(&dset, d_raw, n_raw, w.p, ws, static_cast<p2c_candidate*>(out.p), n_raw,
                                 static_cast<uint64_t*>(st.p), nullptr))
    return rc;
  uint64_t status[2] = {0, 0};
  P2C_HIP_CHECK(hipMemcpy(status, st.p, 16, hipMemcpyDeviceToHost));
  if (status[1] != 0) return p2c::set_error(P2C_EINVAL, p2c_proxy_status_message(status[1]));
  const int64_t c = int64_t(status[0]);
  *cand = static_cast<p2c_candidate*>(std::malloc(sizeof(p2c_candidate) * size_t(c > 0 ? c : 1)));
  if (*cand == nullptr) return p2c::set_error(P2C_ECAPACITY, "host allocation failed");
  if (c > 0) {
    const hipError_t e = hipMemcpy(*cand, out.p, sizeof(p2c_candidate) * size_t(c), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      std::free(*cand);
      *cand = nullptr;
      return p2c::hip_fail(e, "hipMemcpy candidates D2H");
    }
  }
  *n_cand = c;
  return P2C_OK;
}

}  // namespace

int p2c_proxy_generate_host(const p2c_proxy_set* host_set, p2c_raw_hit** raw, int64_t* n_raw, p2c_candidate** cand,
                            int64_t* n_cand, double* stage_ms) {
  using namespace p2c;
  if (int rc = check_set(host_set)) return rc;
  if (raw == nullptr || n_raw == nullptr || cand == nullptr || n_cand == nullptr)
    return set_error(P2C_EINVAL, "output pointers are required");
  *raw = nullptr;
  *cand = nullptr;
  *n_raw = *n_cand = 0;
  hipEvent_t ev[3];
  for (auto& e : ev) P2C_HIP_CHECK(hipEventCreate(&e));
  struct Events {
    hipEvent_t* e;
    ~Events() {
      for (int i = 0; i < 3; ++i) (void)hipEventDestroy(e[i]);
    }
  } guard{ev};
  DeviceSet ds;
  if (int rc = ds.upload(*host_set)) return rc;
  const int64_t n = host_set->n;
  DeviceBlock counts, offsets, part, d_raw, raw_ws;
  const size_t raw_ws_bytes = p2c_proxy_raw_workspace_bytes(n);
  P2C_HIP_CHECK(hipMalloc(&raw_ws.p, raw_ws_bytes));
  P2C_HIP_CHECK(hipMalloc(&counts.p, 4 * size_t(n > 0 ? n : 1)));
  P2C_HIP_CHECK(hipMalloc(&offsets.p, 8 * size_t(n + 1)));
  P2C_HIP_CHECK(hipMalloc(&part.p, 8 * size_t(n / 1024 + 2)));
  P2C_HIP_CHECK(hipEventRecord(ev[0], nullptr));
  uint64_t total = 0;
  if (n > 0) {
    if (int rc = p2c_proxy_raw_count(&ds.d, static_cast<uint32_t*>(counts.p), raw_ws.p, raw_ws_bytes, nullptr))
      return rc;
    if (int rc = p2c_exclusive_scan_u32(static_cast<uint32_t*>(counts.p), n, static_cast<uint64_t*>(offsets.p),
                                        static_cast<uint64_t*>(part.p), nullptr))
      return rc;
    P2C_HIP_CHECK(hipMemcpy(&total, static_cast<uint64_t*>(offsets.p) + n, 8, hipMemcpyDeviceToHost));
  }
  P2C_HIP_CHECK(hipMalloc(&d_raw.p, sizeof(p2c_raw_hit) * size_t(total > 0 ? total : 1)));
  if (total > 0) {
    if (int rc = p2c_proxy_raw_fill(&ds.d, static_cast<const int64_t*>(offsets.p), raw_ws.p, raw_ws_bytes,
                                    static_cast<p2c_raw_hit*>(d_raw.p), int64_t(total), nullptr))
      return rc;
  }
  P2C_HIP_CHECK(hipEventRecord(ev[1], nullptr));
  if (int rc = compact_to_host(ds.d, static_cast<const p2c_raw_hit*>(d_raw.p), int64_t(total), cand, n_cand)) return rc;
  P2C_HIP_CHECK(hipEventRecord(ev[2], nullptr));
  *raw = static_cast<p2c_raw_hit*>(std::malloc(sizeof(p2c_raw_hit) * size_t(total > 0 ? total : 1)));
  if (*raw == nullptr) return set_error(P2C_ECAPACITY, "host allocation failed");
  if (total > 0)
    P2C_HIP_CHECK(hipMemcpy(*raw, d_raw.p, sizeof(p2c_raw_hit) * size_t(total), hipMemcpyDeviceToHost));
  *n_raw = int64_t(total);
  P2C_HIP_CHECK(hipEventSynchronize(ev[2]));
  if (stage_ms != nullptr) {
    float a = 0.f, b = 0.f;
    P2C_HIP_CHECK(hipEventElapsedTime(&a, ev[0], ev[1]));
    P2C_HIP_CHECK(hipEventElapsedTime(&b, ev[1], ev[2]));
    stage_ms[0] = a;
    stage_ms[1] = b;
  }
  return P2C_OK;
}

int p2c_proxy_compact_host(const p2c_proxy_set* host_set, const p2c_raw_hit* raw, int64_t n_raw,
                           p2c_candidate** cand, int64_t* n_cand) {
  using namespace p2c;
  if (int rc = check_set(host_set)) return rc;
  if (cand == nullptr || n_cand == nullptr) return set_error(P2C_EINVAL, "output pointers are required");
  if (n_raw < 0 || (n_raw > 0 && raw == nullptr)) return set_error(P2C_EINVAL, "raw hits are required");
  *cand = nullptr;
  *n_cand = 0;
  DeviceSet ds;
  if (int rc = ds.upload(*host_set)) return rc;
  DeviceBlock d_raw;
  P2C_HIP_CHECK(hipMalloc(&d_raw.p, sizeof(p2c_raw_hit) * size_t(n_raw > 0 ? n_raw : 1)));
  if (n_raw > 0) P2C_HIP_CHECK(hipMemcpy(d_raw.p, raw, sizeof(p2c_raw_hit) * size_t(n_raw), hipMemcpyHostToDevice));
  return compact_to_host(ds.d, static_cast<const p2c_raw_hit*>(d_raw.p), n_raw, cand, n_cand);
}

void p2c_host_free(void* p) { std::free(p); }

const char* p2c_proxy_status_message(uint64_t error_word) {
  if (error_word == 0) return "";
  switch ((error_word - 1) & 0xFu) {
    case p2c::kErrNotValid: return "raw candidate hit is not marked valid";
    case p2c::kErrQueryMissing: return "raw candidate query_id is required";
    case p2c::kErrQueryMismatch: return "raw candidate query_id must match ProxyScene.query_id";
    case p2c::kErrIndexRange: return "raw candidate proxy index is out of range";
    case p2c::kErrSameProxy: return "raw candidate cannot reference the same proxy twice";
    case p2c::kErrPairKey: return "raw candidate pair_key does not match proxy indices";
    case p2c::kErrSameObject: return "raw candidate must reference two different objects";
    case p2c::kErrSlab: return "raw candidate slab_id must match both proxy primitives";
    case p2c::kErrMotion: return "raw candidate motion_bound must be finite and non-negative";
    case p2c::kErrOverflow: return "candidate rt_hit_count overflow during compaction";
    default: return "unknown compaction error";
  }
}

}  // extern "C"
