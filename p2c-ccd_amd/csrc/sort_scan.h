// Device radix sort and exclusive scan shared by the broad phase and the
// proxy-scene compaction (gfx950).
//   radix sort: LSD over 8-bit digits of u32 keys carrying u32 values; per
//     4096-key tile a digit histogram, an exclusive scan of the histograms
//     (digit-major), then a stable scatter ranked with wave ballots and
//     per-wave LDS digit counters.
//   scan: u32 -> u64 exclusive scan in three phases (per-1024 block scans,
//     a scan of the block totals, the add-back), or in one workgroup for
//     short inputs.
#pragma once

#include <utility>

#include "p2c_hip.h"
#include "runtime.h"

namespace p2c {
namespace {

constexpr int kScanBlock = 256;
constexpr int kSortItems = 16;  // keys per lane per radix tile
constexpr int kSortTile = kScanBlock * kSortItems;

// ---- 2. radix sort -------------------------------------------------------
// Where tile `blk`'s count of digit d sits in the histogram: hist[base +
// d * stride].  One segment (nseg == 0): digit-major over all nblk tiles.
// Segments (a segmented sort; seg[k] is the first tile of segment k, seg[nseg]
// the tile count): digit-major within each segment, segment after segment,
// so the exclusive scan of the histogram gives every tile its place inside
// its own segment's stretch of keys and the segments sort independently.
struct HistSlot {
  int64_t base, stride;
};
__device__ inline HistSlot hist_slot(int64_t blk, int64_t nblk, const int32_t* seg, int nseg) {
  if (nseg == 0) return HistSlot{blk, nblk};
  int lo = 0, hi = nseg - 1;  // the last segment starting at or before blk
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (seg[mid] <= blk) lo = mid;
    else hi = mid - 1;
  }
  const int64_t t0 = seg[lo], nt = int64_t(seg[lo + 1]) - t0;
  return HistSlot{256 * t0 + (blk - t0), nt};
}

__global__ __launch_bounds__(kScanBlock) void radix_count_kernel(const uint32_t* __restrict__ keys, int64_t n, int shift,
                                                             uint32_t* __restrict__ hist, int64_t nblk,
                                                             const int32_t* __restrict__ seg, int nseg) {
  __shared__ uint32_t cnt[256];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = int64_t(blockIdx.x) * kSortTile;
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t e = base + int64_t(j) * kScanBlock + threadIdx.x;
    if (e < n) atomicAdd(&cnt[(keys[e] >> shift) & 0xffu], 1u);
  }
  __syncthreads();
  const HistSlot h = hist_slot(blockIdx.x, nblk, seg, nseg);
  hist[h.base + int64_t(threadIdx.x) * h.stride] = cnt[threadIdx.x];
}

// Stable scatter.  Wave w of a tile owns the contiguous chunk
// [base + w*64*kSortItems, +64*kSortItems) and walks it 64 keys per round;
// a key's rank among equal digits in its round comes from 8 ballots, its
// offset in the wave from the wave's LDS digit counters, and the wave's
// offset in the tile from a prefix over the waves' counters.
__global__ __launch_bounds__(kScanBlock) void radix_scatter_kernel(const uint32_t* __restrict__ kin,
                                                               const uint32_t* __restrict__ vin, int64_t n, int shift,
                                                               const uint64_t* __restrict__ hist_off, int64_t nblk,
                                                               const int32_t* __restrict__ seg, int nseg,
                                                               uint32_t* __restrict__ kout,
                                                               uint32_t* __restrict__ vout) {
  constexpr int kWaves = kScanBlock / 64;
  __shared__ uint32_t wcnt[kWaves][256];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < kWaves * 256; i += kScanBlock) (&wcnt[0][0])[i] = 0;
  __syncthreads();
  const int64_t chunk = int64_t(blockIdx.x) * kSortTile + int64_t(w) * 64 * kSortItems;
  uint32_t key[kSortItems], val[kSortItems], pos[kSortItems];
  const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t e = chunk + int64_t(j) * 64 + lane;
    const bool valid = e < n;
    key[j] = valid ? kin[e] : 0u;
    val[j] = valid ? vin[e] : 0u;
    const uint32_t d = (key[j] >> shift) & 0xffu;
    uint64_t same = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t bal = __ballot(valid && ((d >> b) & 1u));
      same &= ((d >> b) & 1u) ? bal : ~bal;
    }
    const uint32_t rank = (uint32_t)__popcll(same & lt);
    const uint32_t base = valid ? wcnt[w][d] : 0u;
    pos[j] = base + rank;
    __builtin_amdgcn_wave_barrier();
    if (valid && rank == 0) wcnt[w][d] = base + (uint32_t)__popcll(same);
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // exclusive prefix over waves, per digit (one thread per digit)
  {
    const int d = threadIdx.x;
    uint32_t run = 0;
#pragma unroll
    for (int ww = 0; ww < kWaves; ++ww) {
      const uint32_t c = wcnt[ww][d];
      wcnt[ww][d] = run;
      run += c;
    }
  }
  __syncthreads();
  const HistSlot h = hist_slot(blockIdx.x, nblk, seg, nseg);
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t e = chunk + int64_t(j) * 64 + lane;
    if (e < n) {
      const uint32_t d = (key[j] >> shift) & 0xffu;
      const uint64_t dst = hist_off[h.base + int64_t(d) * h.stride] + wcnt[w][d] + pos[j];
      kout[dst] = key[j];
      vout[dst] = val[j];
    }
  }
}

// ---- exclusive scan (u32 -> u64), three phases --------------------------
__device__ inline uint64_t wave_inclusive(uint64_t x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  return x;
}

// Block-wide exclusive scan of 4 items per thread (1024 per block).
__device__ inline void block_scan4(uint64_t v[4], uint64_t* block_total) {
  __shared__ uint64_t wsum[kScanBlock / 64];
  uint64_t s = 0;
  uint64_t t[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    t[i] = s;
    s += v[i];
  }
  const uint64_t inc = wave_inclusive(s);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint64_t wbase = 0, tot = 0;
  for (int i = 0; i < kScanBlock / 64; ++i) {
    if (i < w) wbase += wsum[i];
    tot += wsum[i];
  }
  const uint64_t excl = wbase + inc - s;
#pragma unroll
  for (int i = 0; i < 4; ++i) v[i] = excl + t[i];
  *block_total = tot;
  __syncthreads();
}

__global__ __launch_bounds__(kScanBlock) void scan_local_kernel(const uint32_t* __restrict__ in, int64_t n,
                                                            uint64_t* __restrict__ out, uint64_t* __restrict__ part) {
  const int64_t base = int64_t(blockIdx.x) * 1024 + threadIdx.x * 4;
  uint64_t v[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) v[i] = base + i < n ? in[base + i] : 0;
  uint64_t tot;
  block_scan4(v, &tot);
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (base + i < n) out[base + i] = v[i];
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanBlock) void scan_part_kernel(uint64_t* __restrict__ part, int64_t m,
                                                           uint64_t* __restrict__ grand) {
  uint64_t carry = 0;
  for (int64_t c = 0; c < m; c += 1024) {
    const int64_t base = c + threadIdx.x * 4;
    uint64_t v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = base + i < m ? part[base + i] : 0;
    uint64_t tot;
    block_scan4(v, &tot);
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (base + i < m) part[base + i] = carry + v[i];
    carry += tot;
  }
  if (threadIdx.x == 0) *grand = carry;
}

__global__ __launch_bounds__(kScanBlock) void scan_add_kernel(uint64_t* __restrict__ out, int64_t n,
                                                          const uint64_t* __restrict__ part) {
  const int64_t base = int64_t(blockIdx.x) * 1024 + threadIdx.x * 4;
  const uint64_t add = part[blockIdx.x];
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (base + i < n) out[base + i] += add;
}

// One-workgroup scan for short inputs (up to kScanSmallMax items), one
// launch in place of three: for the digit histograms of a sort pass and the
// query counts of small scenes the launches, not the bytes, are the cost.
// A chunk is 16 rows of 1024 items; lane t of the workgroup holds item t of
// every row (coalesced loads and stores), scans the 16 rows within its wave
// step by step (each step's 16 cross-lane reads issued together), and one
// wave scans the 16 x 16 (row, wave) totals in item order.  Acc is the
// running-sum type: u32 where every prefix is known to fit (the digit
// histograms of fewer than 2^32 keys), u64 otherwise.
constexpr int kScanSmallBlock = 1024;
constexpr int kScanSmallRows = 16;
constexpr int64_t kScanSmallChunk = int64_t(kScanSmallBlock) * kScanSmallRows;
constexpr int64_t kScanSmallMax = 4 * kScanSmallChunk;

template <typename Acc>
__global__ __launch_bounds__(kScanSmallBlock) void scan_small_kernel(const uint32_t* __restrict__ in, int64_t n,
                                                                     uint64_t* __restrict__ out,
                                                                     uint64_t* __restrict__ grand) {
  constexpr int kWaves = kScanSmallBlock / 64;
  constexpr int kParts = kScanSmallRows * kWaves;  // (row, wave) totals, row-major = item order
  static_assert(kParts == 4 * 64, "one wave scans the totals, four per lane");
  __shared__ uint64_t part[kParts + 1];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint64_t carry = 0;
  for (int64_t c = 0; c < n; c += kScanSmallChunk) {
    uint32_t v[kScanSmallRows];
    Acc inc[kScanSmallRows];
#pragma unroll
    for (int r = 0; r < kScanSmallRows; ++r) {
      const int64_t e = c + int64_t(r) * kScanSmallBlock + threadIdx.x;
      v[r] = e < n ? in[e] : 0u;
      inc[r] = v[r];
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      Acc y[kScanSmallRows];
#pragma unroll
      for (int r = 0; r < kScanSmallRows; ++r) y[r] = __shfl_up(inc[r], o);
#pragma unroll
      for (int r = 0; r < kScanSmallRows; ++r)
        if (lane >= o) inc[r] += y[r];
    }
    if (lane == 63) {
#pragma unroll
      for (int r = 0; r < kScanSmallRows; ++r) part[r * kWaves + w] = inc[r];
    }
    __syncthreads();
    if (w == 0) {
      uint64_t t[4], sum = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        t[k] = sum;
        sum += part[lane * 4 + k];
      }
      const uint64_t x = wave_inclusive(sum);
      const uint64_t base = carry + x - sum;
#pragma unroll
      for (int k = 0; k < 4; ++k) part[lane * 4 + k] = base + t[k];
      if (lane == 63) part[kParts] = carry + x;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kScanSmallRows; ++r) {
      const int64_t e = c + int64_t(r) * kScanSmallBlock + threadIdx.x;
      if (e < n) out[e] = part[r * kWaves + w] + uint64_t(inc[r] - v[r]);
    }
    carry = part[kParts];
    __syncthreads();  // part is rewritten by the next chunk
  }
  if (grand != nullptr && threadIdx.x == 0) *grand = carry;
}

// Stable LSD sort of (key, value) u32 pairs on the low 8 * passes bits of the
// keys, ping-ponging between (k0, v0) and (k1, v1); an even number of passes
// leaves the result in (k0, v0), an odd number in (k1, v1).  Workspace: hist
// u32[256 * tiles], hist_off u64[256 * tiles], part u64[256 * tiles / 1024 + 2]
// with tiles = ceil(n / kSortTile).  With a segment table (seg: device
// int32[nseg + 1], first tile of each segment, then the tile count; n a
// multiple of kSortTile) each segment of whole tiles is sorted on its own.
inline int radix_sort_passes(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, int64_t n, int passes,
                             uint32_t* hist, uint64_t* hist_off, uint64_t* part, hipStream_t s,
                             const int32_t* seg = nullptr, int nseg = 0) {
  if (n <= 1) return P2C_OK;
  if (n > int64_t(UINT32_MAX)) return set_error(P2C_ECAPACITY, "radix sort: more than 2^32 - 1 keys");
  if (nseg > 0 && n % kSortTile != 0) return set_error(P2C_EINVAL, "radix sort: segments must be whole tiles");
  const int64_t nblk = (n + kSortTile - 1) / kSortTile;
  for (int pass = 0; pass < passes; ++pass) {
    const int shift = pass * 8;
    hipLaunchKernelGGL(radix_count_kernel, dim3(unsigned(nblk)), dim3(kScanBlock), 0, s, k0, n, shift, hist, nblk,
                       seg, nseg);
    const int64_t m = 256 * nblk;
    if (m <= kScanSmallMax) {
      // every prefix of a digit histogram is at most n < 2^32
      hipLaunchKernelGGL(scan_small_kernel<uint32_t>, dim3(1), dim3(kScanSmallBlock), 0, s, hist, m, hist_off,
                         nullptr);
    } else {
      const int64_t sb = (m + 1023) / 1024;
      hipLaunchKernelGGL(scan_local_kernel, dim3(unsigned(sb)), dim3(kScanBlock), 0, s, hist, m, hist_off, part);
      hipLaunchKernelGGL(scan_part_kernel, dim3(1), dim3(kScanBlock), 0, s, part, sb, part + sb);
      hipLaunchKernelGGL(scan_add_kernel, dim3(unsigned(sb)), dim3(kScanBlock), 0, s, hist_off, m, part);
    }
    hipLaunchKernelGGL(radix_scatter_kernel, dim3(unsigned(nblk)), dim3(kScanBlock), 0, s, k0, v0, n, shift, hist_off,
                       nblk, seg, nseg, k1, v1);
    std::swap(k0, k1);
    std::swap(v0, v1);
  }
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

// Full 32-bit sort (4 passes): the result is back in (k0, v0).
inline int radix_sort(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, int64_t n, uint32_t* hist,
                      uint64_t* hist_off, uint64_t* part, hipStream_t s, const int32_t* seg = nullptr,
                      int nseg = 0) {
  return radix_sort_passes(k0, v0, k1, v1, n, 4, hist, hist_off, part, s, seg, nseg);
}

// Exclusive scan of n u32 values into u64 (out[n] = total); part holds
// n / 1024 + 2 u64.
inline int exclusive_scan_u32(const uint32_t* in, int64_t n, uint64_t* out, uint64_t* part, hipStream_t s) {
  if (n <= 0) return P2C_OK;
  if (n <= kScanSmallMax) {
    hipLaunchKernelGGL(scan_small_kernel<uint64_t>, dim3(1), dim3(kScanSmallBlock), 0, s, in, n, out, out + n);
    P2C_HIP_CHECK(hipGetLastError());
    return P2C_OK;
  }
  const int64_t sb = (n + 1023) / 1024;
  hipLaunchKernelGGL(scan_local_kernel, dim3(unsigned(sb)), dim3(kScanBlock), 0, s, in, n, out, part);
  hipLaunchKernelGGL(scan_part_kernel, dim3(1), dim3(kScanBlock), 0, s, part, sb, out + n);
  hipLaunchKernelGGL(scan_add_kernel, dim3(unsigned(sb)), dim3(kScanBlock), 0, s, out, n, part);
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}


}  // namespace
}  // namespace p2c
