// Swept-AABB LBVH broad phase on gfx950: vertex-face and edge-edge candidate
// pairs with the reference's set semantics.
//
// Candidate set (src/tools/scene_object_envelope_native_ti_walltime.cpp):
//   boxes  :256-302  min/max of a primitive's vertices at t0 and t1, then
//                    -/+ thickness (FP64, exact min/max, one rounding each side)
//   VF     :651-671  (vertex v, face f), inclusive overlap (:270-274), object
//                    pair allowed, not (same object and f contains v)
//   EE     :673-696  (edge i < edge j), overlap, allowed, not (same object
//                    and a shared vertex)
//   objects:528-550  object envelopes; distinct objects pair only when their
//                    envelopes overlap, self pairs are allowed
// The reference enumerates with a single-threaded sort-and-sweep (:603-649);
// here the enumeration is an LBVH:
//   1. swept boxes + object envelopes (order-preserving u64 atomics in LDS)
//   2. 30-bit Morton codes of box centres (ordering only: no box is culled
//      by quantisation, every overlap test is the exact FP64 one)
//   3. LSD radix sort, 8-bit digits, stable in-block ranking from wave
//      ballots with per-wave digit counters staged in LDS
//   4. Karras build (distinct 64-bit keys: code << 32 | position) with
//      escape indices for stackless traversal; bottom-up FP64 refit with
//      acquire/release counters
//   5. stackless traversal, one lane per query in Morton order: count
//      (keeping each query's first hit_cache_slots partners in the caller's
//      cache), exclusive scan, fill (deterministic output order): queries
//      with at most that many hits are copied from the cache, only the
//      others traverse again.
//      An edge query visits only the leaves after its own, so each
//      unordered EE pair is found once.
#include <cmath>

#include "p2c_hip.h"
#include "runtime.h"
#include "sort_scan.h"

namespace p2c {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxObjects = 1 << 15;  // the allowed matrix holds kMaxObjects^2 bytes (1 GiB)
constexpr int kLdsObjects = 64;        // envelopes reduced in LDS per workgroup up to this many objects

struct Box6 {
  double mn[3];
  double mx[3];
};

struct Node {  // 64 bytes; internal: left/right children; leaf: left = right = -1, prim
  double mn[3];
  double mx[3];
  int32_t left, right, escape, prim;
};
static_assert(sizeof(Node) == 64, "node layout");

__device__ inline uint64_t f64_order_key(double x) {
  const uint64_t b = __double_as_longlong(x);
  return (b >> 63) ? ~b : (b | (1ull << 63));
}
__device__ inline double f64_from_order_key(uint64_t k) {
  const uint64_t b = (k >> 63) ? (k & ~(1ull << 63)) : ~k;
  return __longlong_as_double(b);
}

__device__ inline void box_reset(Box6& b) {
  b.mn[0] = b.mn[1] = b.mn[2] = __builtin_inf();
  b.mx[0] = b.mx[1] = b.mx[2] = -__builtin_inf();
}
// Extend with std::min / std::max semantics.
__device__ inline void box_extend(Box6& b, const double* p) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    b.mn[k] = (p[k] < b.mn[k]) ? p[k] : b.mn[k];
    b.mx[k] = (b.mx[k] < p[k]) ? p[k] : b.mx[k];
  }
}
__device__ inline void box_expand(Box6& b, double d) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    b.mn[k] -= d;
    b.mx[k] += d;
  }
}
__device__ inline bool overlap(const Box6& a, const double* mn, const double* mx) {
  return a.mn[0] <= mx[0] && mn[0] <= a.mx[0] && a.mn[1] <= mx[1] && mn[1] <= a.mx[1] && a.mn[2] <= mx[2] &&
         mn[2] <= a.mx[2];
}

// ---- workspace ----------------------------------------------------------
struct Layout {
  size_t roots, vbox, fbox, ebox, objkey, bounds, allowed, keys0, vals0, keys1, vals1, hist, fnodes, fparent, fflag, frcs,
      fhi, flo, enodes, eparent, eflag, ercs, ehi, elo, qorder, eorder, counts, resume, offsets, hist_off, scan_part, err,
      eroots, total;
  int64_t sort_cap, hist_blocks;
};

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

Layout make_layout(int64_t nv, int64_t nf, int64_t ne, int32_t nobj) {
  Layout L{};
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o = align256(o + (bytes ? bytes : 1));
    return at;
  };
  const int64_t sn = std::max<int64_t>(std::max(nv, nf), std::max<int64_t>(ne, 1));
  // one sort over the face, edge and vertex keys together
  const int64_t sall = std::max<int64_t>(nf + ne + nv, 1);
  L.sort_cap = sall;
  L.hist_blocks = (sall + kSortTile - 1) / kSortTile;
  L.vbox = take(sizeof(Box6) * nv);
  L.fbox = take(sizeof(Box6) * nf);
  L.ebox = take(sizeof(Box6) * ne);
  L.objkey = take(sizeof(uint64_t) * 6 * size_t(nobj));
  L.bounds = take(sizeof(double) * 6);
  L.allowed = take(size_t(nobj) * nobj);
  L.keys0 = take(4 * sall);
  L.vals0 = take(4 * sall);
  L.keys1 = take(4 * sall);
  L.vals1 = take(4 * sall);
  L.hist = take(4 * 256 * L.hist_blocks);
  L.fnodes = take(sizeof(Node) * std::max<int64_t>(2 * nf - 1, 1));
  L.fparent = take(4 * std::max<int64_t>(2 * nf - 1, 1));
  L.fflag = take(4 * std::max<int64_t>(nf, 1));
  L.frcs = take(4 * std::max<int64_t>(nf, 1));
  L.fhi = take(4 * std::max<int64_t>(nf, 1));
  L.flo = take(4 * std::max<int64_t>(nf, 1));
  L.enodes = take(sizeof(Node) * std::max<int64_t>(2 * ne - 1, 1));
  L.eparent = take(4 * std::max<int64_t>(2 * ne - 1, 1));
  L.eflag = take(4 * std::max<int64_t>(ne, 1));
  L.ercs = take(4 * std::max<int64_t>(ne, 1));
  L.ehi = take(4 * std::max<int64_t>(ne, 1));
  L.elo = take(4 * std::max<int64_t>(ne, 1));
  L.qorder = take(4 * std::max<int64_t>(nv, 1));
  L.eorder = take(4 * std::max<int64_t>(ne, 1));
  L.counts = take(4 * (nv + ne + 1));
  L.resume = take(4 * (nv + ne + 1));
  L.offsets = take(8 * (nv + ne + 1));
  L.hist_off = take(8 * 256 * L.hist_blocks);
  const int64_t scan_items = std::max<int64_t>(nv + ne + 1, 256 * L.hist_blocks);
  L.scan_part = take(8 * (scan_items / 1024 + 2));
  L.err = take(16);
  L.roots = take(4 * (size_t(sn) + 2));  // [0] count, then the nodes whose parent spans blocks (face tree)
  L.eroots = take(4 * (size_t(sn) + 2));  // the same for the edge tree
  L.total = o;
  return L;
}

template <typename T>
T* at(void* ws, size_t off) {
  return reinterpret_cast<T*>(static_cast<char*>(ws) + off);
}

// ---- 1. boxes and object envelopes -------------------------------------
// Resets the object envelopes, the error counters (4 words) and the counts
// of both trees' top-refit root lists.
__global__ __launch_bounds__(kBlock) void objkey_init_kernel(uint64_t* objkey, int32_t nobj, uint32_t* err,
                                                             uint32_t* froots, uint32_t* eroots) {
  const int64_t i0 = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  for (int64_t i = i0; i < int64_t(nobj) * 6; i += int64_t(gridDim.x) * blockDim.x)
    objkey[i] = (i % 6) < 3 ? ~0ull : 0ull;
  if (i0 < 4) err[i0] = 0u;
  if (i0 == 4) froots[0] = 0u;
  if (i0 == 5) eroots[0] = 0u;
}

__global__ __launch_bounds__(kBlock) void vertex_box_kernel(const double* __restrict__ x0,
                                                            const double* __restrict__ x1, int64_t nv, double th,
                                                            const int32_t* __restrict__ vobj, int32_t nobj,
                                                            Box6* __restrict__ vbox, uint64_t* __restrict__ objkey) {
  // up to kLdsObjects objects: envelopes reduced in LDS, then one global
  // atomic per word and workgroup; more objects: global atomics per vertex
  __shared__ unsigned long long lkey[kLdsObjects * 6];
  const bool lds = nobj <= kLdsObjects;
  unsigned long long* key = lds ? lkey : reinterpret_cast<unsigned long long*>(objkey);
  if (lds)
    for (int i = threadIdx.x; i < nobj * 6; i += blockDim.x) lkey[i] = (i % 6) < 3 ? ~0ull : 0ull;
  __syncthreads();
  for (int64_t v = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < nv; v += int64_t(gridDim.x) * blockDim.x) {
    Box6 b;
    box_reset(b);
    box_extend(b, x0 + v * 3);
    box_extend(b, x1 + v * 3);
    box_expand(b, th);
    vbox[v] = b;
    const int32_t o = vobj[v];
    if (o >= 0 && o < nobj) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        atomicMin(&key[int64_t(o) * 6 + k], (unsigned long long)f64_order_key(b.mn[k]));
        atomicMax(&key[int64_t(o) * 6 + 3 + k], (unsigned long long)f64_order_key(b.mx[k]));
      }
    }
  }
  if (!lds) return;
  __syncthreads();
  for (int i = threadIdx.x; i < nobj * 6; i += blockDim.x) {
    if ((i % 6) < 3) atomicMin((unsigned long long*)&objkey[i], lkey[i]);
    else atomicMax((unsigned long long*)&objkey[i], lkey[i]);
  }
}

__device__ inline void object_box(const uint64_t* objkey, int o, double* mn, double* mx) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    mn[k] = f64_from_order_key(objkey[o * 6 + k]);
    mx[k] = f64_from_order_key(objkey[o * 6 + 3 + k]);
  }
}

__global__ __launch_bounds__(kBlock) void allowed_kernel(const uint64_t* __restrict__ objkey, int32_t nobj,
                                                         uint8_t* __restrict__ allowed) {
  const int64_t n2 = int64_t(nobj) * nobj;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n2; i += int64_t(gridDim.x) * blockDim.x) {
    const int a = int(i / nobj), b = int(i % nobj);
    uint8_t ok = 1;
    if (a != b) {
      double amn[3], amx[3], bmn[3], bmx[3];
      object_box(objkey, a, amn, amx);
      object_box(objkey, b, bmn, bmx);
      ok = (amn[0] <= bmx[0] && bmn[0] <= amx[0] && amn[1] <= bmx[1] && bmn[1] <= amx[1] && amn[2] <= bmx[2] &&
            bmn[2] <= amx[2])
               ? 1
               : 0;
    }
    allowed[i] = ok;
  }
}

__device__ inline uint32_t spread10(uint32_t x) {
  x &= 0x3ffu;
  x = (x | (x << 16)) & 0x030000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}

__device__ inline uint32_t morton_of(const Box6& b, const double* lo, const double* inv) {
  uint32_t c[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    double u = (0.5 * (b.mn[k] + b.mx[k]) - lo[k]) * inv[k];
    u = u < 0.0 ? 0.0 : (u > 1.0 ? 1.0 : u);
    c[k] = (uint32_t)(u * 1023.0);
  }
  return (spread10(c[0]) << 2) | (spread10(c[1]) << 1) | spread10(c[2]);
}

// The Morton frame: the union of the object envelopes (lo, 1 / extent per
// axis), reduced once by one workgroup into bounds[6].  Its min/max over the
// same values in any order is exact, so the frame does not depend on the
// reduction order.
__global__ __launch_bounds__(kBlock) void scene_bounds_kernel(const uint64_t* __restrict__ objkey, int32_t nobj,
                                                              double* __restrict__ bounds) {
  __shared__ double red[6][kBlock];
  double lo[3], hi[3];
  for (int k = 0; k < 3; ++k) {
    lo[k] = __builtin_inf();
    hi[k] = -__builtin_inf();
  }
  for (int o = threadIdx.x; o < nobj; o += blockDim.x) {
    double mn[3], mx[3];
    object_box(objkey, o, mn, mx);
    for (int k = 0; k < 3; ++k) {
      lo[k] = mn[k] < lo[k] ? mn[k] : lo[k];
      hi[k] = mx[k] > hi[k] ? mx[k] : hi[k];
    }
  }
  for (int k = 0; k < 3; ++k) {
    red[k][threadIdx.x] = lo[k];
    red[3 + k][threadIdx.x] = hi[k];
  }
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w)
      for (int k = 0; k < 3; ++k) {
        const double a = red[k][threadIdx.x + w], b = red[3 + k][threadIdx.x + w];
        red[k][threadIdx.x] = a < red[k][threadIdx.x] ? a : red[k][threadIdx.x];
        red[3 + k][threadIdx.x] = b > red[3 + k][threadIdx.x] ? b : red[3 + k][threadIdx.x];
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) {
      const double ext = red[3 + k][0] - red[k][0];
      bounds[k] = red[k][0];
      bounds[3 + k] = ext > 0.0 ? 1.0 / ext : 0.0;
    }
}

// Face boxes and keys (items [0, nf)), edge boxes and keys ([nf, nf + ne)),
// vertex keys ([nf + ne, nf + ne + nv)) in one launch; values are the index
// within the kind.  All three key sets are sorted together, tagged with
// their kind in bits 30-31 above the 30-bit Morton code: a stable sort keeps
// each kind contiguous, in the order (and with the tie order) a sort of that
// kind alone would give.
struct PrimKeys {
  const double* x0;
  const double* x1;
  const int32_t* faces;
  const int32_t* edges;
  const Box6* vbox;
  const double* bounds;  // lo[3], inv[3] of scene_bounds_kernel
  Box6* fbox;
  Box6* ebox;
  uint32_t* keys;
  uint32_t* vals;
  int64_t nf, ne, nv;
  double th;
};

__global__ __launch_bounds__(kBlock) void prim_box_kernel(PrimKeys P) {
  double lo[3], inv[3];
  for (int k = 0; k < 3; ++k) {
    lo[k] = P.bounds[k];
    inv[k] = P.bounds[3 + k];
  }
  const int64_t total = P.nf + P.ne + P.nv;
  for (int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < total; g += int64_t(gridDim.x) * blockDim.x) {
    const uint32_t kind = g < P.nf ? 0u : (g < P.nf + P.ne ? 1u : 2u);
    const int64_t i = kind == 0 ? g : (kind == 1 ? g - P.nf : g - P.nf - P.ne);
    Box6 b;
    if (kind == 2) {
      b = P.vbox[i];
    } else {
      box_reset(b);
      const int m = kind == 0 ? 3 : 2;
      const int32_t* idx = kind == 0 ? P.faces : P.edges;
      for (int c = 0; c < m; ++c) {
        const int64_t v = idx[i * m + c];
        box_extend(b, P.x0 + v * 3);
        box_extend(b, P.x1 + v * 3);
      }
      box_expand(b, P.th);
      (kind == 0 ? P.fbox : P.ebox)[i] = b;
    }
    P.keys[g] = morton_of(b, lo, inv) | (kind << 30);
    P.vals[g] = (uint32_t)i;
  }
}

// ---- 3. Karras build, escapes, refit ------------------------------------
__device__ inline int delta(const uint32_t* keys, int64_t n, int64_t i, int64_t j) {
  if (j < 0 || j >= n) return -1;
  const uint64_t a = (uint64_t(keys[i]) << 32) | uint64_t(i);
  const uint64_t b = (uint64_t(keys[j]) << 32) | uint64_t(j);
  return __clzll(a ^ b);
}

__global__ __launch_bounds__(kBlock) void karras_kernel(const uint32_t* __restrict__ keys, int64_t n,
                                                        Node* __restrict__ nodes, int32_t* __restrict__ parent,
                                                        int32_t* __restrict__ rcs, int32_t* __restrict__ hi_of,
                                                        int32_t* __restrict__ lo_of) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n - 1; i += int64_t(gridDim.x) * blockDim.x) {
    const int d = delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1) >= 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int64_t lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int64_t l = 0;
    for (int64_t t = lmax >> 1; t >= 1; t >>= 1)
      if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int64_t j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int64_t s = 0, step = l;
    do {
      step = (step + 1) >> 1;
      if (delta(keys, n, i, i + (s + step) * d) > dnode) s += step;
    } while (step > 1);
    const int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
    const int64_t lo = i < j ? i : j, hi = i < j ? j : i;
    const int32_t left = lo == gamma ? int32_t(n - 1 + gamma) : int32_t(gamma);
    const int32_t right = hi == gamma + 1 ? int32_t(n - 1 + gamma + 1) : int32_t(gamma + 1);
    nodes[i].left = left;
    nodes[i].right = right;
    nodes[i].prim = -1;
    parent[left] = int32_t(i);
    parent[right] = int32_t(i);
    rcs[gamma] = right;
    hi_of[i] = int32_t(hi);
    lo_of[i] = int32_t(lo);
  }
}

__global__ __launch_bounds__(kBlock) void leaves_kernel(const uint32_t* __restrict__ vals, int64_t n,
                                                        const Box6* __restrict__ box, Node* __restrict__ nodes,
                                                        const int32_t* __restrict__ rcs,
                                                        const int32_t* __restrict__ hi_of,
                                                        int32_t* __restrict__ parent, uint32_t* __restrict__ flag) {
  for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < 2 * n - 1;
       p += int64_t(gridDim.x) * blockDim.x) {
    if (p < n - 1) {  // internal node p: escape from its range max; reset refit flag
      const int64_t hi = hi_of[p];
      nodes[p].escape = hi == n - 1 ? -1 : rcs[hi];
      flag[p] = 0;
    } else {
      const int64_t q = p - (n - 1);
      Node leaf;
      const Box6 b = box[vals[q]];
      for (int k = 0; k < 3; ++k) {
        leaf.mn[k] = b.mn[k];
        leaf.mx[k] = b.mx[k];
      }
      leaf.left = leaf.right = -1;
      leaf.escape = q == n - 1 ? -1 : rcs[q];
      leaf.prim = int32_t(vals[q]);
      nodes[p] = leaf;
      if (n == 1) parent[p] = -1;
    }
    if (p == 0 && n > 1) parent[0] = -1;
  }
}

// Bottom-up refit, one leaf per lane.  An internal node's leaf range holds
// its own index (Karras), so a node whose range lies inside the block's 256
// leaves is counted in an LDS flag with workgroup scope; only nodes spanning
// blocks use the global flag with agent scope, whose release also publishes
// the boxes the climbing lane built below it.  The first child to arrive
// stops; the second sees both children's boxes and builds the parent's.
// Refit in two launches, so that no step needs coherence across the chip's
// separate L2s.  1. One block per 256 leaves refits every internal node
// whose leaf range lies in the block (in-block counters in LDS, workgroup
// scope); a thread that reaches a node spanning blocks records the child it
// came from.  2. One workgroup climbs from those children through the nodes
// that span blocks, with per-node counters at workgroup scope.
__global__ __launch_bounds__(kBlock) void refit_kernel(int64_t n, Node* __restrict__ nodes,
                                                       const int32_t* __restrict__ parent,
                                                       const int32_t* __restrict__ lo_of,
                                                       const int32_t* __restrict__ hi_of,
                                                       uint32_t* __restrict__ roots) {
  __shared__ uint32_t local_flag[kBlock];
  local_flag[threadIdx.x] = 0;
  __syncthreads();
  const int64_t b0 = int64_t(blockIdx.x) * kBlock;
  const int64_t b1 = (b0 + kBlock < n ? b0 + kBlock : n) - 1;  // last leaf of the block
  const int64_t q = b0 + threadIdx.x;
  if (q >= n) return;
  int32_t child = int32_t(n - 1 + q);
  int32_t node = parent[child];
  while (node >= 0) {
    if (!(lo_of[node] >= b0 && hi_of[node] <= b1)) {
      const uint32_t at = atomicAdd(&roots[0], 1u);
      roots[2 + at] = uint32_t(child);
      break;
    }
    const uint32_t prev =
        __hip_atomic_fetch_add(&local_flag[node - b0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (prev == 0) break;
    const Node& a = nodes[nodes[node].left];
    const Node& b = nodes[nodes[node].right];
    for (int k = 0; k < 3; ++k) {
      const double amn = a.mn[k], bmn = b.mn[k];
      const double amx = a.mx[k], bmx = b.mx[k];
      nodes[node].mn[k] = bmn < amn ? bmn : amn;
      nodes[node].mx[k] = amx < bmx ? bmx : amx;
    }
    child = node;
    node = parent[node];
  }
}

constexpr int kTopBlock = 1024;

__global__ __launch_bounds__(kTopBlock) void refit_top_kernel(Node* __restrict__ nodes,
                                                              const int32_t* __restrict__ parent,
                                                              const uint32_t* __restrict__ roots,
                                                              uint32_t* __restrict__ flag) {
  const uint32_t total = roots[0];
  for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
    int32_t node = parent[roots[2 + i]];
    while (node >= 0) {
      const uint32_t prev = __hip_atomic_fetch_add(&flag[node], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (prev == 0) break;
      const Node& a = nodes[nodes[node].left];
      const Node& b = nodes[nodes[node].right];
      for (int k = 0; k < 3; ++k) {
        const double amn = a.mn[k], bmn = b.mn[k];
        const double amx = a.mx[k], bmx = b.mx[k];
        nodes[node].mn[k] = bmn < amn ? bmn : amn;
        nodes[node].mx[k] = amx < bmx ? bmx : amx;
      }
      node = parent[node];
    }
  }
}

// The same climb with the arrival flags as bits in LDS (64 KB: trees of up to
// 524,288 internal nodes): an LDS atomic per level instead of an L2 one.
// Workgroup-scope ordering covers the child boxes written to global memory by
// the other lanes of this workgroup, as in refit_top_kernel.
constexpr int kTopLdsWords = 16384;

struct TopTree {
  Node* nodes;
  const int32_t* parent;
  const uint32_t* roots;
  int64_t n_internal;
};

// Block b climbs tree b (faces 0, edges 1): the two trees' top refits in one
// launch, on two CUs.
__global__ __launch_bounds__(kTopBlock) void refit_top_lds_kernel(TopTree t0, TopTree t1) {
  __shared__ uint32_t bits[kTopLdsWords];
  const TopTree& t = blockIdx.x == 0 ? t0 : t1;
  Node* __restrict__ nodes = t.nodes;
  const int32_t* __restrict__ parent = t.parent;
  const int words = int((t.n_internal + 31) / 32);
  for (int i = threadIdx.x; i < words; i += blockDim.x) bits[i] = 0;
  __syncthreads();
  const uint32_t total = t.roots[0];
  for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
    int32_t node = parent[t.roots[2 + i]];
    while (node >= 0) {
      const uint32_t bit = 1u << (node & 31);
      const uint32_t prev =
          __hip_atomic_fetch_or(&bits[node >> 5], bit, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (!(prev & bit)) break;
      const Node& a = nodes[nodes[node].left];
      const Node& b = nodes[nodes[node].right];
      for (int k = 0; k < 3; ++k) {
        const double amn = a.mn[k], bmn = b.mn[k];
        const double amx = a.mx[k], bmx = b.mx[k];
        nodes[node].mn[k] = bmn < amn ? bmn : amn;
        nodes[node].mx[k] = amx < bmx ? bmx : amx;
      }
      node = parent[node];
    }
  }
}

// ---- 4. traversal ------------------------------------------------------
struct VfRule {
  const int32_t* faces;
  const int32_t* fobj;
  const int32_t* vobj;
  const uint8_t* allowed;
  int32_t nobj;
};

struct EeRule {
  const int32_t* edges;
  const int32_t* eobj;
  const uint8_t* allowed;
  int32_t nobj;
};

__device__ inline bool pair_allowed(const uint8_t* allowed, int32_t nobj, int32_t a, int32_t b) {
  if (a < 0 || b < 0 || a >= nobj || b >= nobj) return false;
  return allowed[a * nobj + b] != 0;
}

// Walks the BVH for query box q, calling emit(prim) for every leaf whose box
// overlaps; returns the number of accepted leaves.
// A pre-order walk visits each of the 2n-1 nodes at most once; `max_steps`
// bounds the loop regardless, and an exhausted bound is counted in *err.
// Starts at node `start` with `count0` hits already counted; emit(prim, k,
// next) also gets the node the walk continues at, so a walk can be resumed
// after any hit.
template <typename Accept, typename Emit>
__device__ inline uint32_t traverse(const Node* __restrict__ nodes, int32_t start, uint32_t count0, const Box6& q,
                                    int64_t max_steps, uint32_t* err, Accept accept, Emit emit) {
  uint32_t count = count0;
  int32_t node = start;
  int64_t steps = 0;
  while (node >= 0) {
    if (++steps > max_steps) {
      atomicAdd(err, 1u);
      break;
    }
    const Node& N = nodes[node];
    const double mn[3] = {N.mn[0], N.mn[1], N.mn[2]};
    const double mx[3] = {N.mx[0], N.mx[1], N.mx[2]};
    const int32_t left = N.left, escape = N.escape, prim = N.prim;
    if (overlap(q, mn, mx)) {
      if (left < 0) {
        if (accept(prim)) {
          emit(prim, count, escape);
          ++count;
        }
        node = escape;
      } else {
        node = left;
      }
    } else {
      node = escape;
    }
  }
  return count;
}

struct VfQuery {
  const Node* fnodes;
  int64_t nf;
  const Box6* vbox;
  const uint32_t* qorder;
  int64_t nv;
  VfRule r;
  uint32_t* err;
  uint32_t* counts;
  int32_t* cache;
  int64_t slots;
  int32_t* resume;
  const uint64_t* offsets;
  int32_t* out;
  int64_t cap;
};

// Vertex query s (in Morton order) against the face tree: count pass (kFill
// false: count, cache the first `slots` partners, note where to resume) or
// fill pass for the queries with more partners than the cache holds.
template <bool kFill>
__device__ inline void vf_query_one(const VfQuery& Q, int64_t s) {
  const Node* __restrict__ fnodes = Q.fnodes;
  const int64_t nf = Q.nf, slots = Q.slots, cap = Q.cap;
  const VfRule& r = Q.r;
  const uint64_t* __restrict__ offsets = Q.offsets;
  int32_t* __restrict__ cache = Q.cache;
  int32_t* __restrict__ resume = Q.resume;
  int32_t* __restrict__ out = Q.out;
  uint32_t* __restrict__ counts = Q.counts;
  uint32_t* err = Q.err;
  const uint32_t* __restrict__ qorder = Q.qorder;
  const Box6* __restrict__ vbox = Q.vbox;
  if (kFill && offsets[s + 1] - offsets[s] <= uint64_t(slots)) return;  // copied from the cache
  // a query with more hits resumes after its last cached one
  const int32_t start = kFill && slots > 0 ? resume[s] : 0;
  const uint32_t count0 = kFill && slots > 0 ? uint32_t(slots) : 0u;
  const int32_t v = int32_t(qorder[s]);
  const int32_t ov = r.vobj[v];
  uint32_t c = 0;
  if (ov >= 0 && nf > 0) {
    const Box6 q = vbox[v];
    const uint64_t base = kFill ? offsets[s] : 0;
    c = traverse(
        fnodes, start, count0, q, 2 * nf, err,
        [&](int32_t f) {
          const int32_t of = r.fobj[f];
          if (!pair_allowed(r.allowed, r.nobj, ov, of)) return false;
          if (ov == of) {
            const int32_t* fv = r.faces + int64_t(f) * 3;
            if (fv[0] == v || fv[1] == v || fv[2] == v) return false;
          }
          return true;
        },
        [&](int32_t f, uint32_t k, int32_t next) {
          if (!kFill && int64_t(k) < slots) {
            cache[s * slots + k] = f;
            if (int64_t(k) == slots - 1) resume[s] = next;
          }
          if (kFill && base + k < uint64_t(cap)) {
            out[(base + k) * 2] = v;
            out[(base + k) * 2 + 1] = f;
          }
        });
  }
  if (!kFill) counts[s] = c;
}

template <bool kFill>
__global__ __launch_bounds__(kBlock) void vf_query_kernel(VfQuery Q) {
  const int64_t b0 = xcd_block(blockIdx.x, gridDim.x);  // Morton-neighbouring queries on one XCD
  for (int64_t s = b0 * blockDim.x + threadIdx.x; s < Q.nv; s += int64_t(gridDim.x) * blockDim.x)
    vf_query_one<kFill>(Q, s);
}

struct EeQuery {
  const Node* enodes;
  int64_t ne;
  const Box6* ebox;
  const uint32_t* eorder;
  EeRule r;
  uint32_t* err;
  uint32_t* counts;
  int32_t* cache;
  int64_t slots;
  int32_t* resume;
  const uint64_t* offsets;
  int64_t nv;
  int32_t* out;
  int64_t cap;
};

// Edge query s (in tree order) against the edge tree, count or fill pass as
// for vf_query_one.
template <bool kFill>
__device__ inline void ee_query_one(const EeQuery& Q, int64_t s) {
  const Node* __restrict__ enodes = Q.enodes;
  const int64_t ne = Q.ne, slots = Q.slots, nv = Q.nv, cap = Q.cap;
  const EeRule& r = Q.r;
  const uint64_t* __restrict__ offsets = Q.offsets;
  int32_t* __restrict__ cache = Q.cache;
  int32_t* __restrict__ resume = Q.resume;
  int32_t* __restrict__ out = Q.out;
  uint32_t* __restrict__ counts = Q.counts;
  uint32_t* err = Q.err;
  const uint32_t* __restrict__ eorder = Q.eorder;
  const Box6* __restrict__ ebox = Q.ebox;
  if (kFill && offsets[nv + s + 1] - offsets[nv + s] <= uint64_t(slots)) return;  // copied from the cache
  const int32_t i = int32_t(eorder[s]);
  const Box6 q = ebox[i];
  const int32_t oi = r.eobj[i];
  const int32_t a0 = r.edges[int64_t(i) * 2], a1 = r.edges[int64_t(i) * 2 + 1];
  const uint64_t base = kFill ? offsets[nv + s] - offsets[nv] : 0;
  // Each unordered pair once: only the leaves after this edge's own leaf s
  // (in tree order) are visited -- in the escape-linked pre-order those are
  // exactly the nodes after leaf s -- and the pair is emitted as (min, max).
  const bool resumed = kFill && slots > 0;  // more hits than cached: resume after the last cached one
  const uint32_t c = traverse(
      enodes, resumed ? resume[s] : enodes[ne - 1 + s].escape, resumed ? uint32_t(slots) : 0u, q, 2 * ne, err,
      [&](int32_t j) {
        const int32_t oj = r.eobj[j];
        if (!pair_allowed(r.allowed, r.nobj, oi, oj)) return false;
        if (oi == oj) {
          const int32_t b0 = r.edges[int64_t(j) * 2], b1 = r.edges[int64_t(j) * 2 + 1];
          if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) return false;
        }
        return true;
      },
      [&](int32_t j, uint32_t k, int32_t next) {
        if (!kFill && int64_t(k) < slots) {
          cache[s * slots + k] = j;
          if (int64_t(k) == slots - 1) resume[s] = next;
        }
        if (kFill && base + k < uint64_t(cap)) {
          out[(base + k) * 2] = i < j ? i : j;
          out[(base + k) * 2 + 1] = i < j ? j : i;
        }
      });
  if (!kFill) counts[s] = c;
}

template <bool kFill>
__global__ __launch_bounds__(kBlock) void ee_query_kernel(EeQuery Q) {
  const int64_t b0 = xcd_block(blockIdx.x, gridDim.x);  // Morton-neighbouring queries on one XCD
  for (int64_t s = b0 * blockDim.x + threadIdx.x; s < Q.ne; s += int64_t(gridDim.x) * blockDim.x)
    ee_query_one<kFill>(Q, s);
}

// The count passes of both trees in one launch (vb vertex-query blocks, eb
// edge-query blocks): each XCD (hardware block i goes to XCD i % 8) gets one
// contiguous range of vertex blocks and then one contiguous range of edge
// blocks, so the smaller vertex pass no longer runs alone on a half-empty
// chip and each XCD's L2 still sees one region of each tree.
__global__ __launch_bounds__(kBlock) void count_both_kernel(VfQuery V, EeQuery E, unsigned vb) {
  const unsigned G = gridDim.x, x = blockIdx.x % kXcds, j = blockIdx.x / kXcds;
  const unsigned per = G / kXcds, rem = G % kXcds, vper = vb / kXcds, vrem = vb % kXcds;
  const unsigned vb_x = vper + (x < vrem ? 1u : 0u);
  if (j < vb_x) {
    const int64_t blk = int64_t(x) * vper + (x < vrem ? x : vrem) + j;
    const int64_t s = blk * kBlock + threadIdx.x;
    if (s < V.nv) vf_query_one<false>(V, s);
  } else {
    const int64_t estart = (int64_t(x) * per + (x < rem ? x : rem)) - (int64_t(x) * vper + (x < vrem ? x : vrem));
    const int64_t blk = estart + (j - vb_x);
    const int64_t s = blk * kBlock + threadIdx.x;
    if (s < E.ne) ee_query_one<false>(E, s);
  }
}

// Fill of the cached pairs (a query's first `slots`): one lane per cache slot
// (query-major, so a wave reads whole cache lines), the pair written at the
// query's scanned offset.  Vertex queries are s < nv, edge queries follow.
__global__ __launch_bounds__(kBlock) void cache_copy_kernel(const int32_t* __restrict__ cache,
                                                            const uint64_t* __restrict__ offsets,
                                                            const uint32_t* __restrict__ qorder,
                                                            const uint32_t* __restrict__ eorder, int64_t nv, int64_t ne,
                                                            int64_t slots, int32_t* __restrict__ vf, int64_t vf_cap,
                                                            int32_t* __restrict__ ee, int64_t ee_cap) {
  const int64_t total = (nv + ne) * slots;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t s = t / slots;
    const uint32_t k = uint32_t(t % slots);
    const uint64_t c = offsets[s + 1] - offsets[s];
    if (k >= c) continue;  // the first min(c, slots) pairs of every query come from the cache
    const int32_t p = cache[t];
    if (s < nv) {
      const uint64_t o = offsets[s] + k;
      if (vf == nullptr || o >= uint64_t(vf_cap)) continue;
      vf[o * 2] = int32_t(qorder[s]);
      vf[o * 2 + 1] = p;
    } else {
      const uint64_t o = offsets[s] - offsets[nv] + k;
      if (ee == nullptr || o >= uint64_t(ee_cap)) continue;
      const int32_t i = int32_t(eorder[s - nv]);
      ee[o * 2] = i < p ? i : p;
      ee[o * 2 + 1] = i < p ? p : i;
    }
  }
}

__global__ void totals_kernel(const uint64_t* __restrict__ offsets, int64_t nv, int64_t ne,
                              const uint32_t* __restrict__ err, uint64_t* __restrict__ totals) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    totals[0] = offsets[nv];
    totals[1] = offsets[nv + ne] - offsets[nv];
    totals[2] = *err;
  }
}

__global__ void zero_u32_kernel(uint32_t* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ---- host-side drivers ----------------------------------------------------
int check_mesh(const p2c_bp_mesh* m) {
  if (m == nullptr) return set_error(P2C_EINVAL, "mesh descriptor is required");
  if (m->n_vertices < 0 || m->n_faces < 0 || m->n_edges < 0)
    return set_error(P2C_EINVAL, "mesh sizes must be non-negative");
  if (m->n_objects < 1 || m->n_objects > kMaxObjects)
    return set_error(P2C_EUNSUPPORTED, "object count must be in [1, 32768] per scene");
  if (m->n_faces > (int64_t(1) << 30) || m->n_edges > (int64_t(1) << 30) || m->n_vertices > (int64_t(1) << 30))
    return set_error(P2C_EUNSUPPORTED, "meshes above 2^30 primitives are not supported");
  if (m->hit_cache_slots < 0 || m->hit_cache_slots > 4096)
    return set_error(P2C_EINVAL, "hit_cache_slots must be in [0, 4096]");
  return P2C_OK;
}

}  // namespace
}  // namespace p2c

extern "C" {

size_t p2c_bp_workspace_bytes(int64_t nv, int64_t nf, int64_t ne, int32_t nobj) {
  return p2c::make_layout(nv, nf, ne, nobj).total;
}

int p2c_bp_workspace_layout(int64_t nv, int64_t nf, int64_t ne, int32_t nobj, p2c_bp_layout* out) {
  if (out == nullptr) return p2c::set_error(P2C_EINVAL, "layout output is required");
  if (nv < 0 || nf < 0 || ne < 0 || nobj < 1) return p2c::set_error(P2C_EINVAL, "mesh sizes must be non-negative");
  const p2c::Layout L = p2c::make_layout(nv, nf, ne, nobj);
  *out = p2c_bp_layout{L.fnodes, L.enodes, L.vbox, L.ebox, L.qorder, L.eorder, L.counts, L.resume};
  return P2C_OK;
}

int p2c_bp_build(const p2c_bp_mesh* m, const double* x0, const double* x1, double thickness, void* ws,
                 size_t ws_bytes, void* stream) {
  using namespace p2c;
  if (int rc = check_mesh(m)) return rc;
  if (!std::isfinite(thickness) || thickness < 0.0)
    return set_error(P2C_EINVAL, "thickness must be finite and non-negative");
  const Layout L = make_layout(m->n_vertices, m->n_faces, m->n_edges, m->n_objects);
  if (ws == nullptr || ws_bytes < L.total) return set_error(P2C_ECAPACITY, "broad-phase workspace is too small");
  hipStream_t s = as_stream(stream);
  const int64_t nv = m->n_vertices, nf = m->n_faces, ne = m->n_edges;
  const int32_t nobj = m->n_objects;
  uint64_t* objkey = at<uint64_t>(ws, L.objkey);
  Box6* vbox = at<Box6>(ws, L.vbox);
  hipLaunchKernelGGL(objkey_init_kernel, dim3(grid_for(int64_t(nobj) * 6, kBlock)), dim3(kBlock), 0, s, objkey, nobj,
                     at<uint32_t>(ws, L.err), at<uint32_t>(ws, L.roots), at<uint32_t>(ws, L.eroots));
  if (nv > 0)
    hipLaunchKernelGGL(vertex_box_kernel, dim3(grid_for(nv, kBlock, 2048)), dim3(kBlock), 0, s, x0, x1, nv, thickness,
                       m->vertex_object, nobj, vbox, objkey);
  hipLaunchKernelGGL(allowed_kernel, dim3(grid_for(int64_t(nobj) * nobj, kBlock)), dim3(kBlock), 0, s, objkey, nobj,
                     at<uint8_t>(ws, L.allowed));
  double* bounds = at<double>(ws, L.bounds);
  hipLaunchKernelGGL(scene_bounds_kernel, dim3(1), dim3(kBlock), 0, s, objkey, nobj, bounds);
  uint32_t* k0 = at<uint32_t>(ws, L.keys0);
  uint32_t* v0 = at<uint32_t>(ws, L.vals0);
  uint32_t* k1 = at<uint32_t>(ws, L.keys1);
  uint32_t* v1 = at<uint32_t>(ws, L.vals1);
  uint32_t* hist = at<uint32_t>(ws, L.hist);
  uint64_t* hist_off = at<uint64_t>(ws, L.hist_off);
  uint64_t* part = at<uint64_t>(ws, L.scan_part);
  struct Tree {
    int kind;
    const int32_t* idx;
    int64_t n;
    size_t box, nodes, parent, flag, rcs, hi, lo, roots;
  } trees[2] = {{0, m->faces, nf, L.fbox, L.fnodes, L.fparent, L.fflag, L.frcs, L.fhi, L.flo, L.roots},
                {1, m->edges, ne, L.ebox, L.enodes, L.eparent, L.eflag, L.ercs, L.ehi, L.elo, L.eroots}};
  TopTree lds_top[2];
  int n_lds_top = 0;
  // boxes and tagged keys of every kind, then one sort: faces at 0, edges at nf, vertices at nf + ne
  const int64_t off_of[3] = {0, nf, nf + ne};
  if (nf + ne + nv > 0) {
    const PrimKeys P{x0, x1, m->faces, m->edges, vbox, bounds, at<Box6>(ws, L.fbox), at<Box6>(ws, L.ebox), k0, v0,
                     nf, ne, nv, thickness};
    hipLaunchKernelGGL(prim_box_kernel, dim3(grid_for(nf + ne + nv, kBlock)), dim3(kBlock), 0, s, P);
    if (int rc = radix_sort(k0, v0, k1, v1, nf + ne + nv, hist, hist_off, part, s)) return rc;
  }
  for (const Tree& t : trees) {
    if (t.n == 0) continue;
    uint32_t* tk = k0 + off_of[t.kind];
    uint32_t* tv = v0 + off_of[t.kind];
    Node* nodes = at<Node>(ws, t.nodes);
    int32_t* parent = at<int32_t>(ws, t.parent);
    uint32_t* flag = at<uint32_t>(ws, t.flag);
    if (t.n > 1)
      hipLaunchKernelGGL(karras_kernel, dim3(grid_for(t.n - 1, kBlock)), dim3(kBlock), 0, s, tk, t.n, nodes, parent,
                         at<int32_t>(ws, t.rcs), at<int32_t>(ws, t.hi), at<int32_t>(ws, t.lo));
    hipLaunchKernelGGL(leaves_kernel, dim3(grid_for(2 * t.n - 1, kBlock)), dim3(kBlock), 0, s, tv, t.n,
                       at<Box6>(ws, t.box), nodes, at<int32_t>(ws, t.rcs), at<int32_t>(ws, t.hi), parent, flag);
    if (t.n > 1) {
      // one block per 256 leaves, exactly (the LDS flags cover the block's leaves)
      if ((t.n + kBlock - 1) / kBlock > (int64_t(1) << 20)) return set_error(P2C_ECAPACITY, "too many primitives");
      uint32_t* roots = at<uint32_t>(ws, t.roots);  // count reset by objkey_init_kernel
      hipLaunchKernelGGL(refit_kernel, dim3(unsigned((t.n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, t.n, nodes,
                         parent, at<int32_t>(ws, t.lo), at<int32_t>(ws, t.hi), roots);
      if (t.n - 1 <= int64_t(kTopLdsWords) * 32)
        lds_top[n_lds_top++] = TopTree{nodes, parent, roots, t.n - 1};  // climbed below, both trees at once
      else
        hipLaunchKernelGGL(refit_top_kernel, dim3(1), dim3(kTopBlock), 0, s, nodes, parent, roots, flag);
    }
    // the edge leaf order (sorted edge ids) is also the EE query order
    if (t.kind == 1)
      P2C_HIP_CHECK(hipMemcpyAsync(at<uint32_t>(ws, L.eorder), tv, sizeof(uint32_t) * t.n, hipMemcpyDeviceToDevice, s));
  }
  if (n_lds_top > 0)
    hipLaunchKernelGGL(refit_top_lds_kernel, dim3(unsigned(n_lds_top)), dim3(kTopBlock), 0, s, lds_top[0],
                       n_lds_top > 1 ? lds_top[1] : lds_top[0]);
  // vertex query order: vertices sorted by Morton code of their swept box
  if (nv > 0)
    P2C_HIP_CHECK(hipMemcpyAsync(at<uint32_t>(ws, L.qorder), v0 + off_of[2], sizeof(uint32_t) * nv,
                                 hipMemcpyDeviceToDevice, s));
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

int p2c_bp_count(const p2c_bp_mesh* m, void* ws, uint64_t* totals, void* stream) {
  using namespace p2c;
  if (int rc = check_mesh(m)) return rc;
  const Layout L = make_layout(m->n_vertices, m->n_faces, m->n_edges, m->n_objects);
  hipStream_t s = as_stream(stream);
  const int64_t nv = m->n_vertices, nf = m->n_faces, ne = m->n_edges;
  uint32_t* counts = at<uint32_t>(ws, L.counts);
  uint64_t* offsets = at<uint64_t>(ws, L.offsets);
  const int64_t slots = m->hit_cache != nullptr ? m->hit_cache_slots : 0;
  VfRule vr{m->faces, m->face_object, m->vertex_object, at<uint8_t>(ws, L.allowed), m->n_objects};
  EeRule er{m->edges, m->edge_object, at<uint8_t>(ws, L.allowed), m->n_objects};
  const VfQuery V{at<Node>(ws, L.fnodes), nf, at<Box6>(ws, L.vbox), at<uint32_t>(ws, L.qorder), nv, vr,
                  at<uint32_t>(ws, L.err), counts, m->hit_cache, slots, at<int32_t>(ws, L.resume), nullptr,
                  nullptr, 0};
  const EeQuery E{at<Node>(ws, L.enodes), ne, at<Box6>(ws, L.ebox), at<uint32_t>(ws, L.eorder), er,
                  at<uint32_t>(ws, L.err), counts + nv, m->hit_cache + nv * slots, slots,
                  at<int32_t>(ws, L.resume) + nv, nullptr, nv, nullptr, 0};
  const int64_t vb = (nv + kBlock - 1) / kBlock, eb = (ne + kBlock - 1) / kBlock;
  if (vb > 0 && eb > 0 && vb + eb <= (int64_t(1) << 20)) {
    hipLaunchKernelGGL(count_both_kernel, dim3(unsigned(vb + eb)), dim3(kBlock), 0, s, V, E, unsigned(vb));
  } else {
    if (nv > 0) hipLaunchKernelGGL(vf_query_kernel<false>, dim3(grid_for(nv, kBlock)), dim3(kBlock), 0, s, V);
    if (ne > 0) hipLaunchKernelGGL(ee_query_kernel<false>, dim3(grid_for(ne, kBlock)), dim3(kBlock), 0, s, E);
  }
  if (int rc = exclusive_scan_u32(counts, nv + ne, offsets, at<uint64_t>(ws, L.scan_part), s)) return rc;
  hipLaunchKernelGGL(totals_kernel, dim3(1), dim3(64), 0, s, offsets, nv, ne, at<uint32_t>(ws, L.err), totals);
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

int p2c_bp_emit(const p2c_bp_mesh* m, void* ws, int32_t* vf_pairs, int64_t vf_cap, int32_t* ee_pairs,
                int64_t ee_cap, void* stream) {
  using namespace p2c;
  if (int rc = check_mesh(m)) return rc;
  const Layout L = make_layout(m->n_vertices, m->n_faces, m->n_edges, m->n_objects);
  hipStream_t s = as_stream(stream);
  const int64_t nv = m->n_vertices, nf = m->n_faces, ne = m->n_edges;
  const uint64_t* offsets = at<uint64_t>(ws, L.offsets);
  const int64_t slots = m->hit_cache != nullptr ? m->hit_cache_slots : 0;
  VfRule vr{m->faces, m->face_object, m->vertex_object, at<uint8_t>(ws, L.allowed), m->n_objects};
  EeRule er{m->edges, m->edge_object, at<uint8_t>(ws, L.allowed), m->n_objects};
  if (nv > 0 && vf_pairs != nullptr)
    hipLaunchKernelGGL(vf_query_kernel<true>, dim3(grid_for(nv, kBlock)), dim3(kBlock), 0, s,
                       VfQuery{at<Node>(ws, L.fnodes), nf, at<Box6>(ws, L.vbox), at<uint32_t>(ws, L.qorder), nv, vr,
                               at<uint32_t>(ws, L.err), nullptr, nullptr, slots, at<int32_t>(ws, L.resume), offsets,
                               vf_pairs, vf_cap});
  if (ne > 0 && ee_pairs != nullptr)
    hipLaunchKernelGGL(ee_query_kernel<true>, dim3(grid_for(ne, kBlock)), dim3(kBlock), 0, s,
                       EeQuery{at<Node>(ws, L.enodes), ne, at<Box6>(ws, L.ebox), at<uint32_t>(ws, L.eorder), er,
                               at<uint32_t>(ws, L.err), nullptr, nullptr, slots, at<int32_t>(ws, L.resume) + nv,
                               offsets, nv, ee_pairs, ee_cap});
  if (slots > 0 && nv + ne > 0 && (vf_pairs != nullptr || ee_pairs != nullptr))
    hipLaunchKernelGGL(cache_copy_kernel, dim3(grid_for((nv + ne) * slots, kBlock, 1 << 16)), dim3(kBlock), 0, s,
                       m->hit_cache, offsets, at<uint32_t>(ws, L.qorder), at<uint32_t>(ws, L.eorder), nv, ne, slots,
                       vf_pairs, vf_cap, ee_pairs, ee_cap);
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

}  // extern "C"
