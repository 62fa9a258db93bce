// Swept-AABB LBVH broad phase on gfx950: vertex-face and edge-edge candidate
// pairs with the reference's set semantics.  Types, workspace layout and the
// per-item / per-block bodies of every stage, shared by the one-scene driver
// (lbvh.hip) and the batched driver (lbvh_batch.hip), so that a scene gives
// the same trees and the same pairs in either.
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
#pragma once

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
// A scene's workspace.  `resident` leaves out what a batch keeps in its own
// workspace instead (sort keys, histograms, per-query counts, resume nodes
// and offsets): a resident scene holds its boxes, envelopes, object-pair
// matrix, both trees and its query orders.
struct Layout {
  size_t roots, vbox, fbox, ebox, objkey, bounds, allowed, keys0, vals0, keys1, vals1, hist, fnodes, fparent, fflag, frcs,
      fhi, flo, enodes, eparent, eflag, ercs, ehi, elo, qorder, eorder, counts, resume, offsets, hist_off, scan_part, err,
      eroots, total;
  int64_t sort_cap, hist_blocks;
};

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

inline Layout make_layout(int64_t nv, int64_t nf, int64_t ne, int32_t nobj, bool resident = false) {
  Layout L{};
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o = align256(o + (bytes ? bytes : 1));
    return at;
  };
  auto own = [&](size_t bytes) { return resident ? size_t(0) : take(bytes); };
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
  L.keys0 = own(4 * sall);
  L.vals0 = own(4 * sall);
  L.keys1 = own(4 * sall);
  L.vals1 = own(4 * sall);
  L.hist = own(4 * 256 * L.hist_blocks);
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
  L.counts = own(4 * (nv + ne + 1));
  L.resume = own(4 * (nv + ne + 1));
  L.offsets = own(8 * (nv + ne + 1));
  L.hist_off = own(8 * 256 * L.hist_blocks);
  const int64_t scan_items = std::max<int64_t>(nv + ne + 1, 256 * L.hist_blocks);
  L.scan_part = own(8 * (scan_items / 1024 + 2));
  L.err = take(16);
  L.roots = take(4 * (size_t(sn) + 2));   // [0] count, then the nodes whose parent spans blocks (face tree)
  L.eroots = take(4 * (size_t(sn) + 2));  // the same for the edge tree
  L.total = o;
  return L;
}

template <typename T>
T* at(void* ws, size_t off) {
  return reinterpret_cast<T*>(static_cast<char*>(ws) + off);
}

// ---- 1. boxes and object envelopes -------------------------------------
// Item i of the reset: the object envelopes (6 words per object), the error
// counters (items 0-3) and the counts of both trees' top-refit root lists
// (items 4 and 5); items [0, max(6 nobj, 6)).
__device__ inline void objkey_init_item(uint64_t* objkey, int32_t nobj, uint32_t* err, uint32_t* froots,
                                        uint32_t* eroots, int64_t i) {
  if (i < int64_t(nobj) * 6) objkey[i] = (i % 6) < 3 ? ~0ull : 0ull;
  if (i < 4) err[i] = 0u;
  if (i == 4) froots[0] = 0u;
  if (i == 5) eroots[0] = 0u;
}

// Vertex boxes and object envelopes over the vertices of block b of g (a
// grid-stride loop; every thread of the block calls it).  Up to kLdsObjects
// objects the envelopes are reduced in LDS (lkey, 6 * kLdsObjects words),
// then one global atomic per word and workgroup; more objects: global
// atomics per vertex.  Min/max in any order is exact, so the envelopes do not
// depend on how the vertices are split into blocks.
__device__ inline void vertex_box_block(const double* __restrict__ x0, const double* __restrict__ x1, int64_t nv,
                                        double th, const int32_t* __restrict__ vobj, int32_t nobj,
                                        Box6* __restrict__ vbox, uint64_t* __restrict__ objkey, int64_t b, int64_t g,
                                        unsigned long long* lkey) {
  const bool lds = nobj <= kLdsObjects;
  unsigned long long* key = lds ? lkey : reinterpret_cast<unsigned long long*>(objkey);
  if (lds)
    for (int i = threadIdx.x; i < nobj * 6; i += blockDim.x) lkey[i] = (i % 6) < 3 ? ~0ull : 0ull;
  __syncthreads();
  for (int64_t v = b * blockDim.x + threadIdx.x; v < nv; v += g * blockDim.x) {
    Box6 bx;
    box_reset(bx);
    box_extend(bx, x0 + v * 3);
    box_extend(bx, x1 + v * 3);
    box_expand(bx, th);
    vbox[v] = bx;
    const int32_t o = vobj[v];
    if (o >= 0 && o < nobj) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        atomicMin(&key[int64_t(o) * 6 + k], (unsigned long long)f64_order_key(bx.mn[k]));
        atomicMax(&key[int64_t(o) * 6 + 3 + k], (unsigned long long)f64_order_key(bx.mx[k]));
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

// Entry i = a * nobj + b of the object-pair matrix: 1 when a == b or the
// envelopes of a and b overlap.
__device__ inline void allowed_item(const uint64_t* __restrict__ objkey, int32_t nobj, uint8_t* __restrict__ allowed,
                                    int64_t i) {
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
// axis), reduced by one workgroup of kBlock threads into bounds[6] (red: LDS
// of 6 * kBlock doubles).  Its min/max over the same values in any order is
// exact, so the frame does not depend on the reduction order.
__device__ inline void scene_bounds_block(const uint64_t* __restrict__ objkey, int32_t nobj, double* __restrict__ bounds,
                                          double (*red)[kBlock]) {
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
// vertex keys ([nf + ne, nf + ne + nv)); values are the index within the
// kind.  All three key sets are sorted together, tagged with their kind in
// bits 30-31 above the 30-bit Morton code: a stable sort keeps each kind
// contiguous, in the order (and with the tie order) a sort of that kind
// alone would give.
struct PrimKeys {
  const double* x0;
  const double* x1;
  const int32_t* faces;
  const int32_t* edges;
  const Box6* vbox;
  const double* bounds;  // lo[3], inv[3] of scene_bounds_block
  Box6* fbox;
  Box6* ebox;
  uint32_t* keys;
  uint32_t* vals;
  int64_t nf, ne, nv;
  double th;
};

__device__ inline void prim_item(const PrimKeys& P, const double* lo, const double* inv, int64_t g) {
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

// ---- 3. Karras build, escapes, refit ------------------------------------
__device__ inline int delta(const uint32_t* keys, int64_t n, int64_t i, int64_t j) {
  if (j < 0 || j >= n) return -1;
  const uint64_t a = (uint64_t(keys[i]) << 32) | uint64_t(i);
  const uint64_t b = (uint64_t(keys[j]) << 32) | uint64_t(j);
  return __clzll(a ^ b);
}

// Internal node i in [0, n - 1) of the tree over the n sorted keys.
__device__ inline void karras_node(const uint32_t* __restrict__ keys, int64_t n, int64_t i, Node* __restrict__ nodes,
                                   int32_t* __restrict__ parent, int32_t* __restrict__ rcs,
                                   int32_t* __restrict__ hi_of, int32_t* __restrict__ lo_of) {
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

// Node p in [0, 2n - 1): internal nodes get their escape (from their range
// maximum) and a cleared refit flag; leaves their box, prim and escape.
__device__ inline void leaf_item(const uint32_t* __restrict__ vals, int64_t n, const Box6* __restrict__ box,
                                 Node* __restrict__ nodes, const int32_t* __restrict__ rcs,
                                 const int32_t* __restrict__ hi_of, int32_t* __restrict__ parent,
                                 uint32_t* __restrict__ flag, int64_t p) {
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

__device__ inline void refit_from(Node* __restrict__ nodes, int32_t node) {
  const Node& a = nodes[nodes[node].left];
  const Node& b = nodes[nodes[node].right];
  for (int k = 0; k < 3; ++k) {
    const double amn = a.mn[k], bmn = b.mn[k];
    const double amx = a.mx[k], bmx = b.mx[k];
    nodes[node].mn[k] = bmn < amn ? bmn : amn;
    nodes[node].mx[k] = amx < bmx ? bmx : amx;
  }
}

// Refit in two launches, so that no step needs coherence across the chip's
// separate L2s.  1. One block per 256 leaves refits every internal node
// whose leaf range lies in the block (in-block counters in LDS, workgroup
// scope); a thread that reaches a node spanning blocks records the child it
// came from.  2. One workgroup climbs from those children through the nodes
// that span blocks, with per-node counters at workgroup scope.  An internal
// node's leaf range holds its own index (Karras), so the in-block counters
// are indexed by node - first leaf of the block.  The first child to arrive
// stops; the second sees both children's boxes and builds the parent's.
// Block blk of the first launch (local_flag: LDS of kBlock words).
__device__ inline void refit_block(int64_t n, Node* __restrict__ nodes, const int32_t* __restrict__ parent,
                                   const int32_t* __restrict__ lo_of, const int32_t* __restrict__ hi_of,
                                   uint32_t* __restrict__ roots, int64_t blk, uint32_t* local_flag) {
  local_flag[threadIdx.x] = 0;
  __syncthreads();
  const int64_t b0 = blk * kBlock;
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
    refit_from(nodes, node);
    child = node;
    node = parent[node];
  }
}

constexpr int kTopBlock = 1024;
// The climb's arrival flags as bits in LDS (160 KB, a gfx950 CU's whole LDS:
// trees of up to 1,310,720 internal nodes): an LDS atomic per level instead
// of an L2 one.
constexpr int kTopLdsWords = 40960;

// Second launch, one workgroup of kTopBlock threads per tree: the climb with
// per-node counters in global memory (flag, cleared by leaf_item).
// Workgroup-scope ordering covers the child boxes written by the other lanes
// of this workgroup.
__device__ inline void top_climb_global(Node* __restrict__ nodes, const int32_t* __restrict__ parent,
                                        const uint32_t* __restrict__ roots, uint32_t* __restrict__ flag) {
  const uint32_t total = roots[0];
  for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
    int32_t node = parent[roots[2 + i]];
    while (node >= 0) {
      const uint32_t prev = __hip_atomic_fetch_add(&flag[node], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (prev == 0) break;
      refit_from(nodes, node);
      node = parent[node];
    }
  }
}

// The same climb with the flags as bits in LDS (bits: kTopLdsWords words);
// n_internal <= 32 * kTopLdsWords.
__device__ inline void top_climb_lds(Node* __restrict__ nodes, const int32_t* __restrict__ parent,
                                     const uint32_t* __restrict__ roots, int64_t n_internal, uint32_t* bits) {
  const int words = int((n_internal + 31) / 32);
  for (int i = threadIdx.x; i < words; i += blockDim.x) bits[i] = 0;
  __syncthreads();
  const uint32_t total = roots[0];
  for (uint32_t i = threadIdx.x; i < total; i += blockDim.x) {
    int32_t node = parent[roots[2 + i]];
    while (node >= 0) {
      const uint32_t bit = 1u << (node & 31);
      const uint32_t prev =
          __hip_atomic_fetch_or(&bits[node >> 5], bit, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (!(prev & bit)) break;
      refit_from(nodes, node);
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
