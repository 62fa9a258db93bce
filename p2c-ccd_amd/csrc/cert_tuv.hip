// (t,u,v) inclusion certificate on gfx950 (FP64, no contraction).
//
//   setup       per pair: positions -> TuvPair (error bound, tolerances,
//               split sequence); from a (n,8,3) query array or gathered from
//               mesh vertex arrays through 4 vertex indices per pair.
//   plain       one lane per pair, depth-first walk to completion.  Pairs
//               whose walk exceeds max_itr boxes go to the replay list.
//   persistent  a grid sized to the chip; every lane walks one pair box by
//               box and pulls the next pair the moment its pair finishes
//               (wave-aggregated atomics), so uneven subdivision depth does
//               not idle the other 63 lanes of a wave.
//   replay      the fallback-replay queue: pairs over budget are re-run in
//               the canonical best-first order with a global-memory heap, the
//               order that defines the result when the cap binds.
//
// Replaces the ticcd::vertexFaceCCD / edgeEdgeCCD call sites of the
// reference (src/tools/tight_inclusion_full_query_benchmark.cpp:375-397;
// src/tools/scene_object_envelope_native_ti_walltime.cpp:704-754).
#include <cmath>
#include <mutex>

#include "cert_tuv_math.h"
#include "p2c_hip.h"
#include "runtime.h"

namespace p2c {
namespace {

static_assert(sizeof(TuvPair) == P2C_TUV_PAIR_BYTES, "TuvPair layout");
static_assert(sizeof(p2c_tuv_result) == 32, "p2c_tuv_result layout");

constexpr int kTuvBlock = 256;

// Pair records are stored in tiles of 64 (AoSoA): the 16-byte granule g of
// pair p sits at granule (p / 64) * 18 * 64 + g * 64 + p % 64, so the 64
// lanes of a wave that write (setup) or read (refill) consecutive pairs touch
// 1 KB of contiguous memory per instruction instead of 64 scattered 16-byte
// pieces.  A buffer for n pairs holds ceil(n / 64) * 64 records.
constexpr int kPairTile = 64;
constexpr int kPairGranules = int(sizeof(TuvPair) / 16);
static_assert(sizeof(TuvPair) % 16 == 0, "TuvPair is a whole number of 16-byte granules");

struct alignas(16) Granule {
  uint32_t w[4];
};

__device__ inline const Granule* pair_granules(const TuvPair* base, int64_t p) {
  return reinterpret_cast<const Granule*>(base) + (p / kPairTile) * (int64_t(kPairGranules) * kPairTile) +
         (p % kPairTile);
}

__device__ inline TuvPair load_pair(const TuvPair* __restrict__ base, int64_t p) {
  const Granule* src = pair_granules(base, p);
  Granule g[kPairGranules];
#pragma unroll
  for (int i = 0; i < kPairGranules; ++i) g[i] = src[int64_t(i) * kPairTile];
  TuvPair q;
  __builtin_memcpy(&q, g, sizeof(TuvPair));
  return q;
}

__device__ inline void store_pair(TuvPair* __restrict__ base, int64_t p, const TuvPair& q) {
  Granule g[kPairGranules];
  __builtin_memcpy(g, &q, sizeof(TuvPair));
  Granule* dst = const_cast<Granule*>(pair_granules(base, p));
#pragma unroll
  for (int i = 0; i < kPairGranules; ++i) dst[int64_t(i) * kPairTile] = g[i];
}

// Setup's store: the records are read back once, by the certificate kernels,
// long after (the whole buffer exceeds the caches), so the writes stream past
// the caches (non-temporal).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ inline void store_pair_streaming(TuvPair* __restrict__ base, int64_t p, const TuvPair& q) {
  Granule g[kPairGranules];
  __builtin_memcpy(g, &q, sizeof(TuvPair));
  Granule* dst = const_cast<Granule*>(pair_granules(base, p));
#pragma unroll
  for (int i = 0; i < kPairGranules; ++i) {
    const u32x4 v = {g[i].w[0], g[i].w[1], g[i].w[2], g[i].w[3]};
    __builtin_nontemporal_store(v, reinterpret_cast<u32x4*>(dst + int64_t(i) * kPairTile));
  }
}

__device__ inline void write_result(p2c_tuv_result* out, uint64_t best, uint64_t boxes, uint32_t flags) {
  p2c_tuv_result r;
  r.hit = best != kTuvNoHit ? 1u : 0u;
  r.toi = best != kTuvNoHit ? key_to_f64(best) : __builtin_inf();
  r.boxes = boxes;
  r.flags = flags;
  r.reserved = 0;
  *out = r;
}

// Compact record of a pair given by vertex ids (64 bytes, 4 granules, the
// same 64-pair tiles): the ids and the split sequence.  Positions,
// displacements and error bounds are rebuilt from the vertex arrays when a
// kernel loads the pair; a mesh's arrays stay resident in L2 while its pairs
// stream, so a refill reads 64 bytes from HBM instead of 288.
struct TuvLite {
  int32_t vid[4];
  uint64_t tmask[2];
  uint64_t umask[2];
  uint64_t split_hi;
  uint32_t done_level;
  uint32_t is_ee;
};
static_assert(sizeof(TuvLite) == P2C_TUV_LITE_BYTES, "TuvLite layout");
constexpr int kLiteGranules = int(sizeof(TuvLite) / 16);

__device__ inline const Granule* lite_granules(const TuvLite* base, int64_t p) {
  return reinterpret_cast<const Granule*>(base) + (p / kPairTile) * (int64_t(kLiteGranules) * kPairTile) +
         (p % kPairTile);
}

__device__ inline void store_lite(TuvLite* __restrict__ base, int64_t p, const TuvLite& r) {
  Granule g[kLiteGranules];
  __builtin_memcpy(g, &r, sizeof(TuvLite));
  Granule* dst = const_cast<Granule*>(lite_granules(base, p));
#pragma unroll
  for (int i = 0; i < kLiteGranules; ++i) {
    const u32x4 v = {g[i].w[0], g[i].w[1], g[i].w[2], g[i].w[3]};
    __builtin_nontemporal_store(v, reinterpret_cast<u32x4*>(dst + int64_t(i) * kPairTile));
  }
}

// Where the certificate kernels take pair p from: full records (FullSrc) or
// compact records expanded from the vertex arrays (LiteSrc).  Both give the
// same TuvPair (the split sequence was computed from the same positions),
// so results do not depend on the record form.
struct FullSrc {
  const TuvPair* pairs;
  __device__ TuvPair load(int64_t p) const { return load_pair(pairs, p); }
};

struct LiteSrc {
  const TuvLite* pairs;
  const double* x0;
  const double* x1;
  double ms;
  __device__ TuvPair load(int64_t p) const {
    const Granule* src = lite_granules(pairs, p);
    Granule g[kLiteGranules];
#pragma unroll
    for (int i = 0; i < kLiteGranules; ++i) g[i] = src[int64_t(i) * kPairTile];
    TuvLite r;
    __builtin_memcpy(&r, g, sizeof(TuvLite));
    double v[8][3];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t o = int64_t(r.vid[i]) * 3;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        v[i][k] = x0[o + k];
        v[4 + i][k] = x1[o + k];
      }
    }
    TuvPair q;
    tuv_setup_geometry(&q, v, int(r.is_ee), ms);
    q.inv_tol[0] = q.inv_tol[1] = q.inv_tol[2] = 0.0;  // used only to build the split sequence
    q.tmask[0] = r.tmask[0];
    q.tmask[1] = r.tmask[1];
    q.umask[0] = r.umask[0];
    q.umask[1] = r.umask[1];
    q.split_hi = r.split_hi;
    q.done_level = r.done_level;
    return q;
  }
};

__global__ __launch_bounds__(kTuvBlock) void tuv_setup_positions_kernel(int is_ee, const double* __restrict__ v,
                                                                        int64_t n, double ms, double tol,
                                                                        TuvPair* __restrict__ pairs) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    double x[8][3];
    for (int r = 0; r < 8; ++r)
      for (int k = 0; k < 3; ++k) x[r][k] = v[i * 24 + r * 3 + k];
    TuvPair q;
    tuv_setup(&q, x, is_ee, ms, tol);
    store_pair(pairs, i, q);
  }
}

__global__ __launch_bounds__(kTuvBlock) void tuv_setup_indexed_kernel(int is_ee, const double* __restrict__ x0,
                                                                      const double* __restrict__ x1,
                                                                      const int32_t* __restrict__ idx4, int64_t n,
                                                                      double ms, double tol,
                                                                      TuvPair* __restrict__ pairs) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    double x[8][3];
    for (int r = 0; r < 4; ++r) {
      const int64_t vi = idx4[i * 4 + r];
      for (int k = 0; k < 3; ++k) {
        x[r][k] = x0[vi * 3 + k];
        x[4 + r][k] = x1[vi * 3 + k];
      }
    }
    TuvPair q;
    tuv_setup(&q, x, is_ee, ms, tol);
    store_pair(pairs, i, q);
  }
}

// Pair count of a launch: the host bound, or min(bound, device totals) when
// the counts were produced on the device (VF total, EE total).
__device__ inline int64_t device_count(int64_t n, const uint64_t* totals) {
  if (totals == nullptr) return n;
  const int64_t t = int64_t(totals[0] + totals[1]);
  return t < n ? t : n;
}

// VF pair (v, f) -> (v, f0, f1, f2); EE pair (i, j) -> (i0, i1, j0, j1).
// With device totals, VF writes slots [0, n_vf) and EE [n_vf, n_vf + n_ee).
__global__ __launch_bounds__(kTuvBlock) void tuv_setup_pairs_kernel(int is_ee, const double* __restrict__ x0,
                                                                    const double* __restrict__ x1,
                                                                    const int32_t* __restrict__ topo,
                                                                    const int32_t* __restrict__ pairs, int64_t n,
                                                                    const uint64_t* __restrict__ totals,
                                                                    int64_t out_offset, int64_t out_cap, double ms,
                                                                    double tol, TuvPair* __restrict__ out) {
  int64_t m = n, off = out_offset;
  if (totals != nullptr) {
    const int64_t t = int64_t(totals[is_ee]);
    m = t < n ? t : n;
    off = is_ee ? int64_t(totals[0]) : 0;
  }
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t a = pairs[i * 2], b = pairs[i * 2 + 1];
    int64_t vid[4];
    if (!is_ee) {
      vid[0] = a;
      vid[1] = topo[b * 3];
      vid[2] = topo[b * 3 + 1];
      vid[3] = topo[b * 3 + 2];
    } else {
      vid[0] = topo[a * 2];
      vid[1] = topo[a * 2 + 1];
      vid[2] = topo[b * 2];
      vid[3] = topo[b * 2 + 1];
    }
    double x[8][3];
    for (int r = 0; r < 4; ++r)
      for (int k = 0; k < 3; ++k) {
        x[r][k] = x0[vid[r] * 3 + k];
        x[4 + r][k] = x1[vid[r] * 3 + k];
      }
    TuvPair q;
    tuv_setup(&q, x, is_ee, ms, tol);
    if (off + i < out_cap) store_pair_streaming(out, off + i, q);
  }
}

// The compact record of a pair whose four vertices are rows vid of x0 / x1
// (VF: vertex, then the face's three; EE: edge a's two, then edge b's).
__device__ inline TuvLite lite_record(const double* __restrict__ x0, const double* __restrict__ x1,
                                      const int32_t vid[4], int is_ee, double ms, double tol) {
  double x[8][3];
  for (int r = 0; r < 4; ++r)
    for (int k = 0; k < 3; ++k) {
      x[r][k] = x0[int64_t(vid[r]) * 3 + k];
      x[4 + r][k] = x1[int64_t(vid[r]) * 3 + k];
    }
  TuvPair q;
  tuv_setup(&q, x, is_ee, ms, tol);
  TuvLite r;
  for (int j = 0; j < 4; ++j) r.vid[j] = vid[j];
  for (int w = 0; w < 2; ++w) {
    r.tmask[w] = q.tmask[w];
    r.umask[w] = q.umask[w];
  }
  r.split_hi = q.split_hi;
  r.done_level = q.done_level;
  r.is_ee = q.is_ee;
  return r;
}

// Vertex ids of pair (a, b) through the topology, plus `base`.
__device__ inline void pair_vertices(int is_ee, const int32_t* __restrict__ topo, int64_t a, int64_t b, int64_t base,
                                     int32_t vid[4]) {
  if (!is_ee) {
    vid[0] = int32_t(a + base);
    vid[1] = int32_t(topo[b * 3] + base);
    vid[2] = int32_t(topo[b * 3 + 1] + base);
    vid[3] = int32_t(topo[b * 3 + 2] + base);
  } else {
    vid[0] = int32_t(topo[a * 2] + base);
    vid[1] = int32_t(topo[a * 2 + 1] + base);
    vid[2] = int32_t(topo[b * 2] + base);
    vid[3] = int32_t(topo[b * 2 + 1] + base);
  }
}

// The same pairs as compact records.
__global__ __launch_bounds__(kTuvBlock) void tuv_setup_lite_kernel(int is_ee, const double* __restrict__ x0,
                                                                   const double* __restrict__ x1,
                                                                   const int32_t* __restrict__ topo,
                                                                   const int32_t* __restrict__ pairs, int64_t n,
                                                                   const uint64_t* __restrict__ totals,
                                                                   int64_t out_offset, int64_t out_cap, double ms,
                                                                   double tol, TuvLite* __restrict__ out) {
  int64_t m = n, off = out_offset;
  if (totals != nullptr) {
    const int64_t t = int64_t(totals[is_ee]);
    m = t < n ? t : n;
    off = is_ee ? int64_t(totals[0]) : 0;
  }
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += int64_t(gridDim.x) * blockDim.x) {
    int32_t vid[4];
    pair_vertices(is_ee, topo, pairs[i * 2], pairs[i * 2 + 1], 0, vid);
    const TuvLite r = lite_record(x0, x1, vid, is_ee, ms, tol);
    if (off + i < out_cap) store_lite(out, off + i, r);
  }
}

// A batch of scenes' pair list: pair i belongs to the part j of the last
// offsets[j] <= i (scene j / 2; VF for even j, EE for odd j), its vertex ids
// move to the scene's rows of the batch position arrays.
__global__ __launch_bounds__(kTuvBlock) void tuv_setup_lite_batch_kernel(const double* __restrict__ x0,
                                                                         const double* __restrict__ x1,
                                                                         const p2c_tuv_scene* __restrict__ sc, int S,
                                                                         const int32_t* __restrict__ pairs,
                                                                         const uint64_t* __restrict__ offsets,
                                                                         int64_t n, double ms, double tol,
                                                                         TuvLite* __restrict__ out) {
  const int64_t total = int64_t(offsets[2 * S]);
  const int64_t m = total < n ? total : n;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += int64_t(gridDim.x) * blockDim.x) {
    int lo = 0, hi = 2 * S - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (offsets[mid] <= uint64_t(i)) lo = mid;
      else hi = mid - 1;
    }
    const int is_ee = lo & 1;
    const p2c_tuv_scene& d = sc[lo >> 1];
    int32_t vid[4];
    pair_vertices(is_ee, is_ee ? d.edges : d.faces, pairs[i * 2], pairs[i * 2 + 1], d.vertex_base, vid);
    store_lite(out, i, lite_record(x0, x1, vid, is_ee, ms, tol));
  }
}

__device__ inline void flag_for_replay(uint32_t pair, uint32_t* list, uint32_t* count, uint32_t cap) {
  const uint32_t slot = atomicAdd(count, 1u);
  if (slot < cap) list[slot] = pair;
}

template <typename Src>
__global__ __launch_bounds__(kTuvBlock) void tuv_plain_kernel(Src src, int64_t n_host,
                                                              const uint64_t* __restrict__ totals,
                                                              double tol, uint64_t max_itr,
                                                              p2c_tuv_result* __restrict__ out,
                                                              uint32_t* __restrict__ replay_list,
                                                              uint32_t* __restrict__ replay_count,
                                                              uint32_t replay_cap) {
  const int64_t n = device_count(n_host, totals);
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const TuvPair q = src.load(i);
    uint64_t best = kTuvNoHit, boxes = 0;
    const bool complete = tuv_walk(q, 0, 0, 0, 0, tol, &best, &boxes, max_itr);
    write_result(&out[i], best, boxes, complete ? 0u : P2C_TUV_FLAG_OVER_BUDGET);
    if (!complete) flag_for_replay((uint32_t)i, replay_list, replay_count, replay_cap);
  }
}

// One lane's resumable walk over a pair's box tree.  `ep` packs the box's
// exponents (e_t | e_u << 8 | e_v << 16; its widths are 2^-e).  `ready`: the
// lane stands on a box still to be evaluated that passed box_skipped;
// otherwise it is part-way through moving to the next such box.
struct LaneWalk {
  uint64_t key[3];
  uint64_t best;
  uint32_t boxes;
  uint32_t ep;
  uint32_t level;
  bool ready;
};

__device__ inline int exp_of(uint32_t ep, int d) { return (int)((ep >> (8 * d)) & 0xffu); }

__device__ inline uint32_t pack_exponents(const int e[3]) {
  return (uint32_t)e[0] | ((uint32_t)e[1] << 8) | ((uint32_t)e[2] << 16);
}

__device__ inline void start_walk(LaneWalk& w) {
  w.key[0] = w.key[1] = w.key[2] = 0;
  w.best = kTuvNoHit;
  w.boxes = 0;
  w.ep = 0;
  w.level = 0;
  w.ready = true;
}

// True when the box (and with it its whole subtree) needs no evaluation.
__device__ inline bool box_skipped(const TuvPair& q, const LaneWalk& w) {
  return (!q.is_ee && w.key[1] + w.key[2] > kOne63) || !(w.key[0] < w.best);
}

// Evaluates the box the lane stands on: 0 reject, 1 done, 2 split.
template <typename W>
__device__ inline int evaluate(const TuvPair& q, const W& w, double tol) {
  const int e[3] = {exp_of(w.ep, 0), exp_of(w.ep, 1), exp_of(w.ep, 2)};
  return tuv_classify_at(q, key_to_f64(w.key[0]), key_to_f64(w.key[1]), key_to_f64(w.key[2]), e, w.level, tol);
}

// Moves to the lower child (same lower corner, so it passes box_skipped too).
__device__ inline void descend(const TuvPair& q, LaneWalk& w) {
  w.ep += 1u << (8 * tuv_level_dim(q, w.level));
  ++w.level;
}

// One move of the pre-order walk towards the next box after the current
// box's subtree: climbs one level out of an upper child, or steps from a
// lower child to its upper sibling (ready unless skipped; a skipped sibling
// is climbed out of by the next move).  False when the subtree rooted at
// level `root` is exhausted.  Branch-free apart from the exit.
__device__ inline bool advance_once(const TuvPair& q, LaneWalk& w, uint32_t root) {
  if (w.level <= root) return false;
  const int d = tuv_level_dim(q, w.level - 1);
  const uint64_t bit = kOne63 >> exp_of(w.ep, d);
  const uint64_t b0 = d == 0 ? bit : 0ull, b1 = d == 1 ? bit : 0ull, b2 = d == 2 ? bit : 0ull;
  const bool upper = ((w.key[0] & b0) | (w.key[1] & b1) | (w.key[2] & b2)) != 0ull;
  w.key[0] ^= b0;
  w.key[1] ^= b1;
  w.key[2] ^= b2;
  w.ep -= upper ? (1u << (8 * d)) : 0u;
  w.level -= upper ? 1u : 0u;
  w.ready = !upper && !box_skipped(q, w);
  return true;
}

// Moves until ready; false when the subtree rooted at `root` is exhausted.
__device__ inline bool walk_advance(const TuvPair& q, LaneWalk& w, uint32_t root) {
  while (!w.ready)
    if (!advance_once(q, w, root)) return false;
  return true;
}

// One step of a lane's walk: evaluates the box it stands on (when ready),
// then makes at most kAdvanceMoves moves towards the next box, so one lane's
// long climb costs the wave a few cheap moves per step rather than a serial
// loop.  Returns true when the walk has finished; sets *over (leaving the box
// unevaluated) when the box budget ran out.
#ifndef P2C_ADVANCE_MOVES
#define P2C_ADVANCE_MOVES 2
#endif
constexpr int kAdvanceMoves = P2C_ADVANCE_MOVES;

__device__ inline bool lane_step(const TuvPair& q, LaneWalk& w, double tol, uint64_t budget, bool* over) {
  if (w.ready) {
    if (w.boxes >= budget) {
      *over = true;
      return true;
    }
    ++w.boxes;
    const int outcome = evaluate(q, w, tol);
    if (outcome == 2) {
      descend(q, w);
      return false;
    }
    if (outcome == 1 && w.key[0] < w.best) w.best = w.key[0];
    w.ready = false;
  }
#pragma unroll
  for (int m = 0; m < kAdvanceMoves; ++m) {
    if (!advance_once(q, w, 0)) return true;
    if (w.ready) return false;
  }
  return false;
}

// One step of the jump walk: evaluates the box the lane stands on (when
// ready), then descends or makes one jump.  Returns true when the walk has
// finished; sets *over (leaving the box unevaluated) when the box budget ran
// out.
__device__ inline bool jump_step(const TuvPair& q, JumpWalk& w, double tol, uint64_t budget, bool* over) {
  // levels past 63 are rare: the wave takes the one-word descend and climb
  // unless one of its lanes stands at level 64 or deeper or holds a mark there
  const bool wide = __ballot(w.level >= 64 || w.low[1] != 0) != 0;
  if (w.ready) {
    if (w.boxes >= budget) {
      *over = true;
      return true;
    }
    ++w.boxes;
    const int outcome = jump_classify(q, w, tol);
    if (outcome == 2) {
      if (wide) jump_descend(q, w);
      else jump_descend_lo(q, w);
      return false;
    }
    if (outcome == 1 && w.key[0] < w.best) w.best = w.key[0];
    w.ready = false;
  }
  return !(wide ? jump_next(q, w) : jump_next_lo(q, w));
}

// Pass 1 of the batched form.  A pair whose walk needs more than
// `light_budget` boxes is abandoned here and appended to the heavy list for
// the wave-cooperative pass; its result is written there.
__global__ __launch_bounds__(kTuvBlock) void tuv_persistent_kernel(const TuvPair* __restrict__ pairs,
                                                                   int64_t n_host,
                                                                   const uint64_t* __restrict__ totals,
                                                                   double tol, uint64_t light_budget,
                                                                   p2c_tuv_result* __restrict__ out,
                                                                   unsigned long long* __restrict__ next_pair,
                                                                   uint32_t* __restrict__ heavy_list,
                                                                   uint32_t* __restrict__ heavy_count,
                                                                   uint32_t list_cap) {
  const int64_t n = device_count(n_host, totals);
  const int lane = threadIdx.x & 63;
  int64_t pair = -1;  // -1: needs a pair, -2: none left
  TuvPair q;
  LaneWalk w;
  for (;;) {
    const bool need = pair == -1;
    const uint64_t need_mask = __ballot(need);
    if (need_mask) {
      const int leader = __ffsll((unsigned long long)need_mask) - 1;
      unsigned long long base = 0;
      if (lane == leader) base = atomicAdd(next_pair, (unsigned long long)__popcll(need_mask));
      base = __shfl(base, leader);
      if (need) {
        const uint64_t rank = __popcll(need_mask & ((1ull << lane) - 1ull));
        const int64_t p = (int64_t)(base + rank);
        if (p < n) {
          pair = p;
          q = load_pair(pairs, p);
          start_walk(w);
        } else {
          pair = -2;
        }
      }
    }
    if (__ballot(pair >= 0) == 0) break;
    if (pair >= 0) {
      bool over = false;
      if (lane_step(q, w, tol, light_budget, &over)) {
        if (over) flag_for_replay((uint32_t)pair, heavy_list, heavy_count, list_cap);
        else write_result(&out[pair], w.best, w.boxes, 0u);
        pair = -1;
      }
    }
  }
}

// ---- budgeted rounds ------------------------------------------------------
// Round r walks each listed pair for at most `budget` steps, one lane per
// pair, then either writes the pair's result or saves its walk state (48
// bytes) and appends it to the next round's list.  Round 0 starts every pair
// fresh; the last round hands the survivors to the wave-cooperative pass.
// No atomics or refills in the hot loop, and consecutive lanes read
// consecutive pair records.
struct WalkState {
  uint64_t key[3];
  uint64_t best;
  uint32_t boxes;
  uint16_t level;
  uint8_t e[3];
  uint8_t ready;
};
static_assert(sizeof(WalkState) == 48, "walk state layout");
static_assert(kTuvMaxLevel < 256, "pending sibling counts fit 8 bits");

__device__ inline void save_walk(WalkState* st, const LaneWalk& w) {
  WalkState s;
  s.key[0] = w.key[0];
  s.key[1] = w.key[1];
  s.key[2] = w.key[2];
  s.best = w.best;
  s.boxes = w.boxes;
  s.level = (uint16_t)w.level;
  s.e[0] = (uint8_t)exp_of(w.ep, 0);
  s.e[1] = (uint8_t)exp_of(w.ep, 1);
  s.e[2] = (uint8_t)exp_of(w.ep, 2);
  s.ready = w.ready ? 1 : 0;
  *st = s;
}

__device__ inline void save_walk(WalkState* st, const JumpWalk& w) {
  WalkState s;
  s.key[0] = jump_key64(w.key[0]);
  s.key[1] = jump_key64(w.key[1]);
  s.key[2] = jump_key64(w.key[2]);
  s.best = jump_best64(w.best);
  s.boxes = w.boxes;
  s.level = (uint16_t)w.level;
  s.e[0] = (uint8_t)exp_of(w.ep, 0);
  s.e[1] = (uint8_t)exp_of(w.ep, 1);
  s.e[2] = (uint8_t)exp_of(w.ep, 2);
  s.ready = w.ready ? 1 : 0;
  *st = s;
}

__device__ inline void load_walk(LaneWalk& w, const WalkState& s) {
  w.key[0] = s.key[0];
  w.key[1] = s.key[1];
  w.key[2] = s.key[2];
  w.best = s.best;
  w.boxes = s.boxes;
  w.level = s.level;
  w.ep = (uint32_t)s.e[0] | ((uint32_t)s.e[1] << 8) | ((uint32_t)s.e[2] << 16);
  w.ready = s.ready != 0;
}

__device__ inline void wave_append(uint32_t value, bool want, uint32_t* list, uint32_t* count, uint32_t cap) {
  const uint64_t m = __ballot(want);
  if (m == 0) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
  base = __shfl(base, leader);
  if (want) {
    const uint32_t slot = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    if (slot < cap) list[slot] = value;
  }
}

__global__ __launch_bounds__(kTuvBlock) void tuv_round_kernel(const TuvPair* __restrict__ pairs, int64_t n_host,
                                                              const uint64_t* __restrict__ totals,
                                                              const uint32_t* __restrict__ list_in,
                                                              const uint32_t* __restrict__ count_in,
                                                              WalkState* __restrict__ states, double tol,
                                                              uint32_t budget, uint64_t max_itr,
                                                              p2c_tuv_result* __restrict__ out,
                                                              uint32_t* __restrict__ list_out,
                                                              uint32_t* __restrict__ count_out, int last_round,
                                                              uint32_t* __restrict__ replay_list,
                                                              uint32_t* __restrict__ replay_count,
                                                              uint32_t list_cap) {
  int64_t m;
  if (list_in == nullptr) {
    m = device_count(n_host, totals);
  } else {
    m = *count_in;
    if (m > list_cap) m = list_cap;
  }
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  // the loop bound is the same for every lane of a wave so wave_append's
  // ballots see the whole wave
  for (int64_t base = int64_t(blockIdx.x) * blockDim.x; base < m; base += stride) {
    const int64_t j = base + threadIdx.x;
    const bool valid = j < m;
    uint32_t p = 0;
    bool keep = false, over = false;
    if (valid) {
      p = list_in == nullptr ? (uint32_t)j : list_in[j];
      const TuvPair q = load_pair(pairs, p);
      LaneWalk w;
      if (list_in == nullptr) start_walk(w);
      else load_walk(w, states[p]);
      bool done = false;
      for (uint32_t it = 0; it < budget && !done; ++it) done = lane_step(q, w, tol, max_itr, &over);
      if (done && !over) {
        write_result(&out[p], w.best, w.boxes, 0u);
      } else if (over) {
        write_result(&out[p], w.best, w.boxes, P2C_TUV_FLAG_OVER_BUDGET);
      } else {
        keep = true;
        save_walk(&states[p], w);
      }
    }
    wave_append(p, valid && over, replay_list, replay_count, list_cap);
    wave_append(p, valid && keep, list_out, count_out, list_cap);
  }
}

// ---- stream: one lane per pair, refilled in batches -----------------------
// A lane whose pair has finished idles until `refill_at` lanes of its wave
// are idle; then every idle lane takes the next pair of the wave's claimed
// chunk at once, so one memory latency serves many lanes and busy lanes never
// wait on a refill.  A wave claims kStreamChunk pairs with one atomic.  A
// pair that needs more than `light_budget` boxes saves its walk and goes to
// the heavy list.  Once no pairs are left to claim, a wave whose busy lanes
// number `handover_at` or fewer hands them all to the heavy list too, so a
// few long walks do not hold a nearly empty wave.  Each lane runs the jump
// walk (one box evaluated per step, then the lower child or one jump to the
// next box); pairs the jump walk does not take go to the heavy list unwalked.
// Chunks of 128 pairs: the cloth step's stream kernel 5.50 -> 5.43 ms against
// 256 (64: 5.51, 512: 5.70).
constexpr int kStreamChunk = 128;

template <typename Src>
__global__ __launch_bounds__(kTuvBlock, 3) void tuv_stream_kernel(Src src, int64_t n_host,
                                                               const uint64_t* __restrict__ totals, double tol,
                                                               uint64_t light_budget, uint32_t refill_at,
                                                               uint32_t handover_at,
                                                               p2c_tuv_result* __restrict__ out,
                                                               unsigned long long* __restrict__ cursor,
                                                               WalkState* __restrict__ states,
                                                               uint32_t* __restrict__ heavy_list,
                                                               uint32_t* __restrict__ heavy_count,
                                                               uint32_t list_cap) {
  const int64_t n = device_count(n_host, totals);
  // Two ranges with a cursor each: with device totals the VF pairs [0, split)
  // and the EE pairs [split, n).  A wave starts on its home range (the first
  // split / n of the waves on the VF range) and takes the other only once its
  // own is exhausted, so waves hold one kind of pair apart from the end:
  // the corner evaluation branches on the kind, and a wave holding both
  // kinds runs both branches every step.
  const int64_t split = totals != nullptr ? (int64_t(totals[0]) < n ? int64_t(totals[0]) : n) : n;
  const int lane = threadIdx.x & 63;
  const uint64_t lt = (1ull << lane) - 1ull;
  const int64_t wave = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const int64_t waves = (int64_t(gridDim.x) * blockDim.x) >> 6;
  int range = wave * n < waves * split ? 0 : 1;  // wave-uniform
  bool spent[2] = {false, false};                 // ranges found exhausted (wave-uniform)
  int64_t next = 0, end = 0;  // the wave's claimed chunk (wave-uniform)
  bool drained = false;       // no pairs left to claim (wave-uniform)
  int64_t pair = -1;
  TuvPair q;
  JumpWalk w;
  for (;;) {
    uint64_t idle = __ballot(pair < 0);
    if (!drained && (uint32_t)__popcll(idle) >= refill_at) {
      while (idle != 0) {
        if (next >= end) {
          for (;;) {
            const int64_t lo = range == 0 ? 0 : split, hi = range == 0 ? split : n;
            unsigned long long base = 0;
            if (!(range == 0 ? spent[0] : spent[1])) {
              if (lane == 0) base = atomicAdd(cursor + (range == 0 ? 0 : 3), (unsigned long long)kStreamChunk);
              base = __shfl(base, 0);
              next = lo + (int64_t)base;
              end = next + kStreamChunk < hi ? next + kStreamChunk : hi;
              if (next < hi) break;
            }
            if (range == 0) spent[0] = true;
            else spent[1] = true;
            if (spent[0] && spent[1]) {
              drained = true;
              break;
            }
            range = 1 - range;
          }
          if (drained) break;
        }
        const int64_t avail = end - next;
        const int rank = __popcll(idle & lt);
        if (pair < 0 && rank < avail) {
          pair = next + rank;
          q = src.load(pair);
          start_jump_walk(w);
          if (!jump_walk_fits(q)) {  // keys or split sequence past the jump walk's: to the wave pass
            save_walk(&states[pair], w);
            flag_for_replay((uint32_t)pair, heavy_list, heavy_count, list_cap);
            pair = -1;
          }
        }
        const int64_t took = avail < __popcll(idle) ? avail : (int64_t)__popcll(idle);
        next += took;
        idle = __ballot(pair < 0);
      }
    }
    const uint64_t busy = __ballot(pair >= 0);
    if (busy == 0) break;  // all idle after a refill attempt: nothing left
    if (drained && (uint32_t)__popcll(busy) <= handover_at) {
      if (pair >= 0) {
        save_walk(&states[pair], w);
        flag_for_replay((uint32_t)pair, heavy_list, heavy_count, list_cap);
      }
      break;
    }
    if (pair >= 0) {
      bool over = false;
      if (jump_step(q, w, tol, light_budget, &over)) {
        if (over) {
          save_walk(&states[pair], w);
          flag_for_replay((uint32_t)pair, heavy_list, heavy_count, list_cap);
        } else {
          write_result(&out[pair], jump_best64(w.best), w.boxes, 0u);
        }
        pair = -1;
      }
    }
  }
}

// Orders a wave's LDS accesses before and after this point.
__device__ inline void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// ---- pass 2: one wave per heavy pair ------------------------------------
// The 64 lanes of a wave share one pair's box tree.  Each lane walks a
// subtree depth-first (one evaluation and at most kAdvanceMoves moves per
// step, as in lane_step).  While lanes sit idle, every lane standing on a
// box donates all pending upper siblings on its path (shallowest taken
// first) into the wave's LDS task stack and re-roots its walk at its current
// box; idle lanes take tasks.  The least done t_lo is shared through an LDS
// slot (atomicMin) that every lane prunes with.  Boxes are counted for the
// whole pair, including those a resumed walk already spent; a pair over
// max_itr goes to the canonical replay.  Pairs handed over by the other
// passes resume their saved walk (lane 0 holds it and donates its path at
// once).
constexpr int kWaveQueue = 384;

struct WaveTask {
  uint64_t kt, ku, kv;
  uint32_t level, ep;
};

// Exclusive wave prefix sum of 8-bit values through bit-sliced ballots.
__device__ inline uint32_t wave_prefix_u8(uint32_t v, uint32_t* total) {
  const int lane = threadIdx.x & 63;
  const uint64_t lt = (1ull << lane) - 1ull;
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const uint64_t m = __ballot((v >> b) & 1u);
    pre += (uint32_t)__popcll(m & lt) << b;
    tot += (uint32_t)__popcll(m) << b;
  }
  *total = tot;
  return pre;
}

// Pending upper siblings on the path between level `root` (exponents
// root_ep) and the current box: the path's lower-child choices, i.e. the
// levels minus the key bits set below the root.
__device__ inline uint32_t pending_siblings(const LaneWalk& w, uint32_t root, uint32_t root_ep) {
  uint32_t set = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int hi = 63 - exp_of(root_ep, d);  // key bits below the root's splits on d: [lo, hi)
    const int lo = 63 - exp_of(w.ep, d);
    const uint64_t below_hi = hi >= 64 ? ~0ull : ((1ull << hi) - 1ull);
    const uint64_t below_lo = (1ull << lo) - 1ull;
    set += (uint32_t)__popcll(get3(w.key, d) & (below_hi & ~below_lo));
  }
  return (w.level - root) - set;
}

// A piece of a split pair's box tree: a walk (key, level, exponents, ready)
// confined to the subtree rooted at `root` (exponents root_e).  Pieces of
// pairs the jump walk takes are stored in the same 88-byte slots as eleven
// tagged words (JumpTaskWords).
struct SplitTask {
  uint64_t key[3];
  uint32_t pair;
  uint16_t level, root;
  uint8_t e[3], ready;
  uint8_t root_e[3], pad;
  uint64_t pad2[6];
};
static_assert(sizeof(SplitTask) == 88, "split task layout");

// The wave pass runs the pieces of jump-walk pairs in the same kernel, on
// other waves (of other XCDs, whose L2 caches are not coherent with each
// other).  A piece's words are written and read as device-scope atomics and
// each carries the pass's epoch (31 bits, a new one per call, never 0), so
// that each word shows by itself whether it was written in this pass (the
// slots are reused from call to call); no cache-wide release or acquire is
// needed.  A piece is a lane's walk: ten words, payload in the upper half,
// epoch in the lower (key t, u, v; pair | ready << 31; level | exponents <<
// 8; the giver's best key; the pending-sibling marks in four halves), and a
// control word the queue marks (a slot's waiting wave may leave; the piece
// is done) without touching the piece.
struct JumpTaskWords {
  uint64_t w[11];
};
constexpr int kPieceWords = 10, kControlWord = 10;

struct JumpPiece {
  uint32_t key[3];
  uint32_t pair;
  uint32_t lep;
  uint32_t best;
  uint64_t low[2];
  bool ready;
};
static_assert(sizeof(JumpTaskWords) == sizeof(SplitTask), "jump pieces share the task slots");

__device__ inline uint64_t tag_word(uint32_t payload, uint32_t epoch) { return ((uint64_t)payload << 32) | epoch; }
__device__ inline bool tag_ok(uint64_t v, uint32_t epoch) { return (uint32_t)v == epoch; }

// A split pair's accumulators live in the first bytes of its WalkState slot
// once the walk it held has been resumed: best (u64 at byte 0, atomicMin),
// boxes (u32 at byte 8, atomicAdd), over-budget flag (u32 at byte 12,
// atomicOr).  The wave that lists the pair sets them before it writes any
// piece of the pair, and every other wave finds the slot through a piece's
// pair word, so each of its accesses follows the setting.  The rest of the
// slot is left as it was.
struct SplitAcc {
  unsigned long long best;
  uint32_t boxes;
  uint32_t over;
  uint64_t keep[4];
};
static_assert(sizeof(SplitAcc) == sizeof(WalkState), "split accumulators reuse the walk state slot");

__device__ inline unsigned long long acc_load_best(const SplitAcc* a) {
  return __hip_atomic_load(&a->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint32_t acc_load_over(const SplitAcc* a) {
  return __hip_atomic_load(&a->over, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sets a jump-walk pair's accumulators when the pair is listed (lane 0).
// Each word is set by compare-and-swap and the loop waits for its result,
// so the setting is done before the wave goes on to write the pair's pieces
// (no other wave touches the slot meanwhile: each swap succeeds at once).
__device__ inline void acc_set(SplitAcc* a, uint64_t best, uint32_t boxes) {
  unsigned long long cur = __hip_atomic_load(&a->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (!__hip_atomic_compare_exchange_strong(&a->best, &cur, (unsigned long long)best, __ATOMIC_RELAXED,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
  }
  auto* cw = reinterpret_cast<unsigned long long*>(&a->boxes);  // boxes (low half), over (high half)
  cur = __hip_atomic_load(cw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (!__hip_atomic_compare_exchange_strong(cw, &cur, (unsigned long long)boxes, __ATOMIC_RELAXED,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
  }
}

// The jump-walk queue's counters, each on a 128-byte line of its own (the
// busy waves read `count` and `next` every few steps, waiting waves poll
// their own slots), in the last kHeadSlots task slots: the epoch is kept
// there from call to call, the rest set to 0 by the epoch kernel.
struct alignas(128) QueueLine {
  uint32_t v;
  uint32_t pad[31];
};
struct QueueHead {
  QueueLine epoch, count, next, pending;
};
constexpr uint32_t kHeadSlots = 12;  // 672 bytes: a 128-aligned QueueHead fits
static_assert(sizeof(QueueHead) + 128 <= kHeadSlots * sizeof(SplitTask), "queue head room");

// The wave pass's shared task queue.  Jump-walk pieces take slots [0,
// jump_cap) through task_count (reserved); a waiting wave takes a ticket
// (task_next) and runs the slot of that number once it is written.  The
// other walk's pieces take [jump_cap, task_cap) through lw_count and run in
// the closing task pass only.  `pending` counts the waves still on heavy
// pairs plus the pieces given and not yet done: work comes only from work,
// so once it reads 0 it stays 0 and the waiting waves leave.
struct SplitOut {
  uint64_t split_at;     // boxes a wave takes on a pair or task before it gives pieces away (0: never)
  SplitTask* tasks;
  uint32_t task_cap;     // usable slots
  uint32_t jump_cap;     // slots of the jump-walk pieces
  QueueHead* head;       // also this pass's tag (its epoch line, set before the pass's kernels)
  uint32_t* task_count;  // reserved jump-walk slots
  uint32_t* task_next;   // taken jump-walk slots
  uint32_t* lw_count;    // reserved slots of the other walk (from jump_cap)
  uint32_t* lw_next;     // their cursor in the closing task pass
  uint32_t* pending;
  uint32_t* split_list;  // split pairs
  uint32_t* split_count;
  uint32_t* report;      // the jump-walk piece count, copied here at the end
  SplitAcc* acc;         // per pair (the walk state slots)
};

// The pass's tag: written by the epoch kernel, read-only during the pass.
// Kernels take the queue's descriptor straight from their arguments (a copy
// in registers spilled in the wave pass).
__device__ inline uint32_t pass_epoch(const SplitOut& so) { return so.head->epoch.v; }

// Lane 0 reserves `need` (> 0) slots of `count` below `cap`; returns the
// first or ~0u when there is no room (the caller keeps its work).
__device__ inline uint32_t reserve_slots(uint32_t* count, uint32_t cap, uint32_t need) {
  uint32_t cur = __hip_atomic_load(count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (int tries = 0; tries < 64; ++tries) {
    if (cur + need > cap || cur + need < cur) break;
    if (__hip_atomic_compare_exchange_strong(count, &cur, cur + need, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT))
      return cur;
  }
  return ~0u;
}

// Lane 0 reserves up to `want` jump-walk slots, no more than the tickets
// taken (waves waiting for a piece, or about to), so that donors seeing the
// same waiting waves do not all give to them; returns the first slot and
// sets *got, or ~0u.
__device__ inline uint32_t reserve_for_tickets(const SplitOut& so, uint32_t want, uint32_t* got) {
  uint32_t cur = __hip_atomic_load(so.task_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (int tries = 0; tries < 16; ++tries) {
    uint32_t lim = __hip_atomic_load(so.task_next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    lim = lim < so.jump_cap ? lim : so.jump_cap;
    if (cur >= lim) break;
    const uint32_t n = min(want, lim - cur);
    if (__hip_atomic_compare_exchange_strong(so.task_count, &cur, cur + n, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT)) {
      *got = n;
      return cur;
    }
  }
  return ~0u;
}

// One jump-walk piece, written word by word as device-scope atomics.
__device__ inline void put_jump_task(const SplitOut& so, uint32_t slot, const JumpPiece& pc) {
  auto* w = reinterpret_cast<JumpTaskWords*>(so.tasks + slot)->w;
  const uint32_t payload[kPieceWords] = {pc.key[0], pc.key[1], pc.key[2], pc.pair | (pc.ready ? 1u << 31 : 0u),
                                         pc.lep, pc.best, (uint32_t)pc.low[0], (uint32_t)(pc.low[0] >> 32),
                                         (uint32_t)pc.low[1], (uint32_t)(pc.low[1] >> 32)};
#pragma unroll
  for (int i = 0; i < kPieceWords; ++i)
    __hip_atomic_store(&w[i], tag_word(payload[i], pass_epoch(so)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline JumpPiece unpack_piece(const uint32_t t[kPieceWords]) {
  JumpPiece pc;
  pc.key[0] = t[0];
  pc.key[1] = t[1];
  pc.key[2] = t[2];
  pc.pair = t[3] & 0x7FFFFFFFu;
  pc.ready = (t[3] >> 31) != 0;
  pc.lep = t[4];
  pc.best = t[5];
  pc.low[0] = (uint64_t)t[6] | ((uint64_t)t[7] << 32);
  pc.low[1] = (uint64_t)t[8] | ((uint64_t)t[9] << 32);
  return pc;
}

// Reads a jump-walk piece's payload; false if any word is not yet this pass's.
__device__ inline bool get_jump_task(const SplitOut& so, uint32_t slot, uint32_t out[kPieceWords]) {
  auto* w = reinterpret_cast<JumpTaskWords*>(so.tasks + slot)->w;
  bool ok = true;
#pragma unroll
  for (int i = 0; i < kPieceWords; ++i) {
    const uint64_t v = __hip_atomic_load(&w[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    ok = ok && tag_ok(v, pass_epoch(so));
    out[i] = (uint32_t)(v >> 32);
  }
  return ok;
}

__device__ inline SplitTask make_task(uint32_t pair, const uint64_t k[3], uint32_t level, uint32_t ep, bool ready,
                                      uint32_t root, uint32_t root_ep) {
  SplitTask t;
  t.key[0] = k[0];
  t.key[1] = k[1];
  t.key[2] = k[2];
  t.pair = pair;
  t.level = (uint16_t)level;
  t.root = (uint16_t)root;
  for (int d = 0; d < 3; ++d) {
    t.e[d] = (uint8_t)exp_of(ep, d);
    t.root_e[d] = (uint8_t)exp_of(root_ep, d);
  }
  t.ready = ready ? 1 : 0;
  t.pad = 0;
  return t;
}

// The wave pass over one pair (heavy mode) or one split task (task mode).
// Heavy mode starts from the pair's saved walk (or the root); once the pair
// has taken more than split_at boxes it may split: every pending piece (each
// lane's path siblings, each lane's current subtree or unfinished walk, every
// queued task) becomes a task for the closing task pass, the pair's best and
// box count go to its accumulators, and the pair is listed for the finishing
// pass.  Task mode walks one piece with the pair's shared best: read from the
// accumulator every 16 steps and lowered there on every improvement.  Every form of the
// result is the minimum t_lo over done boxes, so the pieces' order is free.
enum : int { kWaveDone = 0, kWaveOver = 1, kWaveSplit = 2 };

template <bool kTask, bool kSplit>
__device__ int wave_certify(const TuvPair& q, uint32_t p, WaveTask* wq, unsigned long long* best_lds, LaneWalk& w,
                            bool has, uint32_t root, uint32_t root_ep, uint64_t& boxes, double tol,
                            uint64_t max_itr, const SplitOut& so) {
  const int lane = threadIdx.x & 63;
  const uint64_t lt = (1ull << lane) - 1ull;
  int qsize = 0;
  uint32_t step = 0;
  for (;;) {
    // 1. evaluate (a box pruned meanwhile by the shared best is skipped),
    //    then descend, or make at most kAdvanceMoves moves
    bool evaluated = false;
    if (has && w.ready) {
      int outcome = 0;
      if (w.key[0] < w.best) {
        evaluated = true;
        outcome = evaluate(q, w, tol);
        if (outcome == 1 && w.key[0] < w.best) {
          w.best = w.key[0];
          atomicMin(best_lds, (unsigned long long)w.key[0]);
          if (kTask) atomicMin(&so.acc[p].best, (unsigned long long)w.key[0]);
        }
      }
      if (outcome == 2) descend(q, w);
      else w.ready = false;
    }
    if (has && !w.ready) {
#pragma unroll
      for (int m = 0; m < kAdvanceMoves; ++m) {
        if (!advance_once(q, w, root)) {
          has = false;
          break;
        }
        if (w.ready) break;
      }
    }
    boxes += (uint64_t)__popcll(__ballot(evaluated));
    wave_sync_lds();
    {
      uint64_t shared_best = *best_lds;
      if (kTask && (++step & 15u) == 0) {
        const uint64_t g = __hip_atomic_load(&so.acc[p].best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (g < shared_best) shared_best = g;
      }
      if (shared_best < w.best) w.best = shared_best;
    }
    if (boxes > max_itr) return kWaveOver;
    if (kSplit && boxes > so.split_at) {
      // count the pieces: per lane its path siblings and its own walk, plus the queue
      const uint32_t pend = has && w.ready ? pending_siblings(w, root, root_ep) : 0u;
      const uint32_t mine = has ? pend + 1u : 0u;  // at most kTuvMaxLevel + 1 < 256
      uint32_t lanes_tot = 0;
      const uint32_t off = wave_prefix_u8(mine, &lanes_tot);
      const uint32_t need = lanes_tot + (uint32_t)qsize;
      uint32_t base = 0xFFFFFFFFu;
      if (lane == 0 && need > 0) base = reserve_slots(so.lw_count, so.task_cap - so.jump_cap, need);
      base = __shfl(base, 0);
      if (base != 0xFFFFFFFFu) {
        SplitTask* out = so.tasks + so.jump_cap + base;
        if (has) {
          uint32_t slot = off;
          if (w.ready) {
            // path siblings, then the lane's own box as a subtree of its own
            uint64_t k[3] = {w.key[0], w.key[1], w.key[2]};
            uint32_t ep = w.ep;
            for (uint32_t l = w.level; l > root; --l) {
              const int d = tuv_level_dim(q, l - 1);
              const uint64_t bit = kOne63 >> exp_of(ep, d);
              const uint64_t kd = get3(k, d);
              if (!(kd & bit)) {
                uint64_t kk[3] = {k[0], k[1], k[2]};
                put3(kk, d, kd | bit);
                out[slot++] = make_task(p, kk, l, ep, true, l, ep);
              }
              put3(k, d, kd & ~bit);
              ep -= 1u << (8 * d);
            }
            out[slot] = make_task(p, w.key, w.level, w.ep, true, w.level, w.ep);
          } else {
            out[slot] = make_task(p, w.key, w.level, w.ep, false, root, root_ep);
          }
        }
        for (int k = lane; k < qsize; k += 64) {
          const WaveTask t = wq[k];
          const uint64_t kk[3] = {t.kt, t.ku, t.kv};
          out[lanes_tot + (uint32_t)k] = make_task(p, kk, t.level, t.ep, true, t.level, t.ep);
        }
        if (lane == 0) {
          so.acc[p].best = *best_lds < w.best ? *best_lds : w.best;
          so.acc[p].boxes = (uint32_t)boxes;
          so.acc[p].over = 0;
          const uint32_t at = atomicAdd(so.split_count, 1u);
          so.split_list[at] = p;  // at most one entry per listed pair
        }
        return kWaveSplit;
      }
    }
    const int idle = 64 - __popcll(__ballot(has));
    if (idle == 0) continue;
    // 2. donation: lanes standing on a box give away the pending siblings on their path
    if (idle > qsize) {
      // at most kTuvMaxLevel (156) pending siblings: within wave_prefix_u8's range
      const uint32_t pend = has && w.ready ? pending_siblings(w, root, root_ep) : 0u;
      uint32_t tot = 0;
      const uint32_t off = wave_prefix_u8(pend, &tot);
      const bool push = pend > 0 && qsize + (int)(off + pend) <= kWaveQueue;
      if (push) {
        uint64_t k[3] = {w.key[0], w.key[1], w.key[2]};
        uint32_t ep = w.ep;
        int slot = qsize + (int)off;
        for (uint32_t l = w.level; l > root; --l) {  // deepest first: shallowest ends on top
          const int d = tuv_level_dim(q, l - 1);
          const uint64_t bit = kOne63 >> exp_of(ep, d);
          const uint64_t kd = get3(k, d);
          if (!(kd & bit)) {
            WaveTask t;
            t.kt = k[0] | (d == 0 ? bit : 0ull);
            t.ku = k[1] | (d == 1 ? bit : 0ull);
            t.kv = k[2] | (d == 2 ? bit : 0ull);
            t.level = l;
            t.ep = ep;
            wq[slot++] = t;
          }
          put3(k, d, kd & ~bit);
          ep -= 1u << (8 * d);
        }
        root = w.level;
        root_ep = w.ep;
      }
      // lanes whose block did not fit pushed nothing; count what was pushed
      uint32_t pushed_tot = 0;
      wave_prefix_u8(push ? pend : 0u, &pushed_tot);
      qsize += (int)pushed_tot;
      wave_sync_lds();
    }
    // 3. idle lanes take queued subtrees, the top of the stack first
    const uint64_t want = __ballot(!has);
    const int n_take = min(__popcll(want), qsize);
    if (n_take > 0) {
      if (!has) {
        const int rank = __popcll(want & lt);
        if (rank < n_take) {
          const WaveTask t = wq[qsize - 1 - rank];
          w.key[0] = t.kt;
          w.key[1] = t.ku;
          w.key[2] = t.kv;
          w.level = t.level;
          w.ep = t.ep;
          root = t.level;
          root_ep = t.ep;
          w.ready = true;
          has = !box_skipped(q, w);  // a skipped task's whole subtree is skipped
        }
      }
      qsize -= n_take;
      wave_sync_lds();
    }
    if (__ballot(has) == 0 && qsize == 0) return kWaveDone;
  }
}

// ---- wave pass on the jump walk ------------------------------------------
// For pairs the jump walk takes (jump_walk_fits): each lane runs a jump walk
// over a subtree; a lane that finishes its subtree takes one from the wave's
// LDS stack, and while idle lanes outnumber the stacked subtrees every lane
// holding marks gives away its shallowest one (the largest pending subtree).
// The least done t_lo is shared across the wave (and, for split pieces,
// through the pair's accumulator) whenever a lane finds one.  The result is
// the least t_lo over done boxes, whatever the order.
constexpr int kJumpStack = 512;  // stacked subtrees per wave

struct JumpBox {
  uint32_t key[3];
  uint32_t lep;  // level | exponents << 8
};

// A wave's LDS in the wave pass: the old walk's task queue or the jump
// walk's subtree stack (a pair takes one of the two).
union WaveLds {
  WaveTask queue[kWaveQueue];
  JumpBox stack[kJumpStack];
};

__device__ inline uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const uint32_t x = (uint32_t)__shfl_xor((int)v, o);
    v = x < v ? x : v;
  }
  return v;
}

__device__ inline uint32_t best32_of(uint64_t b) { return b == kTuvNoHit ? kJumpNoHit : (uint32_t)(b >> 32); }

// A saved walk (64-bit keys, tuv_walk's position) of a pair the jump walk
// takes, with its marks rebuilt from the keys: lane j checks level j (and
// j + 64).  Every lane returns the same walk.
__device__ inline void jump_resume(const TuvPair& q, const WalkState& st, JumpWalk& w) {
  w.key[0] = (uint32_t)(st.key[0] >> 32);
  w.key[1] = (uint32_t)(st.key[1] >> 32);
  w.key[2] = (uint32_t)(st.key[2] >> 32);
  w.best = best32_of(st.best);
  w.boxes = 0;
  w.ep = (uint32_t)st.e[0] | ((uint32_t)st.e[1] << 8) | ((uint32_t)st.e[2] << 16);
  w.level = st.level;
  w.ready = st.ready != 0;
  const uint32_t lane = threadIdx.x & 63u;
  w.low[0] = __ballot(lane < w.level && jump_pending_at(q, w, lane));
  w.low[1] = __ballot(lane + 64u < w.level && jump_pending_at(q, w, lane + 64u));
}

// Box accounting of one unit of work (a pair or a task), wave-uniform:
// `boxes` evaluated so far (a resumed pair's include those of the stream
// kernel), `flushed` of them already added to the pair's accumulator, and
// `listed` once the pair's accumulators are live.
struct UnitCount {
  uint64_t boxes;
  uint64_t flushed;
  bool listed;
};

__device__ inline uint32_t clamp_u32(uint64_t v) { return v < 0xFFFFFFFFull ? (uint32_t)v : 0xFFFFFFFFu; }

// The shared task queue and a listed pair's accumulators are looked at every
// 32 steps; a wave adds its boxes to the pair's sum once 4096 have gathered.
constexpr uint32_t kLookMask = 31u;
constexpr uint64_t kFlushBoxes = 4096;

// The wave pass over one pair (`active` on lane 0 only, its walk resumed or
// at the root) or one task (a listed pair's piece).  With kSplit, when a look
// finds more waiting waves than queued tasks and the wave has taken split_at
// boxes on this unit, lanes holding marks give their shallowest (at most one
// per lane) to the queue; the first such gift lists the pair.  Returns
// kWaveDone or kWaveOver; w.best is the wave's least done key on return.
template <bool kSplit>
__device__ int wave_certify_jump(const TuvPair& q, uint32_t p, JumpBox* stk, JumpWalk& w, bool active,
                                 UnitCount& uc, double tol, uint64_t max_itr, const SplitOut& so) {
  const int lane = threadIdx.x & 63;
  const uint64_t lt = (1ull << lane) - 1ull;
  int sp = 0;  // stacked subtrees (wave-uniform)
  uint32_t step = 0;
  w.best = wave_min_u32(w.best);
  for (;;) {
    // 1. evaluate, then descend or jump to the next box of the lane's subtree
    const bool wide = __ballot(w.level >= 64 || w.low[1] != 0) != 0;
    bool evaluated = false, found = false;
    if (active) {
      bool finished = !w.ready;
      if (w.ready) {
        if (w.key[0] < w.best) {
          evaluated = true;
          const int outcome = jump_classify(q, w, tol);
          if (outcome == 2) {
            if (wide) jump_descend(q, w);
            else jump_descend_lo(q, w);
          } else {
            finished = true;
            if (outcome == 1) {
              w.best = w.key[0];
              found = true;
            }
          }
        } else {
          finished = true;
        }
      }
      if (finished && !(wide ? jump_next(q, w) : jump_next_lo(q, w))) active = false;
    }
    uc.boxes += (uint64_t)__popcll(__ballot(evaluated));
    // 2. share the least done t_lo: a find is shared across the wave at
    //    once; a look lowers the pair's accumulator with the wave's best and
    //    takes the other waves' (one atomic: a find's own atomic, under the
    //    contention of hundreds of waves on one pair, cost more than a
    //    step), adds the boxes (once kFlushBoxes have gathered; the sum
    //    decides the budget) and reads the task queue
    const bool look = (kSplit || uc.listed) && (++step & kLookMask) == 0;
    const bool any_found = __ballot(found) != 0;
    bool over = uc.boxes > max_itr;
    int deficit = 0;
    if (any_found || look) {
      uint32_t b = wave_min_u32(w.best);
      const bool flush = look && uc.listed && uc.boxes - uc.flushed >= kFlushBoxes;
      if (look) {
        uint32_t g = kJumpNoHit, flags = 0;
        int d = 0;
        if (lane == 0) {
          if (uc.listed) {
            SplitAcc* a = &so.acc[p];
            g = best32_of(atomicMin(&a->best, (unsigned long long)jump_best64(b)));
            if (flush) {
              const uint32_t add = clamp_u32(uc.boxes - uc.flushed);
              const uint64_t tot = (uint64_t)atomicAdd(&a->boxes, add) + add;
              flags = tot > max_itr || acc_load_over(a) != 0 ? 1u : 0u;
            }
          }
          if (kSplit && look) {
            // tickets beyond the written slots: waves waiting for a piece
            const uint32_t count = __hip_atomic_load(so.task_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t next = __hip_atomic_load(so.task_next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            d = next > count ? (int)min(next - count, 64u) : 0;
          }
        }
        g = (uint32_t)__shfl((int)g, 0);
        b = g < b ? g : b;
        if (flush) {
          uc.flushed = uc.boxes;
          over = over || __shfl((int)flags, 0) != 0;
        }
        deficit = __shfl(d, 0);
      }
      w.best = b;
    }
    if (over) return kWaveOver;
    // 3. pieces for waiting waves: every other active lane hands its whole
    //    walk (current box and pending siblings) over, so that the wave's
    //    remaining work about halves whatever the tree's shape; a lone lane
    //    gives its shallowest pending sibling
    if (kSplit && deficit > 0 && uc.boxes >= so.split_at) {
      const uint64_t am = __ballot(active);
      const uint32_t n_active = (uint32_t)__popcll(am);
      const bool lone = n_active == 1;
      const bool lone_can = lone && active && (w.low[0] | w.low[1]) != 0;
      uint32_t give = lone ? (__ballot(lone_can) != 0 ? 1u : 0u) : min(n_active >> 1, (uint32_t)deficit);
      if (give > 0) {
        uint32_t base = ~0u;
        if (lane == 0) {
          base = reserve_for_tickets(so, give, &give);
          if (base != ~0u) atomicAdd(so.pending, give);
        }
        base = (uint32_t)__shfl((int)base, 0);
        give = (uint32_t)__shfl((int)give, 0);
        if (base != ~0u) {
          if (!uc.listed) {
            // the pair's accumulators take the wave's best and boxes so far
            if (lane == 0) {
              acc_set(&so.acc[p], jump_best64(w.best), clamp_u32(uc.boxes));
              const uint32_t at = atomicAdd(so.split_count, 1u);
              so.split_list[at] = p;  // at most one entry per listed pair
            }
            uc.flushed = uc.boxes;
            uc.listed = true;
          }
          const uint32_t rank = (uint32_t)__popcll(am & lt);
          JumpPiece pc;
          pc.pair = p;
          pc.best = w.best;
          if (lone) {
            if (lone_can) {
              uint32_t lvl, ep;
              jump_take_shallowest(q, w, pc.key, &ep, &lvl);
              pc.lep = lvl | (ep << 8);
              pc.low[0] = pc.low[1] = 0;
              pc.ready = true;
              put_jump_task(so, base, pc);
            }
          } else if (active && (rank & 1u) != 0 && (rank >> 1) < give) {
            pc.key[0] = w.key[0];
            pc.key[1] = w.key[1];
            pc.key[2] = w.key[2];
            pc.lep = w.level | (w.ep << 8);
            pc.low[0] = w.low[0];
            pc.low[1] = w.low[1];
            pc.ready = w.ready;
            put_jump_task(so, base + (rank >> 1), pc);
            active = false;
            w.low[0] = w.low[1] = 0;
            w.ready = false;
          }
        }
      }
    }
    // 4. idle lanes beyond the stacked subtrees: lanes with marks give their shallowest
    const uint64_t idle_m = __ballot(!active);
    const int idle = __popcll(idle_m);
    if (idle > sp && sp < kJumpStack) {
      const bool can = active && (w.low[0] | w.low[1]) != 0;
      const uint64_t cm = __ballot(can);
      const int want = min(idle - sp, kJumpStack - sp);
      const int give = min(__popcll(cm), want);
      if (can) {
        const int r = __popcll(cm & lt);
        if (r < give) {
          JumpBox bx;
          uint32_t ep, lvl;
          jump_take_shallowest(q, w, bx.key, &ep, &lvl);
          bx.lep = lvl | (ep << 8);
          stk[sp + r] = bx;
        }
      }
      sp += give;
      wave_sync_lds();
    }
    // 5. idle lanes take stacked subtrees, the top first
    if (idle > 0 && sp > 0) {
      const int take = min(idle, sp);
      if (!active) {
        const int r = __popcll(idle_m & lt);
        if (r < take) {
          const JumpBox bx = stk[sp - 1 - r];
          w.key[0] = bx.key[0];
          w.key[1] = bx.key[1];
          w.key[2] = bx.key[2];
          w.level = bx.lep & 0xffu;
          w.ep = bx.lep >> 8;
          w.low[0] = w.low[1] = 0;
          w.ready = w.key[0] < w.best;
          active = true;
        }
      }
      sp -= take;
      wave_sync_lds();
    }
    if (__ballot(active) == 0 && sp == 0) return kWaveDone;
  }
}

// A listed pair's unit is done (lane 0): its best and its boxes not yet
// added go to the accumulators, and an over-budget unit marks the pair.
__device__ inline void close_unit(const SplitOut& so, uint32_t p, uint32_t best, const UnitCount& uc, bool over) {
  if ((threadIdx.x & 63) == 0) {
    atomicMin(&so.acc[p].best, (unsigned long long)jump_best64(best));
    atomicAdd(&so.acc[p].boxes, clamp_u32(uc.boxes - uc.flushed));
    if (over) atomicOr(&so.acc[p].over, 1u);
  }
}

// One jump-walk piece (a lane's walk), by one wave (kSplit: it may give
// pieces of its own).  A piece of a pair already over budget is dropped: the
// pair goes to the canonical replay.
template <bool kSplit, typename Src>
__device__ void run_jump_task(const Src& src, const SplitOut& so, const JumpPiece& pc, WaveLds& lds, double tol,
                              uint64_t max_itr) {
  const int lane = threadIdx.x & 63;
  const uint32_t p = pc.pair;
  SplitAcc* a = &so.acc[p];
  if (acc_load_over(a) != 0) return;
  const TuvPair q = src.load(p);
  JumpWalk w;
  start_jump_walk(w);
  w.key[0] = pc.key[0];
  w.key[1] = pc.key[1];
  w.key[2] = pc.key[2];
  w.level = pc.lep & 0xffu;
  w.ep = pc.lep >> 8;
  const uint32_t g = best32_of(acc_load_best(a));
  w.best = g < pc.best ? g : pc.best;
  w.ready = pc.ready && w.key[0] < w.best;
  if (lane == 0) {
    w.low[0] = pc.low[0];
    w.low[1] = pc.low[1];
  }
  UnitCount uc{0, 0, true};
  const int r = wave_certify_jump<kSplit>(q, p, lds.stack, w, lane == 0, uc, tol, max_itr, so);
  close_unit(so, p, w.best, uc, r == kWaveOver);
}

// One piece of a 64-bit-key walk (closing task pass only).
template <typename Src>
__device__ void run_lw_task(const Src& src, const SplitOut& so, const SplitTask& t, WaveLds& lds,
                            unsigned long long* best_slot, double tol, uint64_t max_itr) {
  const int lane = threadIdx.x & 63;
  const uint32_t p = t.pair;
  if (so.acc[p].over != 0) return;
  const TuvPair q = src.load(p);
  LaneWalk w;
  start_walk(w);
  w.best = so.acc[p].best;
  w.key[0] = t.key[0];
  w.key[1] = t.key[1];
  w.key[2] = t.key[2];
  w.level = t.level;
  w.ep = (uint32_t)t.e[0] | ((uint32_t)t.e[1] << 8) | ((uint32_t)t.e[2] << 16);
  w.ready = t.ready != 0;
  const uint32_t root_ep = (uint32_t)t.root_e[0] | ((uint32_t)t.root_e[1] << 8) | ((uint32_t)t.root_e[2] << 16);
  const bool has = lane == 0 && (!w.ready || !box_skipped(q, w));
  if (lane == 0) *best_slot = w.best;
  wave_sync_lds();
  uint64_t boxes = 0;
  const int r =
      wave_certify<true, false>(q, p, lds.queue, best_slot, w, has, t.root, root_ep, boxes, tol, max_itr, so);
  if (lane == 0) {
    const uint64_t b = *best_slot < w.best ? *best_slot : w.best;
    atomicMin(&so.acc[p].best, (unsigned long long)b);
    atomicAdd(&so.acc[p].boxes, clamp_u32(boxes));
    if (r == kWaveOver) atomicOr(&so.acc[p].over, 1u);
  }
}

// A wave done with the heavy pairs takes a ticket and runs the jump-walk
// piece of that slot once all ten of its words are this pass's, then takes
// the next ticket.  A waiting wave polls only its own slot, so waiting waves
// load no shared line.  The wave that takes `pending` to 0 (no work left, and
// none can come: work comes only from work) writes a leave mark into the
// first word of every slot a ticket waits on; a wave whose ticket it missed
// (taken meanwhile) sees `pending` at 0 on one of its rare checks.  Marks go
// to the slots' control words and never touch a piece.  The
// pieces and counters are device-scope atomics, read without cache-wide
// maintenance (an acquire invalidates the L2 of the wave's XCD, which a few
// thousand polling waves would keep empty).  A done slot is marked in its
// control word; after kServeSpins polls a waiting wave leaves anyway, and the
// closing task pass runs every written slot not marked done.
constexpr uint32_t kServeSpins = 1u << 20;
constexpr uint32_t kSlotDone = 1u, kSlotLeave = 2u;  // control word marks

__device__ inline void mark_slot(const SplitOut& so, uint32_t slot, uint32_t mark) {
  auto* w = reinterpret_cast<JumpTaskWords*>(so.tasks + slot)->w;
  __hip_atomic_store(&w[kControlWord], tag_word(mark, pass_epoch(so)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Lane 0 took `pending` to 0: leave marks for the waiting tickets.
__device__ inline void release_waiters(const SplitOut& so) {
  const int lane = threadIdx.x & 63;
  uint32_t lo = 0, hi = 0;
  if (lane == 0) {
    lo = __hip_atomic_load(so.task_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    hi = __hip_atomic_load(so.task_next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  lo = (uint32_t)__shfl((int)lo, 0);
  hi = min((uint32_t)__shfl((int)hi, 0), so.jump_cap);
  for (uint32_t t = lo + (uint32_t)lane; t < hi; t += 64) mark_slot(so, t, kSlotLeave);
}

// Lane 0 lowers `pending`; every lane returns whether it reached 0.
__device__ inline bool drop_pending(const SplitOut& so) {
  uint32_t last = 0;
  if ((threadIdx.x & 63) == 0) last = atomicSub(so.pending, 1u) == 1u ? 1u : 0u;
  return __shfl((int)last, 0) != 0;
}

template <typename Src>
__device__ void serve_tasks(const Src& src, const SplitOut& so, WaveLds& lds, double tol, uint64_t max_itr) {
  const int lane = threadIdx.x & 63;
  if (drop_pending(so)) release_waiters(so);  // this wave's heavy pairs are done
  for (;;) {
    uint32_t ticket = 0;
    if (lane == 0) ticket = atomicAdd(so.task_next, 1u);
    ticket = (uint32_t)__shfl((int)ticket, 0);
    if (ticket >= so.jump_cap) return;
    uint32_t t[kPieceWords] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t state = 0;  // 1: piece read, 2: leave
    for (uint32_t spins = 0; state == 0; ++spins) {
      if (lane == 0) {
        const auto* w = reinterpret_cast<JumpTaskWords*>(so.tasks + ticket)->w;
        const uint64_t w0 = __hip_atomic_load(&w[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t wc = __hip_atomic_load(&w[kControlWord], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (tag_ok(w0, pass_epoch(so)) && get_jump_task(so, ticket, t)) {
          state = 1;
        } else if (tag_ok(wc, pass_epoch(so)) && (uint32_t)(wc >> 32) == kSlotLeave) {
          state = 2;
        } else if ((spins & 255u) == 255u &&
                   __hip_atomic_load(so.pending, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u &&
                   __hip_atomic_load(so.task_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= ticket) {
          state = 2;
        }
        if (state == 0 && spins + 1 >= kServeSpins) state = 2;
      }
      state = (uint32_t)__shfl((int)state, 0);
      if (state == 0) __builtin_amdgcn_s_sleep(8);
    }
    if (state == 2) return;
#pragma unroll
    for (int i = 0; i < kPieceWords; ++i) t[i] = (uint32_t)__shfl((int)t[i], 0);
    run_jump_task<true>(src, so, unpack_piece(t), lds, tol, max_itr);
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) mark_slot(so, ticket, kSlotDone);
    if (drop_pending(so)) release_waiters(so);
  }
}

// Three workgroups per CU (12 waves: 3 per SIMD; 147 KB of LDS): at most 168
// VGPRs, a few spilled; the wave pass ran 20 % longer at 2 waves per SIMD.
// With kSplit, a wave done with the heavy pairs serves the task queue.
template <bool kSplit, typename Src>
__global__ __launch_bounds__(kTuvBlock, 3) void tuv_heavy_kernel(Src src,
                                                              const uint32_t* __restrict__ heavy_list,
                                                              const uint32_t* __restrict__ heavy_count,
                                                              uint32_t list_cap, uint32_t* __restrict__ heavy_next,
                                                              const WalkState* __restrict__ resume, double tol,
                                                              uint64_t max_itr, p2c_tuv_result* __restrict__ out,
                                                              uint32_t* __restrict__ replay_list,
                                                              uint32_t* __restrict__ replay_count,
                                                              SplitOut so_in) {
  __shared__ WaveLds lds[kTuvBlock / 64];
  __shared__ unsigned long long best_slot[kTuvBlock / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const SplitOut& so = so_in;
  if constexpr (kSplit) {
    if (lane == 0) atomicAdd(so.pending, 1u);
  }
  uint32_t total = *heavy_count;
  if (total > list_cap) total = list_cap;
  for (;;) {
    uint32_t h = 0;
    if (lane == 0) h = atomicAdd(heavy_next, 1u);
    h = __shfl(h, 0);
    if (h >= total) break;
    const uint32_t p = heavy_list[h];
    const TuvPair q = src.load(p);
    if (jump_walk_fits(q)) {
      JumpWalk w;
      start_jump_walk(w);
      bool active = lane == 0;
      UnitCount uc{0, 0, false};
      if (resume != nullptr) {
        const WalkState st = resume[p];
        uc.boxes = st.boxes;
        jump_resume(q, st, w);
        // every lane at once: lane j + 1 takes the upper sibling at marked
        // level j < 63, lane 0 keeps the current box and the deeper marks
        if (lane == 0) {
          w.low[0] &= 1ull << 63;
        } else {
          const uint32_t j = (uint32_t)lane - 1u;
          uint32_t k[3], ep;
          active = ((w.low[0] >> j) & 1ull) != 0 && jump_sibling_at(q, w, j, k, &ep);
          if (active) {
            w.key[0] = k[0];
            w.key[1] = k[1];
            w.key[2] = k[2];
            w.ep = ep;
            w.level = j + 1u;
            w.ready = w.key[0] < w.best;
          }
          w.low[0] = w.low[1] = 0;
        }
      }
      const int r = wave_certify_jump<kSplit>(q, p, lds[wv].stack, w, active, uc, tol, max_itr, so);
      if (uc.listed) {
        close_unit(so, p, w.best, uc, r == kWaveOver);
      } else if (lane == 0) {
        write_result(&out[p], jump_best64(w.best), uc.boxes, r == kWaveOver ? P2C_TUV_FLAG_OVER_BUDGET : 0u);
        if (r == kWaveOver) flag_for_replay(p, replay_list, replay_count, list_cap);
      }
      __builtin_amdgcn_wave_barrier();
      continue;
    }
    WaveTask* queue_wv = lds[wv].queue;
    LaneWalk w;
    uint64_t boxes = 0;
    start_walk(w);
    if (resume != nullptr) {
      const WalkState st = resume[p];
      boxes = st.boxes;
      w.best = st.best;
      if (lane == 0) load_walk(w, st);
    }
    if (lane == 0) best_slot[wv] = w.best;
    wave_sync_lds();
    const int r =
        wave_certify<false, kSplit>(q, p, queue_wv, &best_slot[wv], w, lane == 0, 0, 0, boxes, tol, max_itr, so);
    if (r != kWaveSplit) {
      const uint64_t final_best = best_slot[wv] < w.best ? best_slot[wv] : w.best;
      if (lane == 0) {
        write_result(&out[p], final_best, boxes, r == kWaveOver ? P2C_TUV_FLAG_OVER_BUDGET : 0u);
        if (r == kWaveOver) flag_for_replay(p, replay_list, replay_count, list_cap);
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  if constexpr (kSplit) serve_tasks(src, so, lds[wv], tol, max_itr);
}

// The closing task pass, one wave per piece, without further pieces: the
// written jump-walk pieces not marked done (normally none; read here as
// plain memory, written before this kernel began), then every piece of the
// 64-bit-key walk.
template <typename Src>
__global__ __launch_bounds__(kTuvBlock, 3) void tuv_task_kernel(Src src, double tol, uint64_t max_itr,
                                                             SplitOut so_in) {
  __shared__ WaveLds lds[kTuvBlock / 64];
  __shared__ unsigned long long best_slot[kTuvBlock / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const SplitOut& so = so_in;
  uint32_t jump_total = *so.task_count;
  if (jump_total > so.jump_cap) jump_total = so.jump_cap;
  const uint32_t waves = gridDim.x * (kTuvBlock / 64);
  for (uint32_t h = blockIdx.x * (kTuvBlock / 64) + (uint32_t)wv; h < jump_total; h += waves) {
    const auto* words = reinterpret_cast<const JumpTaskWords*>(so.tasks + h)->w;
    const uint64_t c = words[kControlWord];
    if (tag_ok(c, pass_epoch(so)) && (uint32_t)(c >> 32) == kSlotDone) continue;
    uint32_t t[kPieceWords];
#pragma unroll
    for (int i = 0; i < kPieceWords; ++i) t[i] = (uint32_t)(words[i] >> 32);
    run_jump_task<false>(src, so, unpack_piece(t), lds[wv], tol, max_itr);
    __builtin_amdgcn_wave_barrier();
  }
  uint32_t lw_total = *so.lw_count;
  if (lw_total > so.task_cap - so.jump_cap) lw_total = so.task_cap - so.jump_cap;
  for (;;) {
    uint32_t h = 0;
    if (lane == 0) h = atomicAdd(so.lw_next, 1u);
    h = __shfl(h, 0);
    if (h >= lw_total) break;
    run_lw_task(src, so, so.tasks[so.jump_cap + h], lds[wv], &best_slot[wv], tol, max_itr);
    __builtin_amdgcn_wave_barrier();
  }
}

// Results of the split pairs once all their tasks are done.
__global__ __launch_bounds__(kTuvBlock) void tuv_split_finish_kernel(SplitOut so_in, uint64_t max_itr,
                                                                     p2c_tuv_result* __restrict__ out,
                                                                     uint32_t* __restrict__ replay_list,
                                                                     uint32_t* __restrict__ replay_count,
                                                                     uint32_t list_cap) {
  const SplitOut& so = so_in;
  if (blockIdx.x == 0 && threadIdx.x == 0) *so.report = *so.task_count;
  uint32_t total = *so.split_count;
  if (total > list_cap) total = list_cap;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    const uint32_t p = so.split_list[i];
    const SplitAcc a = so.acc[p];
    const bool over = a.over != 0 || a.boxes > max_itr;
    write_result(&out[p], a.best, a.boxes, over ? P2C_TUV_FLAG_OVER_BUDGET : 0u);
    if (over) flag_for_replay(p, replay_list, replay_count, list_cap);
  }
}

// The pass's epoch, kept in the queue head between calls: one more than
// the last (31 bits, never 0); the queue's counters start at 0.
__global__ void tuv_split_epoch_kernel(QueueHead* head) {
  if (threadIdx.x == 0) {
    const uint32_t e = (head->epoch.v + 1u) & 0x7FFFFFFFu;
    head->epoch.v = e == 0 ? 1u : e;
    head->count.v = 0;
    head->next.v = 0;
    head->pending.v = 0;
  }
}

// ---- canonical best-first replay --------------------------------------
struct HeapBox {
  uint64_t kt, ku, kv;
  uint32_t level;
  uint32_t pad;
};

__device__ inline bool box_before(const HeapBox& a, const HeapBox& b) {
  if (a.kt != b.kt) return a.kt < b.kt;
  if (a.level != b.level) return a.level > b.level;
  if (a.ku != b.ku) return a.ku < b.ku;
  return a.kv < b.kv;
}

__device__ inline void heap_push(HeapBox* h, int64_t& n, const HeapBox& b) {
  int64_t i = n++;
  while (i > 0) {
    const int64_t p = (i - 1) >> 1;
    if (!box_before(b, h[p])) break;
    h[i] = h[p];
    i = p;
  }
  h[i] = b;
}

__device__ inline HeapBox heap_pop(HeapBox* h, int64_t& n) {
  const HeapBox top = h[0];
  const HeapBox last = h[--n];
  int64_t i = 0;
  for (;;) {
    int64_t c = 2 * i + 1;
    if (c >= n) break;
    if (c + 1 < n && box_before(h[c + 1], h[c])) ++c;
    if (!box_before(h[c], last)) break;
    h[i] = h[c];
    i = c;
  }
  if (n > 0) h[i] = last;
  return top;
}

// Canonical order and caps (see the certificate definition): stops before the
// first box of a new t_lo once more than max_itr boxes of smaller t_lo were
// processed, or at 16 * max_itr boxes, or when the heap is full.
__device__ void tuv_canonical(const TuvPair& q, double tol, int64_t max_itr, HeapBox* heap, int64_t heap_cap,
                              uint64_t* best_out, uint64_t* boxes_out, bool* capped_out) {
  int64_t hn = 0;
  heap_push(heap, hn, HeapBox{0, 0, 0, 0, 0});
  uint64_t best = kTuvNoHit, last_t = kTuvNoHit, cap_key = 0;
  int64_t popped = 0;
  bool capped = false;
  const int64_t hard_cap = 16 * max_itr;
  while (hn > 0) {
    const HeapBox b = heap_pop(heap, hn);
    if (b.kt >= best) break;
    if (b.kt != last_t) {
      last_t = b.kt;
      if (popped > max_itr) {
        capped = true;
        cap_key = b.kt;
        break;
      }
    }
    if (popped >= hard_cap) {
      capped = true;
      cap_key = b.kt;
      break;
    }
    ++popped;
    int e[3];
    tuv_exponents(q, b.level, e);
    const int outcome = tuv_classify(q, b.kt, b.ku, b.kv, e, b.level, tol);
    if (outcome == 0) continue;
    if (outcome == 1) {
      if (b.kt < best) best = b.kt;
      continue;
    }
    const int d = tuv_level_dim(q, b.level);
    const uint64_t half = kOne63 >> (get3(e, d) + 1);
    HeapBox lo = b, hi = b;
    lo.level = hi.level = b.level + 1;
    if (d == 0) hi.kt += half;
    else if (d == 1) hi.ku += half;
    else hi.kv += half;
    const HeapBox ch[2] = {lo, hi};
    for (int c = 0; c < 2; ++c) {
      if (!q.is_ee && ch[c].ku + ch[c].kv > kOne63) continue;
      if (hn >= heap_cap) {
        capped = true;
        cap_key = b.kt;
        goto done;
      }
      heap_push(heap, hn, ch[c]);
    }
  }
done:
  *best_out = capped ? cap_key : best;
  *boxes_out = (uint64_t)popped;
  *capped_out = capped;
}

template <typename Src>
__global__ __launch_bounds__(64) void tuv_replay_kernel(Src src,
                                                        const uint32_t* __restrict__ list,
                                                        const uint32_t* __restrict__ count, uint32_t list_cap,
                                                        double tol, int64_t max_itr, HeapBox* __restrict__ heaps,
                                                        int64_t heap_cap, int64_t slots,
                                                        p2c_tuv_result* __restrict__ out) {
  const int64_t slot = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (slot >= slots) return;
  uint32_t m = *count;
  if (m > list_cap) m = list_cap;
  for (int64_t j = slot; j < (int64_t)m; j += slots) {
    const uint32_t p = list[j];
    const TuvPair q = src.load(p);
    uint64_t best, boxes;
    bool capped;
    tuv_canonical(q, tol, max_itr, heaps + slot * heap_cap, heap_cap, &best, &boxes, &capped);
    write_result(&out[p], best, boxes, P2C_TUV_FLAG_REPLAYED | (capped ? P2C_TUV_FLAG_CAPPED : 0u));
  }
}

int check_tuv(double ms, double tol, double t_max, int64_t max_itr) {
  if (!(tol > 0.0) || !std::isfinite(tol)) return set_error(P2C_EINVAL, "tolerance must be finite and positive");
  if (!(ms >= 0.0) || !std::isfinite(ms)) return set_error(P2C_EINVAL, "minimum separation must be finite and >= 0");
  if (t_max != 1.0) return set_error(P2C_EUNSUPPORTED, "t_max must be 1 (the full step)");
  if (max_itr <= 0) return set_error(P2C_EINVAL, "max_itr must be positive");
  return P2C_OK;
}

// ---- launches over a pair source ----------------------------------------
template <typename Src>
int launch_stream(const Src& src, int64_t n, const uint64_t* totals, double tolerance, int64_t max_itr,
                  int64_t light_budget, int refill_at, int handover_at, p2c_tuv_result* out, void* counters,
                  void* states, uint32_t* heavy_list, uint32_t list_cap, int grid_blocks, void* stream) {
  if (int rc = check_tuv(0.0, tolerance, 1.0, max_itr)) return rc;
  if (n <= 0) return P2C_OK;
  if (grid_blocks <= 0) return set_error(P2C_EINVAL, "grid_blocks must be positive");
  if (light_budget <= 0) return set_error(P2C_EINVAL, "light_budget must be positive");
  if (refill_at < 1 || refill_at > 64) return set_error(P2C_EINVAL, "refill_at must be in 1..64");
  if (handover_at < 0 || handover_at > 64) return set_error(P2C_EINVAL, "handover_at must be in 0..64");
  if (light_budget > max_itr) light_budget = max_itr;
  // counters: [0] next pair (u64), [1] replay count (u32) | heavy count (u32), [2] heavy cursor (u32)
  auto* c64 = static_cast<unsigned long long*>(counters);
  auto* c32 = static_cast<uint32_t*>(counters);
  auto* st = static_cast<WalkState*>(states);
  hipLaunchKernelGGL(tuv_stream_kernel<Src>, dim3(grid_blocks), dim3(kTuvBlock), 0, as_stream(stream), src, n,
                     totals, tolerance, (uint64_t)light_budget, (uint32_t)refill_at, (uint32_t)handover_at, out, c64,
                     st, heavy_list, c32 + 3, list_cap);
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

template <typename Src>
int launch_wave_split(const Src& src, void* counters, void* states, const uint32_t* heavy_list, uint32_t list_cap,
                      double tolerance, int64_t max_itr, p2c_tuv_result* out, uint32_t* replay_list, int64_t split_at,
                      void* tasks, uint32_t task_cap, uint32_t* split_list, int grid_blocks, void* stream) {
  if (int rc = check_tuv(0.0, tolerance, 1.0, max_itr)) return rc;
  if (grid_blocks <= 0) return set_error(P2C_EINVAL, "grid_blocks must be positive");
  if (split_at < 0) return set_error(P2C_EINVAL, "split_at must be non-negative");
  const bool split = split_at > 0;
  if (split && (tasks == nullptr || task_cap < kHeadSlots + 4 || split_list == nullptr))
    return set_error(P2C_EINVAL, "splitting needs a task buffer of at least 20 slots and a split list");
  if (split && list_cap > 0x7FFFFFFFu) return set_error(P2C_EINVAL, "split lists hold pair ids below 2^31");
  auto* c32 = static_cast<uint32_t*>(counters);
  hipStream_t s = as_stream(stream);
  SplitOut so{};
  if (split) {
    // counters (u32): [8] 64-bit-walk pieces, [11] their cursor, [13] jump-walk pieces (reported
    // at the end), [15] split pairs; the last kHeadSlots task slots hold the jump-walk queue's
    // head (its epoch kept between calls); a quarter of the others holds the 64-bit walk's pieces
    so.split_at = (uint64_t)split_at;
    so.tasks = static_cast<SplitTask*>(tasks);
    so.task_cap = task_cap - kHeadSlots;
    so.jump_cap = so.task_cap - so.task_cap / 4;
    const uintptr_t head = (reinterpret_cast<uintptr_t>(so.tasks + so.task_cap) + 127u) & ~uintptr_t(127);
    so.head = reinterpret_cast<QueueHead*>(head);
    so.task_count = &so.head->count.v;
    so.task_next = &so.head->next.v;
    so.report = c32 + 13;
    so.lw_count = c32 + 8;
    so.lw_next = c32 + 11;
    so.pending = &so.head->pending.v;
    so.split_list = split_list;
    so.split_count = c32 + 15;
    so.acc = static_cast<SplitAcc*>(states);
    hipLaunchKernelGGL(tuv_split_epoch_kernel, dim3(1), dim3(64), 0, s, so.head);
  }
  const auto* resume = static_cast<const WalkState*>(states);
  if (split)
    hipLaunchKernelGGL((tuv_heavy_kernel<true, Src>), dim3(grid_blocks), dim3(kTuvBlock), 0, s, src, heavy_list,
                       c32 + 3, list_cap, c32 + 4, resume, tolerance, (uint64_t)max_itr, out, replay_list, c32 + 2, so);
  else
    hipLaunchKernelGGL((tuv_heavy_kernel<false, Src>), dim3(grid_blocks), dim3(kTuvBlock), 0, s, src, heavy_list,
                       c32 + 3, list_cap, c32 + 4, resume, tolerance, (uint64_t)max_itr, out, replay_list, c32 + 2, so);
  if (split) {
    hipLaunchKernelGGL(tuv_task_kernel<Src>, dim3(grid_blocks), dim3(kTuvBlock), 0, s, src, tolerance,
                       (uint64_t)max_itr, so);
    hipLaunchKernelGGL(tuv_split_finish_kernel, dim3(16), dim3(kTuvBlock), 0, s, so, (uint64_t)max_itr, out,
                       replay_list, c32 + 2, list_cap);
  }
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

template <typename Src>
int launch_replay(const Src& src, const uint32_t* replay_list, const uint32_t* replay_count, uint32_t replay_cap,
                  double tolerance, int64_t max_itr, void* heaps, int64_t heap_cap, int64_t slots,
                  p2c_tuv_result* out, void* stream) {
  if (int rc = check_tuv(0.0, tolerance, 1.0, max_itr)) return rc;
  if (heap_cap < 8 || slots <= 0) return set_error(P2C_EINVAL, "replay needs heap_cap >= 8 and slots > 0");
  hipLaunchKernelGGL(tuv_replay_kernel<Src>, dim3((unsigned)((slots + 63) / 64)), dim3(64), 0, as_stream(stream), src,
                     replay_list, replay_count, replay_cap, tolerance, max_itr, static_cast<HeapBox*>(heaps), heap_cap,
                     slots, out);
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

int lite_src(const void* pairs, const p2c_tuv_vertices* verts, LiteSrc* src) {
  if (verts == nullptr || verts->x0 == nullptr || verts->x1 == nullptr)
    return set_error(P2C_EINVAL, "compact pair records need the vertex arrays");
  if (!(verts->ms >= 0.0) || !std::isfinite(verts->ms))
    return set_error(P2C_EINVAL, "minimum separation must be finite and >= 0");
  *src = LiteSrc{static_cast<const TuvLite*>(pairs), verts->x0, verts->x1, verts->ms};
  return P2C_OK;
}

// ---- allocation-free certificate ----------------------------------------
// The stream -> wave -> replay sequence for callers that may not allocate:
// every scratch buffer it needs lives in one file-scope device array.
constexpr int64_t kStaticMaxPairs = int64_t(1) << 18;  // pairs per call
constexpr int64_t kStaticHeapCap = int64_t(1) << 16;   // replay heap entries per slot
constexpr int64_t kStaticSlots = 8;                     // pairs replayed concurrently
constexpr int64_t kStaticLightBudget = 1024;
constexpr int kStaticRefillAt = 8, kStaticHandoverAt = 8, kStaticBlocksPerCu = 3;
constexpr size_t kScratchCountersOff = 0;
constexpr size_t kScratchStatesOff = 256;
constexpr size_t kScratchHeavyOff = kScratchStatesOff + sizeof(WalkState) * size_t(kStaticMaxPairs);
constexpr size_t kScratchReplayOff = kScratchHeavyOff + sizeof(uint32_t) * size_t(kStaticMaxPairs);
constexpr size_t kScratchHeapsOff = kScratchReplayOff + sizeof(uint32_t) * size_t(kStaticMaxPairs);
constexpr size_t kScratchBytes = kScratchHeapsOff + sizeof(HeapBox) * size_t(kStaticHeapCap * kStaticSlots);
static_assert(kScratchHeavyOff % 256 == 0 && kScratchReplayOff % 256 == 0 && kScratchHeapsOff % 256 == 0,
              "scratch sections stay 256-byte aligned");

__global__ void zero_words_kernel(uint64_t* p, int n) {
  const int i = int(threadIdx.x);
  if (i < n) p[i] = 0;
}

}  // namespace
}  // namespace p2c

// One copy per device, held from module load until the process ends.  Host
// code never passes this name to a kernel: the device address comes from
// hipGetSymbolAddress, once per device (scratch_on_current_device).
static __device__ __align__(256) unsigned char scratch[p2c::kScratchBytes];

namespace p2c {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];
int g_scratch_grid[kMaxDevices];

// The current device's copy of `scratch` and the certificate grid for it,
// looked up on the first call on that device with the device current.
int scratch_on_current_device(unsigned char** ptr, int* grid_blocks) {
  int dev = 0;
  P2C_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return set_error(P2C_EUNSUPPORTED, "device ordinal beyond 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (p == nullptr) return set_error(P2C_EHIP, "hipGetSymbolAddress(scratch) returned no address");
    int cus = 0;
    P2C_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (cus <= 0) return set_error(P2C_EHIP, "device reports no compute units");
    g_scratch_grid[dev] = cus * kStaticBlocksPerCu;
    g_scratch_ptr[dev] = p;
  }
  *ptr = static_cast<unsigned char*>(g_scratch_ptr[dev]);
  *grid_blocks = g_scratch_grid[dev];
  return P2C_OK;
}

}  // namespace
}  // namespace p2c

extern "C" {

size_t p2c_tuv_pair_bytes(void) { return sizeof(p2c::TuvPair); }

size_t p2c_tuv_pairs_bytes(int64_t n) {
  if (n < 0) n = 0;
  const int64_t tiles = (n + p2c::kPairTile - 1) / p2c::kPairTile;
  return size_t(tiles) * p2c::kPairTile * sizeof(p2c::TuvPair);
}

size_t p2c_tuv_heap_bytes(int64_t heap_cap) { return sizeof(p2c::HeapBox) * size_t(heap_cap); }

int p2c_tuv_setup_positions(int is_ee, const double* v, int64_t n, double ms, double tolerance, void* pairs,
                            void* stream) {
  if (int rc = p2c::check_tuv(ms, tolerance, 1.0, 1)) return rc;
  if (n <= 0) return P2C_OK;
  hipLaunchKernelGGL(p2c::tuv_setup_positions_kernel, dim3(p2c::grid_for(n, p2c::kTuvBlock)),
                     dim3(p2c::kTuvBlock), 0, p2c::as_stream(stream), is_ee, v, n, ms, tolerance,
                     static_cast<p2c::TuvPair*>(pairs));
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

int p2c_tuv_setup_indexed(int is_ee, const double* x0, const double* x1, const int32_t* idx4, int64_t n,
                          double ms, double tolerance, void* pairs, void* stream) {
  if (int rc = p2c::check_tuv(ms, tolerance, 1.0, 1)) return rc;
  if (n <= 0) return P2C_OK;
  hipLaunchKernelGGL(p2c::tuv_setup_indexed_kernel, dim3(p2c::grid_for(n, p2c::kTuvBlock)),
                     dim3(p2c::kTuvBlock), 0, p2c::as_stream(stream), is_ee, x0, x1, idx4, n, ms, tolerance,
                     static_cast<p2c::TuvPair*>(pairs));
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

int p2c_tuv_setup_pairs(int is_ee, const double* x0, const double* x1, const int32_t* topo, const int32_t* pairs,
                        int64_t n, const uint64_t* totals, int64_t out_offset, int64_t out_cap, double ms,
                        double tolerance, void* tuv_pairs, void* stream) {
  if (int rc = p2c::check_tuv(ms, tolerance, 1.0, 1)) return rc;
  if (n <= 0) return P2C_OK;
  hipLaunchKernelGGL(p2c::tuv_setup_pairs_kernel, dim3(p2c::grid_for(n, p2c::kTuvBlock)), dim3(p2c::kTuvBlock), 0,
                     p2c::as_stream(stream), is_ee, x0, x1, topo, pairs, n, totals, out_offset, out_cap, ms,
                     tolerance, static_cast<p2c::TuvPair*>(tuv_pairs));
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

int p2c_tuv_plain(const void* pairs, int64_t n, const uint64_t* totals, double tolerance, int64_t max_itr,
                  p2c_tuv_result* out, uint32_t* replay_list, uint32_t* replay_count, uint32_t replay_cap,
                  void* stream) {
  if (int rc = p2c::check_tuv(0.0, tolerance, 1.0, max_itr)) return rc;
  if (n <= 0) return P2C_OK;
  hipLaunchKernelGGL(p2c::tuv_plain_kernel<p2c::FullSrc>, dim3(p2c::grid_for(n, p2c::kTuvBlock)),
                     dim3(p2c::kTuvBlock), 0, p2c::as_stream(stream),
                     p2c::FullSrc{static_cast<const p2c::TuvPair*>(pairs)}, n, totals, tolerance, (uint64_t)max_itr,
                     out, replay_list, replay_count, replay_cap);
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

int p2c_tuv_plain_lite(const void* lite_pairs, const p2c_tuv_vertices* verts, int64_t n, const uint64_t* totals,
                       double tolerance, int64_t max_itr, p2c_tuv_result* out, uint32_t* replay_list,
                       uint32_t* replay_count, uint32_t replay_cap, void* stream) {
  if (int rc = p2c::check_tuv(0.0, tolerance, 1.0, max_itr)) return rc;
  p2c::LiteSrc src;
  if (int rc = p2c::lite_src(lite_pairs, verts, &src)) return rc;
  if (n <= 0) return P2C_OK;
  hipLaunchKernelGGL(p2c::tuv_plain_kernel<p2c::LiteSrc>, dim3(p2c::grid_for(n, p2c::kTuvBlock)),
                     dim3(p2c::kTuvBlock), 0, p2c::as_stream(stream), src, n, totals, tolerance, (uint64_t)max_itr,
                     out, replay_list, replay_count, replay_cap);
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

size_t p2c_tuv_state_bytes(void) { return sizeof(p2c::WalkState); }

int p2c_tuv_persistent(const void* pairs, int64_t n, const uint64_t* totals, double tolerance, int64_t max_itr,
                       int64_t light_budget, p2c_tuv_result* out, void* counters, uint32_t* heavy_list,
                       uint32_t* replay_list, uint32_t list_cap, int grid_blocks, void* stream) {
  if (int rc = p2c::check_tuv(0.0, tolerance, 1.0, max_itr)) return rc;
  if (n <= 0) return P2C_OK;
  if (grid_blocks <= 0) return p2c::set_error(P2C_EINVAL, "grid_blocks must be positive");
  if (light_budget <= 0) return p2c::set_error(P2C_EINVAL, "light_budget must be positive");
  if (light_budget > max_itr) light_budget = max_itr;
  // counters: [0] next pair (u64), [1] replay count (u32) | heavy count (u32), [2] heavy cursor (u32)
  auto* c64 = static_cast<unsigned long long*>(counters);
  auto* c32 = static_cast<uint32_t*>(counters);
  hipStream_t s = p2c::as_stream(stream);
  hipLaunchKernelGGL(p2c::tuv_persistent_kernel, dim3(grid_blocks), dim3(p2c::kTuvBlock), 0, s,
                     static_cast<const p2c::TuvPair*>(pairs), n, totals, tolerance, (uint64_t)light_budget, out,
                     c64, heavy_list, c32 + 3, list_cap);
  hipLaunchKernelGGL((p2c::tuv_heavy_kernel<false, p2c::FullSrc>), dim3(grid_blocks), dim3(p2c::kTuvBlock), 0, s,
                     p2c::FullSrc{static_cast<const p2c::TuvPair*>(pairs)}, heavy_list, c32 + 3, list_cap, c32 + 4,
                     nullptr,
                     tolerance, (uint64_t)max_itr, out, replay_list, c32 + 2, p2c::SplitOut{});
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

int p2c_tuv_rounds(const void* pairs, int64_t n, const uint64_t* totals, double tolerance, int64_t max_itr,
                   const uint32_t* budgets, int n_rounds, p2c_tuv_result* out, void* counters, void* states,
                   uint32_t* list_a, uint32_t* list_b, uint32_t* replay_list, uint32_t list_cap, int grid_blocks,
                   void* stream) {
  if (int rc = p2c::check_tuv(0.0, tolerance, 1.0, max_itr)) return rc;
  if (n <= 0) return P2C_OK;
  if (n_rounds < 1 || n_rounds > 8) return p2c::set_error(P2C_EINVAL, "1 to 8 rounds");
  if (grid_blocks <= 0) return p2c::set_error(P2C_EINVAL, "grid_blocks must be positive");
  for (int r = 0; r < n_rounds; ++r)
    if (budgets[r] == 0) return p2c::set_error(P2C_EINVAL, "round budgets must be positive");
  // counters (u32): [2] replay count, [3] heavy count (list of the last round),
  // [4] heavy cursor, [5..12] per-round survivor counts
  auto* c32 = static_cast<uint32_t*>(counters);
  hipStream_t s = p2c::as_stream(stream);
  auto* st = static_cast<p2c::WalkState*>(states);
  const auto* tp = static_cast<const p2c::TuvPair*>(pairs);
  uint32_t* lists[2] = {list_a, list_b};
  for (int r = 0; r < n_rounds; ++r) {
    const bool last = r == n_rounds - 1;
    const uint32_t* lin = r == 0 ? nullptr : lists[(r - 1) & 1];
    const uint32_t* cin = r == 0 ? nullptr : c32 + 5 + (r - 1);
    uint32_t* lout = last ? lists[r & 1] : lists[r & 1];
    uint32_t* cout = last ? c32 + 3 : c32 + 5 + r;
    hipLaunchKernelGGL(p2c::tuv_round_kernel, dim3(grid_blocks), dim3(p2c::kTuvBlock), 0, s, tp, n, totals, lin, cin,
                       st, tolerance, budgets[r], (uint64_t)max_itr, out, lout, cout, last ? 1 : 0, replay_list,
                       c32 + 2, list_cap);
  }
  const uint32_t* heavy = lists[(n_rounds - 1) & 1];
  hipLaunchKernelGGL((p2c::tuv_heavy_kernel<false, p2c::FullSrc>), dim3(grid_blocks), dim3(p2c::kTuvBlock), 0, s,
                     p2c::FullSrc{tp}, heavy, c32 + 3,
                     list_cap, c32 + 4, st, tolerance, (uint64_t)max_itr, out, replay_list, c32 + 2, p2c::SplitOut{});
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

int p2c_tuv_stream(const void* pairs, int64_t n, const uint64_t* totals, double tolerance, int64_t max_itr,
                   int64_t light_budget, int refill_at, int handover_at, p2c_tuv_result* out, void* counters,
                   void* states, uint32_t* heavy_list, uint32_t list_cap, int grid_blocks, void* stream) {
  return p2c::launch_stream(p2c::FullSrc{static_cast<const p2c::TuvPair*>(pairs)}, n, totals, tolerance, max_itr,
                            light_budget, refill_at, handover_at, out, counters, states, heavy_list, list_cap,
                            grid_blocks, stream);
}

int p2c_tuv_wave(const void* pairs, void* counters, const void* states, const uint32_t* heavy_list,
                 uint32_t list_cap, double tolerance, int64_t max_itr, p2c_tuv_result* out, uint32_t* replay_list,
                 int grid_blocks, void* stream) {
  return p2c_tuv_wave_split(pairs, counters, const_cast<void*>(states), heavy_list, list_cap, tolerance, max_itr, out,
                            replay_list, 0, nullptr, 0, nullptr, grid_blocks, stream);
}

size_t p2c_tuv_task_bytes(void) { return sizeof(p2c::SplitTask); }

int p2c_tuv_wave_split(const void* pairs, void* counters, void* states, const uint32_t* heavy_list,
                       uint32_t list_cap, double tolerance, int64_t max_itr, p2c_tuv_result* out,
                       uint32_t* replay_list, int64_t split_at, void* tasks, uint32_t task_cap, uint32_t* split_list,
                       int grid_blocks, void* stream) {
  return p2c::launch_wave_split(p2c::FullSrc{static_cast<const p2c::TuvPair*>(pairs)}, counters, states, heavy_list,
                                list_cap, tolerance, max_itr, out, replay_list, split_at, tasks, task_cap, split_list,
                                grid_blocks, stream);
}

int p2c_tuv_replay(const void* pairs, const uint32_t* replay_list, const uint32_t* replay_count,
                   uint32_t replay_cap, double tolerance, int64_t max_itr, void* heaps, int64_t heap_cap,
                   int64_t slots, p2c_tuv_result* out, void* stream) {
  return p2c::launch_replay(p2c::FullSrc{static_cast<const p2c::TuvPair*>(pairs)}, replay_list, replay_count,
                            replay_cap, tolerance, max_itr, heaps, heap_cap, slots, out, stream);
}

size_t p2c_tuv_lite_bytes(int64_t n) {
  if (n < 0) n = 0;
  return size_t((n + p2c::kPairTile - 1) / p2c::kPairTile) * p2c::kPairTile * sizeof(p2c::TuvLite);
}

int p2c_tuv_setup_pairs_lite(int is_ee, const double* x0, const double* x1, const int32_t* topo,
                             const int32_t* pairs, int64_t n, const uint64_t* totals, int64_t out_offset,
                             int64_t out_cap, double ms, double tolerance, void* lite_pairs, void* stream) {
  if (int rc = p2c::check_tuv(ms, tolerance, 1.0, 1)) return rc;
  if (n <= 0) return P2C_OK;
  hipLaunchKernelGGL(p2c::tuv_setup_lite_kernel, dim3(p2c::grid_for(n, p2c::kTuvBlock)), dim3(p2c::kTuvBlock), 0,
                     p2c::as_stream(stream), is_ee, x0, x1, topo, pairs, n, totals, out_offset, out_cap, ms,
                     tolerance, static_cast<p2c::TuvLite*>(lite_pairs));
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

int p2c_tuv_setup_pairs_lite_batch(const double* x0, const double* x1, const p2c_tuv_scene* scenes,
                                   int32_t n_scenes, const int32_t* pairs, const uint64_t* offsets, int64_t n,
                                   double ms, double tolerance, void* lite_pairs, void* stream) {
  if (int rc = p2c::check_tuv(ms, tolerance, 1.0, 1)) return rc;
  if (n_scenes < 1) return p2c::set_error(P2C_EINVAL, "a batch holds at least one scene");
  if (n <= 0) return P2C_OK;
  if (x0 == nullptr || x1 == nullptr || scenes == nullptr || pairs == nullptr || offsets == nullptr ||
      lite_pairs == nullptr)
    return p2c::set_error(P2C_EINVAL, "batch setup: null pointer");
  hipLaunchKernelGGL(p2c::tuv_setup_lite_batch_kernel, dim3(p2c::grid_for(n, p2c::kTuvBlock)), dim3(p2c::kTuvBlock),
                     0, p2c::as_stream(stream), x0, x1, scenes, int(n_scenes), pairs, offsets, n, ms, tolerance,
                     static_cast<p2c::TuvLite*>(lite_pairs));
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

int p2c_tuv_stream_lite(const void* lite_pairs, const p2c_tuv_vertices* verts, int64_t n, const uint64_t* totals,
                        double tolerance, int64_t max_itr, int64_t light_budget, int refill_at, int handover_at,
                        p2c_tuv_result* out, void* counters, void* states, uint32_t* heavy_list, uint32_t list_cap,
                        int grid_blocks, void* stream) {
  p2c::LiteSrc src;
  if (int rc = p2c::lite_src(lite_pairs, verts, &src)) return rc;
  return p2c::launch_stream(src, n, totals, tolerance, max_itr, light_budget, refill_at, handover_at, out, counters,
                            states, heavy_list, list_cap, grid_blocks, stream);
}

int p2c_tuv_wave_split_lite(const void* lite_pairs, const p2c_tuv_vertices* verts, void* counters, void* states,
                            const uint32_t* heavy_list, uint32_t list_cap, double tolerance, int64_t max_itr,
                            p2c_tuv_result* out, uint32_t* replay_list, int64_t split_at, void* tasks,
                            uint32_t task_cap, uint32_t* split_list, int grid_blocks, void* stream) {
  p2c::LiteSrc src;
  if (int rc = p2c::lite_src(lite_pairs, verts, &src)) return rc;
  return p2c::launch_wave_split(src, counters, states, heavy_list, list_cap, tolerance, max_itr, out, replay_list,
                                split_at, tasks, task_cap, split_list, grid_blocks, stream);
}

int p2c_tuv_replay_lite(const void* lite_pairs, const p2c_tuv_vertices* verts, const uint32_t* replay_list,
                        const uint32_t* replay_count, uint32_t replay_cap, double tolerance, int64_t max_itr,
                        void* heaps, int64_t heap_cap, int64_t slots, p2c_tuv_result* out, void* stream) {
  p2c::LiteSrc src;
  if (int rc = p2c::lite_src(lite_pairs, verts, &src)) return rc;
  return p2c::launch_replay(src, replay_list, replay_count, replay_cap, tolerance, max_itr, heaps, heap_cap, slots,
                            out, stream);
}

size_t p2c_tuv_static_max_pairs(void) { return size_t(p2c::kStaticMaxPairs); }

int p2c_tuv_certify_static(const void* pairs, int64_t n, double tolerance, int64_t max_itr, p2c_tuv_result* out,
                           void* stream) {
  if (int rc = p2c::check_tuv(0.0, tolerance, 1.0, max_itr)) return rc;
  if (n < 0) return p2c::set_error(P2C_EINVAL, "pair count must be non-negative");
  if (n > p2c::kStaticMaxPairs)
    return p2c::set_error(P2C_EINVAL, "p2c_tuv_certify_static takes at most 262144 pairs per call");
  if (n == 0) return P2C_OK;
  if (pairs == nullptr || out == nullptr) return p2c::set_error(P2C_EINVAL, "pair and result buffers are required");
  unsigned char* base = nullptr;
  int grid = 0;
  if (int rc = p2c::scratch_on_current_device(&base, &grid)) return rc;
  void* counters = base + p2c::kScratchCountersOff;
  void* states = base + p2c::kScratchStatesOff;
  auto* heavy = reinterpret_cast<uint32_t*>(base + p2c::kScratchHeavyOff);
  auto* replay = reinterpret_cast<uint32_t*>(base + p2c::kScratchReplayOff);
  void* heaps = base + p2c::kScratchHeapsOff;
  const uint32_t cap = uint32_t(p2c::kStaticMaxPairs);
  // the previous call's counters are still there: clear them with a kernel
  // (no memset node, so the call may also be captured in a graph)
  hipLaunchKernelGGL(p2c::zero_words_kernel, dim3(1), dim3(64), 0, p2c::as_stream(stream),
                     static_cast<uint64_t*>(counters), 8);
  P2C_HIP_CHECK(hipGetLastError());
  if (int rc = p2c_tuv_stream(pairs, n, nullptr, tolerance, max_itr, p2c::kStaticLightBudget, p2c::kStaticRefillAt,
                              p2c::kStaticHandoverAt, out, counters, states, heavy, cap, grid, stream))
    return rc;
  if (int rc = p2c_tuv_wave(pairs, counters, states, heavy, cap, tolerance, max_itr, out, replay, grid, stream))
    return rc;
  return p2c_tuv_replay(pairs, replay, static_cast<const uint32_t*>(counters) + 2, cap, tolerance, max_itr, heaps,
                        p2c::kStaticHeapCap, p2c::kStaticSlots, out, stream);
}

}  // extern "C"
