// The reference-parity certificate walk and its two kernels (gfx950), shared
// by the per-batch entry points (one interval and config per launch) and the
// exact work queue (each queue entry's own interval and config).
//
// A light pass walks each primitive's interval tree in one lane for a few
// nodes; primitives that need more go to a wave pass, one wave per primitive,
// which evaluates up to 64 pending nodes per round and returns the same
// answer and node count as the sequential walk.  The walk is a compile-time
// mode: the CPU engine's (certificate_engine.cpp:300-386), the CUDA twin's
// (certificate_cuda.cu:286-346) or the Python twin's (pure_exact_cpu.py:
// 98-244).  The walk functions are host-callable too (the library's CPU entry
// point runs them).
#pragma once

#include "cert_ref_math.h"
#include "runtime.h"

namespace p2c {
namespace {
constexpr int kRefBlock = 256;
#ifndef P2C_REF_LIGHT_BUDGET
#define P2C_REF_LIGHT_BUDGET 8
#endif
#ifndef P2C_REF_HEAVY_CHUNK
#define P2C_REF_HEAVY_CHUNK 2
#endif
constexpr int kHeavyBlock = 64;      // one wave per over-budget primitive
constexpr int kHeavyChunk = P2C_REF_HEAVY_CHUNK;  // primitives each wave of the wave pass inspects
constexpr int kStackCap = 1024;      // pending nodes of the wave pass (LDS)
constexpr uint32_t kLightBudget = P2C_REF_LIGHT_BUDGET;  // nodes a lane walks before the wave pass takes over
constexpr uint8_t kHeavyMark = 0xFE;   // result status of a primitive left to the wave pass

struct RefCfg {
  double t0, t1, eps_time, eps_space;
  uint32_t max_depth;
};

// The walk, a template parameter of the kernels (kMode):
//   kModeEngine    the CPU engine's (certificate_engine.cpp:300-386): an
//                  interval that may not be split is noted and the search
//                  goes on for a collision;
//   kModeCudaTwin  the CUDA batch twin's (certificate_cuda.cu:286-346):
//                  UNDECIDED on the first such interval in pre-order;
//   kModePython    the Python twin's rules (pure_exact_cpu.py:98-244): EE
//                  denominator guard abs(denom) > 1e-14, motion radius
//                  0.5 * (t1 - t0) * (summed largest displacements), contact
//                  at the first smallest of the three samples.
enum : int { kModeEngine = 0, kModeCudaTwin = 1, kModePython = 2 };

enum : int { kNodeExpand = 0, kNodeCollision = 1, kNodeSeparated = 2, kNodeUndecided = 3 };

// The node's interval, rebuilt from the root by replaying the midpoints the
// recursion computes (bit k of `path` = right half taken at depth k).
P2C_HD void node_interval(const RefCfg& c, uint64_t path, uint32_t depth, double* a, double* b) {
  double x = c.t0, y = c.t1;
  for (uint32_t k = 0; k < depth; ++k) {
    const double m = 0.5 * (x + y);
    if ((path >> k) & 1ull) x = m; else y = m;
  }
  *a = x;
  *b = y;
}

// One node of certificate_engine.cpp:300-386 given the distances at its two
// ends (da at a, db at b): the endpoint/midpoint contact tests in the
// reference's order, then the Lipschitz separation test, then the leaf rule.
template <int kFamily, int kMode>
P2C_HD int eval_node(const RefPrim& q, const double vn[4], const RefCfg& c, double a, double b,
                                uint32_t depth, double da, double db, double* dm, double* toi, double* margin) {
  const double mid = 0.5 * (a + b);
  *dm = ref_distance<kFamily>(q, mid, kMode == kModePython);
  double lb;
  if (kMode == kModePython) {
    // min over (t0, mid, t1) by distance, the first of equals (Python min)
    double tm = a, dmin = da;
    if (*dm < dmin) { tm = mid; dmin = *dm; }
    if (db < dmin) { tm = b; dmin = db; }
    if (dmin <= c.eps_space) { *toi = tm; return kNodeCollision; }
    const double total = kFamily == P2C_FAMILY_POINT_TRIANGLE ? vn[0] + std_max(std_max(vn[1], vn[2]), vn[3])
                                                              : std_max(vn[0], vn[1]) + std_max(vn[2], vn[3]);
    lb = *dm - 0.5 * std_max(0.0, b - a) * total;
  } else {
    if (da <= c.eps_space) { *toi = a; return kNodeCollision; }
    if (*dm <= c.eps_space) { *toi = mid; return kNodeCollision; }
    if (db <= c.eps_space) { *toi = b; return kNodeCollision; }
    lb = *dm - ref_motion_radius_n<kFamily>(vn, a, b);
  }
  if (lb > c.eps_space) {
    *margin = std_max(0.0, lb - c.eps_space);
    return kNodeSeparated;
  }
  if (depth >= c.max_depth || b - a <= c.eps_time) return kNodeUndecided;
  return kNodeExpand;
}

enum : uint32_t { kWalkDone = 0, kWalkCollision = 1, kWalkOverBudget = 2 };

struct RefWalk {
  uint32_t state;
  uint32_t nodes;
  uint64_t path;  // the collision node (or, twin, the unsplittable one): position, interval, contact time
  uint32_t depth;
  double a, b, toi;
  double min_margin;
  bool have_margin, undecided;
  bool limited;  // twin only: the walk ended on an interval that may not be split
  uint32_t max_depth;  // deepest node visited
};

// Pre-order walk (left half first) of the subtree rooted at (path, depth),
// stopping at its first collision or after `budget` nodes.  da and db are the
// distances at the root's two ends.  A left child reuses its parent's start
// and midpoint distances; the right sibling after a finished left subtree
// starts where that subtree's last node ended, so only its end is new.
template <int kFamily, int kMode>
P2C_HD RefWalk walk_subtree(const RefPrim& q, const double vn[4], const RefCfg& c, uint64_t path,
                                uint32_t depth, double da, double db, uint32_t budget) {
  RefWalk w;
  w.state = kWalkDone;
  w.nodes = 0;
  w.path = 0;
  w.depth = 0;
  w.a = w.b = w.toi = 0.0;
  w.min_margin = 0.0;
  w.have_margin = w.undecided = w.limited = false;
  w.max_depth = depth;
  const uint32_t depth0 = depth;
  bool db_known = true;
  for (;;) {
    if (w.nodes == budget) {
      w.state = kWalkOverBudget;
      return w;
    }
    double a, b;
    node_interval(c, path, depth, &a, &b);
    if (!db_known) db = ref_distance<kFamily>(q, b, kMode == kModePython);
    ++w.nodes;
    if (kMode == kModePython) w.max_depth = depth > w.max_depth ? depth : w.max_depth;  // reported by that walk only
    double dm, toi = 0.0, margin = 0.0;
    const int o = eval_node<kFamily, kMode>(q, vn, c, a, b, depth, da, db, &dm, &toi, &margin);
    if (o == kNodeCollision || (kMode == kModeCudaTwin && o == kNodeUndecided)) {
      w.state = kWalkCollision;
      w.limited = o == kNodeUndecided;
      w.path = path;
      w.depth = depth;
      w.a = a;
      w.b = b;
      w.toi = toi;
      return w;
    }
    if (o == kNodeExpand) {  // left child [a, mid]
      ++depth;
      db = dm;
      continue;
    }
    if (o == kNodeSeparated) {
      w.min_margin = w.have_margin ? std_min(w.min_margin, margin) : margin;
      w.have_margin = true;
    } else {
      w.undecided = true;
    }
    while (depth > depth0 && ((path >> (depth - 1)) & 1ull)) {
      path &= ~(1ull << (depth - 1));
      --depth;
    }
    if (depth == depth0) return w;
    path |= 1ull << (depth - 1);  // right sibling [m, b'], m = this node's end
    da = db;
    db_known = false;
  }
}

template <int kFamily>
P2C_HD void write_result(const RefCfg& c, int64_t id_a, int64_t id_b, bool collision, bool limited, double a,
                             double b, double toi, bool undecided, double min_margin, uint32_t nodes,
                             uint32_t max_depth, p2c_prim_result* out) {
  p2c_prim_result r;
  if (collision && limited) {  // twin: UNDECIDED on the unsplittable interval (certificate_cuda.cu:328-335)
    init_result(&r, P2C_STATUS_UNDECIDED, kFamily, a, b);
    r.reason_code = 2;
    r.next_refinement_mode = 1;
  } else if (collision) {
    init_result(&r, P2C_STATUS_COLLISION, kFamily, a, b);
    r.toi_upper = toi;
    r.witness_family = (uint8_t)kFamily;
    r.witness_id_a = id_a;
    r.witness_id_b = id_b;
  } else if (undecided) {
    init_result(&r, P2C_STATUS_UNDECIDED, kFamily, c.t0, c.t1);
    r.reason_code = 2;           // kCertificateReasonMaxSubdivisionDepth
    r.next_refinement_mode = 1;  // kBisectInterval
  } else {
    init_result(&r, P2C_STATUS_SEPARATION, kFamily, c.t0, c.t1);
    r.safe_margin_lb = min_margin;
  }
  r.nodes = nodes;
  r.pad[0] = (uint8_t)max_depth;  // deepest node visited (the Python twin's max_depth)
  *out = r;
}

P2C_HD void vertex_norms(const RefPrim& q, double vn[4]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) vn[k] = norm3(sub3(q.p1[k], q.p0[k]));
}

constexpr uint8_t kSkipMark = 0xFD;    // result status of a launch slot with no primitive

// Where slot i of a launch takes its primitive and its interval and config
// from.  UniformSrc: primitive i, one interval and config for the launch.
// (The exact work queue's source maps slot i to a queue entry's primitive,
// with that entry's interval and config.)  at() returns false for a slot
// with no primitive.
struct UniformSrc {
  const p2c_prim* prims;
  RefCfg c;
  __device__ bool at(int64_t i, const p2c_prim** p, RefCfg* cfg) const {
    *p = prims + i;
    *cfg = c;
    return true;
  }
};

// Light pass: one lane per primitive, at most kLightBudget nodes; a primitive
// that needs more is marked for the wave pass.
template <int kFamily, int kMode, typename Src>
__global__ __launch_bounds__(kRefBlock) void cert_ref_kernel(Src src, int64_t n, uint32_t budget,
                                                             p2c_prim_result* __restrict__ out,
                                                             uint32_t* __restrict__ invalid_count) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const p2c_prim* pp;
    RefCfg c;
    p2c_prim_result r;
    if (!src.at(i, &pp, &c)) {
      out[i].status = kSkipMark;
      continue;
    }
    const p2c_prim p = *pp;
    if (!prim_valid(p)) {
      init_result(&r, P2C_STATUS_INVALID, kFamily, c.t0, c.t1);
      r.status = P2C_STATUS_INVALID;
      r.reason_code = 3;             // kCertificateReasonInvalidInput
      r.next_refinement_mode = 2;    // kRequestGeometry
      if (invalid_count) atomicAdd(invalid_count, 1u);
      out[i] = r;
      continue;
    }
    const RefPrim q = load_prim(p);
    double vn[4];
    vertex_norms(q, vn);
    const double d0 = ref_distance<kFamily>(q, c.t0, kMode == kModePython);
    const double d1 = ref_distance<kFamily>(q, c.t1, kMode == kModePython);
    const RefWalk w = walk_subtree<kFamily, kMode>(q, vn, c, 0, 0, d0, d1, budget);
    if (w.state == kWalkOverBudget) {
      init_result(&r, P2C_STATUS_UNDECIDED, kFamily, c.t0, c.t1);
      r.status = kHeavyMark;
      out[i] = r;
      continue;
    }
    write_result<kFamily>(c, p.id_a, p.id_b, w.state == kWalkCollision, w.limited, w.a, w.b, w.toi, w.undecided,
                          w.min_margin, w.nodes, w.max_depth, &out[i]);
  }
}

// Pre-order position of node (path, depth): bit-reversed path (its first
// branch most significant), then depth, so an ancestor precedes its subtree
// and a left subtree precedes its right sibling.
__device__ inline bool node_after(uint64_t path, uint32_t depth, uint64_t kpath, uint32_t kdepth) {
  const uint64_t x = __builtin_bitreverse64(path), y = __builtin_bitreverse64(kpath);
  return x > y || (x == y && depth > kdepth);
}

struct HeavyStack {
  uint64_t path[kStackCap];
  double da[kStackCap];
  double db[kStackCap];
  uint8_t depth[kStackCap];
};

__device__ inline uint64_t lanes_below(int lane) { return lane == 0 ? 0ull : (~0ull >> (64 - lane)); }

// Wave pass: each wave inspects kHeavyChunk primitives and certifies the
// marked ones among them in turn, one at a time with the whole wave.  Pending nodes sit on an LDS stack
// in pre-order with the leftmost on top; each round the wave takes up to 64
// leftmost nodes, evaluates them (one new distance each: the ends travel with
// the node), and pushes the children of the expanded ones back in order.
//
// Pass 1 finds the first collision in pre-order: when a round finds one,
// every node after it (later lanes, the rest of the stack) is dropped, so the
// stack only ever holds nodes before the best collision.  Without a collision
// nothing is dropped, and the node count, min margin and undecided flag are
// those of the whole tree.  Pass 2, run only after a collision, counts the
// nodes the sequential walk visits: it walks again and never pushes a node
// after that collision.  A full stack falls back to one lane walking the
// leftmost pending subtree in sequence.  Result equal to the sequential walk's,
// node count included.
template <int kFamily, int kMode, typename Src>
__global__ __launch_bounds__(kHeavyBlock) void cert_ref_heavy_kernel(Src src, int64_t n,
                                                                     p2c_prim_result* __restrict__ out) {
  __shared__ HeavyStack st;
  const int lane = threadIdx.x;
  for (int64_t first = int64_t(blockIdx.x) * kHeavyChunk; first < n; first += int64_t(gridDim.x) * kHeavyChunk) {
    const int64_t mine = first + lane;
    uint64_t todo = __ballot(lane < kHeavyChunk && mine < n && out[mine].status == kHeavyMark);
    while (todo) {
      const int pick = __builtin_ctzll(todo);
      todo &= todo - 1;
      const int64_t i = first + pick;
      const p2c_prim* pp;
      RefCfg c;
      src.at(i, &pp, &c);  // a marked slot has a primitive
      const p2c_prim p = *pp;
      const RefPrim q = load_prim(p);
      double vn[4];
      vertex_norms(q, vn);
      const double d0 = ref_distance<kFamily>(q, c.t0, kMode == kModePython);
      const double d1 = ref_distance<kFamily>(q, c.t1, kMode == kModePython);

      bool have_best = false, best_limited = false;
      uint64_t best_path = 0;
      uint32_t best_depth = 0;
      double best_a = 0.0, best_b = 0.0, best_toi = 0.0;
      uint32_t count = 0;  // per lane
      uint32_t maxd = 0;   // per lane: deepest node counted
      double margin_min = 0.0;
      bool have_margin = false, undecided = false;

      for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
          if (!have_best) break;
          count = 0;
          maxd = 0;
        }
        __syncthreads();
        if (lane == 0) {
          st.path[0] = 0;
          st.depth[0] = 0;
          st.da[0] = d0;
          st.db[0] = d1;
        }
        int size = 1;
        __syncthreads();
        while (size > 0) {
          if (size + kHeavyBlock > kStackCap) {
            // no room for a full round: lane 0 walks the leftmost pending
            // subtree to its end (or to its first collision)
            RefWalk w = {};
            if (lane == 0)
              w = walk_subtree<kFamily, kMode>(q, vn, c, st.path[size - 1], st.depth[size - 1], st.da[size - 1],
                                        st.db[size - 1], 0xFFFFFFFFu);
            const bool hit = __shfl(w.state == kWalkCollision ? 1 : 0, 0) != 0;
            if (lane == 0) {
              count += w.nodes;
              if (kMode == kModePython) maxd = w.max_depth > maxd ? w.max_depth : maxd;
              if (pass == 0 && !hit) {
                if (w.have_margin) {
                  margin_min = have_margin ? std_min(margin_min, w.min_margin) : w.min_margin;
                  have_margin = true;
                }
                undecided = undecided || w.undecided;
              }
            }
            if (hit && pass == 0) {
              // the earliest collision so far: everything still pending follows it
              have_best = true;
              best_limited = __shfl(w.limited ? 1 : 0, 0) != 0;
              best_path = __shfl(w.path, 0);
              best_depth = __shfl(w.depth, 0);
              best_a = __shfl(w.a, 0);
              best_b = __shfl(w.b, 0);
              best_toi = __shfl(w.toi, 0);
              size = 0;
            } else {
              size -= 1;
            }
            __syncthreads();
            continue;
          }
          const int take = size < kHeavyBlock ? size : kHeavyBlock;
          const bool act = lane < take;
          uint64_t path = 0;
          uint32_t depth = 0;
          double da = 0.0, db = 0.0;
          if (act) {
            const int at = size - 1 - lane;
            path = st.path[at];
            depth = st.depth[at];
            da = st.da[at];
            db = st.db[at];
          }
          __syncthreads();  // every read of this round before any write
          int o = kNodeUndecided;
          double a = 0.0, b = 0.0, dm = 0.0, toi = 0.0, margin = 0.0;
          if (act) {
            node_interval(c, path, depth, &a, &b);
            o = eval_node<kFamily, kMode>(q, vn, c, a, b, depth, da, db, &dm, &toi, &margin);
          }
          int keep = take;
          int base = size - take;
          if (pass == 0) {
            // twin: an interval that may not be split ends the walk like a collision
            const uint64_t hits = __ballot(act && (o == kNodeCollision || (kMode == kModeCudaTwin && o == kNodeUndecided)));
            if (hits) {
              const int j = __builtin_ctzll(hits);
              keep = j;
              base = 0;
              have_best = true;
              best_limited = __shfl(o == kNodeUndecided ? 1 : 0, j) != 0;
              best_path = __shfl(path, j);
              best_depth = __shfl(depth, j);
              best_a = __shfl(a, j);
              best_b = __shfl(b, j);
              best_toi = __shfl(toi, j);
            }
            if (act && lane < keep) {
              if (o == kNodeSeparated) {
                margin_min = have_margin ? std_min(margin_min, margin) : margin;
                have_margin = true;
              } else if (o == kNodeUndecided) {
                undecided = true;
              }
            }
          }
          if (act && lane <= keep) {  // pass 1's count is used only without a collision
            count += 1;
            if (kMode == kModePython) maxd = depth > maxd ? depth : maxd;
          }
          const bool kept = act && lane < keep && o == kNodeExpand;
          const uint64_t right = path | (1ull << depth);
          bool push_l = kept, push_r = kept;
          if (pass == 1 && kept) {
            push_l = !node_after(path, depth + 1, best_path, best_depth);
            push_r = !node_after(right, depth + 1, best_path, best_depth);
          }
          const uint64_t ml = __ballot(push_l), mr = __ballot(push_r);
          const uint64_t below = lanes_below(lane);
          const int before = __popcll(ml & below) + __popcll(mr & below);
          const int total = __popcll(ml) + __popcll(mr);
          int slot = base + total - 1 - before;  // leftmost on top
          if (push_l) {
            st.path[slot] = path;
            st.depth[slot] = (uint8_t)(depth + 1);
            st.da[slot] = da;
            st.db[slot] = dm;
            --slot;
          }
          if (push_r) {
            st.path[slot] = right;
            st.depth[slot] = (uint8_t)(depth + 1);
            st.da[slot] = dm;
            st.db[slot] = db;
          }
          size = base + total;
          __syncthreads();
        }
      }

      // wave reductions
      for (int m = 32; m >= 1; m >>= 1) {
        count += __shfl_xor(count, m);
        const uint32_t od = __shfl_xor(maxd, m);
        maxd = od > maxd ? od : maxd;
      }
      if (have_best) {
        if (lane == 0)
          write_result<kFamily>(c, p.id_a, p.id_b, true, best_limited, best_a, best_b, best_toi, false, 0.0, count,
                                maxd, &out[i]);
      } else {
        double mm = have_margin ? margin_min : __builtin_inf();
        for (int m = 32; m >= 1; m >>= 1) {
          const double other = __shfl_xor(mm, m);
          mm = std_min(mm, other);
        }
        const bool und = __ballot(undecided) != 0;
        const bool any_margin = __ballot(have_margin) != 0;
        if (lane == 0)
          write_result<kFamily>(c, p.id_a, p.id_b, false, false, 0.0, 0.0, 0.0, und, any_margin ? mm : 0.0, count,
                                maxd, &out[i]);
      }
    }
  }
}

// The light pass and then the wave pass over n slots of src on stream s, for
// the given family (P2C_FAMILY_*) and walk mode.
template <int kFamily, int kMode, typename Src>
void launch_walk_fm(const Src& src, int64_t n, p2c_prim_result* out, uint32_t* invalid_count, hipStream_t s) {
  hipLaunchKernelGGL((cert_ref_kernel<kFamily, kMode, Src>), dim3(grid_for(n, kRefBlock)), dim3(kRefBlock), 0, s,
                     src, n, kLightBudget, out, invalid_count);
  hipLaunchKernelGGL((cert_ref_heavy_kernel<kFamily, kMode, Src>), dim3(grid_for(n, kHeavyChunk)),
                     dim3(kHeavyBlock), 0, s, src, n, out);
}

template <int kFamily, typename Src>
void launch_walk_f(int mode, const Src& src, int64_t n, p2c_prim_result* out, uint32_t* invalid_count,
                   hipStream_t s) {
  if (mode == kModeEngine) launch_walk_fm<kFamily, kModeEngine>(src, n, out, invalid_count, s);
  else if (mode == kModeCudaTwin) launch_walk_fm<kFamily, kModeCudaTwin>(src, n, out, invalid_count, s);
  else launch_walk_fm<kFamily, kModePython>(src, n, out, invalid_count, s);
}

template <typename Src>
void launch_walk(int family, int mode, const Src& src, int64_t n, p2c_prim_result* out, uint32_t* invalid_count,
                 hipStream_t s) {
  if (family == P2C_FAMILY_POINT_TRIANGLE) launch_walk_f<P2C_FAMILY_POINT_TRIANGLE>(mode, src, n, out, invalid_count, s);
  else launch_walk_f<P2C_FAMILY_EDGE_EDGE>(mode, src, n, out, invalid_count, s);
}

}  // namespace
}  // namespace p2c
