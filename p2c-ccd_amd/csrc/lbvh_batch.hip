// Batched driver of the swept-AABB LBVH broad phase (p2c_bp_batch_*): a
// batch of resident scenes built, counted and emitted with one launch per
// stage for the whole batch.
//
// Every stage runs the bodies of lbvh_core.h that the one-scene driver runs,
// on the same scene-local indices, so a scene's trees and pairs do not depend
// on the batch around it.  A launch covers every scene: scene k owns a run of
// whole workgroups of each stage (a block table per stage, found by a binary
// search on the block index, uniform per workgroup), sized as the one-scene
// launch would size it.  The keys of all scenes are sorted in one segmented
// radix sort: scene k's faces, edges and vertices fill a segment of whole
// 4096-key tiles (padded with keys that sort last), and each segment sorts
// on its own.  Per-query counts, resume nodes, offsets and the hit cache are
// batch arrays, scene after scene (vertex queries, then edge queries), so one
// exclusive scan places scene k's VF pairs and then its EE pairs in one pair
// list, scene after scene.
#include <algorithm>
#include <vector>

#include "lbvh_core.h"

namespace p2c {
namespace {

constexpr int kMaxBatchScenes = 1024;

// Device descriptor of one scene of a batch: its inputs and the regions of
// its resident workspace.
struct BScene {
  const double* x0;
  const double* x1;
  const int32_t* faces;
  const int32_t* edges;
  const int32_t* vobj;
  const int32_t* fobj;
  const int32_t* eobj;
  Box6* vbox;
  Box6* fbox;
  Box6* ebox;
  uint64_t* objkey;
  double* bounds;
  uint8_t* allowed;
  Node* fnodes;
  int32_t* fparent;
  uint32_t* fflag;
  int32_t* frcs;
  int32_t* fhi;
  int32_t* flo;
  Node* enodes;
  int32_t* eparent;
  uint32_t* eflag;
  int32_t* ercs;
  int32_t* ehi;
  int32_t* elo;
  uint32_t* qorder;
  uint32_t* eorder;
  uint32_t* err;
  uint32_t* froots;
  uint32_t* eroots;
  int64_t nv, nf, ne;
  int64_t key0;  // first key of the scene's sort segment (a multiple of kSortTile)
  int64_t q0;    // first query slot: vertex queries [q0, q0 + nv), edge queries [q0 + nv, q0 + nv + ne)
  double th;
  int32_t nobj, reserved;
};

// Stages and the parts each scene has in them (a part is a run of blocks).
enum Stage { kInit, kVbox, kObj, kPrim, kKarras, kLeaves, kRefit, kTop, kQuery, kCopy, kStages };
constexpr int kParts[kStages] = {1, 1, 2, 1, 2, 3, 2, 2, 2, 1};

inline int64_t blocks_of(int64_t items, int64_t per = kBlock) { return items > 0 ? (items + per - 1) / per : 0; }

// Part j of stage st for block L: the last part whose first block is <= L
// (empty parts share their start with the next one and are skipped).
__device__ inline int find_part(const int32_t* __restrict__ st, int n, int32_t L) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (st[mid] <= L) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

struct Tab {
  const int32_t* st;  // starts[S * parts + 1]
  int n;              // S * parts
};

__global__ __launch_bounds__(kBlock) void batch_init_kernel(const BScene* __restrict__ sc, Tab t) {
  const int j = find_part(t.st, t.n, blockIdx.x);
  const BScene& d = sc[j];
  const int64_t i = int64_t(blockIdx.x - t.st[j]) * kBlock + threadIdx.x;
  if (i < std::max<int64_t>(int64_t(d.nobj) * 6, 6)) objkey_init_item(d.objkey, d.nobj, d.err, d.froots, d.eroots, i);
}

__global__ __launch_bounds__(kBlock) void batch_vbox_kernel(const BScene* __restrict__ sc, Tab t) {
  __shared__ unsigned long long lkey[kLdsObjects * 6];
  const int j = find_part(t.st, t.n, blockIdx.x);
  const BScene& d = sc[j];
  vertex_box_block(d.x0, d.x1, d.nv, d.th, d.vobj, d.nobj, d.vbox, d.objkey, blockIdx.x - t.st[j],
                   t.st[j + 1] - t.st[j], lkey);
}

// part 0: the Morton frame (one block); part 1: the object-pair matrix
__global__ __launch_bounds__(kBlock) void batch_obj_kernel(const BScene* __restrict__ sc, Tab t) {
  __shared__ double red[6][kBlock];
  const int j = find_part(t.st, t.n, blockIdx.x);
  const BScene& d = sc[j >> 1];
  if ((j & 1) == 0) {
    scene_bounds_block(d.objkey, d.nobj, d.bounds, red);
  } else {
    const int64_t i = int64_t(blockIdx.x - t.st[j]) * kBlock + threadIdx.x;
    if (i < int64_t(d.nobj) * d.nobj) allowed_item(d.objkey, d.nobj, d.allowed, i);
  }
}

// Keys of the scene's segment: its primitives, then padding keys that sort
// after every primitive key (kind 3 is unused).
__global__ __launch_bounds__(kBlock) void batch_prim_kernel(const BScene* __restrict__ sc, Tab t,
                                                            uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int j = find_part(t.st, t.n, blockIdx.x);
  const BScene& d = sc[j];
  const int64_t g = int64_t(blockIdx.x - t.st[j]) * kBlock + threadIdx.x;
  if (g < d.nf + d.ne + d.nv) {
    const PrimKeys P{d.x0,  d.x1,   d.faces,        d.edges,        d.vbox, d.bounds, d.fbox,
                     d.ebox, keys + d.key0, vals + d.key0, d.nf, d.ne, d.nv, d.th};
    double lo[3], inv[3];
    for (int k = 0; k < 3; ++k) {
      lo[k] = d.bounds[k];
      inv[k] = d.bounds[3 + k];
    }
    prim_item(P, lo, inv, g);
  } else {
    keys[d.key0 + g] = 0xffffffffu;
    vals[d.key0 + g] = 0u;
  }
}

// part 0: face tree; part 1: edge tree
__global__ __launch_bounds__(kBlock) void batch_karras_kernel(const BScene* __restrict__ sc, Tab t,
                                                              const uint32_t* __restrict__ keys) {
  const int j = find_part(t.st, t.n, blockIdx.x);
  const BScene& d = sc[j >> 1];
  const int64_t i = int64_t(blockIdx.x - t.st[j]) * kBlock + threadIdx.x;
  if ((j & 1) == 0) {
    if (i < d.nf - 1) karras_node(keys + d.key0, d.nf, i, d.fnodes, d.fparent, d.frcs, d.fhi, d.flo);
  } else {
    if (i < d.ne - 1) karras_node(keys + d.key0 + d.nf, d.ne, i, d.enodes, d.eparent, d.ercs, d.ehi, d.elo);
  }
}

// part 0: face tree nodes; part 1: edge tree nodes and the edge query order
// (the edge leaf order); part 2: the vertex query order
__global__ __launch_bounds__(kBlock) void batch_leaves_kernel(const BScene* __restrict__ sc, Tab t,
                                                              const uint32_t* __restrict__ vals) {
  const int j = find_part(t.st, t.n, blockIdx.x);
  const BScene& d = sc[j / 3];
  const int part = j % 3;
  const int64_t p = int64_t(blockIdx.x - t.st[j]) * kBlock + threadIdx.x;
  if (part == 0) {
    if (p < 2 * d.nf - 1) leaf_item(vals + d.key0, d.nf, d.fbox, d.fnodes, d.frcs, d.fhi, d.fparent, d.fflag, p);
  } else if (part == 1) {
    const uint32_t* ev = vals + d.key0 + d.nf;
    if (p < 2 * d.ne - 1) leaf_item(ev, d.ne, d.ebox, d.enodes, d.ercs, d.ehi, d.eparent, d.eflag, p);
    if (p < d.ne) d.eorder[p] = ev[p];
  } else {
    if (p < d.nv) d.qorder[p] = vals[d.key0 + d.nf + d.ne + p];
  }
}

// One block per 256 leaves of each tree, exactly (part 0 faces, 1 edges).
__global__ __launch_bounds__(kBlock) void batch_refit_kernel(const BScene* __restrict__ sc, Tab t) {
  __shared__ uint32_t local_flag[kBlock];
  const int j = find_part(t.st, t.n, blockIdx.x);
  const BScene& d = sc[j >> 1];
  const int64_t blk = blockIdx.x - t.st[j];
  if ((j & 1) == 0) refit_block(d.nf, d.fnodes, d.fparent, d.flo, d.fhi, d.froots, blk, local_flag);
  else refit_block(d.ne, d.enodes, d.eparent, d.elo, d.ehi, d.eroots, blk, local_flag);
}

// One workgroup per tree with more than one leaf: the climb through the
// nodes that span refit blocks, flags in LDS when they fit.
__global__ __launch_bounds__(kTopBlock) void batch_top_kernel(const BScene* __restrict__ sc, Tab t) {
  __shared__ uint32_t bits[kTopLdsWords];
  const int j = find_part(t.st, t.n, blockIdx.x);
  const BScene& d = sc[j >> 1];
  const bool face = (j & 1) == 0;
  const int64_t n = face ? d.nf : d.ne;
  Node* nodes = face ? d.fnodes : d.enodes;
  const int32_t* parent = face ? d.fparent : d.eparent;
  const uint32_t* roots = face ? d.froots : d.eroots;
  if (n - 1 <= int64_t(kTopLdsWords) * 32) top_climb_lds(nodes, parent, roots, n - 1, bits);
  else top_climb_global(nodes, parent, roots, face ? d.fflag : d.eflag);
}

struct QueryArgs {
  uint32_t* counts;
  int32_t* cache;
  int64_t slots;
  int32_t* resume;
  const uint64_t* offsets;
  int32_t* pairs;
  int64_t cap;
};

// Count (kFill false) or fill pass of every scene's vertex queries (part 0)
// and edge queries (part 1).  The block table runs scene after scene, so the
// XCD mapping gives each XCD's L2 one run of neighbouring queries.
template <bool kFill>
__global__ __launch_bounds__(kBlock) void batch_query_kernel(const BScene* __restrict__ sc, Tab t, QueryArgs a) {
  const int32_t L = int32_t(xcd_block(blockIdx.x, gridDim.x));
  const int j = find_part(t.st, t.n, L);
  const BScene& d = sc[j >> 1];
  const int64_t s = int64_t(L - t.st[j]) * kBlock + threadIdx.x;
  const int64_t q0 = d.q0;
  if ((j & 1) == 0) {
    if (s >= d.nv) return;
    const VfQuery V{d.fnodes,
                    d.nf,
                    d.vbox,
                    d.qorder,
                    d.nv,
                    VfRule{d.faces, d.fobj, d.vobj, d.allowed, d.nobj},
                    d.err,
                    a.counts + q0,
                    a.cache + q0 * a.slots,
                    a.slots,
                    a.resume + q0,
                    a.offsets + q0,
                    a.pairs,
                    a.cap};
    vf_query_one<kFill>(V, s);
  } else {
    if (s >= d.ne) return;
    // the fill writes the scene's EE pairs relative to their first slot
    const int64_t ee0 = kFill ? int64_t(a.offsets[q0 + d.nv]) : 0;
    const EeQuery E{d.enodes,
                    d.ne,
                    d.ebox,
                    d.eorder,
                    EeRule{d.edges, d.eobj, d.allowed, d.nobj},
                    d.err,
                    a.counts + q0 + d.nv,
                    a.cache + (q0 + d.nv) * a.slots,
                    a.slots,
                    a.resume + q0 + d.nv,
                    a.offsets + q0,
                    d.nv,
                    kFill ? a.pairs + 2 * ee0 : nullptr,
                    kFill ? (a.cap > ee0 ? a.cap - ee0 : 0) : 0};
    ee_query_one<kFill>(E, s);
  }
}

// The cached pairs of every scene: one lane per cache slot of the scene's
// queries (vertex queries, then edge queries), as cache_copy_kernel does for
// one scene, written at the query's place in the batch's pair list.
__global__ __launch_bounds__(kBlock) void batch_copy_kernel(const BScene* __restrict__ sc, Tab t, QueryArgs a) {
  const int j = find_part(t.st, t.n, blockIdx.x);
  const BScene& d = sc[j];
  const int64_t tt = int64_t(blockIdx.x - t.st[j]) * kBlock + threadIdx.x;
  const int64_t slots = a.slots;
  if (tt >= (d.nv + d.ne) * slots) return;
  const int64_t s = tt / slots;
  const uint32_t k = uint32_t(tt % slots);
  const int64_t q = d.q0 + s;
  const uint64_t c = a.offsets[q + 1] - a.offsets[q];
  if (k >= c) return;
  const uint64_t o = a.offsets[q] + k;
  if (o >= uint64_t(a.cap)) return;
  const int32_t p = a.cache[q * slots + k];
  if (s < d.nv) {
    a.pairs[o * 2] = int32_t(d.qorder[s]);
    a.pairs[o * 2 + 1] = p;
  } else {
    const int32_t i = int32_t(d.eorder[s - d.nv]);
    a.pairs[o * 2] = i < p ? i : p;
    a.pairs[o * 2 + 1] = i < p ? p : i;
  }
}

// Per scene the first VF and first EE slot; then the total and the summed
// step-bound count.
__global__ __launch_bounds__(kBlock) void batch_totals_kernel(const BScene* __restrict__ sc, int S,
                                                              const uint64_t* __restrict__ offsets, int64_t Q,
                                                              uint64_t* __restrict__ out) {
  for (int k = threadIdx.x; k < S; k += blockDim.x) {
    const BScene& d = sc[k];
    out[2 * k] = Q > 0 ? offsets[d.q0] : 0;
    out[2 * k + 1] = Q > 0 ? offsets[d.q0 + d.nv] : 0;
  }
  if (threadIdx.x == 0) {
    uint64_t bad = 0;
    for (int k = 0; k < S; ++k) bad += sc[k].err[0];
    out[2 * S] = Q > 0 ? offsets[Q] : 0;
    out[2 * S + 1] = bad;
  }
}

// ---- host plan ------------------------------------------------------------
// The device image of a batch: BScene[S], the sort's segment table (S + 1
// first tiles), then each stage's block table (S * parts + 1 starts).
struct Plan {
  std::vector<char> blob;
  int32_t S = 0;
  size_t seg_off = 0, tab_off[kStages] = {};
  int32_t blocks[kStages] = {};
  int64_t tiles = 0, queries = 0;
};

struct BatchLayout {
  size_t plan, keys0, vals0, keys1, vals1, hist, hist_off, part, counts, resume, offsets, cache, total;
};

inline size_t plan_bytes(int32_t S) {
  size_t b = align256(sizeof(BScene) * size_t(S)) + align256(4 * size_t(S + 1));
  for (int st = 0; st < kStages; ++st) b += align256(4 * (size_t(S) * kParts[st] + 1));
  return b;
}

BatchLayout batch_layout(int32_t S, int64_t tiles, int64_t Q, int64_t slots) {
  BatchLayout B{};
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o = align256(o + (bytes ? bytes : 1));
    return at;
  };
  const int64_t keys = tiles * kSortTile;
  B.plan = take(plan_bytes(S));
  B.keys0 = take(4 * size_t(keys));
  B.vals0 = take(4 * size_t(keys));
  B.keys1 = take(4 * size_t(keys));
  B.vals1 = take(4 * size_t(keys));
  B.hist = take(4 * 256 * size_t(tiles));
  B.hist_off = take(8 * 256 * size_t(tiles));
  const int64_t scan_items = std::max<int64_t>(Q + 1, 256 * tiles);
  B.part = take(8 * size_t(scan_items / 1024 + 2));
  B.counts = take(4 * size_t(Q + 1));
  B.resume = take(4 * size_t(Q + 1));
  B.offsets = take(8 * size_t(Q + 1));
  B.cache = take(4 * size_t(Q) * size_t(slots));
  B.total = o;
  return B;
}

int check_batch(const p2c_bp_scene* sc, int32_t S, int64_t slots) {
  if (sc == nullptr) return set_error(P2C_EINVAL, "scene array is required");
  if (S < 1 || S > kMaxBatchScenes) return set_error(P2C_EINVAL, "a batch holds 1 to 1024 scenes");
  if (slots < 0 || slots > 4096) return set_error(P2C_EINVAL, "hit_cache_slots must be in [0, 4096]");
  for (int32_t k = 0; k < S; ++k) {
    p2c_bp_mesh m = sc[k].mesh;
    m.hit_cache = nullptr;
    m.hit_cache_slots = 0;
    if (int rc = check_mesh(&m)) return rc;
    if (!std::isfinite(sc[k].thickness) || sc[k].thickness < 0.0)
      return set_error(P2C_EINVAL, "thickness must be finite and non-negative");
    const size_t need = make_layout(m.n_vertices, m.n_faces, m.n_edges, m.n_objects, true).total;
    if (sc[k].workspace == nullptr || sc[k].workspace_bytes < need)
      return set_error(P2C_ECAPACITY, "scene " + std::to_string(k) + ": scene workspace is too small");
    if (m.n_vertices > 0 && (sc[k].x0 == nullptr || sc[k].x1 == nullptr))
      return set_error(P2C_EINVAL, "scene positions are required");
  }
  // the scenes of a batch build their trees side by side: no workspace twice
  std::vector<const void*> ws(S);
  for (int32_t k = 0; k < S; ++k) ws[k] = sc[k].workspace;
  std::sort(ws.begin(), ws.end());
  if (std::adjacent_find(ws.begin(), ws.end()) != ws.end())
    return set_error(P2C_EINVAL, "a scene workspace appears twice in the batch");
  return P2C_OK;
}

// Sizes only (tiles, queries): enough for the batch workspace size.
void batch_sizes(const p2c_bp_scene* sc, int32_t S, int64_t* tiles, int64_t* queries) {
  int64_t t = 0, q = 0;
  for (int32_t k = 0; k < S; ++k) {
    const p2c_bp_mesh& m = sc[k].mesh;
    t += blocks_of(m.n_faces + m.n_edges + m.n_vertices, kSortTile);
    q += m.n_vertices + m.n_edges;
  }
  *tiles = t;
  *queries = q;
}

int make_plan(const p2c_bp_scene* sc, int32_t S, int64_t slots, Plan* P) {
  P->S = S;
  P->blob.assign(plan_bytes(S), 0);
  char* base = P->blob.data();
  auto* bs = reinterpret_cast<BScene*>(base);
  size_t o = align256(sizeof(BScene) * size_t(S));
  P->seg_off = o;
  auto* seg = reinterpret_cast<int32_t*>(base + o);
  o += align256(4 * size_t(S + 1));
  int32_t* tab[kStages];
  for (int st = 0; st < kStages; ++st) {
    P->tab_off[st] = o;
    tab[st] = reinterpret_cast<int32_t*>(base + o);
    o += align256(4 * (size_t(S) * kParts[st] + 1));
  }
  int64_t tile = 0, q = 0;
  int64_t run[kStages] = {};
  for (int32_t k = 0; k < S; ++k) {
    const p2c_bp_mesh& m = sc[k].mesh;
    const int64_t nv = m.n_vertices, nf = m.n_faces, ne = m.n_edges;
    const int32_t nobj = m.n_objects;
    const Layout L = make_layout(nv, nf, ne, nobj, true);
    void* ws = sc[k].workspace;
    BScene& d = bs[k];
    d.x0 = sc[k].x0;
    d.x1 = sc[k].x1;
    d.faces = m.faces;
    d.edges = m.edges;
    d.vobj = m.vertex_object;
    d.fobj = m.face_object;
    d.eobj = m.edge_object;
    d.vbox = at<Box6>(ws, L.vbox);
    d.fbox = at<Box6>(ws, L.fbox);
    d.ebox = at<Box6>(ws, L.ebox);
    d.objkey = at<uint64_t>(ws, L.objkey);
    d.bounds = at<double>(ws, L.bounds);
    d.allowed = at<uint8_t>(ws, L.allowed);
    d.fnodes = at<Node>(ws, L.fnodes);
    d.fparent = at<int32_t>(ws, L.fparent);
    d.fflag = at<uint32_t>(ws, L.fflag);
    d.frcs = at<int32_t>(ws, L.frcs);
    d.fhi = at<int32_t>(ws, L.fhi);
    d.flo = at<int32_t>(ws, L.flo);
    d.enodes = at<Node>(ws, L.enodes);
    d.eparent = at<int32_t>(ws, L.eparent);
    d.eflag = at<uint32_t>(ws, L.eflag);
    d.ercs = at<int32_t>(ws, L.ercs);
    d.ehi = at<int32_t>(ws, L.ehi);
    d.elo = at<int32_t>(ws, L.elo);
    d.qorder = at<uint32_t>(ws, L.qorder);
    d.eorder = at<uint32_t>(ws, L.eorder);
    d.err = at<uint32_t>(ws, L.err);
    d.froots = at<uint32_t>(ws, L.roots);
    d.eroots = at<uint32_t>(ws, L.eroots);
    d.nv = nv;
    d.nf = nf;
    d.ne = ne;
    d.key0 = tile * kSortTile;
    d.q0 = q;
    d.th = sc[k].thickness;
    d.nobj = nobj;
    const int64_t nt = blocks_of(nf + ne + nv, kSortTile);
    seg[k] = int32_t(tile);
    // blocks of each part of each stage, as the one-scene launches size them
    const int64_t parts[kStages][3] = {
        {blocks_of(std::max<int64_t>(int64_t(nobj) * 6, 6)), 0, 0},
        {blocks_of(nv), 0, 0},
        {1, blocks_of(int64_t(nobj) * nobj), 0},
        {nt * (kSortTile / kBlock), 0, 0},
        {nf > 1 ? blocks_of(nf - 1) : 0, ne > 1 ? blocks_of(ne - 1) : 0, 0},
        {nf > 0 ? blocks_of(2 * nf - 1) : 0, ne > 0 ? blocks_of(2 * ne - 1) : 0, blocks_of(nv)},
        {nf > 1 ? blocks_of(nf) : 0, ne > 1 ? blocks_of(ne) : 0, 0},
        {nf > 1 ? 1 : 0, ne > 1 ? 1 : 0, 0},
        {blocks_of(nv), blocks_of(ne), 0},
        {blocks_of((nv + ne) * slots), 0, 0},
    };
    for (int st = 0; st < kStages; ++st)
      for (int p = 0; p < kParts[st]; ++p) {
        tab[st][k * kParts[st] + p] = int32_t(run[st]);
        run[st] += parts[st][p];
        if (run[st] > int64_t(INT32_MAX) / kBlock)
          return set_error(P2C_ECAPACITY, "batch too large for one launch per stage");
      }
    tile += nt;
    q += nv + ne;
  }
  seg[S] = int32_t(tile);
  for (int st = 0; st < kStages; ++st) {
    tab[st][int64_t(S) * kParts[st]] = int32_t(run[st]);
    P->blocks[st] = int32_t(run[st]);
  }
  P->tiles = tile;
  P->queries = q;
  return P2C_OK;
}

inline Tab tab_of(const Plan& P, void* dplan, int st) {
  return Tab{at<const int32_t>(dplan, P.tab_off[st]), P.S * kParts[st]};
}

int prepare(const p2c_bp_scene* sc, int32_t S, int64_t slots, void* bws, size_t bws_bytes, Plan* P,
            BatchLayout* B, hipStream_t s) {
  if (int rc = check_batch(sc, S, slots)) return rc;
  if (int rc = make_plan(sc, S, slots, P)) return rc;
  *B = batch_layout(S, P->tiles, P->queries, slots);
  if (bws == nullptr || bws_bytes < B->total) return set_error(P2C_ECAPACITY, "batch workspace is too small");
  return upload_table(P->blob.data(), P->blob.size(), at<char>(bws, B->plan), s);
}

}  // namespace
}  // namespace p2c

extern "C" {

size_t p2c_bp_scene_workspace_bytes(int64_t nv, int64_t nf, int64_t ne, int32_t nobj) {
  if (nv < 0 || nf < 0 || ne < 0 || nobj < 1) return 0;
  return p2c::make_layout(nv, nf, ne, nobj, true).total;
}

size_t p2c_bp_batch_workspace_bytes(const p2c_bp_scene* scenes, int32_t n_scenes, int64_t slots) {
  using namespace p2c;
  if (scenes == nullptr || n_scenes < 1 || n_scenes > kMaxBatchScenes || slots < 0) return 0;
  int64_t tiles = 0, queries = 0;
  batch_sizes(scenes, n_scenes, &tiles, &queries);
  return batch_layout(n_scenes, tiles, queries, slots).total;
}

int p2c_bp_batch_count(const p2c_bp_scene* scenes, int32_t S, int64_t slots, void* bws, size_t bws_bytes,
                       uint64_t* offsets_out, void* stream) {
  using namespace p2c;
  if (offsets_out == nullptr) return set_error(P2C_EINVAL, "offsets output is required");
  hipStream_t s = as_stream(stream);
  Plan P;
  BatchLayout B;
  if (int rc = prepare(scenes, S, slots, bws, bws_bytes, &P, &B, s)) return rc;
  void* dplan = at<char>(bws, B.plan);
  const BScene* sc = at<const BScene>(dplan, 0);
  auto launch = [&](int st, auto kernel, int threads, auto... args) {
    if (P.blocks[st] > 0)
      hipLaunchKernelGGL(kernel, dim3(unsigned(P.blocks[st])), dim3(threads), 0, s, sc, tab_of(P, dplan, st), args...);
  };
  uint32_t* k0 = at<uint32_t>(bws, B.keys0);
  uint32_t* v0 = at<uint32_t>(bws, B.vals0);
  launch(kInit, batch_init_kernel, kBlock);
  launch(kVbox, batch_vbox_kernel, kBlock);
  launch(kObj, batch_obj_kernel, kBlock);
  launch(kPrim, batch_prim_kernel, kBlock, k0, v0);
  if (P.tiles > 0)
    if (int rc = radix_sort(k0, v0, at<uint32_t>(bws, B.keys1), at<uint32_t>(bws, B.vals1), P.tiles * kSortTile,
                            at<uint32_t>(bws, B.hist), at<uint64_t>(bws, B.hist_off), at<uint64_t>(bws, B.part), s,
                            at<const int32_t>(dplan, P.seg_off), S))
      return rc;
  launch(kKarras, batch_karras_kernel, kBlock, static_cast<const uint32_t*>(k0));
  launch(kLeaves, batch_leaves_kernel, kBlock, static_cast<const uint32_t*>(v0));
  launch(kRefit, batch_refit_kernel, kBlock);
  launch(kTop, batch_top_kernel, kTopBlock);
  uint32_t* counts = at<uint32_t>(bws, B.counts);
  uint64_t* offsets = at<uint64_t>(bws, B.offsets);
  const QueryArgs qa{counts,  at<int32_t>(bws, B.cache), slots, at<int32_t>(bws, B.resume), offsets, nullptr, 0};
  launch(kQuery, batch_query_kernel<false>, kBlock, qa);
  if (int rc = exclusive_scan_u32(counts, P.queries, offsets, at<uint64_t>(bws, B.part), s)) return rc;
  hipLaunchKernelGGL(batch_totals_kernel, dim3(1), dim3(kBlock), 0, s, sc, S, static_cast<const uint64_t*>(offsets),
                     P.queries, offsets_out);
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

int p2c_bp_batch_emit(const p2c_bp_scene* scenes, int32_t S, int64_t slots, void* bws, size_t bws_bytes,
                      int32_t* pairs, int64_t cap, void* stream) {
  using namespace p2c;
  if (pairs == nullptr && cap > 0) return set_error(P2C_EINVAL, "pair output is required");
  hipStream_t s = as_stream(stream);
  Plan P;
  BatchLayout B;
  if (int rc = prepare(scenes, S, slots, bws, bws_bytes, &P, &B, s)) return rc;
  if (pairs == nullptr || cap <= 0) return P2C_OK;
  void* dplan = at<char>(bws, B.plan);
  const BScene* sc = at<const BScene>(dplan, 0);
  const QueryArgs qa{nullptr, at<int32_t>(bws, B.cache), slots, at<int32_t>(bws, B.resume),
                     at<uint64_t>(bws, B.offsets), pairs, cap};
  if (P.blocks[kQuery] > 0)
    hipLaunchKernelGGL(batch_query_kernel<true>, dim3(unsigned(P.blocks[kQuery])), dim3(kBlock), 0, s, sc,
                       tab_of(P, dplan, kQuery), qa);
  if (slots > 0 && P.blocks[kCopy] > 0)
    hipLaunchKernelGGL(batch_copy_kernel, dim3(unsigned(P.blocks[kCopy])), dim3(kBlock), 0, s, sc,
                       tab_of(P, dplan, kCopy), qa);
  P2C_HIP_CHECK(hipGetLastError());
  return P2C_OK;
}

}  // extern "C"
