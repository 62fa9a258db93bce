// One-scene driver of the swept-AABB LBVH broad phase (p2c_bp_build,
// p2c_bp_count, p2c_bp_emit): the stage bodies of lbvh_core.h launched over
// one mesh.
#include "lbvh_core.h"

namespace p2c {
namespace {

__global__ __launch_bounds__(kBlock) void objkey_init_kernel(uint64_t* objkey, int32_t nobj, uint32_t* err,
                                                             uint32_t* froots, uint32_t* eroots) {
  const int64_t n = std::max<int64_t>(int64_t(nobj) * 6, 6);
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    objkey_init_item(objkey, nobj, err, froots, eroots, i);
}

__global__ __launch_bounds__(kBlock) void vertex_box_kernel(const double* __restrict__ x0,
                                                            const double* __restrict__ x1, int64_t nv, double th,
                                                            const int32_t* __restrict__ vobj, int32_t nobj,
                                                            Box6* __restrict__ vbox, uint64_t* __restrict__ objkey) {
  __shared__ unsigned long long lkey[kLdsObjects * 6];
  vertex_box_block(x0, x1, nv, th, vobj, nobj, vbox, objkey, blockIdx.x, gridDim.x, lkey);
}

__global__ __launch_bounds__(kBlock) void allowed_kernel(const uint64_t* __restrict__ objkey, int32_t nobj,
                                                         uint8_t* __restrict__ allowed) {
  const int64_t n2 = int64_t(nobj) * nobj;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n2; i += int64_t(gridDim.x) * blockDim.x)
    allowed_item(objkey, nobj, allowed, i);
}

__global__ __launch_bounds__(kBlock) void scene_bounds_kernel(const uint64_t* __restrict__ objkey, int32_t nobj,
                                                              double* __restrict__ bounds) {
  __shared__ double red[6][kBlock];
  scene_bounds_block(objkey, nobj, bounds, red);
}

__global__ __launch_bounds__(kBlock) void prim_box_kernel(PrimKeys P) {
  double lo[3], inv[3];
  for (int k = 0; k < 3; ++k) {
    lo[k] = P.bounds[k];
    inv[k] = P.bounds[3 + k];
  }
  const int64_t total = P.nf + P.ne + P.nv;
  for (int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; g < total; g += int64_t(gridDim.x) * blockDim.x)
    prim_item(P, lo, inv, g);
}

__global__ __launch_bounds__(kBlock) void karras_kernel(const uint32_t* __restrict__ keys, int64_t n,
                                                        Node* __restrict__ nodes, int32_t* __restrict__ parent,
                                                        int32_t* __restrict__ rcs, int32_t* __restrict__ hi_of,
                                                        int32_t* __restrict__ lo_of) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n - 1; i += int64_t(gridDim.x) * blockDim.x)
    karras_node(keys, n, i, nodes, parent, rcs, hi_of, lo_of);
}

__global__ __launch_bounds__(kBlock) void leaves_kernel(const uint32_t* __restrict__ vals, int64_t n,
                                                        const Box6* __restrict__ box, Node* __restrict__ nodes,
                                                        const int32_t* __restrict__ rcs,
                                                        const int32_t* __restrict__ hi_of,
                                                        int32_t* __restrict__ parent, uint32_t* __restrict__ flag) {
  for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p < 2 * n - 1;
       p += int64_t(gridDim.x) * blockDim.x)
    leaf_item(vals, n, box, nodes, rcs, hi_of, parent, flag, p);
}

// One block per 256 leaves, exactly.
__global__ __launch_bounds__(kBlock) void refit_kernel(int64_t n, Node* __restrict__ nodes,
                                                       const int32_t* __restrict__ parent,
                                                       const int32_t* __restrict__ lo_of,
                                                       const int32_t* __restrict__ hi_of,
                                                       uint32_t* __restrict__ roots) {
  __shared__ uint32_t local_flag[kBlock];
  refit_block(n, nodes, parent, lo_of, hi_of, roots, blockIdx.x, local_flag);
}

__global__ __launch_bounds__(kTopBlock) void refit_top_kernel(Node* __restrict__ nodes,
                                                              const int32_t* __restrict__ parent,
                                                              const uint32_t* __restrict__ roots,
                                                              uint32_t* __restrict__ flag) {
  top_climb_global(nodes, parent, roots, flag);
}

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
  top_climb_lds(t.nodes, t.parent, t.roots, t.n_internal, bits);
}

template <bool kFill>
__global__ __launch_bounds__(kBlock) void vf_query_kernel(VfQuery Q) {
  const int64_t b0 = xcd_block(blockIdx.x, gridDim.x);  // Morton-neighbouring queries on one XCD
  for (int64_t s = b0 * blockDim.x + threadIdx.x; s < Q.nv; s += int64_t(gridDim.x) * blockDim.x)
    vf_query_one<kFill>(Q, s);
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
