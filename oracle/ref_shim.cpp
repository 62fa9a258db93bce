// Oracle (test infrastructure only): a C-ABI shim over the reference's own
// CPU certificate code, compiled from the reference sources where they lie
// (src/cpp/certificate/certificate_engine.cpp, src/cpp/common/validators.cpp)
// into oracle/_ref/libp2c_ref.so by oracle/Makefile.  Used in this container
// to write golden vectors (tests/golden/make_ref_golden.py) and to validate
// the C restatement in oracle/ref_cert.c.  Never part of the shipped library.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "certificate/certificate_engine.h"
#include "common/validators.h"
#include "oracle.h"

namespace {

p2cccd::LinearVertexTrajectory ToTraj(const orc_traj& t) {
  p2cccd::LinearVertexTrajectory out;
  out.feature_id = t.feature_id;
  for (int k = 0; k < 3; ++k) {
    out.position_t0[k] = t.p0[k];
    out.position_t1[k] = t.p1[k];
  }
  return out;
}

p2cccd::PointTriangleIntervalPrimitive ToPt(const orc_prim& p) {
  p2cccd::PointTriangleIntervalPrimitive out;
  out.point_id = p.id_a;
  out.triangle_id = p.id_b;
  out.point = ToTraj(p.v[0]);
  out.triangle_v0 = ToTraj(p.v[1]);
  out.triangle_v1 = ToTraj(p.v[2]);
  out.triangle_v2 = ToTraj(p.v[3]);
  return out;
}

p2cccd::EdgeEdgeIntervalPrimitive ToEe(const orc_prim& p) {
  p2cccd::EdgeEdgeIntervalPrimitive out;
  out.edge_a_id = p.id_a;
  out.edge_b_id = p.id_b;
  out.edge_a0 = ToTraj(p.v[0]);
  out.edge_a1 = ToTraj(p.v[1]);
  out.edge_b0 = ToTraj(p.v[2]);
  out.edge_b1 = ToTraj(p.v[3]);
  return out;
}

void FromPrim(const p2cccd::PrimitiveIntervalResult& r, orc_result* o) {
  std::memset(o, 0, sizeof(*o));
  o->status = static_cast<int32_t>(r.status);
  o->covered_feature_mask = r.covered_feature_mask;
  o->interval_t0 = r.interval_t0;
  o->interval_t1 = r.interval_t1;
  o->toi_upper = r.toi_upper;
  o->safe_margin_lb = r.safe_margin_lb;
  o->witness_family = r.witness_family;
  o->reason_code = r.reason_code;
  o->witness_id_a = r.witness_id_a;
  o->witness_id_b = r.witness_id_b;
  o->next_refinement_mode = static_cast<int32_t>(r.next_refinement_mode);
}

}  // namespace

extern "C" {

// Per-primitive interval certificate.  Returns 0, or 1 + index of the first
// primitive the reference rejected (its Status message is not carried).
int64_t ref_eval_batch(int family, const orc_prim* prims, int64_t n, double t0, double t1,
                       double eps_time, double eps_space, int max_depth, orc_result* out) {
  p2cccd::CertificateEngineConfig cfg;
  cfg.eps_time = eps_time;
  cfg.eps_space = eps_space;
  cfg.max_subdivision_depth = static_cast<std::uint16_t>(max_depth);
  for (int64_t i = 0; i < n; ++i) {
    p2cccd::PrimitiveIntervalResult r;
    p2cccd::Status st = family == 1 ? p2cccd::EvaluatePointTriangleInterval(ToPt(prims[i]), t0, t1, cfg, &r)
                                    : p2cccd::EvaluateEdgeEdgeInterval(ToEe(prims[i]), t0, t1, cfg, &r);
    if (!st.ok) return 1 + i;
    FromPrim(r, &out[i]);
  }
  return 0;
}

// CertificateEngine::Evaluate over one work item holding n_pt point-triangle
// and n_ee edge-edge primitives.  out_fields: status, toi_upper, margin,
// witness_family, witness_a, witness_b, covered_mask, reason, refinement.
int ref_engine_eval(uint32_t family_mask, double t0, double t1, const orc_prim* pts, int64_t n_pt,
                    const orc_prim* ees, int64_t n_ee, double eps_time, double eps_space,
                    int max_depth, double* out_fields) {
  p2cccd::ExactCertificateQuery q;
  q.work_item.work_item_id = 1;
  q.work_item.parent_candidate_id = 1;
  q.work_item.query_id = 1;
  q.work_item.interval_t0 = t0;
  q.work_item.interval_t1 = t1;
  q.work_item.feature_family_mask = family_mask;
  q.config.eps_time = eps_time;
  q.config.eps_space = eps_space;
  q.config.max_subdivision_depth = static_cast<std::uint16_t>(max_depth);
  for (int64_t i = 0; i < n_pt; ++i) q.point_triangle_primitives.push_back(ToPt(pts[i]));
  for (int64_t i = 0; i < n_ee; ++i) q.edge_edge_primitives.push_back(ToEe(ees[i]));
  p2cccd::CertificateEngine engine;
  p2cccd::CertificateResult r;
  p2cccd::Status st = engine.Evaluate(q, &r);
  if (!st.ok) return -1;
  out_fields[0] = static_cast<double>(r.status);
  out_fields[1] = r.toi_upper;
  out_fields[2] = r.safe_margin_lb;
  out_fields[3] = r.witness_family;
  out_fields[4] = static_cast<double>(r.witness_id_a);
  out_fields[5] = static_cast<double>(r.witness_id_b);
  out_fields[6] = r.covered_feature_mask;
  out_fields[7] = r.reason_code;
  out_fields[8] = static_cast<double>(r.next_refinement_mode);
  return 0;
}

}  // extern "C"
