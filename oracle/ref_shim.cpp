// Oracle (test infrastructure only): a C-ABI shim over the reference's own
// CPU certificate code, compiled from the reference sources where they lie
// (src/cpp/certificate/certificate_engine.cpp, src/cpp/common/validators.cpp)
// into oracle/_ref/libp2c_ref.so by oracle/Makefile.  Used to write golden
// vectors (tests/golden/make_ref_golden.py, make_queue_golden.py) and to
// validate the C restatements in oracle/ref_cert.c and oracle/ref_queue.c.  Never part of the shipped library.
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


void FromCert(const p2cccd::CertificateResult& r, orc_certificate* o) {
  std::memset(o, 0, sizeof(*o));
  o->work_item_id = r.work_item_id;
  o->query_id = r.query_id;
  o->interval_t0 = r.interval_t0;
  o->interval_t1 = r.interval_t1;
  o->toi_upper = r.toi_upper;
  o->safe_margin_lb = r.safe_margin_lb;
  o->witness_id_a = r.witness_id_a;
  o->witness_id_b = r.witness_id_b;
  o->eps_time = r.eps_time;
  o->eps_space = r.eps_space;
  o->covered_feature_mask = r.covered_feature_mask;
  o->reason_code = r.reason_code;
  o->status = static_cast<uint8_t>(r.status);
  o->witness_family = r.witness_family;
  o->next_refinement_mode = static_cast<uint8_t>(r.next_refinement_mode);
}

p2cccd::CertificateResult ToCert(const orc_certificate& o) {
  p2cccd::CertificateResult r;
  r.work_item_id = o.work_item_id;
  r.query_id = o.query_id;
  r.status = static_cast<p2cccd::CertificateStatus>(o.status);
  r.interval_t0 = o.interval_t0;
  r.interval_t1 = o.interval_t1;
  r.toi_upper = o.toi_upper;
  r.safe_margin_lb = o.safe_margin_lb;
  r.witness_family = o.witness_family;
  r.witness_id_a = o.witness_id_a;
  r.witness_id_b = o.witness_id_b;
  r.covered_feature_mask = o.covered_feature_mask;
  r.eps_time = o.eps_time;
  r.eps_space = o.eps_space;
  r.reason_code = o.reason_code;
  r.next_refinement_mode = static_cast<p2cccd::CertificateRefinementMode>(o.next_refinement_mode);
  return r;
}

void FromRow(const p2cccd::AuditLogRow& r, orc_audit_row* o) {
  std::memset(o, 0, sizeof(*o));
  o->event_id = r.event_id;
  o->query_id = r.query_id;
  o->candidate_id = r.candidate_id;
  o->work_item_id = r.work_item_id;
  o->interval_t0 = r.interval_t0;
  o->interval_t1 = r.interval_t1;
  o->timestamp_us = r.timestamp_us;
  o->aux_value0 = r.aux_value0;
  o->aux_value1 = r.aux_value1;
  o->action = r.action;
  o->depth = r.depth;
  o->stage = static_cast<uint8_t>(r.stage);
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

// ProcessExactWorkQueueCpu over n entries (each an ExactCertificateQuery
// built from the entry and its primitive ranges).  Returns 0, or -1 with the
// reference's message copied into msg (msg_cap bytes).
int ref_process_queue(const orc_queue_entry* entries, int64_t n, const orc_prim* pts, const orc_prim* ees,
                      uint64_t first_event_id, uint64_t first_timestamp_us, int emit, orc_certificate* certs,
                      orc_audit_row* rows, char* msg, int64_t msg_cap) {
  std::vector<p2cccd::ExactCertificateQuery> queue(static_cast<std::size_t>(n));
  for (int64_t i = 0; i < n; ++i) {
    const orc_queue_entry& e = entries[i];
    p2cccd::ExactCertificateQuery& q = queue[static_cast<std::size_t>(i)];
    q.work_item.work_item_id = e.work_item_id;
    q.work_item.parent_candidate_id = e.parent_candidate_id;
    q.work_item.query_id = e.query_id;
    q.work_item.slab_id = e.slab_id;
    q.work_item.patch_a_id = e.patch_a_id;
    q.work_item.patch_b_id = e.patch_b_id;
    q.work_item.interval_t0 = e.interval_t0;
    q.work_item.interval_t1 = e.interval_t1;
    q.work_item.feature_family_mask = e.feature_family_mask;
    q.work_item.topk_feature_ids_offset = e.topk_feature_ids_offset;
    q.work_item.depth = e.depth;
    q.work_item.priority_score = e.priority_score;
    q.work_item.source = static_cast<p2cccd::ProposalSource>(e.source);
    q.config.eps_time = e.eps_time;
    q.config.eps_space = e.eps_space;
    q.config.max_subdivision_depth = static_cast<std::uint16_t>(e.max_subdivision_depth);
    for (int64_t k = 0; k < e.pt_count; ++k) q.point_triangle_primitives.push_back(ToPt(pts[e.pt_begin + k]));
    for (int64_t k = 0; k < e.ee_count; ++k) q.edge_edge_primitives.push_back(ToEe(ees[e.ee_begin + k]));
  }
  p2cccd::ExactWorkQueueConfig cfg;
  cfg.first_event_id = first_event_id;
  cfg.first_timestamp_us = first_timestamp_us;
  cfg.emit_dequeue_events = emit != 0;
  p2cccd::ExactWorkQueueResult res;
  p2cccd::Status st = p2cccd::ProcessExactWorkQueueCpu(queue, cfg, &res);
  if (!st.ok) {
    if (msg_cap > 0) {
      std::strncpy(msg, st.message.c_str(), static_cast<std::size_t>(msg_cap - 1));
      msg[msg_cap - 1] = 0;
    }
    return -1;
  }
  for (std::size_t i = 0; i < res.certificates.size(); ++i) FromCert(res.certificates[i], &certs[i]);
  for (std::size_t i = 0; i < res.audit_log.size(); ++i) FromRow(res.audit_log[i], &rows[i]);
  return 0;
}

// GenerateConservativeRefinementWorkItems for each entry in order, with the
// running first child id; children as entries with the parent's config and
// ranges.  Returns the child count, or -1 with the message.
int64_t ref_refine_queue(const orc_queue_entry* entries, const orc_certificate* certs, int64_t n,
                         uint64_t first_child, uint32_t max_child_depth, double min_interval_width,
                         orc_queue_entry* out, int64_t cap, char* msg, int64_t msg_cap) {
  p2cccd::ExactRefinementConfig rc;
  rc.max_child_depth = static_cast<std::uint16_t>(max_child_depth);
  rc.min_interval_width = min_interval_width;
  int64_t count = 0;
  for (int64_t i = 0; i < n; ++i) {
    const orc_queue_entry& e = entries[i];
    p2cccd::ExactWorkItem parent;
    parent.work_item_id = e.work_item_id;
    parent.parent_candidate_id = e.parent_candidate_id;
    parent.query_id = e.query_id;
    parent.slab_id = e.slab_id;
    parent.patch_a_id = e.patch_a_id;
    parent.patch_b_id = e.patch_b_id;
    parent.interval_t0 = e.interval_t0;
    parent.interval_t1 = e.interval_t1;
    parent.feature_family_mask = e.feature_family_mask;
    parent.topk_feature_ids_offset = e.topk_feature_ids_offset;
    parent.depth = e.depth;
    parent.priority_score = e.priority_score;
    parent.source = static_cast<p2cccd::ProposalSource>(e.source);
    p2cccd::CertificateResult c = ToCert(certs[i]);
    rc.first_child_work_item_id = first_child + static_cast<std::uint64_t>(count);
    std::vector<p2cccd::ExactWorkItem> children;
    p2cccd::Status st = p2cccd::GenerateConservativeRefinementWorkItems(parent, c, rc, &children);
    if (!st.ok) {
      if (msg_cap > 0) {
        std::strncpy(msg, st.message.c_str(), static_cast<std::size_t>(msg_cap - 1));
        msg[msg_cap - 1] = 0;
      }
      return -1;
    }
    for (const auto& ch : children) {
      if (count < cap) {
        orc_queue_entry o = e;
        o.work_item_id = ch.work_item_id;
        o.parent_candidate_id = ch.parent_candidate_id;
        o.query_id = ch.query_id;
        o.slab_id = ch.slab_id;
        o.patch_a_id = ch.patch_a_id;
        o.patch_b_id = ch.patch_b_id;
        o.interval_t0 = ch.interval_t0;
        o.interval_t1 = ch.interval_t1;
        o.feature_family_mask = ch.feature_family_mask;
        o.topk_feature_ids_offset = ch.topk_feature_ids_offset;
        o.depth = ch.depth;
        o.priority_score = ch.priority_score;
        o.source = static_cast<uint8_t>(ch.source);
        out[count] = o;
      }
      ++count;
    }
  }
  return count;
}

}  // extern "C"
