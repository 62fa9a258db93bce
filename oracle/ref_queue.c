/*
 * Oracle (test infrastructure only): the exact work queue restated in plain
 * C over the restated primitive certificate (orc_ref_eval, ref_cert.c).
 *
 *   orc_queue_process  ProcessExactWorkQueueCpu (certificate_engine.cpp:
 *                      593-661) with CertificateEngine::Evaluate (:729-824),
 *                      MakeBaseCertificateResult / ApplyPrimitiveCollision /
 *                      ApplyUndecided / MakeInvalidInputCertificate
 *                      (:388-424), AuditActionForCertificate and MakeAuditRow
 *                      (:436-468), and the validators it calls
 *                      (validators.cpp:207-315)
 *   orc_queue_refine   GenerateConservativeRefinementWorkItems (:663-720)
 *                      per entry with a running first child id
 *
 * Pinned against the compiled reference (oracle/_ref) and its golden vectors.
 */
#include <math.h>
#include <stddef.h>
#include <string.h>

#include "oracle.h"

static void base_cert(const orc_queue_entry* e, orc_certificate* c) { /* :388-402 */
  memset(c, 0, sizeof(*c));
  c->work_item_id = e->work_item_id;
  c->query_id = e->query_id;
  c->interval_t0 = e->interval_t0;
  c->interval_t1 = e->interval_t1;
  c->toi_upper = e->interval_t1;
  c->safe_margin_lb = 0.0;
  c->witness_id_a = -1;
  c->witness_id_b = -1;
  c->eps_time = e->eps_time;
  c->eps_space = e->eps_space;
  c->status = 2;
  c->reason_code = 1;
  c->next_refinement_mode = 2;
}

static uint8_t mode_for(uint16_t reason) { /* RefinementModeForReason */
  if (reason == 1 || reason == 3) return 2;
  if (reason == 2) return 1;
  return 3;
}

static void undecided(uint16_t reason, orc_certificate* c) { /* :417-424 */
  c->status = 2;
  c->covered_feature_mask = 0;
  c->safe_margin_lb = 0.0;
  c->reason_code = reason == 0 ? 2 : reason;
  c->next_refinement_mode = mode_for(c->reason_code);
}

/* CertificateEngine::Evaluate over one entry: 0, or -1 when it fails (an
 * invalid config or an invalid primitive reached before any collision). */
static int evaluate(const orc_queue_entry* e, const orc_prim* pts, const orc_prim* ees, orc_certificate* c) {
  if (!isfinite(e->eps_time) || e->eps_time <= 0.0 || !isfinite(e->eps_space) || e->eps_space <= 0.0) return -1;
  base_cert(e, c);
  double margin = INFINITY;
  uint32_t covered = 0;
  int any = 0, und = 0;
  uint16_t reason = 0;
  for (int f = 0; f < 2; ++f) {
    if (!(e->feature_family_mask & (1u << f))) continue;
    const int64_t count = f == 0 ? e->pt_count : e->ee_count;
    const orc_prim* prims = f == 0 ? pts + e->pt_begin : ees + e->ee_begin;
    if (count == 0) {
      und = 1;
      reason = 1;
    }
    for (int64_t k = 0; k < count; ++k) {
      any = 1;
      orc_result r;
      if (orc_ref_eval(f + 1, &prims[k], e->interval_t0, e->interval_t1, e->eps_time, e->eps_space,
                       (int)e->max_subdivision_depth, &r) != 0)
        return -1;
      if (r.status == 0) { /* ApplyPrimitiveCollision, :404-415 */
        c->status = 0;
        c->toi_upper = r.toi_upper;
        c->safe_margin_lb = 0.0;
        c->witness_family = (uint8_t)r.witness_family;
        c->witness_id_a = r.witness_id_a;
        c->witness_id_b = r.witness_id_b;
        c->covered_feature_mask = 0;
        c->reason_code = 0;
        c->next_refinement_mode = 0;
        return 0;
      }
      if (r.status == 1) {
        covered |= r.covered_feature_mask;
        margin = r.safe_margin_lb < margin ? r.safe_margin_lb : margin;
      } else {
        und = 1;
        reason = (uint16_t)r.reason_code;
      }
    }
  }
  if (!any) {
    undecided(1, c);
  } else if (und) {
    undecided(reason, c);
  } else {
    c->status = 1;
    c->safe_margin_lb = isfinite(margin) ? margin : 0.0;
    c->covered_feature_mask = covered;
    c->reason_code = 0;
    c->next_refinement_mode = 0;
  }
  return 0;
}

static const char* work_item_error(const orc_queue_entry* e) { /* validators.cpp:207-230 */
  if (e->work_item_id == 0) return "ExactWorkItem.work_item_id is required";
  if (e->parent_candidate_id == 0) return "ExactWorkItem.parent_candidate_id is required";
  if (e->query_id == 0) return "ExactWorkItem.query_id is required";
  if (!isfinite(e->interval_t0)) return "interval_t0 must be finite";
  if (!isfinite(e->interval_t1)) return "interval_t1 must be finite";
  if (e->interval_t0 < 0.0 || e->interval_t1 > 1.0 || e->interval_t0 > e->interval_t1)
    return "interval must satisfy 0 <= interval_t0 <= interval_t1 <= 1";
  if (e->feature_family_mask == 0) return "ExactWorkItem.feature_family_mask is required";
  if (!isfinite(e->priority_score)) return "ExactWorkItem.priority_score must be finite";
  if (e->source > 2) return "ExactWorkItem.source is invalid";
  return NULL;
}

static const char* cert_error(const orc_certificate* c) { /* validators.cpp:232-293 */
  if (c->work_item_id == 0) return "CertificateResult.work_item_id is required";
  if (c->query_id == 0) return "CertificateResult.query_id is required";
  if (c->status > 2) return "CertificateResult.status is invalid";
  if (!isfinite(c->interval_t0)) return "interval_t0 must be finite";
  if (!isfinite(c->interval_t1)) return "interval_t1 must be finite";
  if (c->interval_t0 < 0.0 || c->interval_t1 > 1.0 || c->interval_t0 > c->interval_t1)
    return "interval must satisfy 0 <= interval_t0 <= interval_t1 <= 1";
  if (!isfinite(c->toi_upper)) return "CertificateResult.toi_upper must be finite";
  if (!isfinite(c->safe_margin_lb)) return "CertificateResult.safe_margin_lb must be finite";
  if (!isfinite(c->eps_time)) return "CertificateResult.eps_time must be finite";
  if (!(c->eps_time > 0.0)) return "CertificateResult.eps_time must be positive";
  if (!isfinite(c->eps_space)) return "CertificateResult.eps_space must be finite";
  if (!(c->eps_space > 0.0)) return "CertificateResult.eps_space must be positive";
  if (c->next_refinement_mode > 3) return "CertificateResult.next_refinement_mode is invalid";
  if (c->status == 0) {
    if (c->toi_upper < c->interval_t0 || c->toi_upper > c->interval_t1)
      return "CertificateResult.toi_upper must lie inside the certified interval";
    if (c->witness_id_a < 0 || c->witness_id_b < 0) return "CertificateResult collision witnesses are required";
    if (c->next_refinement_mode != 0) return "CertificateResult collision cannot request refinement";
  }
  if (c->status == 1) {
    if (c->safe_margin_lb < 0.0) return "CertificateResult.safe_margin_lb must be non-negative";
    if (c->covered_feature_mask == 0) return "CertificateResult.covered_feature_mask is required for separation";
    if (c->next_refinement_mode != 0) return "CertificateResult separation cannot request refinement";
  }
  if (c->status == 2 && c->reason_code == 0) return "CertificateResult.reason_code is required for undecided results";
  if (c->status == 2 && c->next_refinement_mode == 0)
    return "CertificateResult.next_refinement_mode is required for undecided results";
  return NULL;
}

static const char* row_error(const orc_audit_row* r) { /* validators.cpp:295-315 */
  if (r->event_id == 0) return "AuditLogRow.event_id is required";
  if (r->query_id == 0) return "AuditLogRow.query_id is required";
  if (r->stage > 4) return "AuditLogRow.stage is invalid";
  if (!isfinite(r->interval_t0)) return "interval_t0 must be finite";
  if (!isfinite(r->interval_t1)) return "interval_t1 must be finite";
  if (r->interval_t0 < 0.0 || r->interval_t1 > 1.0 || r->interval_t0 > r->interval_t1)
    return "interval must satisfy 0 <= interval_t0 <= interval_t1 <= 1";
  if (r->timestamp_us == 0) return "AuditLogRow.timestamp_us is required";
  if (!isfinite(r->aux_value0)) return "AuditLogRow.aux_value0 must be finite";
  if (!isfinite(r->aux_value1)) return "AuditLogRow.aux_value1 must be finite";
  return NULL;
}

static void make_row(uint64_t event, uint64_t ts, const orc_queue_entry* e, uint8_t stage, uint16_t action,
                     double aux0, double aux1, orc_audit_row* r) { /* :450-468 */
  memset(r, 0, sizeof(*r));
  r->event_id = event;
  r->query_id = e->query_id;
  r->candidate_id = e->parent_candidate_id;
  r->work_item_id = e->work_item_id;
  r->stage = stage;
  r->action = action;
  r->depth = e->depth;
  r->interval_t0 = e->interval_t0;
  r->interval_t1 = e->interval_t1;
  r->timestamp_us = ts;
  r->aux_value0 = aux0;
  r->aux_value1 = aux1;
}

int orc_queue_process(const orc_queue_entry* entries, int64_t n, const orc_prim* pts, const orc_prim* ees,
                      uint64_t first_event_id, uint64_t first_timestamp_us, int emit, orc_certificate* certs,
                      orc_audit_row* rows, const char** msg) {
  *msg = NULL;
  if (first_event_id == 0) {
    *msg = "ExactWorkQueueConfig.first_event_id must be non-zero";
    return -1;
  }
  if (first_timestamp_us == 0) {
    *msg = "ExactWorkQueueConfig.first_timestamp_us must be non-zero";
    return -1;
  }
  uint64_t event = first_event_id, ts = first_timestamp_us;
  int64_t nrow = 0;
  for (int64_t i = 0; i < n; ++i) {
    const orc_queue_entry* e = &entries[i];
    if ((*msg = work_item_error(e))) return -1;
    if (emit) {
      make_row(event++, ts++, e, 2, 1, (double)e->priority_score, (double)e->feature_family_mask, &rows[nrow]);
      if ((*msg = row_error(&rows[nrow]))) return -1;
      ++nrow;
    }
    orc_certificate c;
    if (evaluate(e, pts, ees, &c) != 0) { /* MakeInvalidInputCertificate, :426-430 */
      base_cert(e, &c);
      undecided(3, &c);
    }
    if ((*msg = cert_error(&c))) return -1;
    const uint16_t action = c.status == 0 ? 2 : c.status == 1 ? 3 : (c.reason_code == 3 ? 5 : 4);
    make_row(event++, ts++, e, 4, action, c.status == 1 ? c.safe_margin_lb : c.toi_upper,
             (double)c.next_refinement_mode, &rows[nrow]);
    if ((*msg = row_error(&rows[nrow]))) return -1;
    ++nrow;
    certs[i] = c;
  }
  /* ValidateExactWorkQueueCoverage (:553-591): one certificate per entry, so
   * only a repeated (query_id, work_item_id) can fail it; quadratic scan */
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < i; ++j)
      if (entries[i].query_id == entries[j].query_id && entries[i].work_item_id == entries[j].work_item_id) {
        *msg = "exact work queue contains duplicate work_item_id for a query";
        return -1;
      }
  return 0;
}

int64_t orc_queue_refine(const orc_queue_entry* entries, const orc_certificate* certs, int64_t n,
                         uint64_t first_child, uint32_t max_child_depth, double min_interval_width,
                         orc_queue_entry* out, int64_t cap) {
  int64_t count = 0;
  for (int64_t i = 0; i < n; ++i) {
    const orc_queue_entry* p = &entries[i];
    const orc_certificate* c = &certs[i];
    if ((uint32_t)p->depth >= max_child_depth) continue;
    double t[3];
    int k = 0;
    if (c->status == 0) {
      const double hi = c->toi_upper < p->interval_t0 ? p->interval_t0
                        : (p->interval_t1 < c->toi_upper ? p->interval_t1 : c->toi_upper);
      t[0] = p->interval_t0;
      t[1] = hi;
      if (t[1] - t[0] > min_interval_width) k = 1;
    } else if (c->status == 2 && c->next_refinement_mode == 1 &&
               p->interval_t1 - p->interval_t0 > min_interval_width) {
      t[0] = p->interval_t0;
      t[1] = 0.5 * (p->interval_t0 + p->interval_t1);
      t[2] = p->interval_t1;
      k = 2;
    }
    for (int j = 0; j < k; ++j, ++count) {
      if (count >= cap) continue;
      orc_queue_entry ch = *p;
      ch.work_item_id = first_child + (uint64_t)count;
      ch.interval_t0 = t[j];
      ch.interval_t1 = t[j + 1];
      ch.depth = (uint16_t)(p->depth + 1);
      ch.source = 1;
      out[count] = ch;
    }
  }
  return count;
}
