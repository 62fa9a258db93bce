// Implementation of the reference C++ certificate interface over libp2c_hip.
//
// Single-query aggregation, coverage guard, per-parent refinement children and
// validators restate src/cpp/certificate/certificate_engine.cpp:388-824 and
// src/cpp/common/validators.cpp:160-315 (same rules, same messages);
// primitive certificates come from the gfx950 kernels, and the work queue
// (certificates and audit rows of every entry) from the library's device
// queue pass.
#include "certificate_hip.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <map>
#include <set>
#include <tuple>
#include <utility>

#include "p2c_hip.h"

namespace p2cccd {
namespace {

Status FromRc(int rc) {
  if (rc == P2C_OK) return Status::Ok();
  return Status::Error(p2c_last_error());
}

p2c_traj ToC(const LinearVertexTrajectory& t) {
  p2c_traj o;
  o.feature_id = t.feature_id;
  for (int k = 0; k < 3; ++k) {
    o.position_t0[k] = t.position_t0[k];
    o.position_t1[k] = t.position_t1[k];
  }
  return o;
}

p2c_prim Pack(const PointTriangleIntervalPrimitive& p) {
  p2c_prim o;
  o.id_a = p.point_id;
  o.id_b = p.triangle_id;
  o.v[0] = ToC(p.point);
  o.v[1] = ToC(p.triangle_v0);
  o.v[2] = ToC(p.triangle_v1);
  o.v[3] = ToC(p.triangle_v2);
  return o;
}

p2c_prim Pack(const EdgeEdgeIntervalPrimitive& p) {
  p2c_prim o;
  o.id_a = p.edge_a_id;
  o.id_b = p.edge_b_id;
  o.v[0] = ToC(p.edge_a0);
  o.v[1] = ToC(p.edge_a1);
  o.v[2] = ToC(p.edge_b0);
  o.v[3] = ToC(p.edge_b1);
  return o;
}

PrimitiveIntervalResult Unpack(const p2c_prim_result& r) {
  PrimitiveIntervalResult o;
  o.status = static_cast<CertificateStatus>(r.status);
  o.covered_feature_mask = r.covered_feature_mask;
  o.interval_t0 = r.interval_t0;
  o.interval_t1 = r.interval_t1;
  o.toi_upper = r.toi_upper;
  o.safe_margin_lb = r.safe_margin_lb;
  o.witness_family = r.witness_family;
  o.witness_id_a = r.witness_id_a;
  o.witness_id_b = r.witness_id_b;
  o.reason_code = r.reason_code;
  o.next_refinement_mode = static_cast<CertificateRefinementMode>(r.next_refinement_mode);
  return o;
}

p2c_cert_config ToC(const CertificateEngineConfig& c) {
  p2c_cert_config o;
  o.eps_time = c.eps_time;
  o.eps_space = c.eps_space;
  o.max_subdivision_depth = c.max_subdivision_depth;
  o.flags = 0;
  return o;
}

bool PrimValid(const p2c_prim& p) {
  if (p.id_a < 0 || p.id_b < 0) return false;
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 3; ++k)
      if (!std::isfinite(p.v[i].position_t0[k]) || !std::isfinite(p.v[i].position_t1[k]))
        return false;
  return true;
}

const char* InvalidPrimMessage(int family, const p2c_prim& p) {
  const bool pt = family == P2C_FAMILY_POINT_TRIANGLE;
  if (p.id_a < 0 || p.id_b < 0)
    return pt ? "point-triangle primitive ids are required" : "edge-edge primitive ids are required";
  return pt ? "point-triangle primitive trajectories must be finite"
            : "edge-edge primitive trajectories must be finite";
}

// ---- validators.cpp:13-72 helpers --------------------------------------
Status Require(bool c, const char* m) { return c ? Status::Ok() : Status::Error(m); }
Status RequireNonZero(std::uint64_t v, const char* f) {
  return v != 0 ? Status::Ok() : Status::Error(std::string(f) + " is required");
}
Status RequireFinite(double v, const char* f) {
  return std::isfinite(v) ? Status::Ok() : Status::Error(std::string(f) + " must be finite");
}
Status RequirePositive(double v, const char* f) {
  if (auto s = RequireFinite(v, f); !s.ok) return s;
  return v > 0.0 ? Status::Ok() : Status::Error(std::string(f) + " must be positive");
}
Status ValidateUnitInterval(double t0, double t1) {
  if (auto s = RequireFinite(t0, "interval_t0"); !s.ok) return s;
  if (auto s = RequireFinite(t1, "interval_t1"); !s.ok) return s;
  if (t0 < 0.0 || t1 > 1.0 || t0 > t1)
    return Status::Error("interval must satisfy 0 <= interval_t0 <= interval_t1 <= 1");
  return Status::Ok();
}
bool KnownSource(ProposalSource v) { return static_cast<int>(v) <= 2; }
bool KnownStatus(CertificateStatus v) { return static_cast<int>(v) <= 2; }
bool KnownMode(CertificateRefinementMode v) { return static_cast<int>(v) <= 3; }
bool KnownStage(AuditStage v) { return static_cast<int>(v) <= 4; }

// ---- certificate_engine.cpp:224-232, 279-298, 388-477 -------------------
Status ValidateConfig(const CertificateEngineConfig& c) {
  if (!std::isfinite(c.eps_time) || c.eps_time <= 0.0)
    return Status::Error("CertificateEngineConfig.eps_time must be finite and positive");
  if (!std::isfinite(c.eps_space) || c.eps_space <= 0.0)
    return Status::Error("CertificateEngineConfig.eps_space must be finite and positive");
  return Status::Ok();
}

CertificateRefinementMode ModeForReason(std::uint16_t reason) {
  switch (reason) {
    case kCertificateReasonMissingGeometry:
      return CertificateRefinementMode::kRequestGeometry;
    case kCertificateReasonMaxSubdivisionDepth:
      return CertificateRefinementMode::kBisectInterval;
    case kCertificateReasonInvalidInput:
      return CertificateRefinementMode::kRequestGeometry;
    default:
      return CertificateRefinementMode::kEscalatePrecision;
  }
}

CertificateResult BaseResult(const ExactWorkItem& w, const CertificateEngineConfig& c) {
  CertificateResult r;
  r.work_item_id = w.work_item_id;
  r.query_id = w.query_id;
  r.interval_t0 = w.interval_t0;
  r.interval_t1 = w.interval_t1;
  r.toi_upper = w.interval_t1;
  r.eps_time = c.eps_time;
  r.eps_space = c.eps_space;
  r.status = CertificateStatus::kUndecided;
  r.reason_code = kCertificateReasonMissingGeometry;
  r.next_refinement_mode = CertificateRefinementMode::kRequestGeometry;
  return r;
}

void ApplyCollision(const PrimitiveIntervalResult& p, CertificateResult* r) {
  r->status = CertificateStatus::kCollision;
  r->toi_upper = p.toi_upper;
  r->safe_margin_lb = 0.0;
  r->witness_family = p.witness_family;
  r->witness_id_a = p.witness_id_a;
  r->witness_id_b = p.witness_id_b;
  r->covered_feature_mask = 0U;
  r->reason_code = kCertificateReasonNone;
  r->next_refinement_mode = CertificateRefinementMode::kNone;
}

void ApplyUndecided(std::uint16_t reason, CertificateResult* r) {
  r->status = CertificateStatus::kUndecided;
  r->covered_feature_mask = 0U;
  r->safe_margin_lb = 0.0;
  r->reason_code = reason == kCertificateReasonNone ? kCertificateReasonMaxSubdivisionDepth : reason;
  r->next_refinement_mode = ModeForReason(r->reason_code);
}

// A batch of primitives sharing (family, interval, config), certified in one
// gfx950 launch.  Invalid primitives are not sent; they keep status INVALID.
struct LaunchKey {
  int family;
  double t0, t1, eps_time, eps_space;
  int depth;
  bool operator<(const LaunchKey& o) const {
    return std::tie(family, t0, t1, eps_time, eps_space, depth) <
           std::tie(o.family, o.t0, o.t1, o.eps_time, o.eps_space, o.depth);
  }
};

struct PrimSlot {
  p2c_prim prim;
  p2c_prim_result result;
  bool valid;
};

Status RunLaunches(std::vector<PrimSlot>* slots, const std::vector<LaunchKey>& keys) {
  std::map<LaunchKey, std::vector<std::size_t>> groups;
  for (std::size_t i = 0; i < slots->size(); ++i)
    if ((*slots)[i].valid) groups[keys[i]].push_back(i);
  for (const auto& [key, idx] : groups) {
    std::vector<p2c_prim> in(idx.size());
    std::vector<p2c_prim_result> out(idx.size());
    for (std::size_t j = 0; j < idx.size(); ++j) in[j] = (*slots)[idx[j]].prim;
    p2c_cert_config cfg{key.eps_time, key.eps_space, static_cast<std::uint32_t>(key.depth), 0};
    if (auto s = FromRc(p2c_cert_ref_batch_host(key.family, in.data(), static_cast<std::int64_t>(in.size()),
                                                 key.t0, key.t1, &cfg, out.data()));
        !s.ok)
      return s;
    for (std::size_t j = 0; j < idx.size(); ++j) (*slots)[idx[j]].result = out[j];
  }
  return Status::Ok();
}

// CertificateEngine::Evaluate aggregation (certificate_engine.cpp:729-824)
// over already-certified primitive slots [pt_begin, pt_end) / [ee_begin, ee_end).
Status Aggregate(const ExactCertificateQuery& q, const std::vector<PrimSlot>& slots,
                 std::size_t pt_begin, std::size_t ee_begin, CertificateResult* result) {
  CertificateResult cert = BaseResult(q.work_item, q.config);
  double margin = std::numeric_limits<double>::infinity();
  std::uint32_t mask = 0U;
  bool any = false, undecided = false;
  std::uint16_t reason = kCertificateReasonNone;
  const std::uint32_t fam_mask = q.work_item.feature_family_mask;
  auto visit = [&](bool enabled, std::size_t begin, std::size_t count, int family,
                   bool* done) -> Status {
    if (!enabled) return Status::Ok();
    if (count == 0) {
      undecided = true;
      reason = kCertificateReasonMissingGeometry;
    }
    for (std::size_t i = 0; i < count; ++i) {
      any = true;
      const PrimSlot& s = slots[begin + i];
      if (!s.valid) return Status::Error(InvalidPrimMessage(family, s.prim));
      const PrimitiveIntervalResult r = Unpack(s.result);
      if (r.status == CertificateStatus::kCollision) {
        ApplyCollision(r, &cert);
        *done = true;
        return Status::Ok();
      }
      if (r.status == CertificateStatus::kSeparation) {
        mask |= r.covered_feature_mask;
        margin = std::min(margin, r.safe_margin_lb);
      } else {
        undecided = true;
        reason = r.reason_code;
      }
    }
    return Status::Ok();
  };
  bool done = false;
  if (auto s = visit((fam_mask & kFeatureFamilyPointTriangle) != 0U, pt_begin,
                     q.point_triangle_primitives.size(), P2C_FAMILY_POINT_TRIANGLE, &done);
      !s.ok)
    return s;
  if (!done) {
    if (auto s = visit((fam_mask & kFeatureFamilyEdgeEdge) != 0U, ee_begin,
                       q.edge_edge_primitives.size(), P2C_FAMILY_EDGE_EDGE, &done);
        !s.ok)
      return s;
  }
  if (!done) {
    if (!any) {
      ApplyUndecided(kCertificateReasonMissingGeometry, &cert);
    } else if (undecided) {
      ApplyUndecided(reason, &cert);
    } else {
      cert.status = CertificateStatus::kSeparation;
      cert.safe_margin_lb = std::isfinite(margin) ? margin : 0.0;
      cert.covered_feature_mask = mask;
      cert.reason_code = kCertificateReasonNone;
      cert.next_refinement_mode = CertificateRefinementMode::kNone;
    }
  }
  *result = cert;
  return Status::Ok();
}

// Certifies every primitive of every query whose config is valid, in as few
// gfx950 launches as there are distinct (family, interval, config) keys.
// begin offsets per query: pt at 2*i, ee at 2*i+1.
Status CertifyQueries(const std::vector<ExactCertificateQuery>& queue, std::vector<PrimSlot>* slots,
                      std::vector<std::size_t>* begins) {
  std::vector<LaunchKey> keys;
  begins->assign(queue.size() * 2, 0);
  for (std::size_t qi = 0; qi < queue.size(); ++qi) {
    const ExactCertificateQuery& q = queue[qi];
    const bool cfg_ok = ValidateConfig(q.config).ok;
    auto add = [&](int family, const p2c_prim& p, bool enabled) {
      PrimSlot s;
      s.prim = p;
      s.valid = PrimValid(p);
      s.result = p2c_prim_result{};
      s.result.status = P2C_STATUS_INVALID;
      if (!(cfg_ok && enabled)) s.valid = false;
      slots->push_back(s);
      keys.push_back(LaunchKey{family, q.work_item.interval_t0, q.work_item.interval_t1,
                               q.config.eps_time, q.config.eps_space, q.config.max_subdivision_depth});
    };
    const bool pt_on = (q.work_item.feature_family_mask & kFeatureFamilyPointTriangle) != 0U;
    const bool ee_on = (q.work_item.feature_family_mask & kFeatureFamilyEdgeEdge) != 0U;
    (*begins)[2 * qi] = slots->size();
    for (const auto& p : q.point_triangle_primitives) add(P2C_FAMILY_POINT_TRIANGLE, Pack(p), pt_on);
    (*begins)[2 * qi + 1] = slots->size();
    for (const auto& p : q.edge_edge_primitives) add(P2C_FAMILY_EDGE_EDGE, Pack(p), ee_on);
  }
  return RunLaunches(slots, keys);
}

Status EvaluateQuery(const ExactCertificateQuery& query, CertificateResult* result) {
  if (result == nullptr) return Status::Error("certificate result output pointer is null");
  if (auto s = ValidateConfig(query.config); !s.ok) return s;
  if (auto s = ValidateExactWorkItem(query.work_item); !s.ok) return s;
  std::vector<PrimSlot> slots;
  std::vector<std::size_t> begins;
  if (auto s = CertifyQueries({query}, &slots, &begins); !s.ok) return s;
  return Aggregate(query, slots, begins[0], begins[1], result);
}

template <typename Prim>
Status RunBatch(int family, const std::vector<Prim>& prims, double t0, double t1,
                const CertificateEngineConfig& config, std::vector<PrimitiveIntervalResult>* results) {
  if (results == nullptr) return Status::Error("CUDA exact result output pointer is null");
  results->clear();
  if (prims.empty()) return Status::Ok();
  std::vector<p2c_prim> in;
  in.reserve(prims.size());
  for (const Prim& p : prims) in.push_back(Pack(p));
  std::vector<p2c_prim_result> out(prims.size());
  p2c_cert_config cfg = ToC(config);
  cfg.flags = P2C_CERT_CUDA_TWIN;  // the CUDA twin's walk (certificate_cuda.cu:286-346)
  if (auto s = FromRc(p2c_cert_ref_batch_host(family, in.data(), static_cast<std::int64_t>(in.size()),
                                               t0, t1, &cfg, out.data()));
      !s.ok)
    return s;
  results->reserve(out.size());
  for (const auto& r : out) results->push_back(Unpack(r));
  return Status::Ok();
}

// ComparePrimitiveResults (certificate_cuda.cu:554-583)
Status Compare(const PrimitiveIntervalResult& cpu, const PrimitiveIntervalResult& gpu, double eps,
               const char* label) {
  const std::string l(label);
  if (cpu.status != gpu.status) return Status::Error(l + " status mismatch");
  if (cpu.covered_feature_mask != gpu.covered_feature_mask)
    return Status::Error(l + " covered feature mask mismatch");
  if (cpu.status == CertificateStatus::kCollision) {
    if (std::abs(cpu.toi_upper - gpu.toi_upper) > eps) return Status::Error(l + " TOI mismatch");
    if (cpu.witness_family != gpu.witness_family || cpu.witness_id_a != gpu.witness_id_a ||
        cpu.witness_id_b != gpu.witness_id_b)
      return Status::Error(l + " witness mismatch");
  }
  if (cpu.status == CertificateStatus::kSeparation &&
      std::abs(cpu.safe_margin_lb - gpu.safe_margin_lb) > eps)
    return Status::Error(l + " safe margin mismatch");
  if (cpu.status == CertificateStatus::kUndecided &&
      (cpu.reason_code != gpu.reason_code || cpu.next_refinement_mode != gpu.next_refinement_mode))
    return Status::Error(l + " undecided refinement mismatch");
  return Status::Ok();
}

}  // namespace

// ---- validators (validators.cpp:178-292) ---------------------------------
Status ValidateExactWorkItem(const ExactWorkItem& item) {
  if (auto s = RequireNonZero(item.work_item_id, "ExactWorkItem.work_item_id"); !s.ok) return s;
  if (auto s = RequireNonZero(item.parent_candidate_id, "ExactWorkItem.parent_candidate_id"); !s.ok)
    return s;
  if (auto s = RequireNonZero(item.query_id, "ExactWorkItem.query_id"); !s.ok) return s;
  if (auto s = ValidateUnitInterval(item.interval_t0, item.interval_t1); !s.ok) return s;
  if (item.feature_family_mask == 0) return Status::Error("ExactWorkItem.feature_family_mask is required");
  if (auto s = RequireFinite(item.priority_score, "ExactWorkItem.priority_score"); !s.ok) return s;
  return Require(KnownSource(item.source), "ExactWorkItem.source is invalid");
}

Status ValidateCertificateResult(const CertificateResult& r) {
  if (auto s = RequireNonZero(r.work_item_id, "CertificateResult.work_item_id"); !s.ok) return s;
  if (auto s = RequireNonZero(r.query_id, "CertificateResult.query_id"); !s.ok) return s;
  if (!KnownStatus(r.status)) return Status::Error("CertificateResult.status is invalid");
  if (auto s = ValidateUnitInterval(r.interval_t0, r.interval_t1); !s.ok) return s;
  if (auto s = RequireFinite(r.toi_upper, "CertificateResult.toi_upper"); !s.ok) return s;
  if (auto s = RequireFinite(r.safe_margin_lb, "CertificateResult.safe_margin_lb"); !s.ok) return s;
  if (auto s = RequirePositive(r.eps_time, "CertificateResult.eps_time"); !s.ok) return s;
  if (auto s = RequirePositive(r.eps_space, "CertificateResult.eps_space"); !s.ok) return s;
  if (!KnownMode(r.next_refinement_mode))
    return Status::Error("CertificateResult.next_refinement_mode is invalid");
  if (r.status == CertificateStatus::kCollision) {
    if (r.toi_upper < r.interval_t0 || r.toi_upper > r.interval_t1)
      return Status::Error("CertificateResult.toi_upper must lie inside the certified interval");
    if (r.witness_id_a < 0 || r.witness_id_b < 0)
      return Status::Error("CertificateResult collision witnesses are required");
    if (r.next_refinement_mode != CertificateRefinementMode::kNone)
      return Status::Error("CertificateResult collision cannot request refinement");
  }
  if (r.status == CertificateStatus::kSeparation) {
    if (r.safe_margin_lb < 0.0) return Status::Error("CertificateResult.safe_margin_lb must be non-negative");
    if (r.covered_feature_mask == 0)
      return Status::Error("CertificateResult.covered_feature_mask is required for separation");
    if (r.next_refinement_mode != CertificateRefinementMode::kNone)
      return Status::Error("CertificateResult separation cannot request refinement");
  }
  if (r.status == CertificateStatus::kUndecided && r.reason_code == 0)
    return Status::Error("CertificateResult.reason_code is required for undecided results");
  if (r.status == CertificateStatus::kUndecided && r.next_refinement_mode == CertificateRefinementMode::kNone)
    return Status::Error("CertificateResult.next_refinement_mode is required for undecided results");
  return Status::Ok();
}

Status ValidateAuditLogRow(const AuditLogRow& row) {
  if (auto s = RequireNonZero(row.event_id, "AuditLogRow.event_id"); !s.ok) return s;
  if (auto s = RequireNonZero(row.query_id, "AuditLogRow.query_id"); !s.ok) return s;
  if (!KnownStage(row.stage)) return Status::Error("AuditLogRow.stage is invalid");
  if (auto s = ValidateUnitInterval(row.interval_t0, row.interval_t1); !s.ok) return s;
  if (row.timestamp_us == 0) return Status::Error("AuditLogRow.timestamp_us is required");
  if (auto s = RequireFinite(row.aux_value0, "AuditLogRow.aux_value0"); !s.ok) return s;
  return RequireFinite(row.aux_value1, "AuditLogRow.aux_value1");
}

// ---- primitive certificates ---------------------------------------------
Status EvaluatePointTriangleInterval(const PointTriangleIntervalPrimitive& primitive, double t0,
                                     double t1, const CertificateEngineConfig& config,
                                     PrimitiveIntervalResult* result) {
  if (result == nullptr) return Status::Error("point-triangle interval result pointer is null");
  const p2c_prim p = Pack(primitive);
  const p2c_cert_config c = ToC(config);
  p2c_prim_result r;
  if (auto s = FromRc(p2c_cert_ref_eval_cpu(P2C_FAMILY_POINT_TRIANGLE, &p, t0, t1, &c, &r)); !s.ok)
    return s;
  *result = Unpack(r);
  return Status::Ok();
}

Status EvaluateEdgeEdgeInterval(const EdgeEdgeIntervalPrimitive& primitive, double t0, double t1,
                                const CertificateEngineConfig& config, PrimitiveIntervalResult* result) {
  if (result == nullptr) return Status::Error("edge-edge interval result pointer is null");
  const p2c_prim p = Pack(primitive);
  const p2c_cert_config c = ToC(config);
  p2c_prim_result r;
  if (auto s = FromRc(p2c_cert_ref_eval_cpu(P2C_FAMILY_EDGE_EDGE, &p, t0, t1, &c, &r)); !s.ok) return s;
  *result = Unpack(r);
  return Status::Ok();
}

// ---- queue, coverage, refinement (certificate_engine.cpp:553-720) -------
Status ValidateExactWorkQueueCoverage(const std::vector<ExactCertificateQuery>& queue,
                                      const ExactWorkQueueResult& result) {
  if (result.processed_count != queue.size())
    return Status::Error("exact work queue processed_count must match input size");
  if (result.certificates.size() != queue.size())
    return Status::Error("each exact work item must produce exactly one certificate");
  std::set<std::pair<std::uint64_t, std::uint64_t>> expected;
  for (const auto& q : queue) {
    if (auto s = ValidateExactWorkItem(q.work_item); !s.ok) return s;
    if (!expected.insert({q.work_item.query_id, q.work_item.work_item_id}).second)
      return Status::Error("exact work queue contains duplicate work_item_id for a query");
  }
  std::set<std::pair<std::uint64_t, std::uint64_t>> covered;
  for (const auto& c : result.certificates) {
    if (auto s = ValidateCertificateResult(c); !s.ok) return s;
    const auto key = std::make_pair(c.query_id, c.work_item_id);
    if (expected.count(key) == 0) return Status::Error("certificate references a work item that was not queued");
    if (!covered.insert(key).second) return Status::Error("exact work item produced more than one certificate");
  }
  if (covered.size() != expected.size())
    return Status::Error("at least one exact work item disappeared without certificate coverage");
  return Status::Ok();
}

// ProcessExactWorkQueueCpu (certificate_engine.cpp:593-661) on the device:
// the queue is packed into p2c_queue_entry records and two primitive arrays,
// and p2c_exact_queue_host computes every certificate and audit row in one
// pass of kernels, failing with the message the reference's loop would give
// first.
Status ProcessExactWorkQueueCpu(const std::vector<ExactCertificateQuery>& queue,
                                const ExactWorkQueueConfig& config, ExactWorkQueueResult* result) {
  if (result == nullptr) return Status::Error("exact work queue result output pointer is null");
  std::vector<p2c_queue_entry> entries(queue.size());
  std::vector<p2c_prim> pts, ees;
  for (std::size_t i = 0; i < queue.size(); ++i) {
    const ExactCertificateQuery& q = queue[i];
    const ExactWorkItem& w = q.work_item;
    p2c_queue_entry& e = entries[i];
    e = p2c_queue_entry{};
    e.work_item_id = w.work_item_id;
    e.parent_candidate_id = w.parent_candidate_id;
    e.query_id = w.query_id;
    e.slab_id = w.slab_id;
    e.patch_a_id = w.patch_a_id;
    e.patch_b_id = w.patch_b_id;
    e.feature_family_mask = w.feature_family_mask;
    e.interval_t0 = w.interval_t0;
    e.interval_t1 = w.interval_t1;
    e.topk_feature_ids_offset = w.topk_feature_ids_offset;
    e.depth = w.depth;
    e.source = static_cast<std::uint8_t>(w.source);
    e.priority_score = w.priority_score;
    e.max_subdivision_depth = q.config.max_subdivision_depth;
    e.eps_time = q.config.eps_time;
    e.eps_space = q.config.eps_space;
    e.pt_begin = static_cast<std::int64_t>(pts.size());
    e.pt_count = static_cast<std::int64_t>(q.point_triangle_primitives.size());
    e.ee_begin = static_cast<std::int64_t>(ees.size());
    e.ee_count = static_cast<std::int64_t>(q.edge_edge_primitives.size());
    for (const auto& p : q.point_triangle_primitives) pts.push_back(Pack(p));
    for (const auto& p : q.edge_edge_primitives) ees.push_back(Pack(p));
  }
  const p2c_queue_config qc{config.first_event_id, config.first_timestamp_us,
                            config.emit_dequeue_events ? 1u : 0u, 0u};
  std::vector<p2c_certificate> certs(queue.size());
  std::vector<p2c_audit_row> rows(queue.size() * (config.emit_dequeue_events ? 2U : 1U));
  if (auto s = FromRc(p2c_exact_queue_host(entries.data(), static_cast<std::int64_t>(entries.size()), pts.data(),
                                           static_cast<std::int64_t>(pts.size()), ees.data(),
                                           static_cast<std::int64_t>(ees.size()), &qc, nullptr, certs.data(),
                                           rows.data(), nullptr, 0, nullptr));
      !s.ok)
    return s;
  ExactWorkQueueResult out;
  out.certificates.reserve(certs.size());
  for (const p2c_certificate& c : certs) {
    CertificateResult r;
    r.work_item_id = c.work_item_id;
    r.query_id = c.query_id;
    r.status = static_cast<CertificateStatus>(c.status);
    r.interval_t0 = c.interval_t0;
    r.interval_t1 = c.interval_t1;
    r.toi_upper = c.toi_upper;
    r.safe_margin_lb = c.safe_margin_lb;
    r.witness_family = c.witness_family;
    r.witness_id_a = c.witness_id_a;
    r.witness_id_b = c.witness_id_b;
    r.covered_feature_mask = c.covered_feature_mask;
    r.eps_time = c.eps_time;
    r.eps_space = c.eps_space;
    r.reason_code = c.reason_code;
    r.next_refinement_mode = static_cast<CertificateRefinementMode>(c.next_refinement_mode);
    out.certificates.push_back(r);
  }
  out.audit_log.reserve(rows.size());
  for (const p2c_audit_row& a : rows) {
    AuditLogRow r;
    r.event_id = a.event_id;
    r.query_id = a.query_id;
    r.candidate_id = a.candidate_id;
    r.work_item_id = a.work_item_id;
    r.stage = static_cast<AuditStage>(a.stage);
    r.action = a.action;
    r.depth = a.depth;
    r.interval_t0 = a.interval_t0;
    r.interval_t1 = a.interval_t1;
    r.timestamp_us = a.timestamp_us;
    r.aux_value0 = a.aux_value0;
    r.aux_value1 = a.aux_value1;
    out.audit_log.push_back(r);
  }
  out.processed_count = queue.size();
  *result = std::move(out);
  return Status::Ok();
}

Status GenerateConservativeRefinementWorkItems(const ExactWorkItem& parent,
                                               const CertificateResult& certificate,
                                               const ExactRefinementConfig& config,
                                               std::vector<ExactWorkItem>* children) {
  if (children == nullptr) return Status::Error("refinement child output pointer is null");
  children->clear();
  if (auto s = ValidateExactWorkItem(parent); !s.ok) return s;
  if (auto s = ValidateCertificateResult(certificate); !s.ok) return s;
  if (certificate.work_item_id != parent.work_item_id || certificate.query_id != parent.query_id)
    return Status::Error("certificate must correspond to the parent exact work item");
  if (config.first_child_work_item_id == 0)
    return Status::Error("ExactRefinementConfig.first_child_work_item_id must be non-zero");
  if (!std::isfinite(config.min_interval_width) || config.min_interval_width <= 0.0)
    return Status::Error("ExactRefinementConfig.min_interval_width must be finite and positive");
  if (parent.depth >= config.max_child_depth) return Status::Ok();
  auto child = [&](std::uint64_t id, double t0, double t1) {
    ExactWorkItem c = parent;
    c.work_item_id = id;
    c.interval_t0 = t0;
    c.interval_t1 = t1;
    c.depth = static_cast<std::uint16_t>(parent.depth + 1);
    c.source = ProposalSource::kRefined;
    return c;
  };
  if (certificate.status == CertificateStatus::kCollision) {
    const double t0 = parent.interval_t0;
    const double t1 = std::clamp(certificate.toi_upper, parent.interval_t0, parent.interval_t1);
    if (t1 - t0 > config.min_interval_width) children->push_back(child(config.first_child_work_item_id, t0, t1));
    return Status::Ok();
  }
  if (certificate.status == CertificateStatus::kUndecided &&
      certificate.next_refinement_mode == CertificateRefinementMode::kBisectInterval) {
    if (parent.interval_t1 - parent.interval_t0 > config.min_interval_width) {
      const double mid = 0.5 * (parent.interval_t0 + parent.interval_t1);
      children->push_back(child(config.first_child_work_item_id, parent.interval_t0, mid));
      children->push_back(child(config.first_child_work_item_id + 1, mid, parent.interval_t1));
    }
  }
  return Status::Ok();
}

CertificateResult CertificateEngine::Evaluate(const ExactWorkItem& work_item) const {
  CertificateEngineConfig config;
  CertificateResult r = BaseResult(work_item, config);
  r.reason_code = kCertificateReasonMissingGeometry;
  return r;
}

Status CertificateEngine::Evaluate(const ExactCertificateQuery& query, CertificateResult* result) const {
  return EvaluateQuery(query, result);
}

// ---- gfx950 batch (certificate_cuda.h:10-26) -----------------------------
bool IsCudaExactBuilt() { return true; }

Status EvaluatePointTriangleBatchCuda(const std::vector<PointTriangleIntervalPrimitive>& primitives,
                                      double t0, double t1, const CertificateEngineConfig& config,
                                      std::vector<PrimitiveIntervalResult>* results) {
  return RunBatch(P2C_FAMILY_POINT_TRIANGLE, primitives, t0, t1, config, results);
}

Status EvaluateEdgeEdgeBatchCuda(const std::vector<EdgeEdgeIntervalPrimitive>& primitives, double t0,
                                 double t1, const CertificateEngineConfig& config,
                                 std::vector<PrimitiveIntervalResult>* results) {
  return RunBatch(P2C_FAMILY_EDGE_EDGE, primitives, t0, t1, config, results);
}

Status CrossCheckCpuCudaExact(const std::vector<PointTriangleIntervalPrimitive>& pts,
                              const std::vector<EdgeEdgeIntervalPrimitive>& ees, double t0, double t1,
                              const CertificateEngineConfig& config, double eps_cert) {
  if (!std::isfinite(eps_cert) || eps_cert < 0.0)
    return Status::Error("eps_cert must be finite and non-negative");
  std::vector<PrimitiveIntervalResult> gpu;
  if (auto s = EvaluatePointTriangleBatchCuda(pts, t0, t1, config, &gpu); !s.ok) return s;
  for (std::size_t i = 0; i < pts.size(); ++i) {
    PrimitiveIntervalResult cpu;
    if (auto s = EvaluatePointTriangleInterval(pts[i], t0, t1, config, &cpu); !s.ok) return s;
    if (auto s = Compare(cpu, gpu[i], eps_cert, "point-triangle CPU/CUDA cross-check"); !s.ok) return s;
  }
  if (auto s = EvaluateEdgeEdgeBatchCuda(ees, t0, t1, config, &gpu); !s.ok) return s;
  for (std::size_t i = 0; i < ees.size(); ++i) {
    PrimitiveIntervalResult cpu;
    if (auto s = EvaluateEdgeEdgeInterval(ees[i], t0, t1, config, &cpu); !s.ok) return s;
    if (auto s = Compare(cpu, gpu[i], eps_cert, "edge-edge CPU/CUDA cross-check"); !s.ok) return s;
  }
  return Status::Ok();
}

}  // namespace p2cccd
