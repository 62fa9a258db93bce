// C++ certificate interface of the reference, kept signature-for-signature so
// existing C++ callers drop in, implemented over the gfx950 C-ABI
// (include/p2c_hip.h).  Layouts follow
//   src/cpp/common/status.h:7-19, src/cpp/common/runtime_contracts.h:8-114,
//   src/cpp/certificate/certificate_engine.h:12-127,
//   src/cpp/certificate/certificate_cuda.h:10-26.
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace p2cccd {

struct Status {
  bool ok = true;
  std::string message;
  static Status Ok() { return Status{}; }
  static Status Error(std::string msg) {
    Status s;
    s.ok = false;
    s.message = std::move(msg);
    return s;
  }
};

enum class ProxyType : std::uint8_t { kUnknown = 0, kSweptAabb = 1, kCapsule = 2 };
enum class ProposalSource : std::uint8_t { kRaw = 0, kRefined = 1, kFallback = 2 };
enum class CertificateStatus : std::uint8_t { kCollision = 0, kSeparation = 1, kUndecided = 2 };
enum class CertificateRefinementMode : std::uint8_t {
  kNone = 0,
  kBisectInterval = 1,
  kRequestGeometry = 2,
  kEscalatePrecision = 3,
};
enum class AuditStage : std::uint8_t { kRt = 0, kProposal = 1, kExact = 2, kRefine = 3, kCertify = 4 };

enum ExactFeatureFamilyMask : std::uint32_t {
  kFeatureFamilyPointTriangle = 1U << 0U,
  kFeatureFamilyEdgeEdge = 1U << 1U,
};
enum CertificateReasonCode : std::uint16_t {
  kCertificateReasonNone = 0,
  kCertificateReasonMissingGeometry = 1,
  kCertificateReasonMaxSubdivisionDepth = 2,
  kCertificateReasonInvalidInput = 3,
};
enum ExactAuditAction : std::uint16_t {
  kExactAuditDequeued = 1,
  kExactAuditCollision = 2,
  kExactAuditSeparation = 3,
  kExactAuditUndecided = 4,
  kExactAuditInvalidInput = 5,
};

struct ExactWorkItem {
  std::uint64_t work_item_id = 0;
  std::uint64_t parent_candidate_id = 0;
  std::uint64_t query_id = 0;
  std::uint32_t slab_id = 0;
  std::uint32_t patch_a_id = 0;
  std::uint32_t patch_b_id = 0;
  double interval_t0 = 0.0;
  double interval_t1 = 1.0;
  std::uint32_t feature_family_mask = 0;
  std::uint32_t topk_feature_ids_offset = 0;
  std::uint16_t depth = 0;
  float priority_score = 0.0f;
  ProposalSource source = ProposalSource::kRaw;
};

struct CertificateResult {
  std::uint64_t work_item_id = 0;
  std::uint64_t query_id = 0;
  CertificateStatus status = CertificateStatus::kUndecided;
  double interval_t0 = 0.0;
  double interval_t1 = 1.0;
  double toi_upper = 1.0;
  double safe_margin_lb = 0.0;
  std::uint8_t witness_family = 0;
  std::int64_t witness_id_a = -1;
  std::int64_t witness_id_b = -1;
  std::uint32_t covered_feature_mask = 0;
  double eps_time = 1.0e-4;
  double eps_space = 1.0e-6;
  std::uint16_t reason_code = 0;
  CertificateRefinementMode next_refinement_mode = CertificateRefinementMode::kNone;
};

struct AuditLogRow {
  std::uint64_t event_id = 0;
  std::uint64_t query_id = 0;
  std::uint64_t candidate_id = 0;
  std::uint64_t work_item_id = 0;
  AuditStage stage = AuditStage::kRt;
  std::uint16_t action = 0;
  std::uint16_t depth = 0;
  double interval_t0 = 0.0;
  double interval_t1 = 1.0;
  std::uint64_t timestamp_us = 0;
  double aux_value0 = 0.0;
  double aux_value1 = 0.0;
};

struct LinearVertexTrajectory {
  std::int64_t feature_id = -1;
  std::array<double, 3> position_t0{0.0, 0.0, 0.0};
  std::array<double, 3> position_t1{0.0, 0.0, 0.0};
};

struct PointTriangleIntervalPrimitive {
  std::int64_t point_id = -1;
  std::int64_t triangle_id = -1;
  LinearVertexTrajectory point;
  LinearVertexTrajectory triangle_v0;
  LinearVertexTrajectory triangle_v1;
  LinearVertexTrajectory triangle_v2;
};

struct EdgeEdgeIntervalPrimitive {
  std::int64_t edge_a_id = -1;
  std::int64_t edge_b_id = -1;
  LinearVertexTrajectory edge_a0;
  LinearVertexTrajectory edge_a1;
  LinearVertexTrajectory edge_b0;
  LinearVertexTrajectory edge_b1;
};

struct CertificateEngineConfig {
  double eps_time = 1.0e-4;
  double eps_space = 1.0e-6;
  std::uint16_t max_subdivision_depth = 32;
};

struct ExactCertificateQuery {
  ExactWorkItem work_item;
  CertificateEngineConfig config;
  std::vector<PointTriangleIntervalPrimitive> point_triangle_primitives;
  std::vector<EdgeEdgeIntervalPrimitive> edge_edge_primitives;
};

struct PrimitiveIntervalResult {
  CertificateStatus status = CertificateStatus::kUndecided;
  std::uint32_t covered_feature_mask = 0;
  double interval_t0 = 0.0;
  double interval_t1 = 1.0;
  double toi_upper = 1.0;
  double safe_margin_lb = 0.0;
  std::uint8_t witness_family = 0;
  std::int64_t witness_id_a = -1;
  std::int64_t witness_id_b = -1;
  std::uint16_t reason_code = kCertificateReasonNone;
  CertificateRefinementMode next_refinement_mode = CertificateRefinementMode::kNone;
};

struct ExactWorkQueueConfig {
  std::uint64_t first_event_id = 1;
  std::uint64_t first_timestamp_us = 1;
  bool emit_dequeue_events = true;
};

struct ExactWorkQueueResult {
  std::vector<CertificateResult> certificates;
  std::vector<AuditLogRow> audit_log;
  std::uint64_t processed_count = 0;
};

struct ExactRefinementConfig {
  std::uint64_t first_child_work_item_id = 1;
  std::uint16_t max_child_depth = 64;
  double min_interval_width = 1.0e-4;
};

// validators.h:8-13
Status ValidateExactWorkItem(const ExactWorkItem& item);
Status ValidateCertificateResult(const CertificateResult& result);
Status ValidateAuditLogRow(const AuditLogRow& row);

// certificate_engine.h:101-127 (primitive certificates run on the host CPU
// from the same source the gfx950 kernel is built from)
Status EvaluatePointTriangleInterval(const PointTriangleIntervalPrimitive& primitive,
                                     double interval_t0, double interval_t1,
                                     const CertificateEngineConfig& config,
                                     PrimitiveIntervalResult* result);
Status EvaluateEdgeEdgeInterval(const EdgeEdgeIntervalPrimitive& primitive, double interval_t0,
                                double interval_t1, const CertificateEngineConfig& config,
                                PrimitiveIntervalResult* result);
// Work queue: every primitive of every queued query is certified in batched
// gfx950 launches, then aggregated with the engine's rules on the host.
Status ProcessExactWorkQueueCpu(const std::vector<ExactCertificateQuery>& work_queue,
                                const ExactWorkQueueConfig& config, ExactWorkQueueResult* result);
Status ValidateExactWorkQueueCoverage(const std::vector<ExactCertificateQuery>& work_queue,
                                      const ExactWorkQueueResult& result);
Status GenerateConservativeRefinementWorkItems(const ExactWorkItem& parent,
                                               const CertificateResult& certificate,
                                               const ExactRefinementConfig& config,
                                               std::vector<ExactWorkItem>* children);

class CertificateEngine {
 public:
  CertificateResult Evaluate(const ExactWorkItem& work_item) const;
  Status Evaluate(const ExactCertificateQuery& query, CertificateResult* result) const;
};

// certificate_cuda.h:10-26 (the gfx950 batch)
bool IsCudaExactBuilt();
Status EvaluatePointTriangleBatchCuda(const std::vector<PointTriangleIntervalPrimitive>& primitives,
                                      double interval_t0, double interval_t1,
                                      const CertificateEngineConfig& config,
                                      std::vector<PrimitiveIntervalResult>* results);
Status EvaluateEdgeEdgeBatchCuda(const std::vector<EdgeEdgeIntervalPrimitive>& primitives,
                                 double interval_t0, double interval_t1,
                                 const CertificateEngineConfig& config,
                                 std::vector<PrimitiveIntervalResult>* results);
Status CrossCheckCpuCudaExact(const std::vector<PointTriangleIntervalPrimitive>& point_triangles,
                              const std::vector<EdgeEdgeIntervalPrimitive>& edge_edges,
                              double interval_t0, double interval_t1,
                              const CertificateEngineConfig& config, double eps_cert);

}  // namespace p2cccd
