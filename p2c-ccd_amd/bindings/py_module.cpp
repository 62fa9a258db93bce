// p2cccd_cpp: the reference's pybind11 module name and certificate surface
// (src/cpp/bindings/py_module.cpp:1326-1357, 1358-1431, 1715-1796, 1831-1851,
// 1958-2014), backed by the gfx950 library.  Errors raise ValueError with the
// Status message, as the reference's ThrowIfError does (:38-42).
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "certificate_hip.h"
#include "p2c_hip.h"

namespace py = pybind11;
using namespace p2cccd;

namespace {

void ThrowIfError(const Status& s) {
  if (!s.ok) throw py::value_error(s.message);
}

template <typename Prim>
std::vector<PrimitiveIntervalResult> BatchOrThrow(
    Status (*fn)(const std::vector<Prim>&, double, double, const CertificateEngineConfig&,
                 std::vector<PrimitiveIntervalResult>*),
    const std::vector<Prim>& prims, double t0, double t1, const CertificateEngineConfig& cfg) {
  std::vector<PrimitiveIntervalResult> out;
  Status s;
  {
    py::gil_scoped_release release;
    s = fn(prims, t0, t1, cfg, &out);
  }
  ThrowIfError(s);
  return out;
}

py::dict BindingStatus() {
  py::dict d;
  d["cuda_exact_built"] = py::bool_(IsCudaExactBuilt());
  d["host_batch_exact_api"] = py::bool_(true);
  d["device_pointer_abi"] = py::bool_(true);
  d["backend_name"] = py::str("hip_exact_gfx950");
  d["device_available"] = py::bool_(p2c_device_available() != 0);
  d["safety_policy"] = py::str(
      "Host-owned batch APIs copy in and out as before; the device-pointer C-ABI "
      "(libp2c_hip) takes caller-owned device buffers and a caller stream and never "
      "allocates on the hot path.");
  return d;
}

}  // namespace

PYBIND11_MODULE(p2cccd_cpp, m) {
  m.doc() = "P2CCCD certificate bindings backed by hand-written gfx950 (MI355X) kernels.";

  py::enum_<ProxyType>(m, "ProxyType")
      .value("UNKNOWN", ProxyType::kUnknown)
      .value("SWEPT_AABB", ProxyType::kSweptAabb)
      .value("CAPSULE", ProxyType::kCapsule);
  py::enum_<ProposalSource>(m, "ProposalSource")
      .value("RAW", ProposalSource::kRaw)
      .value("REFINED", ProposalSource::kRefined)
      .value("FALLBACK", ProposalSource::kFallback);
  py::enum_<CertificateStatus>(m, "CertificateStatus")
      .value("COLLISION", CertificateStatus::kCollision)
      .value("SEPARATION", CertificateStatus::kSeparation)
      .value("UNDECIDED", CertificateStatus::kUndecided)
      .value("Collision", CertificateStatus::kCollision)
      .value("Separation", CertificateStatus::kSeparation)
      .value("Undecided", CertificateStatus::kUndecided);
  py::enum_<CertificateRefinementMode>(m, "CertificateRefinementMode")
      .value("NONE", CertificateRefinementMode::kNone)
      .value("BISECT_INTERVAL", CertificateRefinementMode::kBisectInterval)
      .value("REQUEST_GEOMETRY", CertificateRefinementMode::kRequestGeometry)
      .value("ESCALATE_PRECISION", CertificateRefinementMode::kEscalatePrecision);
  py::enum_<AuditStage>(m, "AuditStage")
      .value("RT", AuditStage::kRt)
      .value("PROPOSAL", AuditStage::kProposal)
      .value("EXACT", AuditStage::kExact)
      .value("REFINE", AuditStage::kRefine)
      .value("CERTIFY", AuditStage::kCertify);

  py::class_<ExactWorkItem>(m, "ExactWorkItem")
      .def(py::init<>())
      .def_readwrite("work_item_id", &ExactWorkItem::work_item_id)
      .def_readwrite("parent_candidate_id", &ExactWorkItem::parent_candidate_id)
      .def_readwrite("query_id", &ExactWorkItem::query_id)
      .def_readwrite("slab_id", &ExactWorkItem::slab_id)
      .def_readwrite("patch_a_id", &ExactWorkItem::patch_a_id)
      .def_readwrite("patch_b_id", &ExactWorkItem::patch_b_id)
      .def_readwrite("interval_t0", &ExactWorkItem::interval_t0)
      .def_readwrite("interval_t1", &ExactWorkItem::interval_t1)
      .def_readwrite("feature_family_mask", &ExactWorkItem::feature_family_mask)
      .def_readwrite("topk_feature_ids_offset", &ExactWorkItem::topk_feature_ids_offset)
      .def_readwrite("depth", &ExactWorkItem::depth)
      .def_readwrite("priority_score", &ExactWorkItem::priority_score)
      .def_readwrite("source", &ExactWorkItem::source);

  py::class_<CertificateResult>(m, "CertificateResult")
      .def(py::init<>())
      .def_readwrite("work_item_id", &CertificateResult::work_item_id)
      .def_readwrite("query_id", &CertificateResult::query_id)
      .def_readwrite("status", &CertificateResult::status)
      .def_readwrite("interval_t0", &CertificateResult::interval_t0)
      .def_readwrite("interval_t1", &CertificateResult::interval_t1)
      .def_readwrite("toi_upper", &CertificateResult::toi_upper)
      .def_readwrite("safe_margin_lb", &CertificateResult::safe_margin_lb)
      .def_readwrite("witness_family", &CertificateResult::witness_family)
      .def_readwrite("witness_id_a", &CertificateResult::witness_id_a)
      .def_readwrite("witness_id_b", &CertificateResult::witness_id_b)
      .def_readwrite("covered_feature_mask", &CertificateResult::covered_feature_mask)
      .def_readwrite("eps_time", &CertificateResult::eps_time)
      .def_readwrite("eps_space", &CertificateResult::eps_space)
      .def_readwrite("reason_code", &CertificateResult::reason_code)
      .def_readwrite("next_refinement_mode", &CertificateResult::next_refinement_mode);

  py::class_<AuditLogRow>(m, "AuditLogRow")
      .def(py::init<>())
      .def_readwrite("event_id", &AuditLogRow::event_id)
      .def_readwrite("query_id", &AuditLogRow::query_id)
      .def_readwrite("candidate_id", &AuditLogRow::candidate_id)
      .def_readwrite("work_item_id", &AuditLogRow::work_item_id)
      .def_readwrite("stage", &AuditLogRow::stage)
      .def_readwrite("action", &AuditLogRow::action)
      .def_readwrite("depth", &AuditLogRow::depth)
      .def_readwrite("interval_t0", &AuditLogRow::interval_t0)
      .def_readwrite("interval_t1", &AuditLogRow::interval_t1)
      .def_readwrite("timestamp_us", &AuditLogRow::timestamp_us)
      .def_readwrite("aux_value0", &AuditLogRow::aux_value0)
      .def_readwrite("aux_value1", &AuditLogRow::aux_value1);

  py::class_<LinearVertexTrajectory>(m, "LinearVertexTrajectory")
      .def(py::init<>())
      .def_readwrite("feature_id", &LinearVertexTrajectory::feature_id)
      .def_readwrite("position_t0", &LinearVertexTrajectory::position_t0)
      .def_readwrite("position_t1", &LinearVertexTrajectory::position_t1);

  py::class_<PointTriangleIntervalPrimitive>(m, "PointTriangleIntervalPrimitive")
      .def(py::init<>())
      .def_readwrite("point_id", &PointTriangleIntervalPrimitive::point_id)
      .def_readwrite("triangle_id", &PointTriangleIntervalPrimitive::triangle_id)
      .def_readwrite("point", &PointTriangleIntervalPrimitive::point)
      .def_readwrite("triangle_v0", &PointTriangleIntervalPrimitive::triangle_v0)
      .def_readwrite("triangle_v1", &PointTriangleIntervalPrimitive::triangle_v1)
      .def_readwrite("triangle_v2", &PointTriangleIntervalPrimitive::triangle_v2);

  py::class_<EdgeEdgeIntervalPrimitive>(m, "EdgeEdgeIntervalPrimitive")
      .def(py::init<>())
      .def_readwrite("edge_a_id", &EdgeEdgeIntervalPrimitive::edge_a_id)
      .def_readwrite("edge_b_id", &EdgeEdgeIntervalPrimitive::edge_b_id)
      .def_readwrite("edge_a0", &EdgeEdgeIntervalPrimitive::edge_a0)
      .def_readwrite("edge_a1", &EdgeEdgeIntervalPrimitive::edge_a1)
      .def_readwrite("edge_b0", &EdgeEdgeIntervalPrimitive::edge_b0)
      .def_readwrite("edge_b1", &EdgeEdgeIntervalPrimitive::edge_b1);

  py::class_<CertificateEngineConfig>(m, "CertificateEngineConfig")
      .def(py::init<>())
      .def_readwrite("eps_time", &CertificateEngineConfig::eps_time)
      .def_readwrite("eps_space", &CertificateEngineConfig::eps_space)
      .def_readwrite("max_subdivision_depth", &CertificateEngineConfig::max_subdivision_depth);

  py::class_<ExactCertificateQuery>(m, "ExactCertificateQuery")
      .def(py::init<>())
      .def_readwrite("work_item", &ExactCertificateQuery::work_item)
      .def_readwrite("config", &ExactCertificateQuery::config)
      .def_readwrite("point_triangle_primitives", &ExactCertificateQuery::point_triangle_primitives)
      .def_readwrite("edge_edge_primitives", &ExactCertificateQuery::edge_edge_primitives);

  py::class_<PrimitiveIntervalResult>(m, "PrimitiveIntervalResult")
      .def(py::init<>())
      .def_readwrite("status", &PrimitiveIntervalResult::status)
      .def_readwrite("covered_feature_mask", &PrimitiveIntervalResult::covered_feature_mask)
      .def_readwrite("interval_t0", &PrimitiveIntervalResult::interval_t0)
      .def_readwrite("interval_t1", &PrimitiveIntervalResult::interval_t1)
      .def_readwrite("toi_upper", &PrimitiveIntervalResult::toi_upper)
      .def_readwrite("safe_margin_lb", &PrimitiveIntervalResult::safe_margin_lb)
      .def_readwrite("witness_family", &PrimitiveIntervalResult::witness_family)
      .def_readwrite("witness_id_a", &PrimitiveIntervalResult::witness_id_a)
      .def_readwrite("witness_id_b", &PrimitiveIntervalResult::witness_id_b)
      .def_readwrite("reason_code", &PrimitiveIntervalResult::reason_code)
      .def_readwrite("next_refinement_mode", &PrimitiveIntervalResult::next_refinement_mode);

  py::class_<ExactWorkQueueConfig>(m, "ExactWorkQueueConfig")
      .def(py::init<>())
      .def_readwrite("first_event_id", &ExactWorkQueueConfig::first_event_id)
      .def_readwrite("first_timestamp_us", &ExactWorkQueueConfig::first_timestamp_us)
      .def_readwrite("emit_dequeue_events", &ExactWorkQueueConfig::emit_dequeue_events);

  py::class_<ExactWorkQueueResult>(m, "ExactWorkQueueResult")
      .def(py::init<>())
      .def_readwrite("certificates", &ExactWorkQueueResult::certificates)
      .def_readwrite("audit_log", &ExactWorkQueueResult::audit_log)
      .def_readwrite("processed_count", &ExactWorkQueueResult::processed_count);

  py::class_<ExactRefinementConfig>(m, "ExactRefinementConfig")
      .def(py::init<>())
      .def_readwrite("first_child_work_item_id", &ExactRefinementConfig::first_child_work_item_id)
      .def_readwrite("max_child_depth", &ExactRefinementConfig::max_child_depth)
      .def_readwrite("min_interval_width", &ExactRefinementConfig::min_interval_width);

  py::class_<CertificateEngine>(m, "CertificateEngine")
      .def(py::init<>())
      .def("evaluate_work_item",
           static_cast<CertificateResult (CertificateEngine::*)(const ExactWorkItem&) const>(
               &CertificateEngine::Evaluate))
      .def("evaluate_query", [](const CertificateEngine& engine, const ExactCertificateQuery& q) {
        CertificateResult r;
        ThrowIfError(engine.Evaluate(q, &r));
        return r;
      });

  m.attr("FEATURE_FAMILY_POINT_TRIANGLE") = py::int_(static_cast<std::uint32_t>(kFeatureFamilyPointTriangle));
  m.attr("FEATURE_FAMILY_EDGE_EDGE") = py::int_(static_cast<std::uint32_t>(kFeatureFamilyEdgeEdge));
  m.attr("CERTIFICATE_REASON_NONE") = py::int_(static_cast<std::uint16_t>(kCertificateReasonNone));
  m.attr("CERTIFICATE_REASON_MISSING_GEOMETRY") =
      py::int_(static_cast<std::uint16_t>(kCertificateReasonMissingGeometry));
  m.attr("CERTIFICATE_REASON_MAX_SUBDIVISION_DEPTH") =
      py::int_(static_cast<std::uint16_t>(kCertificateReasonMaxSubdivisionDepth));
  m.attr("CERTIFICATE_REASON_INVALID_INPUT") =
      py::int_(static_cast<std::uint16_t>(kCertificateReasonInvalidInput));
  m.attr("EXACT_AUDIT_DEQUEUED") = py::int_(static_cast<std::uint16_t>(kExactAuditDequeued));
  m.attr("EXACT_AUDIT_COLLISION") = py::int_(static_cast<std::uint16_t>(kExactAuditCollision));
  m.attr("EXACT_AUDIT_SEPARATION") = py::int_(static_cast<std::uint16_t>(kExactAuditSeparation));
  m.attr("EXACT_AUDIT_UNDECIDED") = py::int_(static_cast<std::uint16_t>(kExactAuditUndecided));
  m.attr("EXACT_AUDIT_INVALID_INPUT") = py::int_(static_cast<std::uint16_t>(kExactAuditInvalidInput));
  m.attr("CUDA_DEVICE_POINTER_ABI_ENABLED") = py::bool_(true);

  m.def("validate_exact_work_item", [](const ExactWorkItem& i) { ThrowIfError(ValidateExactWorkItem(i)); return true; },
        py::arg("item"));
  m.def("validate_certificate_result",
        [](const CertificateResult& r) { ThrowIfError(ValidateCertificateResult(r)); return true; },
        py::arg("result"));
  m.def("validate_audit_log_row",
        [](const AuditLogRow& r) { ThrowIfError(ValidateAuditLogRow(r)); return true; }, py::arg("row"));

  m.def("evaluate_point_triangle_interval",
        [](const PointTriangleIntervalPrimitive& p, double t0, double t1, const CertificateEngineConfig& c) {
          PrimitiveIntervalResult r;
          ThrowIfError(EvaluatePointTriangleInterval(p, t0, t1, c, &r));
          return r;
        },
        py::arg("primitive"), py::arg("interval_t0"), py::arg("interval_t1"), py::arg("config"));
  m.def("evaluate_edge_edge_interval",
        [](const EdgeEdgeIntervalPrimitive& p, double t0, double t1, const CertificateEngineConfig& c) {
          PrimitiveIntervalResult r;
          ThrowIfError(EvaluateEdgeEdgeInterval(p, t0, t1, c, &r));
          return r;
        },
        py::arg("primitive"), py::arg("interval_t0"), py::arg("interval_t1"), py::arg("config"));
  m.def("is_cuda_exact_built", &IsCudaExactBuilt);
  m.def("cuda_binding_status", &BindingStatus);
  m.def("evaluate_point_triangle_batch_cuda",
        [](const std::vector<PointTriangleIntervalPrimitive>& p, double t0, double t1,
           const CertificateEngineConfig& c) { return BatchOrThrow(&EvaluatePointTriangleBatchCuda, p, t0, t1, c); },
        py::arg("primitives"), py::arg("interval_t0"), py::arg("interval_t1"), py::arg("config"));
  m.def("evaluate_edge_edge_batch_cuda",
        [](const std::vector<EdgeEdgeIntervalPrimitive>& p, double t0, double t1,
           const CertificateEngineConfig& c) { return BatchOrThrow(&EvaluateEdgeEdgeBatchCuda, p, t0, t1, c); },
        py::arg("primitives"), py::arg("interval_t0"), py::arg("interval_t1"), py::arg("config"));
  m.def("cross_check_cpu_cuda_exact",
        [](const std::vector<PointTriangleIntervalPrimitive>& pts,
           const std::vector<EdgeEdgeIntervalPrimitive>& ees, double t0, double t1,
           const CertificateEngineConfig& c, double eps) {
          Status s;
          {
            py::gil_scoped_release release;
            s = CrossCheckCpuCudaExact(pts, ees, t0, t1, c, eps);
          }
          ThrowIfError(s);
          return true;
        },
        py::arg("point_triangles"), py::arg("edge_edges"), py::arg("interval_t0"),
        py::arg("interval_t1"), py::arg("config"), py::arg("eps_cert"));
  m.def("evaluate_certificate_query_cpu",
        [](const ExactCertificateQuery& q) {
          CertificateEngine e;
          CertificateResult r;
          ThrowIfError(e.Evaluate(q, &r));
          return r;
        },
        py::arg("query"));
  m.def("process_exact_work_queue_cpu",
        [](const std::vector<ExactCertificateQuery>& queue, const ExactWorkQueueConfig& cfg) {
          ExactWorkQueueResult r;
          Status s;
          {
            py::gil_scoped_release release;
            s = ProcessExactWorkQueueCpu(queue, cfg, &r);
          }
          ThrowIfError(s);
          return r;
        },
        py::arg("work_queue"), py::arg("config"));
  m.def("validate_exact_work_queue_coverage",
        [](const std::vector<ExactCertificateQuery>& queue, const ExactWorkQueueResult& r) {
          ThrowIfError(ValidateExactWorkQueueCoverage(queue, r));
          return true;
        },
        py::arg("work_queue"), py::arg("result"));
  m.def("generate_conservative_refinement_work_items",
        [](const ExactWorkItem& parent, const CertificateResult& cert, const ExactRefinementConfig& cfg) {
          std::vector<ExactWorkItem> children;
          ThrowIfError(GenerateConservativeRefinementWorkItems(parent, cert, cfg, &children));
          return children;
        },
        py::arg("parent"), py::arg("certificate"), py::arg("config"));
  m.def("validate_audit_log_rows",
        [](const std::vector<AuditLogRow>& rows) {
          for (const auto& r : rows) ThrowIfError(ValidateAuditLogRow(r));
          return true;
        },
        py::arg("rows"));
  m.def("audit_log_rows_for_query",
        [](const std::vector<AuditLogRow>& rows, std::uint64_t query_id) {
          std::vector<AuditLogRow> out;
          for (const auto& r : rows)
            if (r.query_id == query_id) out.push_back(r);
          return out;
        },
        py::arg("rows"), py::arg("query_id"));
}
