// p2cccd_cpp (compiled as its core, p2cccd_cpp._p2cccd_cpp): the reference's
// pybind11 module name and certificate surface
// (src/cpp/bindings/py_module.cpp:1326-1357, 1358-1431, 1715-1796, 1831-1851,
// 1958-2014), backed by the gfx950 library.  Errors raise ValueError with the
// Status message, as the reference's ThrowIfError does (:38-42).
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "certificate_hip.h"
#include "p2c_hip.h"
#include "proxy_hip.h"

namespace py = pybind11;
using namespace p2cccd;

void BindPolicy(py::module_& m);    // py_policy.cpp
void BindDatasets(py::module_& m);  // datasets_io.cpp
void BindReplay(py::module_& m);    // bp_replay.cpp

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

// Backend names of ParseCandidateBackend (py_module.cpp:87-95) plus this
// build's own ("hip_exact"); every accepted name runs the gfx950 all-pairs
// proxy kernel with the CPU reference's FP64 semantics.
void CheckCandidateBackend(const std::string& name) {
  static const char* const kNames[] = {"cpu_reference", "cpu_reference_rt", "optix", "optix_compatible",
                                       "optix_rt", "hip_exact"};
  for (const char* n : kNames)
    if (name == n) return;
  throw py::value_error("unsupported candidate backend: " + name);
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

// An external query batch given as the reference's Python objects (anything
// with .queries whose items carry query_id, source_query_index, family "vf" /
// "ee", vertices_t0, vertices_t1, ground_truth_collides and box_pair), parsed
// with the reference's rules and messages (py_module.cpp:118-181).
std::array<std::array<double, 3>, 4> ParseVertices(const py::handle& vertices, const char* field) {
  const py::sequence seq = py::cast<py::sequence>(vertices);
  if (py::len(seq) != 4) throw py::value_error(std::string(field) + " must contain four vertices");
  std::array<std::array<double, 3>, 4> out{};
  for (std::size_t i = 0; i < 4; ++i) {
    const py::sequence v = py::cast<py::sequence>(seq[i]);
    if (py::len(v) != 3) throw py::value_error(std::string(field) + " vertices must be 3D");
    for (std::size_t a = 0; a < 3; ++a) out[i][a] = py::cast<double>(v[a]);
  }
  return out;
}

std::vector<ExternalBatchQuery> ParseExternalBatch(const py::object& batch) {
  const py::sequence items = py::cast<py::sequence>(batch.attr("queries"));
  std::vector<ExternalBatchQuery> out;
  out.reserve(py::len(items));
  for (const py::handle& item : items) {
    ExternalBatchQuery q;
    q.source_query_id = py::cast<std::uint64_t>(item.attr("query_id"));
    q.source_query_index = py::cast<std::uint32_t>(item.attr("source_query_index"));
    const std::string fam = py::str(item.attr("family"));
    if (fam == "vf") q.family = ExternalQueryFamily::kVertexFace;
    else if (fam == "ee") q.family = ExternalQueryFamily::kEdgeEdge;
    else throw py::value_error("unsupported external CCD query family: " + fam);
    q.vertices_t0 = ParseVertices(item.attr("vertices_t0"), "vertices_t0");
    q.vertices_t1 = ParseVertices(item.attr("vertices_t1"), "vertices_t1");
    const py::object gt = item.attr("ground_truth_collides");
    if (!gt.is_none()) {
      q.has_ground_truth = true;
      q.ground_truth_collides = py::cast<bool>(gt);
    }
    const py::object pair = item.attr("box_pair");
    if (!pair.is_none()) {
      const py::sequence bp = py::cast<py::sequence>(pair);
      if (py::len(bp) != 2) throw py::value_error("box_pair must contain exactly two object ids");
      q.has_box_pair = true;
      q.box_pair = {py::cast<std::uint32_t>(bp[0]), py::cast<std::uint32_t>(bp[1])};
    }
    out.push_back(q);
  }
  return out;
}

}  // namespace

PYBIND11_MODULE(_p2cccd_cpp, m) {
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
  // ---- proxy-scene candidates (py_module.cpp:1577-1660, 1870-1882) ----
  py::class_<Aabb>(m, "Aabb")
      .def(py::init<>())
      .def_readwrite("min", &Aabb::min)
      .def_readwrite("max", &Aabb::max);
  py::class_<PatchMotionBound>(m, "PatchMotionBound")
      .def(py::init<>())
      .def_readwrite("patch_id", &PatchMotionBound::patch_id)
      .def_readwrite("t0", &PatchMotionBound::t0)
      .def_readwrite("t1", &PatchMotionBound::t1)
      .def_readwrite("center_t0", &PatchMotionBound::center_t0)
      .def_readwrite("center_t1", &PatchMotionBound::center_t1)
      .def_readwrite("translation_bound", &PatchMotionBound::translation_bound)
      .def_readwrite("rotation_angle", &PatchMotionBound::rotation_angle)
      .def_readwrite("center_rotation_bound", &PatchMotionBound::center_rotation_bound)
      .def_readwrite("surface_rotation_bound", &PatchMotionBound::surface_rotation_bound)
      .def_readwrite("radial_motion_bound", &PatchMotionBound::radial_motion_bound)
      .def_readwrite("conservative_radius", &PatchMotionBound::conservative_radius);
  py::class_<CapsuleProxy>(m, "CapsuleProxy")
      .def(py::init<>())
      .def_readwrite("patch_id", &CapsuleProxy::patch_id)
      .def_readwrite("endpoint0", &CapsuleProxy::endpoint0)
      .def_readwrite("endpoint1", &CapsuleProxy::endpoint1)
      .def_readwrite("radius", &CapsuleProxy::radius)
      .def_readwrite("motion_bound", &CapsuleProxy::motion_bound);
  py::class_<ProxyPrimitive>(m, "ProxyPrimitive")
      .def(py::init<>())
      .def_readwrite("proxy_id", &ProxyPrimitive::proxy_id)
      .def_readwrite("object_id", &ProxyPrimitive::object_id)
      .def_readwrite("patch_id", &ProxyPrimitive::patch_id)
      .def_readwrite("slab_id", &ProxyPrimitive::slab_id)
      .def_readwrite("motion_segment_id", &ProxyPrimitive::motion_segment_id)
      .def_readwrite("proxy_type", &ProxyPrimitive::proxy_type)
      .def_readwrite("t0", &ProxyPrimitive::t0)
      .def_readwrite("t1", &ProxyPrimitive::t1)
      .def_readwrite("bounds", &ProxyPrimitive::bounds)
      .def_readwrite("capsule", &ProxyPrimitive::capsule)
      .def_readwrite("motion_bound", &ProxyPrimitive::motion_bound);
  py::class_<ProxyScene>(m, "ProxyScene")
      .def(py::init<>())
      .def_readwrite("query_id", &ProxyScene::query_id)
      .def_readwrite("primitives", &ProxyScene::primitives);
  py::class_<RawCandidateHit>(m, "RawCandidateHit")
      .def(py::init<>())
      .def_readwrite("query_id", &RawCandidateHit::query_id)
      .def_readwrite("pair_key", &RawCandidateHit::pair_key)
      .def_readwrite("proxy_a_index", &RawCandidateHit::proxy_a_index)
      .def_readwrite("proxy_b_index", &RawCandidateHit::proxy_b_index)
      .def_readwrite("slab_id", &RawCandidateHit::slab_id)
      .def_readwrite("rt_hit_count", &RawCandidateHit::rt_hit_count)
      .def_readwrite("flags", &RawCandidateHit::flags)
      .def_readwrite("motion_bound", &RawCandidateHit::motion_bound);
  py::class_<RawCandidateBuffer>(m, "RawCandidateBuffer")
      .def(py::init<>())
      .def_readwrite("hits", &RawCandidateBuffer::hits);
  py::class_<CandidateRecord>(m, "CandidateRecord")
      .def(py::init<>())
      .def_readwrite("schema_version", &CandidateRecord::schema_version)
      .def_readwrite("candidate_id", &CandidateRecord::candidate_id)
      .def_readwrite("query_id", &CandidateRecord::query_id)
      .def_readwrite("slab_id", &CandidateRecord::slab_id)
      .def_readwrite("object_a_id", &CandidateRecord::object_a_id)
      .def_readwrite("patch_a_id", &CandidateRecord::patch_a_id)
      .def_readwrite("object_b_id", &CandidateRecord::object_b_id)
      .def_readwrite("patch_b_id", &CandidateRecord::patch_b_id)
      .def_readwrite("proxy_type_a", &CandidateRecord::proxy_type_a)
      .def_readwrite("proxy_type_b", &CandidateRecord::proxy_type_b)
      .def_readwrite("rt_hit_count", &CandidateRecord::rt_hit_count)
      .def_readwrite("motion_bound", &CandidateRecord::motion_bound)
      .def_readwrite("proxy_features_offset", &CandidateRecord::proxy_features_offset)
      .def_readwrite("flags", &CandidateRecord::flags);
  py::class_<RtCandidateTiming>(m, "RtCandidateTiming")
      .def(py::init<>())
      .def_readwrite("build_ms", &RtCandidateTiming::build_ms)
      .def_readwrite("update_ms", &RtCandidateTiming::update_ms)
      .def_readwrite("trace_ms", &RtCandidateTiming::trace_ms)
      .def_readwrite("compact_ms", &RtCandidateTiming::compact_ms)
      .def_readwrite("stats_ms", &RtCandidateTiming::stats_ms)
      .def_readwrite("total_ms", &RtCandidateTiming::total_ms);
  py::class_<CandidateDensityStats>(m, "CandidateDensityStats")
      .def(py::init<>())
      .def_readwrite("schema_version", &CandidateDensityStats::schema_version)
      .def_readwrite("query_id", &CandidateDensityStats::query_id)
      .def_readwrite("proxy_count", &CandidateDensityStats::proxy_count)
      .def_readwrite("object_count", &CandidateDensityStats::object_count)
      .def_readwrite("slab_count", &CandidateDensityStats::slab_count)
      .def_readwrite("cross_object_same_slab_pair_count", &CandidateDensityStats::cross_object_same_slab_pair_count)
      .def_readwrite("raw_hit_count", &CandidateDensityStats::raw_hit_count)
      .def_readwrite("compact_candidate_count", &CandidateDensityStats::compact_candidate_count)
      .def_readwrite("raw_hits_per_proxy", &CandidateDensityStats::raw_hits_per_proxy)
      .def_readwrite("candidates_per_proxy", &CandidateDensityStats::candidates_per_proxy)
      .def_readwrite("candidates_per_slab", &CandidateDensityStats::candidates_per_slab)
      .def_readwrite("aabb_overlap_ratio", &CandidateDensityStats::aabb_overlap_ratio)
      .def_readwrite("avg_rt_hits_per_candidate", &CandidateDensityStats::avg_rt_hits_per_candidate)
      .def_readwrite("timing", &CandidateDensityStats::timing)
      .def_readwrite("backend_name", &CandidateDensityStats::backend_name);
  py::class_<CandidateGenerationResult>(m, "CandidateGenerationResult")
      .def(py::init<>())
      .def_readwrite("backend_name", &CandidateGenerationResult::backend_name)
      .def_readwrite("raw_buffer", &CandidateGenerationResult::raw_buffer)
      .def_readwrite("candidates", &CandidateGenerationResult::candidates)
      .def_readwrite("timing", &CandidateGenerationResult::timing)
      .def_readwrite("density", &CandidateGenerationResult::density);
  py::class_<PoseSample>(m, "PoseSample")
      .def(py::init<>())
      .def_readwrite("translation", &PoseSample::translation)
      .def_readwrite("rotation_xyzw", &PoseSample::rotation_xyzw);
  py::class_<MotionSegment>(m, "MotionSegment")
      .def(py::init<>())
      .def_readwrite("t0", &MotionSegment::t0)
      .def_readwrite("t1", &MotionSegment::t1)
      .def_readwrite("pose_t0", &MotionSegment::pose_t0)
      .def_readwrite("pose_t1", &MotionSegment::pose_t1);
  py::class_<Patch>(m, "Patch")
      .def(py::init<>())
      .def_readwrite("patch_id", &Patch::patch_id)
      .def_readwrite("triangle_ids", &Patch::triangle_ids)
      .def_readwrite("triangle_count", &Patch::triangle_count)
      .def_readwrite("area", &Patch::area)
      .def_readwrite("local_center", &Patch::local_center)
      .def_readwrite("radius", &Patch::radius);
  py::class_<ProxyObjectBuildInput>(m, "ProxyObjectBuildInput")
      .def(py::init<>())
      .def_readwrite("object_id", &ProxyObjectBuildInput::object_id)
      .def_readwrite("proxy_type", &ProxyObjectBuildInput::proxy_type)
      .def_readwrite("patches", &ProxyObjectBuildInput::patches)
      .def_readwrite("motion_segments", &ProxyObjectBuildInput::motion_segments)
      .def_readwrite("slabs_per_motion_segment", &ProxyObjectBuildInput::slabs_per_motion_segment)
      .def_readwrite("eps_proxy", &ProxyObjectBuildInput::eps_proxy);
  py::class_<ProxySceneBuildInput>(m, "ProxySceneBuildInput")
      .def(py::init<>())
      .def_readwrite("query_id", &ProxySceneBuildInput::query_id)
      .def_readwrite("objects", &ProxySceneBuildInput::objects);
  py::class_<RuntimeQueryIdMapping>(m, "RuntimeQueryIdMapping")
      .def(py::init<>())
      .def_readwrite("source_query_id", &RuntimeQueryIdMapping::source_query_id)
      .def_readwrite("runtime_query_id", &RuntimeQueryIdMapping::runtime_query_id);
  py::class_<ExternalBatchCandidateResult>(m, "ExternalBatchCandidateResult")
      .def(py::init<>())
      .def_readwrite("backend_name", &ExternalBatchCandidateResult::backend_name)
      .def_readwrite("timing", &ExternalBatchCandidateResult::timing)
      .def_readwrite("primitive_count", &ExternalBatchCandidateResult::primitive_count)
      .def_readwrite("raw_hit_count", &ExternalBatchCandidateResult::raw_hit_count)
      .def_readwrite("compact_candidate_count", &ExternalBatchCandidateResult::compact_candidate_count)
      .def_readwrite("candidate_recall", &ExternalBatchCandidateResult::candidate_recall)
      .def_readwrite("candidates", &ExternalBatchCandidateResult::candidates)
      .def_readwrite("runtime_query_ids", &ExternalBatchCandidateResult::runtime_query_ids);
  m.def("build_proxy_scene",
        [](const ProxySceneBuildInput& input) {
          ProxyScene scene;
          ThrowIfError(BuildProxyScene(input, &scene));
          return scene;
        },
        py::arg("input"));
  m.def("generate_candidates_for_external_batch",
        [](const py::object& batch, const std::string& backend_name, bool allow_optix_cpu_fallback) {
          CheckCandidateBackend(backend_name);
          CandidateGeneratorConfig config;
          config.backend = backend_name.rfind("optix", 0) == 0 ? CandidateBackend::kOptix : CandidateBackend::kCpuReference;
          config.allow_optix_cpu_fallback = allow_optix_cpu_fallback;
          const std::vector<ExternalBatchQuery> queries = ParseExternalBatch(batch);
          ExternalBatchCandidateResult result;
          Status s;
          {
            py::gil_scoped_release release;
            s = GenerateCandidatesForExternalBatch(queries, config, &result);
          }
          ThrowIfError(s);
          return result;
        },
        py::arg("batch"), py::arg("backend_name") = "cpu_reference", py::arg("allow_optix_cpu_fallback") = false);
  m.def("validate_proxy_scene",
        [](const ProxyScene& scene) {
          ThrowIfError(ValidateProxyScene(scene));
          return true;
        },
        py::arg("scene"));
  m.def("generate_candidates_for_proxy_scene",
        [](const ProxyScene& scene, std::uint64_t query_id, const std::string& backend_name,
           bool allow_optix_cpu_fallback) {
          (void)allow_optix_cpu_fallback;  // no fallback exists: every backend name runs the device path
          CheckCandidateBackend(backend_name);
          CandidateGenerationResult result;
          Status s;
          {
            py::gil_scoped_release release;
            s = GenerateCandidatesHip(scene, query_id == 0 ? scene.query_id : query_id, &result);
          }
          ThrowIfError(s);
          return result;
        },
        py::arg("scene"), py::arg("query_id") = 0, py::arg("backend_name") = "cpu_reference",
        py::arg("allow_optix_cpu_fallback") = false);
  m.def("generate_raw_candidates_cpu",
        [](const ProxyScene& scene, std::uint64_t query_id) {
          // The reference's name; the hits come from the device path, identical
          // to the CPU loop's in content and order.
          CandidateGenerationResult result;
          Status s;
          {
            py::gil_scoped_release release;
            s = GenerateCandidatesHip(scene, query_id == 0 ? scene.query_id : query_id, &result);
          }
          ThrowIfError(s);
          return result.raw_buffer;
        },
        py::arg("scene"), py::arg("query_id") = 0);
  m.def("compact_raw_candidates",
        [](const ProxyScene& scene, const RawCandidateBuffer& raw_buffer) {
          std::vector<CandidateRecord> out;
          Status s;
          {
            py::gil_scoped_release release;
            s = CompactRawCandidatesHip(scene, raw_buffer, &out);
          }
          ThrowIfError(s);
          return out;
        },
        py::arg("scene"), py::arg("raw_buffer"));
  BindPolicy(m);
  BindDatasets(m);
  BindReplay(m);
  m.def("audit_log_rows_for_query",
        [](const std::vector<AuditLogRow>& rows, std::uint64_t query_id) {
          std::vector<AuditLogRow> out;
          for (const auto& r : rows)
            if (r.query_id == query_id) out.push_back(r);
          return out;
        },
        py::arg("rows"), py::arg("query_id"));
}
