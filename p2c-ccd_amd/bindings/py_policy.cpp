// p2cccd_cpp: proposal scheduling and dense-group early stop under the
// reference's pybind names and argument conventions (src/cpp/bindings/
// py_module.cpp:577-645, 757-880, 1035-1311, 1375-1500, 1912-1945), plus two
// column-array entries (_schedule_runtime_columns, _dense_group_columns) that
// the numpy-level p2cccd.policy module calls.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <chrono>
#include <cmath>
#include <string>

#include "policy_hip.h"

namespace py = pybind11;
using namespace p2cccd;

namespace {

void ThrowIfError(const Status& s) {
  if (!s.ok) throw py::value_error(s.message);
}

template <typename T>
using Arr = py::array_t<T, py::array::c_style | py::array::forcecast>;

template <typename T>
Arr<T> Column(const py::dict& d, const char* key, py::ssize_t n) {
  Arr<T> a = Arr<T>::ensure(d[key]);
  if (!a) throw py::value_error(std::string("missing or invalid array field: ") + key);
  if (a.ndim() != 1 || a.shape(0) != n)
    throw py::value_error(std::string(key) + " must have shape [" + std::to_string(n) + "]");
  return a;
}

template <typename T>
Arr<T> Matrix(const py::dict& d, const char* key, py::ssize_t n, py::ssize_t m) {
  Arr<T> a = Arr<T>::ensure(d[key]);
  if (!a) throw py::value_error(std::string("missing or invalid array field: ") + key);
  if (a.ndim() != 2 || a.shape(0) != n || a.shape(1) != m)
    throw py::value_error(std::string(key) + " must have shape [" + std::to_string(n) + ", " + std::to_string(m) + "]");
  return a;
}

template <std::size_t N>
std::array<float, N> Floats(const py::handle& h, const char* field) {
  const py::sequence v = py::cast<py::sequence>(h);
  if (py::len(v) != static_cast<py::ssize_t>(N))
    throw py::value_error(std::string(field) + " must have length " + std::to_string(N));
  std::array<float, N> out{};
  for (std::size_t i = 0; i < N; ++i) out[i] = py::cast<float>(v[i]);
  return out;
}

// Duck-typed record parsers: pybind objects, dataclasses or namespaces.
CandidateRecord ParseCandidate(const py::handle& o) {
  CandidateRecord c;
  c.schema_version = py::cast<std::uint32_t>(o.attr("schema_version"));
  c.candidate_id = py::cast<std::uint64_t>(o.attr("candidate_id"));
  c.query_id = py::cast<std::uint64_t>(o.attr("query_id"));
  c.slab_id = py::cast<std::uint32_t>(o.attr("slab_id"));
  c.object_a_id = py::cast<std::uint32_t>(o.attr("object_a_id"));
  c.patch_a_id = py::cast<std::uint32_t>(o.attr("patch_a_id"));
  c.object_b_id = py::cast<std::uint32_t>(o.attr("object_b_id"));
  c.patch_b_id = py::cast<std::uint32_t>(o.attr("patch_b_id"));
  c.proxy_type_a = static_cast<ProxyType>(py::cast<int>(py::int_(o.attr("proxy_type_a"))));
  c.proxy_type_b = static_cast<ProxyType>(py::cast<int>(py::int_(o.attr("proxy_type_b"))));
  c.rt_hit_count = py::cast<std::uint32_t>(o.attr("rt_hit_count"));
  const py::sequence mb = py::cast<py::sequence>(o.attr("motion_bound"));
  if (py::len(mb) != 4) throw py::value_error("CandidateRecord.motion_bound must have length 4");
  for (std::size_t k = 0; k < 4; ++k) c.motion_bound[k] = py::cast<float>(mb[k]);
  c.proxy_features_offset = py::cast<std::uint32_t>(o.attr("proxy_features_offset"));
  c.flags = py::cast<std::uint32_t>(o.attr("flags"));
  return c;
}

ProposalFeatureRow ParseRow(const py::handle& o) {
  ProposalFeatureRow r;
  r.schema_version = py::cast<std::uint32_t>(o.attr("schema_version"));
  r.query_id = py::cast<std::uint64_t>(o.attr("query_id"));
  r.candidate_id = py::cast<std::uint64_t>(o.attr("candidate_id"));
  r.slab_id = py::cast<std::uint32_t>(o.attr("slab_id"));
  r.object_a_id = py::cast<std::uint32_t>(o.attr("object_a_id"));
  r.patch_a_id = py::cast<std::uint32_t>(o.attr("patch_a_id"));
  r.object_b_id = py::cast<std::uint32_t>(o.attr("object_b_id"));
  r.patch_b_id = py::cast<std::uint32_t>(o.attr("patch_b_id"));
  r.features = Floats<kProposalFeatureDimension>(o.attr("features"), "ProposalFeatureRow.features");
  r.interval_targets =
      Floats<kProposalIntervalBinCount>(o.attr("interval_targets"), "ProposalFeatureRow.interval_targets");
  r.family_targets = Floats<kProposalFamilyCount>(o.attr("family_targets"), "ProposalFeatureRow.family_targets");
  r.priority_target = py::cast<float>(o.attr("priority_target"));
  r.cost_target = py::cast<float>(o.attr("cost_target"));
  r.uncertainty_target = py::cast<float>(o.attr("uncertainty_target"));
  r.target_mask = py::cast<std::uint32_t>(o.attr("target_mask"));
  return r;
}

ProposalOutput ParseOutput(const py::handle& o) {
  ProposalOutput p;
  p.candidate_id = py::cast<std::uint64_t>(o.attr("candidate_id"));
  p.interval_scores = Floats<kProposalIntervalBinCount>(o.attr("interval_scores"), "ProposalOutput.interval_scores");
  p.family_scores = Floats<kProposalFamilyCount>(o.attr("family_scores"), "ProposalOutput.family_scores");
  p.priority_score = py::cast<float>(o.attr("priority_score"));
  p.cost_score = py::cast<float>(o.attr("cost_score"));
  p.uncertainty_score = py::cast<float>(o.attr("uncertainty_score"));
  return p;
}

template <typename T, typename F>
std::vector<T> ParseAll(const py::sequence& seq, F fn) {
  std::vector<T> out;
  out.reserve(py::len(seq));
  for (const py::handle& h : seq) out.push_back(fn(h));
  return out;
}

// Feature rows and proposal outputs from the reference's array dictionaries
// (py_module.cpp:792-917): rows in feature_arrays, predictions aligned to them.
std::vector<ProposalFeatureRow> RowsFromArrays(const py::dict& a) {
  const auto schema = py::cast<std::uint32_t>(a["schema_version"]);
  const auto cid = Column<std::uint64_t>(a, "candidate_id", py::len(a["query_id"]));
  const py::ssize_t n = cid.shape(0);
  const auto qid = Column<std::uint64_t>(a, "query_id", n);
  const auto slab = Column<std::uint32_t>(a, "slab_id", n);
  const auto oa = Column<std::uint32_t>(a, "object_a_id", n);
  const auto pa = Column<std::uint32_t>(a, "patch_a_id", n);
  const auto ob = Column<std::uint32_t>(a, "object_b_id", n);
  const auto pb = Column<std::uint32_t>(a, "patch_b_id", n);
  const auto feat = Matrix<float>(a, "features", n, kProposalFeatureDimension);
  const auto it = Matrix<float>(a, "interval_targets", n, kProposalIntervalBinCount);
  const auto ft = Matrix<float>(a, "family_targets", n, kProposalFamilyCount);
  const auto prt = Column<float>(a, "priority_target", n);
  const auto ct = Column<float>(a, "cost_target", n);
  const auto ut = Column<float>(a, "uncertainty_target", n);
  const auto tm = Column<std::uint32_t>(a, "target_mask", n);
  std::vector<ProposalFeatureRow> rows(static_cast<std::size_t>(n));
  for (py::ssize_t i = 0; i < n; ++i) {
    ProposalFeatureRow& r = rows[static_cast<std::size_t>(i)];
    r.schema_version = schema;
    r.query_id = qid.at(i);
    r.candidate_id = cid.at(i);
    r.slab_id = slab.at(i);
    r.object_a_id = oa.at(i);
    r.patch_a_id = pa.at(i);
    r.object_b_id = ob.at(i);
    r.patch_b_id = pb.at(i);
    for (std::size_t k = 0; k < kProposalFeatureDimension; ++k) r.features[k] = feat.at(i, k);
    for (std::size_t k = 0; k < kProposalIntervalBinCount; ++k) r.interval_targets[k] = it.at(i, k);
    for (std::size_t k = 0; k < kProposalFamilyCount; ++k) r.family_targets[k] = ft.at(i, k);
    r.priority_target = prt.at(i);
    r.cost_target = ct.at(i);
    r.uncertainty_target = ut.at(i);
    r.target_mask = tm.at(i);
  }
  return rows;
}

std::vector<ProposalOutput> OutputsFromArrays(const py::dict& feature, const py::dict& pred) {
  const auto cid = Column<std::uint64_t>(feature, "candidate_id", py::len(feature["query_id"]));
  const py::ssize_t n = cid.shape(0);
  const auto is = Matrix<float>(pred, "interval_scores", n, kProposalIntervalBinCount);
  const auto fs = Matrix<float>(pred, "family_scores", n, kProposalFamilyCount);
  const auto pr = Column<float>(pred, "priority_score", n);
  const auto co = Column<float>(pred, "cost_score", n);
  const auto un = Column<float>(pred, "uncertainty_score", n);
  std::vector<ProposalOutput> out(static_cast<std::size_t>(n));
  for (py::ssize_t i = 0; i < n; ++i) {
    ProposalOutput& o = out[static_cast<std::size_t>(i)];
    o.candidate_id = cid.at(i);
    for (std::size_t k = 0; k < kProposalIntervalBinCount; ++k) o.interval_scores[k] = is.at(i, k);
    for (std::size_t k = 0; k < kProposalFamilyCount; ++k) o.family_scores[k] = fs.at(i, k);
    o.priority_score = pr.at(i);
    o.cost_score = co.at(i);
    o.uncertainty_score = un.at(i);
  }
  return out;
}

py::tuple Schedule(const std::vector<CandidateRecord>& cands, const std::vector<ProposalFeatureRow>& rows,
                   const std::vector<ProposalOutput>& outs, const FamilyMasksByQuery& masks,
                   const ProposalSchedulingConfig& cfg) {
  std::vector<ExactWorkItem> q;
  ProposalScheduleStats st;
  Status s;
  {
    py::gil_scoped_release release;
    s = ScheduleRuntimeExactWorkItemsFromProposals(cands, rows, outs, masks, cfg, &q, &st);
  }
  ThrowIfError(s);
  return py::make_tuple(q, st);
}

py::dict StatsDict(const DenseGroupEarlyStopStats& s) {
  py::dict d;
  for (auto [k, v] : {std::pair<const char*, std::uint64_t>{"candidate_count", s.candidate_count},
                      {"group_count", s.group_count},
                      {"positive_group_count", s.positive_group_count},
                      {"no_proposal_exact_calls", s.no_proposal_exact_calls},
                      {"learned_exact_calls", s.learned_exact_calls},
                      {"learned_fallback_calls", s.learned_fallback_calls},
                      {"learned_interval_hit_count", s.learned_interval_hit_count},
                      {"learned_interval_miss_count", s.learned_interval_miss_count},
                      {"tp", s.tp},
                      {"tn", s.tn},
                      {"fp", s.fp},
                      {"fn", s.fn},
                      {"reordered_count", s.reordered_count},
                      {"high_uncertainty_group_count", s.high_uncertainty_group_count}})
    d[k] = py::int_(v);
  d["no_proposal_exact_work"] = py::float_(s.no_proposal_exact_work);
  d["learned_exact_work"] = py::float_(s.learned_exact_work);
  d["schedule_ms"] = py::float_(s.schedule_ms);
  d["exact_ms"] = py::float_(s.exact_ms);
  return d;
}

DenseGroupEarlyStopStats RunDense(const std::vector<DenseGroupCandidateInput>& in, const DenseGroupEarlyStopConfig& c) {
  DenseGroupEarlyStopStats st;
  Status s;
  {
    py::gil_scoped_release release;
    s = RunDenseGroupExactEarlyStop(in, c, &st);
  }
  ThrowIfError(s);
  return st;
}

}  // namespace

void BindPolicy(py::module_& m) {
  py::class_<ProposalFeatureRow>(m, "ProposalFeatureRow")
      .def(py::init<>())
      .def_readwrite("schema_version", &ProposalFeatureRow::schema_version)
      .def_readwrite("query_id", &ProposalFeatureRow::query_id)
      .def_readwrite("candidate_id", &ProposalFeatureRow::candidate_id)
      .def_readwrite("slab_id", &ProposalFeatureRow::slab_id)
      .def_readwrite("object_a_id", &ProposalFeatureRow::object_a_id)
      .def_readwrite("patch_a_id", &ProposalFeatureRow::patch_a_id)
      .def_readwrite("object_b_id", &ProposalFeatureRow::object_b_id)
      .def_readwrite("patch_b_id", &ProposalFeatureRow::patch_b_id)
      .def_readwrite("features", &ProposalFeatureRow::features)
      .def_readwrite("interval_targets", &ProposalFeatureRow::interval_targets)
      .def_readwrite("family_targets", &ProposalFeatureRow::family_targets)
      .def_readwrite("priority_target", &ProposalFeatureRow::priority_target)
      .def_readwrite("cost_target", &ProposalFeatureRow::cost_target)
      .def_readwrite("uncertainty_target", &ProposalFeatureRow::uncertainty_target)
      .def_readwrite("target_mask", &ProposalFeatureRow::target_mask);
  py::class_<ProposalOutput>(m, "ProposalOutput")
      .def(py::init<>())
      .def_readwrite("candidate_id", &ProposalOutput::candidate_id)
      .def_readwrite("interval_scores", &ProposalOutput::interval_scores)
      .def_readwrite("family_scores", &ProposalOutput::family_scores)
      .def_readwrite("priority_score", &ProposalOutput::priority_score)
      .def_readwrite("cost_score", &ProposalOutput::cost_score)
      .def_readwrite("uncertainty_score", &ProposalOutput::uncertainty_score);
  py::class_<ProposalSchedulingConfig>(m, "ProposalSchedulingConfig")
      .def(py::init<>())
      .def_readwrite("first_work_item_id", &ProposalSchedulingConfig::first_work_item_id)
      .def_readwrite("conservative_feature_family_mask", &ProposalSchedulingConfig::conservative_feature_family_mask)
      .def_readwrite("fallback_interval_t0", &ProposalSchedulingConfig::fallback_interval_t0)
      .def_readwrite("fallback_interval_t1", &ProposalSchedulingConfig::fallback_interval_t1)
      .def_readwrite("family_score_threshold", &ProposalSchedulingConfig::family_score_threshold)
      .def_readwrite("uncertainty_fallback_threshold", &ProposalSchedulingConfig::uncertainty_fallback_threshold)
      .def_readwrite("ood_abs_feature_threshold", &ProposalSchedulingConfig::ood_abs_feature_threshold)
      .def_readwrite("preserve_candidate_order", &ProposalSchedulingConfig::preserve_candidate_order);
  py::class_<ProposalScheduleStats>(m, "ProposalScheduleStats")
      .def(py::init<>())
      .def_readwrite("raw_candidate_count", &ProposalScheduleStats::raw_candidate_count)
      .def_readwrite("proposal_output_count", &ProposalScheduleStats::proposal_output_count)
      .def_readwrite("work_item_count", &ProposalScheduleStats::work_item_count)
      .def_readwrite("fallback_count", &ProposalScheduleStats::fallback_count)
      .def_readwrite("missing_proposal_fallback_count", &ProposalScheduleStats::missing_proposal_fallback_count)
      .def_readwrite("invalid_proposal_fallback_count", &ProposalScheduleStats::invalid_proposal_fallback_count)
      .def_readwrite("ood_fallback_count", &ProposalScheduleStats::ood_fallback_count)
      .def_readwrite("high_uncertainty_fallback_count", &ProposalScheduleStats::high_uncertainty_fallback_count)
      .def_readwrite("reordered_count", &ProposalScheduleStats::reordered_count)
      .def_readwrite("monotonic_safe", &ProposalScheduleStats::monotonic_safe);

  m.def("schedule_runtime_exact_work_items",
        [](const py::sequence& candidates, const py::sequence& rows, const py::sequence& outputs,
           const FamilyMasksByQuery& masks, const ProposalSchedulingConfig& config) {
          return Schedule(ParseAll<CandidateRecord>(candidates, ParseCandidate),
                          ParseAll<ProposalFeatureRow>(rows, ParseRow), ParseAll<ProposalOutput>(outputs, ParseOutput),
                          masks, config);
        },
        py::arg("candidates"), py::arg("rows"), py::arg("proposal_outputs"),
        py::arg("conservative_family_masks_by_query_id"), py::arg("config"));
  m.def("schedule_runtime_exact_work_items_from_arrays",
        [](const py::sequence& candidates, const py::dict& feature_arrays, const py::dict& prediction_arrays,
           const FamilyMasksByQuery& masks, const ProposalSchedulingConfig& config) {
          auto cands = ParseAll<CandidateRecord>(candidates, ParseCandidate);
          auto rows = RowsFromArrays(feature_arrays);
          auto outs = OutputsFromArrays(feature_arrays, prediction_arrays);
          return Schedule(cands, rows, outs, masks, config);
        },
        py::arg("candidates"), py::arg("feature_arrays"), py::arg("prediction_arrays"),
        py::arg("conservative_family_masks_by_query_id"), py::arg("config"));
  m.def("schedule_runtime_exact_work_items_from_proposal_arrays",
        [](const py::dict& feature_arrays, const py::dict& prediction_arrays, const FamilyMasksByQuery& masks,
           const ProposalSchedulingConfig& config) {
          auto rows = RowsFromArrays(feature_arrays);
          auto outs = OutputsFromArrays(feature_arrays, prediction_arrays);
          if (rows.size() != outs.size())
            throw py::value_error("feature row and proposal output array lengths must match");
          std::vector<ExactWorkItem> q;
          ProposalScheduleStats st;
          Status s;
          {
            py::gil_scoped_release release;
            s = ScheduleRuntimeExactWorkItemsFromAlignedProposals(rows, outs, masks, config, &q, &st);
          }
          ThrowIfError(s);
          return py::make_tuple(q, st);
        },
        py::arg("feature_arrays"), py::arg("prediction_arrays"), py::arg("conservative_family_masks_by_query_id"),
        py::arg("config"));
  m.def("run_native_dense_group_exact_early_stop",
        [](const py::dict& feature_arrays, const py::dict& prediction_arrays, float uncertainty_fallback_threshold,
           std::uint32_t representative_attempt_limit, double interval_miss_penalty_scale, bool preserve_input_order) {
          // inputs of RunNativeDenseGroupExactEarlyStopOrThrow (py_module.cpp:1179-1248): the proposed
          // interval is the best of the 8 interval bins; the narrow cost is the full cost scaled by
          // cost_target; the oracle trace carries collision (0), full cost (5) and contact interval (6, 7)
          const auto t0 = std::chrono::steady_clock::now();
          const auto qid = Column<std::uint64_t>(feature_arrays, "query_id", py::len(feature_arrays["query_id"]));
          const py::ssize_t n = qid.shape(0);
          const auto cid = Column<std::uint64_t>(feature_arrays, "candidate_id", n);
          const auto cost_target = Column<float>(feature_arrays, "cost_target", n);
          const auto trace = Matrix<double>(feature_arrays, "oracle_trace", n, 8);
          const auto bins = Matrix<float>(prediction_arrays, "interval_scores", n, kProposalIntervalBinCount);
          const auto prio = Column<float>(prediction_arrays, "priority_score", n);
          const auto cost = Column<float>(prediction_arrays, "cost_score", n);
          const auto unc = Column<float>(prediction_arrays, "uncertainty_score", n);
          std::vector<DenseGroupCandidateInput> in(static_cast<std::size_t>(n));
          const double width = 1.0 / static_cast<double>(kProposalIntervalBinCount);
          for (py::ssize_t i = 0; i < n; ++i) {
            py::ssize_t best = 0;
            float best_score = bins.at(i, 0);
            for (py::ssize_t k = 1; k < static_cast<py::ssize_t>(kProposalIntervalBinCount); ++k)
              if (bins.at(i, k) > best_score) {
                best_score = bins.at(i, k);
                best = k;
              }
            DenseGroupCandidateInput& c = in[static_cast<std::size_t>(i)];
            c.query_id = qid.at(i);
            c.candidate_id = cid.at(i);
            c.full_exact_cost = trace.at(i, 5);
            const double ratio = cost_target.at(i);
            c.narrow_exact_cost = std::isfinite(ratio) && ratio > 0.0
                                      ? std::max(1.0e-12, c.full_exact_cost * static_cast<double>(cost_target.at(i)))
                                      : c.full_exact_cost;
            c.interval_t0 = width * static_cast<double>(best);
            c.interval_t1 = std::min(1.0, c.interval_t0 + width);
            c.contact_t0 = trace.at(i, 6);
            c.contact_t1 = trace.at(i, 7);
            c.priority_score = prio.at(i);
            c.cost_score = cost.at(i);
            c.uncertainty_score = unc.at(i);
            c.candidate_collides = trace.at(i, 0) > 0.5;
          }
          const auto t1 = std::chrono::steady_clock::now();
          DenseGroupEarlyStopConfig cfg;
          cfg.uncertainty_fallback_threshold = uncertainty_fallback_threshold;
          cfg.representative_attempt_limit = representative_attempt_limit;
          cfg.interval_miss_penalty_scale = interval_miss_penalty_scale;
          cfg.preserve_input_order = preserve_input_order;
          const DenseGroupEarlyStopStats s = RunDense(in, cfg);
          const double parse_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
          py::dict d = StatsDict(s);
          const double pos = static_cast<double>(s.positive_group_count);
          d["first_positive_rank_mean"] = py::float_(s.positive_group_count == 0 ? 0.0 : s.first_positive_rank_sum / pos);
          d["cost_weighted_first_positive_mean"] =
              py::float_(s.positive_group_count == 0 ? 0.0 : s.cost_before_first_positive_sum / pos);
          d["parse_ms"] = py::float_(parse_ms);
          d["native_driver_ms"] = py::float_(s.total_ms);
          d["native_total_ms"] = py::float_(parse_ms + s.total_ms);
          d["exact_call_reduction"] = py::float_(
              s.no_proposal_exact_calls == 0
                  ? 0.0
                  : 1.0 - static_cast<double>(s.learned_exact_calls) / static_cast<double>(s.no_proposal_exact_calls));
          d["exact_work_reduction"] =
              py::float_(s.no_proposal_exact_work <= 0.0 ? 0.0 : 1.0 - s.learned_exact_work / s.no_proposal_exact_work);
          return d;
        },
        py::arg("feature_arrays"), py::arg("prediction_arrays"), py::arg("uncertainty_fallback_threshold") = 0.75F,
        py::arg("representative_attempt_limit") = 3U, py::arg("interval_miss_penalty_scale") = 0.22,
        py::arg("preserve_input_order") = false);

  // ---- column entries for p2cccd.policy (numpy structured arrays) -------------
  m.def("_schedule_runtime_columns",
        [](const py::dict& c, const py::dict& r, const py::dict& o, const FamilyMasksByQuery& masks,
           const ProposalSchedulingConfig& config) {
          const py::ssize_t n = py::len(c["candidate_id"]);
          const auto cid = Column<std::uint64_t>(c, "candidate_id", n);
          const auto qid = Column<std::uint64_t>(c, "query_id", n);
          const auto u32 = [&](const char* k) { return Column<std::uint32_t>(c, k, n); };
          const auto sv = u32("schema_version"), slab = u32("slab_id"), oa = u32("object_a_id"),
                     pa = u32("patch_a_id"), ob = u32("object_b_id"), pb = u32("patch_b_id"),
                     ta = u32("proxy_type_a"), tb = u32("proxy_type_b"), rt = u32("rt_hit_count");
          const auto mb = Matrix<float>(c, "motion_bound", n, 4);
          std::vector<CandidateRecord> cands(static_cast<std::size_t>(n));
          for (py::ssize_t i = 0; i < n; ++i) {
            CandidateRecord& x = cands[static_cast<std::size_t>(i)];
            x.schema_version = sv.at(i);
            x.candidate_id = cid.at(i);
            x.query_id = qid.at(i);
            x.slab_id = slab.at(i);
            x.object_a_id = oa.at(i);
            x.patch_a_id = pa.at(i);
            x.object_b_id = ob.at(i);
            x.patch_b_id = pb.at(i);
            x.proxy_type_a = static_cast<ProxyType>(ta.at(i));
            x.proxy_type_b = static_cast<ProxyType>(tb.at(i));
            x.rt_hit_count = rt.at(i);
            for (int k = 0; k < 4; ++k) x.motion_bound[k] = mb.at(i, k);
          }
          const py::ssize_t nr = py::len(r["candidate_id"]);
          const auto rcid = Column<std::uint64_t>(r, "candidate_id", nr);
          const auto feat = Matrix<float>(r, "features", nr, kProposalFeatureDimension);
          std::vector<ProposalFeatureRow> rows(static_cast<std::size_t>(nr));
          for (py::ssize_t i = 0; i < nr; ++i) {
            rows[static_cast<std::size_t>(i)].candidate_id = rcid.at(i);
            for (std::size_t k = 0; k < kProposalFeatureDimension; ++k)
              rows[static_cast<std::size_t>(i)].features[k] = feat.at(i, k);
          }
          py::dict od;
          od["query_id"] = o["candidate_id"];
          od["candidate_id"] = o["candidate_id"];
          std::vector<ProposalOutput> outs = OutputsFromArrays(od, o);
          std::vector<ExactWorkItem> q;
          ProposalScheduleStats st;
          Status s;
          {
            py::gil_scoped_release release;
            s = ScheduleRuntimeExactWorkItemsFromProposals(cands, rows, outs, masks, config, &q, &st);
          }
          ThrowIfError(s);
          const py::ssize_t m = static_cast<py::ssize_t>(q.size());
          Arr<std::uint64_t> wid(m), parent(m), query(m);
          Arr<std::uint32_t> wslab(m), wpa(m), wpb(m), fam(m), topk(m);
          Arr<double> it0(m), it1(m);
          Arr<std::uint16_t> depth(m);
          Arr<float> prio(m);
          Arr<std::uint8_t> src(m);
          for (py::ssize_t i = 0; i < m; ++i) {
            const ExactWorkItem& w = q[static_cast<std::size_t>(i)];
            wid.mutable_at(i) = w.work_item_id;
            parent.mutable_at(i) = w.parent_candidate_id;
            query.mutable_at(i) = w.query_id;
            wslab.mutable_at(i) = w.slab_id;
            wpa.mutable_at(i) = w.patch_a_id;
            wpb.mutable_at(i) = w.patch_b_id;
            it0.mutable_at(i) = w.interval_t0;
            it1.mutable_at(i) = w.interval_t1;
            fam.mutable_at(i) = w.feature_family_mask;
            topk.mutable_at(i) = w.topk_feature_ids_offset;
            depth.mutable_at(i) = w.depth;
            prio.mutable_at(i) = w.priority_score;
            src.mutable_at(i) = static_cast<std::uint8_t>(w.source);
          }
          py::dict items;
          items["work_item_id"] = wid;
          items["parent_candidate_id"] = parent;
          items["query_id"] = query;
          items["slab_id"] = wslab;
          items["patch_a_id"] = wpa;
          items["patch_b_id"] = wpb;
          items["interval_t0"] = it0;
          items["interval_t1"] = it1;
          items["feature_family_mask"] = fam;
          items["topk_feature_ids_offset"] = topk;
          items["depth"] = depth;
          items["priority_score"] = prio;
          items["source"] = src;
          py::dict sd;
          sd["raw_candidate_count"] = st.raw_candidate_count;
          sd["proposal_output_count"] = st.proposal_output_count;
          sd["work_item_count"] = st.work_item_count;
          sd["fallback_count"] = st.fallback_count;
          sd["missing_proposal_fallback_count"] = st.missing_proposal_fallback_count;
          sd["invalid_proposal_fallback_count"] = st.invalid_proposal_fallback_count;
          sd["ood_fallback_count"] = st.ood_fallback_count;
          sd["high_uncertainty_fallback_count"] = st.high_uncertainty_fallback_count;
          sd["reordered_count"] = st.reordered_count;
          sd["monotonic_safe"] = st.monotonic_safe;
          return py::make_tuple(items, sd);
        },
        py::arg("candidates"), py::arg("rows"), py::arg("outputs"), py::arg("masks"), py::arg("config"));
  m.def("_dense_group_columns",
        [](const py::dict& c, float uncertainty_fallback_threshold, std::uint32_t representative_attempt_limit,
           double interval_miss_penalty_scale, bool preserve_input_order) {
          const py::ssize_t n = py::len(c["query_id"]);
          const auto qid = Column<std::uint64_t>(c, "query_id", n);
          const auto cid = Column<std::uint64_t>(c, "candidate_id", n);
          const auto f64 = [&](const char* k) { return Column<double>(c, k, n); };
          const auto f32 = [&](const char* k) { return Column<float>(c, k, n); };
          const auto full = f64("full_exact_cost"), narrow = f64("narrow_exact_cost"), i0 = f64("interval_t0"),
                     i1 = f64("interval_t1"), c0 = f64("contact_t0"), c1 = f64("contact_t1");
          const auto pr = f32("priority_score"), co = f32("cost_score"), un = f32("uncertainty_score");
          const auto hit = Column<std::uint8_t>(c, "candidate_collides", n);
          std::vector<DenseGroupCandidateInput> in(static_cast<std::size_t>(n));
          for (py::ssize_t i = 0; i < n; ++i) {
            DenseGroupCandidateInput& x = in[static_cast<std::size_t>(i)];
            x.query_id = qid.at(i);
            x.candidate_id = cid.at(i);
            x.full_exact_cost = full.at(i);
            x.narrow_exact_cost = narrow.at(i);
            x.interval_t0 = i0.at(i);
            x.interval_t1 = i1.at(i);
            x.contact_t0 = c0.at(i);
            x.contact_t1 = c1.at(i);
            x.priority_score = pr.at(i);
            x.cost_score = co.at(i);
            x.uncertainty_score = un.at(i);
            x.candidate_collides = hit.at(i) != 0;
          }
          DenseGroupEarlyStopConfig cfg;
          cfg.uncertainty_fallback_threshold = uncertainty_fallback_threshold;
          cfg.representative_attempt_limit = representative_attempt_limit;
          cfg.interval_miss_penalty_scale = interval_miss_penalty_scale;
          cfg.preserve_input_order = preserve_input_order;
          const DenseGroupEarlyStopStats s = RunDense(in, cfg);
          py::dict d = StatsDict(s);
          d["first_positive_rank_sum"] = py::float_(s.first_positive_rank_sum);
          d["cost_before_first_positive_sum"] = py::float_(s.cost_before_first_positive_sum);
          return d;
        },
        py::arg("candidates"), py::arg("uncertainty_fallback_threshold"), py::arg("representative_attempt_limit"),
        py::arg("interval_miss_penalty_scale"), py::arg("preserve_input_order"));
}
