// GenerateCandidatesForExternalBatch over the gfx950 C-ABI.
//
// The host part is what the reference does in order and cannot be done out
// of order without changing its answers: per query, in input order, the
// finiteness checks, the uniqueness of source ids, the runtime query id (the
// source id, else index + 1, bumped past ids already taken) and the x shift
// of the query's slab in the batched proxy scene (a running FP64 sum of the
// queries' x extents plus a 1e-2 margin,
// src/cpp/rt_candidate/external_batch_candidate.cpp:128-229).  The device
// part (p2c_external_batch_host) computes each query's shifted swept boxes,
// the overlap test, the candidate records and their ids in ascending runtime
// id order (:295-312).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <unordered_set>

#include "p2c_hip.h"
#include "proxy_hip.h"

namespace p2cccd {
namespace {

using Clock = std::chrono::steady_clock;

constexpr double kSlabMarginX = 1.0e-2;

double ElapsedMs(Clock::time_point start) {
  return std::chrono::duration<double, std::milli>(Clock::now() - start).count();
}

bool FinitePoints(const std::array<std::array<double, 3>, 4>& v) {
  for (const auto& p : v)
    if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2])) return false;
  return true;
}

// Least and greatest x over the query's points of both sides, at t0 and t1
// (the two boxes' x extents joined).
void XRange(const ExternalBatchQuery& q, double* lo, double* hi) {
  const bool vf = q.family == ExternalQueryFamily::kVertexFace;
  double side_lo[2], side_hi[2];
  for (int s = 0; s < 2; ++s) {
    double a = HUGE_VAL, b = -HUGE_VAL;
    const int first = s == 0 ? 0 : (vf ? 1 : 2);
    const int last = s == 0 ? (vf ? 0 : 1) : 3;
    for (int v = first; v <= last; ++v)
      for (const auto* p : {&q.vertices_t0[v], &q.vertices_t1[v]}) {
        a = std::min(a, (*p)[0]);
        b = std::max(b, (*p)[0]);
      }
    side_lo[s] = a;
    side_hi[s] = b;
  }
  *lo = std::min(side_lo[0], side_lo[1]);
  *hi = std::max(side_hi[0], side_hi[1]);
}

}  // namespace

Status GenerateCandidatesForExternalBatch(const std::vector<ExternalBatchQuery>& queries,
                                          const CandidateGeneratorConfig& config,
                                          ExternalBatchCandidateResult* result) {
  (void)config;  // every backend name runs the device path
  if (result == nullptr) return Status::Error("external batch candidate result output pointer is null");
  if (queries.empty()) return Status::Error("external batch candidate generation requires at least one query");
  const Clock::time_point build_start = Clock::now();
  const std::size_t n = queries.size();
  std::vector<p2c_external_query> packed(n);
  std::vector<RuntimeQueryIdMapping> ids(n);
  std::unordered_set<std::uint64_t> seen, taken;
  double cursor = 0.0;
  for (std::size_t i = 0; i < n; ++i) {
    const ExternalBatchQuery& q = queries[i];
    if (!FinitePoints(q.vertices_t0)) return Status::Error("external batch query vertices_t0 must be finite");
    if (!FinitePoints(q.vertices_t1)) return Status::Error("external batch query vertices_t1 must be finite");
    if (!seen.insert(q.source_query_id).second)
      return Status::Error("external batch source_query_id values must be unique");
    std::uint64_t rid = q.source_query_id > 0 ? q.source_query_id : static_cast<std::uint64_t>(i + 1);
    while (taken.count(rid) != 0) ++rid;
    taken.insert(rid);
    ids[i] = {q.source_query_id, rid};
    double lo = 0.0, hi = 0.0;
    XRange(q, &lo, &hi);
    p2c_external_query& e = packed[i];
    std::memcpy(e.vertices_t0, q.vertices_t0.data(), sizeof(e.vertices_t0));
    std::memcpy(e.vertices_t1, q.vertices_t1.data(), sizeof(e.vertices_t1));
    e.x_offset = cursor - lo;
    e.runtime_query_id = rid;
    e.source_query_index = q.source_query_index;
    e.family = q.family == ExternalQueryFamily::kVertexFace ? 0u : 1u;
    e.object_a_id = q.has_box_pair ? q.box_pair[0] : 1u;
    e.object_b_id = q.has_box_pair ? q.box_pair[1] : 2u;
    cursor += (hi - lo) + kSlabMarginX;
  }
  // candidates are numbered in ascending runtime id (the reference sorts by
  // (query_id, patch_a_id, patch_b_id); runtime ids are unique)
  std::vector<std::uint32_t> order(n);
  std::iota(order.begin(), order.end(), 0u);
  std::sort(order.begin(), order.end(), [&](std::uint32_t a, std::uint32_t b) {
    return packed[a].runtime_query_id < packed[b].runtime_query_id;
  });
  const double build_ms = ElapsedMs(build_start);

  p2c_candidate* cand = nullptr;
  std::int64_t n_cand = 0;
  std::uint64_t status[4] = {0, 0, 0, 0};
  double trace_ms = 0.0;
  if (p2c_external_batch_host(packed.data(), order.data(), static_cast<std::int64_t>(n), &cand, &n_cand, status,
                              &trace_ms) != P2C_OK)
    return Status::Error(p2c_last_error());
  struct Free {
    p2c_candidate* p;
    ~Free() { p2c_host_free(p); }
  } guard{cand};
  constexpr std::uint64_t kNone = ~std::uint64_t(0);
  if (status[1] != kNone) return Status::Error("proxy bounds contain a non-finite value");
  if (status[2] != kNone) return Status::Error("raw candidate motion_bound must be finite and non-negative");
  if (status[3] != kNone) return Status::Error("CandidateRecord.candidate_id is required");

  const Clock::time_point compact_start = Clock::now();
  ExternalBatchCandidateResult r;
  r.backend_name = kHipCandidateBackend;
  r.candidates.resize(static_cast<std::size_t>(n_cand));
  static_assert(sizeof(CandidateRecord) == sizeof(p2c_candidate), "CandidateRecord layout");
  if (n_cand > 0) std::memcpy(r.candidates.data(), cand, sizeof(CandidateRecord) * std::size_t(n_cand));
  std::vector<char> active(n, 0);
  for (const CandidateRecord& c : r.candidates) active[c.slab_id] = 1;  // slab_id is the query's index
  std::uint64_t positive = 0, covered = 0;
  for (std::size_t i = 0; i < n; ++i) {
    if (!queries[i].has_ground_truth || !queries[i].ground_truth_collides) continue;
    ++positive;
    covered += active[i] ? 1 : 0;
  }
  r.primitive_count = 2 * n;
  r.raw_hit_count = static_cast<std::uint64_t>(n_cand);
  r.compact_candidate_count = static_cast<std::uint64_t>(n_cand);
  r.candidate_recall = positive == 0 ? 1.0 : static_cast<double>(covered) / static_cast<double>(positive);
  r.runtime_query_ids = std::move(ids);
  r.timing.build_ms = build_ms;
  r.timing.update_ms = 0.0;
  r.timing.trace_ms = trace_ms;  // device: boxes, overlap, compaction, ids
  r.timing.compact_ms = ElapsedMs(compact_start);
  r.timing.stats_ms = 0.0;
  r.timing.total_ms = r.timing.build_ms + r.timing.update_ms + r.timing.trace_ms + r.timing.compact_ms;
  *result = std::move(r);
  return Status::Ok();
}

}  // namespace p2cccd
