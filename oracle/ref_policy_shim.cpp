// Oracle (test infrastructure only): a C-ABI shim over the reference's own
// proposal scheduling policy and dense-group early stop, compiled from the
// reference sources where they lie (src/cpp/proposal/proposal_policy.cpp,
// proposal_features.cpp, src/cpp/common/group_scheduler.cpp) into
// oracle/_ref/libp2c_ref.so by oracle/Makefile.  Used in this container to
// write golden vectors (tests/golden/make_policy_golden.py).  Never part of
// the shipped library.
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "common/group_scheduler.h"
#include "proposal/proposal_policy.h"

namespace {
void CopyError(const p2cccd::Status& st, char* err, int errlen) {
  if (err != nullptr && errlen > 0) {
    std::strncpy(err, st.message.c_str(), static_cast<size_t>(errlen - 1));
    err[errlen - 1] = 0;
  }
}
}  // namespace

// Candidates: ids, query, slab, objects, patches, types, rt_hit_count, motion.
// Rows: candidate ids and 32 features.  Outputs: candidate ids, 8 interval
// scores, 8 family scores, priority, cost, uncertainty.  Work queue out: per
// item (work_item_id, parent_candidate_id, query_id, slab, patch_a, patch_b,
// mask, source) as u64[8], interval as f64[2], priority as f32.
extern "C" int ref_schedule_runtime(int64_t n, const uint64_t* cand_id, const uint64_t* query_id, const uint32_t* slab,
                                    const uint32_t* obj_a, const uint32_t* patch_a, const uint32_t* obj_b,
                                    const uint32_t* patch_b, const uint32_t* type_a, const uint32_t* type_b,
                                    const uint32_t* rt, const float* motion4, int64_t nrows, const uint64_t* row_id,
                                    const float* row_feat32, int64_t nout, const uint64_t* out_id, const float* out_iv8,
                                    const float* out_fam8, const float* out_prio, const float* out_cost,
                                    const float* out_unc, int64_t nmask, const uint64_t* mask_query,
                                    const uint32_t* mask_val, const double* cfg_d /* t0, t1 */,
                                    const float* cfg_f /* family thr, unc thr, ood thr */,
                                    const uint64_t* cfg_u /* first id, conservative mask, preserve order */,
                                    uint64_t* wq_u8, double* wq_iv, float* wq_prio, uint64_t* stats10, char* err,
                                    int errlen) {
  std::vector<p2cccd::CandidateRecord> cands(static_cast<size_t>(n));
  for (int64_t i = 0; i < n; ++i) {
    auto& c = cands[static_cast<size_t>(i)];
    c.candidate_id = cand_id[i];
    c.query_id = query_id[i];
    c.slab_id = slab[i];
    c.object_a_id = obj_a[i];
    c.patch_a_id = patch_a[i];
    c.object_b_id = obj_b[i];
    c.patch_b_id = patch_b[i];
    c.proxy_type_a = static_cast<p2cccd::ProxyType>(type_a[i]);
    c.proxy_type_b = static_cast<p2cccd::ProxyType>(type_b[i]);
    c.rt_hit_count = rt[i];
    for (int k = 0; k < 4; ++k) c.motion_bound[k] = motion4[4 * i + k];
  }
  std::vector<p2cccd::ProposalFeatureRow> rows(static_cast<size_t>(nrows));
  for (int64_t i = 0; i < nrows; ++i) {
    rows[static_cast<size_t>(i)].candidate_id = row_id[i];
    for (int k = 0; k < 32; ++k) rows[static_cast<size_t>(i)].features[k] = row_feat32[32 * i + k];
  }
  std::vector<p2cccd::ProposalOutput> outs(static_cast<size_t>(nout));
  for (int64_t i = 0; i < nout; ++i) {
    auto& o = outs[static_cast<size_t>(i)];
    o.candidate_id = out_id[i];
    for (int k = 0; k < 8; ++k) {
      o.interval_scores[k] = out_iv8[8 * i + k];
      o.family_scores[k] = out_fam8[8 * i + k];
    }
    o.priority_score = out_prio[i];
    o.cost_score = out_cost[i];
    o.uncertainty_score = out_unc[i];
  }
  std::unordered_map<std::uint64_t, std::uint32_t> masks;
  for (int64_t i = 0; i < nmask; ++i) masks[mask_query[i]] = mask_val[i];
  p2cccd::ProposalSchedulingConfig cfg;
  cfg.fallback_interval_t0 = cfg_d[0];
  cfg.fallback_interval_t1 = cfg_d[1];
  cfg.family_score_threshold = cfg_f[0];
  cfg.uncertainty_fallback_threshold = cfg_f[1];
  cfg.ood_abs_feature_threshold = cfg_f[2];
  cfg.first_work_item_id = cfg_u[0];
  cfg.conservative_feature_family_mask = static_cast<uint32_t>(cfg_u[1]);
  cfg.preserve_candidate_order = cfg_u[2] != 0;
  std::vector<p2cccd::ExactWorkItem> queue;
  p2cccd::ProposalScheduleStats st;
  const auto status =
      p2cccd::ScheduleRuntimeExactWorkItemsFromProposals(cands, rows, outs, masks, cfg, &queue, &st);
  if (!status.ok) {
    CopyError(status, err, errlen);
    return -1;
  }
  for (size_t i = 0; i < queue.size(); ++i) {
    const auto& w = queue[i];
    uint64_t* u = wq_u8 + 8 * i;
    u[0] = w.work_item_id;
    u[1] = w.parent_candidate_id;
    u[2] = w.query_id;
    u[3] = w.slab_id;
    u[4] = w.patch_a_id;
    u[5] = w.patch_b_id;
    u[6] = w.feature_family_mask;
    u[7] = static_cast<uint64_t>(w.source);
    wq_iv[2 * i] = w.interval_t0;
    wq_iv[2 * i + 1] = w.interval_t1;
    wq_prio[i] = w.priority_score;
  }
  const uint64_t s[10] = {st.raw_candidate_count, st.proposal_output_count, st.work_item_count, st.fallback_count,
                          st.missing_proposal_fallback_count, st.invalid_proposal_fallback_count,
                          st.ood_fallback_count, st.high_uncertainty_fallback_count, st.reordered_count,
                          st.monotonic_safe ? 1u : 0u};
  std::memcpy(stats10, s, sizeof(s));
  return 0;
}

// Dense-group early stop: inputs as parallel arrays; stats out as
// u64[12] (counts) and f64[4] (work sums), timing excluded.
extern "C" int ref_dense_group(int64_t n, const uint64_t* query_id, const uint64_t* cand_id, const double* d6,
                               const float* f3, const uint8_t* collides, const float* cfg_f /* unc thr */,
                               const uint64_t* cfg_u /* attempt limit, preserve order */, double penalty,
                               uint64_t* su, double* sd, char* err, int errlen) {
  std::vector<p2cccd::DenseGroupCandidateInput> in(static_cast<size_t>(n));
  for (int64_t i = 0; i < n; ++i) {
    auto& c = in[static_cast<size_t>(i)];
    c.query_id = query_id[i];
    c.candidate_id = cand_id[i];
    c.full_exact_cost = d6[6 * i + 0];
    c.narrow_exact_cost = d6[6 * i + 1];
    c.interval_t0 = d6[6 * i + 2];
    c.interval_t1 = d6[6 * i + 3];
    c.contact_t0 = d6[6 * i + 4];
    c.contact_t1 = d6[6 * i + 5];
    c.priority_score = f3[3 * i + 0];
    c.cost_score = f3[3 * i + 1];
    c.uncertainty_score = f3[3 * i + 2];
    c.candidate_collides = collides[i] != 0;
  }
  p2cccd::DenseGroupEarlyStopConfig cfg;
  cfg.uncertainty_fallback_threshold = cfg_f[0];
  cfg.representative_attempt_limit = static_cast<uint32_t>(cfg_u[0]);
  cfg.preserve_input_order = cfg_u[1] != 0;
  cfg.interval_miss_penalty_scale = penalty;
  p2cccd::DenseGroupEarlyStopStats st;
  const auto status = p2cccd::RunDenseGroupExactEarlyStop(in, cfg, &st);
  if (!status.ok) {
    CopyError(status, err, errlen);
    return -1;
  }
  const uint64_t u[12] = {st.candidate_count, st.group_count, st.positive_group_count, st.no_proposal_exact_calls,
                          st.learned_exact_calls, st.learned_fallback_calls, st.learned_interval_hit_count,
                          st.learned_interval_miss_count, st.tp, st.tn, st.fp, st.fn};
  std::memcpy(su, u, sizeof(u));
  su[12] = st.reordered_count;
  su[13] = st.high_uncertainty_group_count;
  sd[0] = st.no_proposal_exact_work;
  sd[1] = st.learned_exact_work;
  sd[2] = st.first_positive_rank_sum;
  sd[3] = st.cost_before_first_positive_sum;
  return 0;
}
