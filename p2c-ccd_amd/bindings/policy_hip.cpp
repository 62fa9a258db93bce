// Proposal scheduling and the dense-group early stop, host bookkeeping.
//
// Schedule (src/cpp/proposal/proposal_policy.cpp:595-742): one work item per
// candidate.  A candidate falls back to the conservative item (its query's
// family mask, rt_hit_count as priority, the fallback interval) when its
// proposal is missing, invalid, out of distribution (no feature row, or a
// feature non-finite or above the magnitude bound) or too uncertain; otherwise
// the item adds the predicted families and takes the proposal's priority.
// Items are stably sorted by descending priority unless the order is kept,
// numbered, validated and checked for conservation, each check with the
// reference's message and in its order.
//
// Dense-group stop (src/cpp/common/group_scheduler.cpp:60-239): groups are
// runs of equal query id after a stable sort (priority descending, then cost
// ascending, then input order); each group certifies representatives until a
// contact inside the proposed interval, the representative limit, or a
// too-uncertain candidate, which replays the rest in full.  Calls, work,
// first-positive ranks and the confusion counts accumulate in the reference's
// order, so every sum is the same double.
#include "policy_hip.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <numeric>
#include <set>
#include <string>

namespace p2cccd {
namespace {

using Clock = std::chrono::steady_clock;
constexpr std::uint32_t kConservativeMask = kFeatureFamilyPointTriangle | kFeatureFamilyEdgeEdge;

double Ms(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); }

template <std::size_t N>
bool AllFinite(const std::array<float, N>& v) {
  for (float x : v)
    if (!std::isfinite(x)) return false;
  return true;
}

bool ProposalHasNonFinite(const ProposalOutput& o) {
  return !AllFinite(o.interval_scores) || !AllFinite(o.family_scores) || !std::isfinite(o.priority_score) ||
         !std::isfinite(o.cost_score) || !std::isfinite(o.uncertainty_score);
}

bool RowOutOfDistribution(const ProposalFeatureRow& row, const ProposalSchedulingConfig& c) {
  for (float x : row.features)
    if (!std::isfinite(x) || std::abs(x) > c.ood_abs_feature_threshold) return true;
  return false;
}

std::uint32_t BaseMask(const FamilyMasksByQuery& masks, std::uint64_t query_id, const ProposalSchedulingConfig& c) {
  const auto it = masks.find(query_id);
  return (it != masks.end() && it->second != 0U) ? it->second : c.conservative_feature_family_mask;
}

std::uint32_t PredictedMask(const ProposalOutput& o, const ProposalSchedulingConfig& c) {
  return (o.family_scores[0] >= c.family_score_threshold ? kFeatureFamilyPointTriangle : 0U) |
         (o.family_scores[1] >= c.family_score_threshold ? kFeatureFamilyEdgeEdge : 0U);
}

// One candidate's item before numbering; updates the fallback statistics.
ExactWorkItem MakeItem(std::uint64_t candidate_id, std::uint64_t query_id, std::uint32_t slab, std::uint32_t pa,
                       std::uint32_t pb, float raw_priority, const ProposalFeatureRow* row,
                       const ProposalOutput* out, const FamilyMasksByQuery& masks, const ProposalSchedulingConfig& c,
                       bool count_missing, ProposalScheduleStats* st) {
  const bool missing = out == nullptr;
  const bool invalid = out != nullptr && (ProposalHasNonFinite(*out) || !ValidateProposalOutput(*out).ok);
  const bool ood = row == nullptr || RowOutOfDistribution(*row, c);
  const bool uncertain =
      out != nullptr && std::isfinite(out->uncertainty_score) && out->uncertainty_score >= c.uncertainty_fallback_threshold;
  const bool fallback = missing || invalid || ood || uncertain;
  st->fallback_count += fallback ? 1U : 0U;
  if (count_missing) st->missing_proposal_fallback_count += missing ? 1U : 0U;
  st->invalid_proposal_fallback_count += invalid ? 1U : 0U;
  st->ood_fallback_count += ood ? 1U : 0U;
  st->high_uncertainty_fallback_count += uncertain ? 1U : 0U;
  const std::uint32_t base = BaseMask(masks, query_id, c);
  ExactWorkItem item;
  item.parent_candidate_id = candidate_id;
  item.query_id = query_id;
  item.slab_id = slab;
  item.patch_a_id = pa;
  item.patch_b_id = pb;
  item.interval_t0 = c.fallback_interval_t0;
  item.interval_t1 = c.fallback_interval_t1;
  item.feature_family_mask = base;
  item.topk_feature_ids_offset = 0U;
  item.depth = 0U;
  item.priority_score = raw_priority;
  item.source = ProposalSource::kFallback;
  if (!fallback) {
    item.feature_family_mask = base | PredictedMask(*out, c);
    item.priority_score = out->priority_score;
    item.source = ProposalSource::kRefined;
  }
  return item;
}

// Stable descending priority order when any neighbour pair is out of order.
void SortByPriority(std::vector<ExactWorkItem>* q) {
  bool sorted = true;
  for (std::size_t i = 1; i < q->size() && sorted; ++i) sorted = !((*q)[i - 1].priority_score < (*q)[i].priority_score);
  if (!sorted)
    std::stable_sort(q->begin(), q->end(),
                     [](const ExactWorkItem& a, const ExactWorkItem& b) { return a.priority_score > b.priority_score; });
}

double PositiveCost(double v, double fallback) { return (!std::isfinite(v) || v <= 0.0) ? fallback : v; }

}  // namespace

Status ValidateCandidateRecord(const CandidateRecord& r) {
  if (r.schema_version != 1) return Status::Error("CandidateRecord.schema_version is unsupported");
  if (r.candidate_id == 0) return Status::Error("CandidateRecord.candidate_id is required");
  if (r.query_id == 0) return Status::Error("CandidateRecord.query_id is required");
  const auto concrete = [](ProxyType t) { return t == ProxyType::kSweptAabb || t == ProxyType::kCapsule; };
  if (!concrete(r.proxy_type_a)) return Status::Error("CandidateRecord.proxy_type_a must be a concrete proxy type");
  if (!concrete(r.proxy_type_b)) return Status::Error("CandidateRecord.proxy_type_b must be a concrete proxy type");
  if (r.rt_hit_count == 0) return Status::Error("CandidateRecord.rt_hit_count is required");
  for (float x : r.motion_bound)
    if (!std::isfinite(x) || x < 0.0F)
      return Status::Error("CandidateRecord.motion_bound must contain finite non-negative values");
  return Status::Ok();
}

Status ValidateProposalOutput(const ProposalOutput& o) {
  if (o.candidate_id == 0) return Status::Error("ProposalOutput.candidate_id is required");
  if (!AllFinite(o.interval_scores)) return Status::Error("ProposalOutput.interval_scores contains a non-finite value");
  if (!AllFinite(o.family_scores)) return Status::Error("ProposalOutput.family_scores contains a non-finite value");
  if (!std::isfinite(o.priority_score)) return Status::Error("ProposalOutput.priority_score must be finite");
  if (!std::isfinite(o.cost_score)) return Status::Error("ProposalOutput.cost_score must be finite");
  if (!(o.cost_score >= 0.0F)) return Status::Error("ProposalOutput.cost_score must be non-negative");
  if (!std::isfinite(o.uncertainty_score)) return Status::Error("ProposalOutput.uncertainty_score must be finite");
  if (!(o.uncertainty_score >= 0.0F)) return Status::Error("ProposalOutput.uncertainty_score must be non-negative");
  return Status::Ok();
}

std::string SchedulingConfigError(const ProposalSchedulingConfig& c) {
  if (c.first_work_item_id == 0) return "ProposalSchedulingConfig.first_work_item_id must be non-zero";
  if (c.conservative_feature_family_mask == 0)
    return "ProposalSchedulingConfig.conservative_feature_family_mask must be non-zero";
  if ((c.conservative_feature_family_mask & ~kConservativeMask) != 0)
    return "ProposalSchedulingConfig.conservative_feature_family_mask has unknown bits";
  if (!std::isfinite(c.fallback_interval_t0) || !std::isfinite(c.fallback_interval_t1) || c.fallback_interval_t0 < 0.0 ||
      c.fallback_interval_t1 > 1.0 || c.fallback_interval_t0 > c.fallback_interval_t1)
    return "ProposalSchedulingConfig fallback interval must lie in [0, 1]";
  if (!std::isfinite(c.family_score_threshold)) return "ProposalSchedulingConfig.family_score_threshold must be finite";
  if (!std::isfinite(c.uncertainty_fallback_threshold) || c.uncertainty_fallback_threshold < 0.0F)
    return "ProposalSchedulingConfig.uncertainty_fallback_threshold must be non-negative";
  if (!std::isfinite(c.ood_abs_feature_threshold) || c.ood_abs_feature_threshold <= 0.0F)
    return "ProposalSchedulingConfig.ood_abs_feature_threshold must be positive";
  return "";
}

Status ScheduleRuntimeExactWorkItemsFromProposals(const std::vector<CandidateRecord>& candidates,
                                                  const std::vector<ProposalFeatureRow>& rows,
                                                  const std::vector<ProposalOutput>& outputs,
                                                  const FamilyMasksByQuery& masks,
                                                  const ProposalSchedulingConfig& config,
                                                  std::vector<ExactWorkItem>* work_queue, ProposalScheduleStats* stats) {
  if (work_queue == nullptr) return Status::Error("runtime exact work queue pointer is null");
  if (const std::string e = SchedulingConfigError(config); !e.empty()) return Status::Error(e);
  std::map<std::uint64_t, const ProposalFeatureRow*> row_of;
  for (const ProposalFeatureRow& r : rows) {
    if (r.candidate_id == 0U) return Status::Error("ProposalFeatureRow.candidate_id is required");
    if (!row_of.emplace(r.candidate_id, &r).second)
      return Status::Error("proposal runtime feature rows contain duplicate candidate_id");
  }
  std::map<std::uint64_t, const ProposalOutput*> out_of;
  for (const ProposalOutput& o : outputs) {
    if (o.candidate_id == 0U) return Status::Error("ProposalOutput.candidate_id is required");
    if (!out_of.emplace(o.candidate_id, &o).second)
      return Status::Error("proposal runtime outputs contain duplicate candidate_id");
  }
  ProposalScheduleStats st;
  st.raw_candidate_count = candidates.size();
  st.proposal_output_count = outputs.size();
  st.work_item_count = candidates.size();
  std::vector<ExactWorkItem> q;
  q.reserve(candidates.size());
  for (const CandidateRecord& c : candidates) {
    if (auto s = ValidateCandidateRecord(c); !s.ok) return s;
    const auto r = row_of.find(c.candidate_id);
    const auto o = out_of.find(c.candidate_id);
    q.push_back(MakeItem(c.candidate_id, c.query_id, c.slab_id, c.patch_a_id, c.patch_b_id,
                         static_cast<float>(c.rt_hit_count), r == row_of.end() ? nullptr : r->second,
                         o == out_of.end() ? nullptr : o->second, masks, config, true, &st));
  }
  if (!config.preserve_candidate_order) SortByPriority(&q);
  for (std::size_t i = 0; i < q.size(); ++i) {
    q[i].work_item_id = config.first_work_item_id + i;
    if (q[i].parent_candidate_id != candidates[i].candidate_id) ++st.reordered_count;
    if (auto s = ValidateExactWorkItem(q[i]); !s.ok) return s;
  }
  std::set<std::uint64_t> ids;
  for (const CandidateRecord& c : candidates)
    if (!ids.insert(c.candidate_id).second) return Status::Error("runtime candidates contain duplicate candidate_id");
  std::set<std::uint64_t> parents, items;
  for (const ExactWorkItem& it : q) {
    if (ids.count(it.parent_candidate_id) == 0)
      return Status::Error("runtime scheduled work item references an unknown candidate");
    if (!parents.insert(it.parent_candidate_id).second)
      return Status::Error("runtime scheduled work queue duplicates a parent candidate");
    if (!items.insert(it.work_item_id).second)
      return Status::Error("runtime scheduled work queue contains duplicate work_item_id");
    if ((it.feature_family_mask & kConservativeMask) == 0U)
      return Status::Error("runtime scheduled work item has no exact feature family enabled");
    if ((it.feature_family_mask & ~kConservativeMask) != 0U)
      return Status::Error("runtime scheduled work item has unknown exact feature-family bits");
  }
  st.monotonic_safe = true;
  *work_queue = std::move(q);
  if (stats != nullptr) *stats = st;
  return Status::Ok();
}

Status ScheduleRuntimeExactWorkItemsFromAlignedProposals(const std::vector<ProposalFeatureRow>& rows,
                                                         const std::vector<ProposalOutput>& outputs,
                                                         const FamilyMasksByQuery& masks,
                                                         const ProposalSchedulingConfig& config,
                                                         std::vector<ExactWorkItem>* work_queue,
                                                         ProposalScheduleStats* stats) {
  if (work_queue == nullptr) return Status::Error("runtime exact work queue pointer is null");
  if (rows.size() != outputs.size()) return Status::Error("feature row and proposal output array lengths must match");
  if (const std::string e = SchedulingConfigError(config); !e.empty()) return Status::Error(e);
  ProposalScheduleStats st;
  st.raw_candidate_count = rows.size();
  st.proposal_output_count = outputs.size();
  st.work_item_count = rows.size();
  std::vector<ExactWorkItem> q;
  q.reserve(rows.size());
  for (std::size_t i = 0; i < rows.size(); ++i) {
    const ProposalFeatureRow& r = rows[i];
    if (r.candidate_id == 0U || outputs[i].candidate_id != r.candidate_id)
      return Status::Error("proposal arrays must preserve candidate_id order");
    // the row's feature 3 (rt_hit_count of the candidate) is the fallback priority
    q.push_back(MakeItem(r.candidate_id, r.query_id, r.slab_id, r.patch_a_id, r.patch_b_id, r.features[3], &r,
                         &outputs[i], masks, config, false, &st));
  }
  if (!config.preserve_candidate_order) SortByPriority(&q);
  for (std::size_t i = 0; i < q.size(); ++i) {
    q[i].work_item_id = config.first_work_item_id + i;
    if (q[i].parent_candidate_id != rows[i].candidate_id) ++st.reordered_count;
    if (auto s = ValidateExactWorkItem(q[i]); !s.ok) return s;
  }
  st.monotonic_safe = true;
  *work_queue = std::move(q);
  if (stats != nullptr) *stats = st;
  return Status::Ok();
}

Status RunDenseGroupExactEarlyStop(const std::vector<DenseGroupCandidateInput>& cands,
                                   const DenseGroupEarlyStopConfig& config, DenseGroupEarlyStopStats* stats) {
  if (stats == nullptr) return Status::Error("DenseGroupEarlyStopStats pointer is null");
  if (config.representative_attempt_limit == 0U) return Status::Error("representative_attempt_limit must be positive");
  if (!std::isfinite(config.uncertainty_fallback_threshold) || config.uncertainty_fallback_threshold < 0.0F)
    return Status::Error("uncertainty_fallback_threshold must be finite and non-negative");
  if (!std::isfinite(config.interval_miss_penalty_scale) || config.interval_miss_penalty_scale < 0.0)
    return Status::Error("interval_miss_penalty_scale must be finite and non-negative");
  DenseGroupEarlyStopStats s;
  s.candidate_count = cands.size();
  const Clock::time_point t_total = Clock::now();
  for (const DenseGroupCandidateInput& c : cands) {
    if (c.query_id == 0U || c.candidate_id == 0U) return Status::Error("dense group candidate ids must be non-zero");
    const bool finite = std::isfinite(c.full_exact_cost) && std::isfinite(c.narrow_exact_cost) &&
                        std::isfinite(c.interval_t0) && std::isfinite(c.interval_t1) && std::isfinite(c.contact_t0) &&
                        std::isfinite(c.contact_t1) && std::isfinite(c.priority_score) && std::isfinite(c.cost_score) &&
                        std::isfinite(c.uncertainty_score);
    if (!finite) return Status::Error("dense group candidate contains a non-finite field");
  }
  const Clock::time_point t_sched = Clock::now();
  std::vector<std::size_t> order(cands.size());
  std::iota(order.begin(), order.end(), std::size_t{0});
  std::stable_sort(order.begin(), order.end(), [&](std::size_t a, std::size_t b) {
    const DenseGroupCandidateInput& x = cands[a];
    const DenseGroupCandidateInput& y = cands[b];
    if (x.query_id != y.query_id) return x.query_id < y.query_id;
    if (!config.preserve_input_order) {
      if (x.priority_score != y.priority_score) return x.priority_score > y.priority_score;
      if (x.cost_score != y.cost_score) return x.cost_score < y.cost_score;
    }
    return a < b;
  });
  for (std::size_t i = 0; i < order.size(); ++i) s.reordered_count += order[i] != i ? 1U : 0U;
  const Clock::time_point t_exact = Clock::now();
  s.schedule_ms = Ms(t_sched, t_exact);
  const double pen = config.interval_miss_penalty_scale;
  for (std::size_t b = 0; b < order.size();) {
    std::size_t e = b + 1;
    while (e < order.size() && cands[order[e]].query_id == cands[order[b]].query_id) ++e;
    ++s.group_count;
    bool truth = false;
    double group_work = 0.0;
    for (std::size_t k = b; k < e; ++k) {
      truth = truth || cands[order[k]].candidate_collides;
      group_work += PositiveCost(cands[order[k]].full_exact_cost, 1.0);
    }
    s.positive_group_count += truth ? 1U : 0U;
    s.no_proposal_exact_calls += e - b;
    s.no_proposal_exact_work += group_work;
    bool predicted = false, replayed = false;
    double before = 0.0;  // work before the first positive
    std::uint32_t attempts = 0;
    for (std::size_t k = b; k < e; ++k) {
      const DenseGroupCandidateInput& c = cands[order[k]];
      ++attempts;
      const double full = PositiveCost(c.full_exact_cost, 1.0);
      const double narrow = PositiveCost(c.narrow_exact_cost, full);
      if (c.uncertainty_score >= config.uncertainty_fallback_threshold) {
        // full fallback replay of this and every later candidate of the group
        replayed = true;
        for (std::size_t j = k; j < e; ++j) {
          const DenseGroupCandidateInput& f = cands[order[j]];
          ++s.learned_exact_calls;
          ++s.learned_fallback_calls;
          const double cost = PositiveCost(f.full_exact_cost, 1.0);
          s.learned_exact_work += cost;
          if (f.candidate_collides) {
            predicted = true;
            s.first_positive_rank_sum += static_cast<double>(j - b + 1);
            s.cost_before_first_positive_sum += before + cost;
            break;
          }
          before += cost;
        }
        break;
      }
      ++s.learned_exact_calls;
      s.learned_exact_work += narrow;
      if (!c.candidate_collides) {
        before += narrow;
        continue;
      }
      if (std::max(c.interval_t0, c.contact_t0) <= std::min(c.interval_t1, c.contact_t1)) {
        ++s.learned_interval_hit_count;
        predicted = true;
        s.first_positive_rank_sum += static_cast<double>(k - b + 1);
        s.cost_before_first_positive_sum += before + narrow;
        break;
      }
      ++s.learned_interval_miss_count;
      s.learned_exact_work += full * pen;
      if (attempts >= config.representative_attempt_limit) {
        ++s.learned_exact_calls;
        ++s.learned_fallback_calls;
        s.learned_exact_work += full;
        predicted = true;
        s.first_positive_rank_sum += static_cast<double>(k - b + 1);
        s.cost_before_first_positive_sum += before + narrow + full;
        break;
      }
      before += narrow + (full * pen);
    }
    s.high_uncertainty_group_count += replayed ? 1U : 0U;
    if (truth && predicted) ++s.tp;
    else if (!truth && !predicted) ++s.tn;
    else if (predicted) ++s.fp;
    else ++s.fn;
    b = e;
  }
  const Clock::time_point t_end = Clock::now();
  s.exact_ms = Ms(t_exact, t_end);
  s.total_ms = Ms(t_total, t_end);
  *stats = s;
  return Status::Ok();
}

}  // namespace p2cccd
