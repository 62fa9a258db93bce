// Proposal policy and dense-group early stop of the reference, kept type for
// type and message for message so C++ and pybind callers drop in:
//   src/cpp/proposal/proposal_features.h:27-43 (ProposalFeatureRow),
//   src/cpp/common/runtime_contracts.h:58-65 (ProposalOutput),
//   src/cpp/proposal/proposal_policy.h:13-36 (config, stats),
//   src/cpp/common/group_scheduler.h:10-57 (dense-group types).
// Both are host bookkeeping over per-candidate records: the schedule turns
// candidates plus proposal outputs into the exact work queue the device
// certifies; the dense-group stop evaluates group-wise any-hit orders over
// certificate outcomes (the GPU certificate's, in this build's tests).
#pragma once

#include <array>
#include <cstdint>
#include <unordered_map>
#include <vector>

#include "certificate_hip.h"
#include "proxy_hip.h"

namespace p2cccd {

inline constexpr std::size_t kProposalFeatureDimension = 32;
inline constexpr std::size_t kProposalIntervalBinCount = 8;
inline constexpr std::size_t kProposalFamilyCount = 8;

struct ProposalFeatureRow {
  std::uint32_t schema_version = 1;
  std::uint64_t query_id = 0;
  std::uint64_t candidate_id = 0;
  std::uint32_t slab_id = 0;
  std::uint32_t object_a_id = 0;
  std::uint32_t patch_a_id = 0;
  std::uint32_t object_b_id = 0;
  std::uint32_t patch_b_id = 0;
  std::array<float, kProposalFeatureDimension> features{};
  std::array<float, kProposalIntervalBinCount> interval_targets{};
  std::array<float, kProposalFamilyCount> family_targets{};
  float priority_target = 0.0F;
  float cost_target = 0.0F;
  float uncertainty_target = 0.0F;
  std::uint32_t target_mask = 0;
};

struct ProposalOutput {
  std::uint64_t candidate_id = 0;
  std::array<float, kProposalIntervalBinCount> interval_scores{};
  std::array<float, kProposalFamilyCount> family_scores{};
  float priority_score = 0.0F;
  float cost_score = 0.0F;
  float uncertainty_score = 0.0F;
};

struct ProposalSchedulingConfig {
  std::uint64_t first_work_item_id = 1;
  std::uint32_t conservative_feature_family_mask = kFeatureFamilyPointTriangle | kFeatureFamilyEdgeEdge;
  double fallback_interval_t0 = 0.0;
  double fallback_interval_t1 = 1.0;
  float family_score_threshold = 0.5F;
  float uncertainty_fallback_threshold = 0.95F;
  float ood_abs_feature_threshold = 1.0e5F;
  bool preserve_candidate_order = false;
};

struct ProposalScheduleStats {
  std::uint64_t raw_candidate_count = 0;
  std::uint64_t proposal_output_count = 0;
  std::uint64_t work_item_count = 0;
  std::uint64_t fallback_count = 0;
  std::uint64_t missing_proposal_fallback_count = 0;
  std::uint64_t invalid_proposal_fallback_count = 0;
  std::uint64_t ood_fallback_count = 0;
  std::uint64_t high_uncertainty_fallback_count = 0;
  std::uint64_t reordered_count = 0;
  bool monotonic_safe = false;
};

struct DenseGroupCandidateInput {
  std::uint64_t query_id = 0;
  std::uint64_t candidate_id = 0;
  double full_exact_cost = 1.0;
  double narrow_exact_cost = 1.0;
  double interval_t0 = 0.0;
  double interval_t1 = 1.0;
  double contact_t0 = 0.0;
  double contact_t1 = 0.0;
  float priority_score = 0.0F;
  float cost_score = 0.0F;
  float uncertainty_score = 0.0F;
  bool candidate_collides = false;
};

struct DenseGroupEarlyStopConfig {
  float uncertainty_fallback_threshold = 0.75F;
  std::uint32_t representative_attempt_limit = 3U;
  double interval_miss_penalty_scale = 0.22;
  bool preserve_input_order = false;
};

struct DenseGroupEarlyStopStats {
  std::uint64_t candidate_count = 0;
  std::uint64_t group_count = 0;
  std::uint64_t positive_group_count = 0;
  std::uint64_t no_proposal_exact_calls = 0;
  double no_proposal_exact_work = 0.0;
  std::uint64_t learned_exact_calls = 0;
  std::uint64_t learned_fallback_calls = 0;
  std::uint64_t learned_interval_hit_count = 0;
  std::uint64_t learned_interval_miss_count = 0;
  double learned_exact_work = 0.0;
  double first_positive_rank_sum = 0.0;
  double cost_before_first_positive_sum = 0.0;
  std::uint64_t tp = 0;
  std::uint64_t tn = 0;
  std::uint64_t fp = 0;
  std::uint64_t fn = 0;
  std::uint64_t reordered_count = 0;
  std::uint64_t high_uncertainty_group_count = 0;
  double schedule_ms = 0.0;
  double exact_ms = 0.0;
  double total_ms = 0.0;
};

using FamilyMasksByQuery = std::unordered_map<std::uint64_t, std::uint32_t>;

// ValidateCandidateRecord / ValidateProposalOutput (validators.cpp:160-205).
Status ValidateCandidateRecord(const CandidateRecord& record);
Status ValidateProposalOutput(const ProposalOutput& output);
// IsValidSchedulingConfig (proposal_policy.cpp:87-122): "" when valid, else the message.
std::string SchedulingConfigError(const ProposalSchedulingConfig& config);

// ScheduleRuntimeExactWorkItemsFromProposals (proposal_policy.cpp:595-742).
Status ScheduleRuntimeExactWorkItemsFromProposals(const std::vector<CandidateRecord>& candidates,
                                                  const std::vector<ProposalFeatureRow>& rows,
                                                  const std::vector<ProposalOutput>& proposal_outputs,
                                                  const FamilyMasksByQuery& conservative_family_masks_by_query_id,
                                                  const ProposalSchedulingConfig& config,
                                                  std::vector<ExactWorkItem>* work_queue,
                                                  ProposalScheduleStats* stats);
// The row-aligned array form of the reference's pybind layer
// (py_module.cpp:1088-1177): row i and output i describe one candidate.
Status ScheduleRuntimeExactWorkItemsFromAlignedProposals(const std::vector<ProposalFeatureRow>& rows,
                                                         const std::vector<ProposalOutput>& proposal_outputs,
                                                         const FamilyMasksByQuery& conservative_family_masks_by_query_id,
                                                         const ProposalSchedulingConfig& config,
                                                         std::vector<ExactWorkItem>* work_queue,
                                                         ProposalScheduleStats* stats);
// RunDenseGroupExactEarlyStop (group_scheduler.cpp:60-239).
Status RunDenseGroupExactEarlyStop(const std::vector<DenseGroupCandidateInput>& candidates,
                                   const DenseGroupEarlyStopConfig& config, DenseGroupEarlyStopStats* stats);

}  // namespace p2cccd
