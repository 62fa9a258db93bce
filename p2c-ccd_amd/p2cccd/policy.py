"""Proposal policy: turning candidates plus proposal outputs into the exact
work queue, and the dense-group early-stop evaluation.

schedule_runtime_exact_work_items follows ScheduleRuntimeExactWorkItemsFromProposals
(src/cpp/proposal/proposal_policy.cpp:595-742): a candidate falls back when
its proposal is missing, invalid (non-finite or failing the output checks),
out of distribution (a missing feature row, or a feature that is non-finite
or above the configured magnitude) or too uncertain; a fallback item keeps
the conservative family mask and takes rt_hit_count as its priority, a
refined item adds the predicted families and takes the proposal's priority.
Items are stably sorted by descending priority unless the order is kept,
then numbered, validated and checked for conservation (one item per
candidate, no unknown family bits), with the reference's messages.

run_dense_group_exact_early_stop follows RunDenseGroupExactEarlyStop
(src/cpp/common/group_scheduler.cpp:60-239): per query group in priority
order (ties: lower cost, then input order), certify representatives until a
contact inside the proposed interval, a representative limit, or a
high-uncertainty candidate sends the group to the full fallback replay
(:144-162); it reports exact calls, work, first-positive ranks and the
confusion counts.  dense_group_from_certificates builds its inputs from the
GPU certificate's own outcomes (contact flag, TOI, boxes as cost).

Both run in the native module (p2cccd_cpp: ScheduleRuntimeExactWorkItemsFromProposals
and RunDenseGroupExactEarlyStop of bindings/policy_hip.cpp, also exposed under
the reference's pybind names schedule_runtime_exact_work_items* and
run_native_dense_group_exact_early_stop); this module is their numpy façade:
structured arrays in, structured arrays and dictionaries out.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

FAMILY_PT, FAMILY_EE = 1, 2
CONSERVATIVE_MASK = FAMILY_PT | FAMILY_EE
SOURCE_RAW, SOURCE_REFINED, SOURCE_FALLBACK = 0, 1, 2

ROW_DTYPE = np.dtype([("candidate_id", "<u8"), ("features", "<f4", (32,))])
OUTPUT_DTYPE = np.dtype([("candidate_id", "<u8"), ("interval_scores", "<f4", (8,)), ("family_scores", "<f4", (8,)),
                         ("priority_score", "<f4"), ("cost_score", "<f4"), ("uncertainty_score", "<f4")])
WORK_ITEM_DTYPE = np.dtype([("work_item_id", "<u8"), ("parent_candidate_id", "<u8"), ("query_id", "<u8"),
                            ("slab_id", "<u4"), ("patch_a_id", "<u4"), ("patch_b_id", "<u4"),
                            ("interval_t0", "<f8"), ("interval_t1", "<f8"), ("feature_family_mask", "<u4"),
                            ("topk_feature_ids_offset", "<u4"), ("depth", "<u2"), ("priority_score", "<f4"),
                            ("source", "u1")])
DENSE_DTYPE = np.dtype([("query_id", "<u8"), ("candidate_id", "<u8"), ("full_exact_cost", "<f8"),
                        ("narrow_exact_cost", "<f8"), ("interval_t0", "<f8"), ("interval_t1", "<f8"),
                        ("contact_t0", "<f8"), ("contact_t1", "<f8"), ("priority_score", "<f4"),
                        ("cost_score", "<f4"), ("uncertainty_score", "<f4"), ("candidate_collides", "?")])


@dataclass
class SchedulingConfig:
    first_work_item_id: int = 1
    conservative_feature_family_mask: int = CONSERVATIVE_MASK
    fallback_interval_t0: float = 0.0
    fallback_interval_t1: float = 1.0
    family_score_threshold: float = 0.5
    uncertainty_fallback_threshold: float = 0.95
    ood_abs_feature_threshold: float = 1.0e5
    preserve_candidate_order: bool = False


def _cpp():
    import p2cccd_cpp

    return p2cccd_cpp


def _columns(a: np.ndarray) -> dict:
    return {name: np.ascontiguousarray(a[name]) for name in a.dtype.names}


def schedule_runtime_exact_work_items(candidates: np.ndarray, rows: np.ndarray, outputs: np.ndarray,
                                      family_masks_by_query: dict | None = None,
                                      config: SchedulingConfig | None = None):
    """(work items WORK_ITEM_DTYPE[n], statistics dict); ValueError with the
    reference's message when an input or the result fails its checks."""
    cfg = config or SchedulingConfig()
    m = _cpp()
    c = m.ProposalSchedulingConfig()
    c.first_work_item_id = int(cfg.first_work_item_id)
    c.conservative_feature_family_mask = int(cfg.conservative_feature_family_mask)
    c.fallback_interval_t0, c.fallback_interval_t1 = float(cfg.fallback_interval_t0), float(cfg.fallback_interval_t1)
    c.family_score_threshold = float(cfg.family_score_threshold)
    c.uncertainty_fallback_threshold = float(cfg.uncertainty_fallback_threshold)
    c.ood_abs_feature_threshold = float(cfg.ood_abs_feature_threshold)
    c.preserve_candidate_order = bool(cfg.preserve_candidate_order)
    masks = {int(k): int(v) for k, v in (family_masks_by_query or {}).items()}
    cols, stats = m._schedule_runtime_columns(_columns(candidates), _columns(rows), _columns(outputs), masks, c)
    items = np.zeros(len(cols["work_item_id"]), WORK_ITEM_DTYPE)
    for name in WORK_ITEM_DTYPE.names:
        items[name] = cols[name]
    return items, stats


@dataclass
class DenseGroupConfig:
    uncertainty_fallback_threshold: float = 0.75
    representative_attempt_limit: int = 3
    interval_miss_penalty_scale: float = 0.22
    preserve_input_order: bool = False


def _positive(v: float, fallback: float) -> float:
    return v if (np.isfinite(v) and v > 0.0) else fallback


def run_dense_group_exact_early_stop(cands: np.ndarray, config: DenseGroupConfig | None = None) -> dict:
    """Statistics of RunDenseGroupExactEarlyStop over DENSE_DTYPE inputs."""
    cfg = config or DenseGroupConfig()
    cols = _columns(cands)
    cols["candidate_collides"] = cols["candidate_collides"].astype(np.uint8)
    return _cpp()._dense_group_columns(cols, float(cfg.uncertainty_fallback_threshold),
                                       int(cfg.representative_attempt_limit), float(cfg.interval_miss_penalty_scale),
                                       bool(cfg.preserve_input_order))


def dense_group_from_certificates(query_id: np.ndarray, hit: np.ndarray, toi: np.ndarray, boxes: np.ndarray,
                                  priority: np.ndarray, cost: np.ndarray | None = None,
                                  uncertainty: np.ndarray | None = None, interval: tuple = (0.0, 1.0)) -> np.ndarray:
    """Dense-group inputs from GPU certificate outcomes: a candidate collides
    when its certificate found a contact; the contact interval is [toi, toi];
    the exact cost of a candidate is the number of boxes its certificate took
    (the narrow cost equals the full one: the certificate already runs on the
    whole step)."""
    n = len(hit)
    out = np.zeros(n, DENSE_DTYPE)
    out["query_id"] = query_id
    out["candidate_id"] = np.arange(1, n + 1, dtype=np.uint64)
    out["full_exact_cost"] = np.maximum(boxes.astype(np.float64), 1.0)
    out["narrow_exact_cost"] = out["full_exact_cost"]
    out["interval_t0"], out["interval_t1"] = interval
    hit = hit.astype(bool)
    out["contact_t0"] = np.where(hit, toi, 0.0)
    out["contact_t1"] = np.where(hit, toi, 0.0)
    out["priority_score"] = priority
    out["cost_score"] = 0.0 if cost is None else cost
    out["uncertainty_score"] = 0.0 if uncertainty is None else uncertainty
    out["candidate_collides"] = hit
    return out
