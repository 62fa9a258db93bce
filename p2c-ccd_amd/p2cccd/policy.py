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

All arrays are numpy; both functions are host-side bookkeeping over
per-candidate data.
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


def _check_config(c: SchedulingConfig) -> None:
    f32 = np.float32
    if c.first_work_item_id == 0:
        raise ValueError("ProposalSchedulingConfig.first_work_item_id must be non-zero")
    if c.conservative_feature_family_mask == 0:
        raise ValueError("ProposalSchedulingConfig.conservative_feature_family_mask must be non-zero")
    if c.conservative_feature_family_mask & ~CONSERVATIVE_MASK:
        raise ValueError("ProposalSchedulingConfig.conservative_feature_family_mask has unknown bits")
    t0, t1 = c.fallback_interval_t0, c.fallback_interval_t1
    if not (np.isfinite(t0) and np.isfinite(t1)) or t0 < 0.0 or t1 > 1.0 or t0 > t1:
        raise ValueError("ProposalSchedulingConfig fallback interval must lie in [0, 1]")
    if not np.isfinite(f32(c.family_score_threshold)):
        raise ValueError("ProposalSchedulingConfig.family_score_threshold must be finite")
    u = f32(c.uncertainty_fallback_threshold)
    if not np.isfinite(u) or u < 0:
        raise ValueError("ProposalSchedulingConfig.uncertainty_fallback_threshold must be non-negative")
    o = f32(c.ood_abs_feature_threshold)
    if not np.isfinite(o) or o <= 0:
        raise ValueError("ProposalSchedulingConfig.ood_abs_feature_threshold must be positive")


def _check_candidates(c: np.ndarray) -> None:
    """ValidateCandidateRecord (src/cpp/common/validators.cpp:160-181), first failure wins."""
    for i in range(len(c)):
        if c["schema_version"][i] != 1:
            raise ValueError("CandidateRecord.schema_version is unsupported")
        if c["candidate_id"][i] == 0:
            raise ValueError("CandidateRecord.candidate_id is required")
        if c["query_id"][i] == 0:
            raise ValueError("CandidateRecord.query_id is required")
        if c["proxy_type_a"][i] not in (1, 2):
            raise ValueError("CandidateRecord.proxy_type_a must be a concrete proxy type")
        if c["proxy_type_b"][i] not in (1, 2):
            raise ValueError("CandidateRecord.proxy_type_b must be a concrete proxy type")
        if c["rt_hit_count"][i] == 0:
            raise ValueError("CandidateRecord.rt_hit_count is required")
        m = c["motion_bound"][i]
        if not np.isfinite(m).all() or (m < 0).any():
            raise ValueError("CandidateRecord.motion_bound must contain finite non-negative values")


def _output_valid(o) -> bool:
    """HasNonFiniteProposal is false and ValidateProposalOutput passes."""
    vals = np.concatenate([o["interval_scores"], o["family_scores"],
                           [o["priority_score"], o["cost_score"], o["uncertainty_score"]]])
    return bool(np.isfinite(vals).all() and o["candidate_id"] != 0 and o["cost_score"] >= 0
                and o["uncertainty_score"] >= 0)


def schedule_runtime_exact_work_items(candidates: np.ndarray, rows: np.ndarray, outputs: np.ndarray,
                                      family_masks_by_query: dict | None = None,
                                      config: SchedulingConfig | None = None):
    cfg = config or SchedulingConfig()
    _check_config(cfg)
    masks = family_masks_by_query or {}
    row_of: dict = {}
    for r in rows:
        cid = int(r["candidate_id"])
        if cid == 0:
            raise ValueError("ProposalFeatureRow.candidate_id is required")
        if cid in row_of:
            raise ValueError("proposal runtime feature rows contain duplicate candidate_id")
        row_of[cid] = r
    out_of: dict = {}
    for o in outputs:
        cid = int(o["candidate_id"])
        if cid == 0:
            raise ValueError("ProposalOutput.candidate_id is required")
        if cid in out_of:
            raise ValueError("proposal runtime outputs contain duplicate candidate_id")
        out_of[cid] = o
    _check_candidates(candidates)
    n = len(candidates)
    items = np.zeros(n, WORK_ITEM_DTYPE)
    stats = dict(raw_candidate_count=n, proposal_output_count=len(outputs), work_item_count=n, fallback_count=0,
                 missing_proposal_fallback_count=0, invalid_proposal_fallback_count=0, ood_fallback_count=0,
                 high_uncertainty_fallback_count=0, reordered_count=0, monotonic_safe=False)
    thr_fam = np.float32(cfg.family_score_threshold)
    thr_unc = np.float32(cfg.uncertainty_fallback_threshold)
    thr_ood = np.float32(cfg.ood_abs_feature_threshold)
    for i, c in enumerate(candidates):
        cid = int(c["candidate_id"])
        row, out = row_of.get(cid), out_of.get(cid)
        missing = out is None
        invalid = out is not None and not _output_valid(out)
        ood = row is None or bool((~np.isfinite(row["features"]) | (np.abs(row["features"]) > thr_ood)).any())
        high_unc = out is not None and bool(np.isfinite(out["uncertainty_score"])) and \
            out["uncertainty_score"] >= thr_unc
        fallback = missing or invalid or ood or high_unc
        stats["fallback_count"] += fallback
        stats["missing_proposal_fallback_count"] += missing
        stats["invalid_proposal_fallback_count"] += invalid
        stats["ood_fallback_count"] += ood
        stats["high_uncertainty_fallback_count"] += high_unc
        q = int(c["query_id"])
        base = masks.get(q, 0) or cfg.conservative_feature_family_mask
        it = items[i]
        it["parent_candidate_id"] = cid
        it["query_id"] = q
        it["slab_id"], it["patch_a_id"], it["patch_b_id"] = c["slab_id"], c["patch_a_id"], c["patch_b_id"]
        it["interval_t0"], it["interval_t1"] = cfg.fallback_interval_t0, cfg.fallback_interval_t1
        it["feature_family_mask"] = base
        it["priority_score"] = np.float32(c["rt_hit_count"])
        it["source"] = SOURCE_FALLBACK
        if not fallback:
            pred = (FAMILY_PT if out["family_scores"][0] >= thr_fam else 0) | \
                (FAMILY_EE if out["family_scores"][1] >= thr_fam else 0)
            it["feature_family_mask"] = base | pred
            it["priority_score"] = out["priority_score"]
            it["source"] = SOURCE_REFINED
    if not cfg.preserve_candidate_order and n >= 2 and bool((items["priority_score"][:-1] <
                                                             items["priority_score"][1:]).any()):
        items = items[np.argsort(-items["priority_score"], kind="stable")]
    items["work_item_id"] = cfg.first_work_item_id + np.arange(n, dtype=np.uint64)
    stats["reordered_count"] = int((items["parent_candidate_id"] != candidates["candidate_id"]).sum()) if n else 0
    for it in items:
        if it["priority_score"] != it["priority_score"] or not np.isfinite(it["priority_score"]):
            raise ValueError("ExactWorkItem.priority_score must be finite")
    ids = candidates["candidate_id"]
    if len(np.unique(ids)) != n:
        raise ValueError("runtime candidates contain duplicate candidate_id")
    fam = items["feature_family_mask"]
    if ((fam & CONSERVATIVE_MASK) == 0).any():
        raise ValueError("runtime scheduled work item has no exact feature family enabled")
    if (fam & ~np.uint32(CONSERVATIVE_MASK)).any():
        raise ValueError("runtime scheduled work item has unknown exact feature-family bits")
    stats["monotonic_safe"] = True
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
    cfg = config or DenseGroupConfig()
    if cfg.representative_attempt_limit == 0:
        raise ValueError("representative_attempt_limit must be positive")
    thr = np.float32(cfg.uncertainty_fallback_threshold)
    if not np.isfinite(thr) or thr < 0:
        raise ValueError("uncertainty_fallback_threshold must be finite and non-negative")
    pen = cfg.interval_miss_penalty_scale
    if not np.isfinite(pen) or pen < 0:
        raise ValueError("interval_miss_penalty_scale must be finite and non-negative")
    n = len(cands)
    if n and ((cands["query_id"] == 0) | (cands["candidate_id"] == 0)).any():
        raise ValueError("dense group candidate ids must be non-zero")
    for f in ("full_exact_cost", "narrow_exact_cost", "interval_t0", "interval_t1", "contact_t0", "contact_t1",
              "priority_score", "cost_score", "uncertainty_score"):
        if n and not np.isfinite(cands[f]).all():
            raise ValueError("dense group candidate contains a non-finite field")
    idx = np.arange(n)
    if cfg.preserve_input_order:
        order = np.lexsort((idx, cands["query_id"]))
    else:
        order = np.lexsort((idx, cands["cost_score"], -cands["priority_score"].astype(np.float64), cands["query_id"]))
    s = dict(candidate_count=n, group_count=0, positive_group_count=0, no_proposal_exact_calls=0,
             no_proposal_exact_work=0.0, learned_exact_calls=0, learned_fallback_calls=0,
             learned_interval_hit_count=0, learned_interval_miss_count=0, learned_exact_work=0.0,
             first_positive_rank_sum=0.0, cost_before_first_positive_sum=0.0, tp=0, tn=0, fp=0, fn=0,
             reordered_count=int((order != idx).sum()), high_uncertainty_group_count=0)
    c = cands[order]
    q = c["query_id"]
    starts = np.flatnonzero(np.r_[True, q[1:] != q[:-1]]) if n else np.zeros(0, int)
    ends = np.r_[starts[1:], n] if n else starts
    for b, e in zip(starts, ends):
        g = c[b:e]
        s["group_count"] += 1
        truth = bool(g["candidate_collides"].any())
        s["positive_group_count"] += truth
        s["no_proposal_exact_calls"] += int(e - b)
        group_work = 0.0  # summed per group, then added (the reference's order)
        for x in g["full_exact_cost"]:
            group_work += _positive(float(x), 1.0)
        s["no_proposal_exact_work"] += group_work
        predicted, resolved, used_unc = False, False, False
        work = 0.0
        attempts = 0
        for k in range(e - b):
            cand = g[k]
            attempts += 1
            full = _positive(float(cand["full_exact_cost"]), 1.0)
            narrow = _positive(float(cand["narrow_exact_cost"]), full)
            if cand["uncertainty_score"] >= thr:  # full fallback from here on
                used_unc = True
                for j in range(k, e - b):
                    fc = g[j]
                    s["learned_exact_calls"] += 1
                    s["learned_fallback_calls"] += 1
                    cost = _positive(float(fc["full_exact_cost"]), 1.0)
                    s["learned_exact_work"] += cost
                    if fc["candidate_collides"]:
                        predicted = True
                        s["first_positive_rank_sum"] += float(j + 1)
                        s["cost_before_first_positive_sum"] += work + cost
                        break
                    work += cost
                resolved = True
                break
            s["learned_exact_calls"] += 1
            s["learned_exact_work"] += narrow
            if cand["candidate_collides"]:
                if max(cand["interval_t0"], cand["contact_t0"]) <= min(cand["interval_t1"], cand["contact_t1"]):
                    s["learned_interval_hit_count"] += 1
                    predicted = resolved = True
                    s["first_positive_rank_sum"] += float(k + 1)
                    s["cost_before_first_positive_sum"] += work + narrow
                    break
                s["learned_interval_miss_count"] += 1
                s["learned_exact_work"] += full * pen
                if attempts >= cfg.representative_attempt_limit:
                    s["learned_exact_calls"] += 1
                    s["learned_fallback_calls"] += 1
                    s["learned_exact_work"] += full
                    predicted = resolved = True
                    s["first_positive_rank_sum"] += float(k + 1)
                    s["cost_before_first_positive_sum"] += work + narrow + full
                    break
                work += narrow + full * pen
                continue
            work += narrow
        s["high_uncertainty_group_count"] += used_unc
        if truth and predicted:
            s["tp"] += 1
        elif not truth and not predicted:
            s["tn"] += 1
        elif predicted:
            s["fp"] += 1
        else:
            s["fn"] += 1
    return s


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
