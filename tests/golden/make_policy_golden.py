"""Writes tests/golden/policy_golden.npz from the reference's own compiled
ScheduleRuntimeExactWorkItemsFromProposals (src/cpp/proposal/proposal_policy.cpp
:595-742) and RunDenseGroupExactEarlyStop (src/cpp/common/group_scheduler.cpp:
60-239), built in place into oracle/_ref by `make -C oracle ref`.

Run in the container that holds the reference:  python tests/golden/make_policy_golden.py
Inputs are seeded: candidates with every fallback cause (missing, invalid,
out-of-distribution and high-uncertainty proposals), per-query family
masks, tied priorities, kept and sorted orders; dense groups of 1-20
candidates with contacts inside and outside the proposed intervals.  Error
cases store the reference's message.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path[:0] = [str(ROOT), str(ROOT / "p2c-ccd_amd")]

import oracle  # noqa: E402
from p2cccd import policy, proxy  # noqa: E402

OUT = Path(__file__).resolve().parent / "policy_golden.npz"


def make_case(rng, n):
    c = np.zeros(n, proxy.CANDIDATE_DTYPE)
    c["schema_version"] = 1
    c["candidate_id"] = rng.permutation(np.arange(1, n + 1)) * 3
    c["query_id"] = rng.integers(1, 5, n)
    c["slab_id"] = rng.integers(0, 3, n)
    c["object_a_id"], c["object_b_id"] = rng.integers(1, 4, n), rng.integers(4, 8, n)
    c["patch_a_id"], c["patch_b_id"] = rng.integers(0, 9, n), rng.integers(0, 9, n)
    c["proxy_type_a"], c["proxy_type_b"] = rng.integers(1, 3, n), rng.integers(1, 3, n)
    c["rt_hit_count"] = rng.integers(1, 4, n)
    c["motion_bound"] = rng.uniform(0, 1, (n, 4))
    keep = rng.random(n) > 0.1
    rows = np.zeros(int(keep.sum()), policy.ROW_DTYPE)
    rows["candidate_id"] = c["candidate_id"][keep]
    rows["features"] = rng.normal(0, 10, (len(rows), 32))
    rows["features"][rng.random(len(rows)) < 0.05, 3] = 3e5  # out of distribution
    has = rng.random(n) > 0.15
    outs = np.zeros(int(has.sum()), policy.OUTPUT_DTYPE)
    outs["candidate_id"] = c["candidate_id"][has]
    outs["interval_scores"] = rng.uniform(0, 1, (len(outs), 8))
    outs["family_scores"] = rng.uniform(0, 1, (len(outs), 8))
    outs["priority_score"] = rng.integers(0, 6, len(outs)) / 4.0  # ties
    outs["cost_score"] = rng.uniform(0, 2, len(outs))
    outs["uncertainty_score"] = rng.uniform(0, 1, len(outs))
    bad = rng.random(len(outs))
    outs["interval_scores"][bad < 0.04, 2] = np.nan
    outs["cost_score"][(bad >= 0.04) & (bad < 0.08)] = -1.0
    masks = {1: 1, 2: 0, 3: 2}
    return c, rows, outs, masks


def main():
    rng = np.random.default_rng(77)
    data = {}
    k = 0
    for n, cfgkw in ((40, {}), (200, {"preserve_candidate_order": True}), (150, {"family_score_threshold": 0.8,
                     "uncertainty_fallback_threshold": 0.5, "first_work_item_id": 1000})):
        c, rows, outs, masks = make_case(rng, n)
        cfg = policy.SchedulingConfig(**cfgkw).__dict__
        wu, wi, wp, st = oracle.ref_schedule_runtime(c, rows, outs, masks, cfg)
        data.update({f"p{k}_cand": c, f"p{k}_rows": rows, f"p{k}_outs": outs,
                     f"p{k}_mask_q": np.array(list(masks), np.uint64), f"p{k}_mask_v": np.array(list(masks.values()),
                                                                                                 np.uint32),
                     f"p{k}_cfg": np.array([repr(cfgkw)]), f"p{k}_wu": wu, f"p{k}_wi": wi, f"p{k}_wp": wp,
                     f"p{k}_stats": st})
        print("policy case", k, n, "fallbacks", st[3])
        k += 1
    data["p_count"] = np.array(k)
    # error cases
    errs = []
    c, rows, outs, masks = make_case(rng, 10)
    r2 = np.concatenate([rows, rows[:1]])
    o2 = outs.copy()
    o2["candidate_id"][0] = 0
    c2 = c.copy()
    c2["rt_hit_count"][3] = 0
    for name, args in (("dup_rows", (c, r2, outs, masks, {})), ("zero_output", (c, rows, o2, masks, {})),
                       ("bad_rt", (c2, rows, outs, masks, {})), ("bad_cfg", (c, rows, outs, masks, {"first_work_item_id": 0})),
                       ("bad_interval", (c, rows, outs, masks, {"fallback_interval_t1": 1.5}))):
        cfg = policy.SchedulingConfig(**args[4]).__dict__
        try:
            oracle.ref_schedule_runtime(args[0], args[1], args[2], args[3], cfg)
            msg = ""
        except ValueError as e:
            msg = str(e)
        data[f"e_{name}_cand"], data[f"e_{name}_rows"], data[f"e_{name}_outs"] = args[0], args[1], args[2]
        data[f"e_{name}_cfg"] = np.array([repr(args[4])])
        data[f"e_{name}_msg"] = np.array([msg])
        errs.append(name)
        print("error case", name, "->", msg)
    data["e_names"] = np.array(errs)
    # dense groups
    k = 0
    for n, cfgkw in ((300, {}), (300, {"preserve_input_order": True, "representative_attempt_limit": 1}),
                     (500, {"uncertainty_fallback_threshold": 0.3, "interval_miss_penalty_scale": 0.5})):
        d = np.zeros(n, policy.DENSE_DTYPE)
        d["query_id"] = np.sort(rng.integers(1, n // 8, n))[rng.permutation(n)]
        d["candidate_id"] = np.arange(1, n + 1)
        d["full_exact_cost"] = rng.uniform(-0.5, 3, n)
        d["narrow_exact_cost"] = rng.uniform(-0.5, 2, n)
        d["interval_t0"] = rng.uniform(0, 0.5, n)
        d["interval_t1"] = d["interval_t0"] + rng.uniform(0, 0.5, n)
        d["contact_t0"] = rng.uniform(0, 1, n)
        d["contact_t1"] = d["contact_t0"] + rng.uniform(0, 0.1, n)
        d["priority_score"] = rng.integers(0, 5, n) / 4.0
        d["cost_score"] = rng.integers(0, 3, n) / 2.0
        d["uncertainty_score"] = rng.uniform(0, 1, n)
        d["candidate_collides"] = rng.random(n) < 0.2
        cfg = {**policy.DenseGroupConfig().__dict__, **cfgkw}
        su, sd = oracle.ref_dense_group(d, cfg)
        data.update({f"d{k}_in": d, f"d{k}_cfg": np.array([repr(cfgkw)]), f"d{k}_su": su, f"d{k}_sd": sd})
        print("dense case", k, su[:4])
        k += 1
    data["d_count"] = np.array(k)
    np.savez_compressed(OUT, **data)


if __name__ == "__main__":
    main()
