"""The proposal policy and the dense-group early stop against the reference's
compiled ScheduleRuntimeExactWorkItemsFromProposals and
RunDenseGroupExactEarlyStop (golden vectors, tests/golden/policy_golden.npz,
written by tests/golden/make_policy_golden.py): every work-queue field, every
statistic and every error message."""
import ast
from pathlib import Path

import numpy as np
import pytest

from p2cccd import policy

G = np.load(Path(__file__).resolve().parent / "golden" / "policy_golden.npz")


def _cfg(key, cls):
    return cls(**ast.literal_eval(str(G[key][0])))


@pytest.mark.parametrize("k", range(int(G["p_count"])))
def test_schedule_matches_reference(k):
    masks = dict(zip(G[f"p{k}_mask_q"].tolist(), G[f"p{k}_mask_v"].tolist()))
    items, st = policy.schedule_runtime_exact_work_items(G[f"p{k}_cand"], G[f"p{k}_rows"], G[f"p{k}_outs"], masks,
                                                         _cfg(f"p{k}_cfg", policy.SchedulingConfig))
    wu, wi, wp = G[f"p{k}_wu"], G[f"p{k}_wi"], G[f"p{k}_wp"]
    cols = ("work_item_id", "parent_candidate_id", "query_id", "slab_id", "patch_a_id", "patch_b_id",
            "feature_family_mask", "source")
    for j, f in enumerate(cols):
        assert np.array_equal(items[f].astype(np.uint64), wu[:, j]), f
    assert np.array_equal(items["interval_t0"], wi[:, 0]) and np.array_equal(items["interval_t1"], wi[:, 1])
    assert np.array_equal(items["priority_score"], wp)
    names = ("raw_candidate_count", "proposal_output_count", "work_item_count", "fallback_count",
             "missing_proposal_fallback_count", "invalid_proposal_fallback_count", "ood_fallback_count",
             "high_uncertainty_fallback_count", "reordered_count", "monotonic_safe")
    assert [int(st[n]) for n in names] == G[f"p{k}_stats"].tolist()


@pytest.mark.parametrize("name", [str(n) for n in G["e_names"]])
def test_errors_match_reference(name):
    msg = str(G[f"e_{name}_msg"][0])
    assert msg
    with pytest.raises(ValueError) as e:
        policy.schedule_runtime_exact_work_items(G[f"e_{name}_cand"], G[f"e_{name}_rows"], G[f"e_{name}_outs"],
                                                 {1: 1, 2: 0, 3: 2}, _cfg(f"e_{name}_cfg", policy.SchedulingConfig))
    assert str(e.value) == msg


@pytest.mark.parametrize("k", range(int(G["d_count"])))
def test_dense_group_early_stop_matches_reference(k):
    st = policy.run_dense_group_exact_early_stop(G[f"d{k}_in"], _cfg(f"d{k}_cfg", policy.DenseGroupConfig))
    su, sd = G[f"d{k}_su"], G[f"d{k}_sd"]
    counts = ("candidate_count", "group_count", "positive_group_count", "no_proposal_exact_calls",
              "learned_exact_calls", "learned_fallback_calls", "learned_interval_hit_count",
              "learned_interval_miss_count", "tp", "tn", "fp", "fn", "reordered_count",
              "high_uncertainty_group_count")
    assert [int(st[c]) for c in counts] == su.tolist()
    sums = ("no_proposal_exact_work", "learned_exact_work", "first_positive_rank_sum", "cost_before_first_positive_sum")
    assert [st[s] for s in sums] == sd.tolist()
