"""The reference's pybind names for the proposal schedule and the dense-group
stop (src/cpp/bindings/py_module.cpp:1912-1945) on the host:
- schedule_runtime_exact_work_items over record objects equals the golden
  work queues of the compiled reference (tests/golden/policy_golden.npz) and
  raises its messages;
- the array forms (_from_arrays, _from_proposal_arrays) equal the object form
  on aligned inputs, where the fallback priority (feature 3) is the
  candidate's rt_hit_count;
- run_native_dense_group_exact_early_stop equals RunDenseGroupExactEarlyStop
  on the inputs its array conversion defines (best interval bin, narrow cost =
  full cost x cost_target), and the dense-group statistics match the golden
  ones."""
import ast
from pathlib import Path
from types import SimpleNamespace as NS

import numpy as np
import pytest

from p2cccd import policy

cpp = pytest.importorskip("p2cccd_cpp")
G = np.load(Path(__file__).resolve().parent / "golden" / "policy_golden.npz")


def _cfg(key):
    d = ast.literal_eval(str(G[key][0]))
    c = cpp.ProposalSchedulingConfig()
    for k, v in d.items():
        setattr(c, k, v)
    return c


def _cand_objects(c):
    return [NS(schema_version=int(r["schema_version"]), candidate_id=int(r["candidate_id"]), query_id=int(r["query_id"]),
               slab_id=int(r["slab_id"]), object_a_id=int(r["object_a_id"]), patch_a_id=int(r["patch_a_id"]),
               object_b_id=int(r["object_b_id"]), patch_b_id=int(r["patch_b_id"]),
               proxy_type_a=int(r["proxy_type_a"]), proxy_type_b=int(r["proxy_type_b"]),
               rt_hit_count=int(r["rt_hit_count"]), motion_bound=r["motion_bound"].tolist(),
               proxy_features_offset=0, flags=3) for r in c]


def _row_objects(rows, cands):
    q = {int(c["candidate_id"]): c for c in cands}
    out = []
    for r in rows:
        c = q.get(int(r["candidate_id"]))
        out.append(NS(schema_version=1, query_id=int(c["query_id"]) if c is not None else 1,
                      candidate_id=int(r["candidate_id"]), slab_id=0, object_a_id=0, patch_a_id=0, object_b_id=0,
                      patch_b_id=0, features=r["features"].tolist(), interval_targets=[0.0] * 8,
                      family_targets=[0.0] * 8, priority_target=0.0, cost_target=0.0, uncertainty_target=0.0,
                      target_mask=0))
    return out


def _output_objects(outs):
    return [NS(candidate_id=int(o["candidate_id"]), interval_scores=o["interval_scores"].tolist(),
               family_scores=o["family_scores"].tolist(), priority_score=float(o["priority_score"]),
               cost_score=float(o["cost_score"]), uncertainty_score=float(o["uncertainty_score"])) for o in outs]


@pytest.mark.parametrize("k", range(int(G["p_count"])))
def test_object_form_matches_reference(k):
    cands = G[f"p{k}_cand"]
    masks = dict(zip(G[f"p{k}_mask_q"].tolist(), G[f"p{k}_mask_v"].tolist()))
    items, st = cpp.schedule_runtime_exact_work_items(_cand_objects(cands), _row_objects(G[f"p{k}_rows"], cands),
                                                      _output_objects(G[f"p{k}_outs"]), masks, _cfg(f"p{k}_cfg"))
    wu, wi, wp = G[f"p{k}_wu"], G[f"p{k}_wi"], G[f"p{k}_wp"]
    got = np.array([[it.work_item_id, it.parent_candidate_id, it.query_id, it.slab_id, it.patch_a_id, it.patch_b_id,
                     it.feature_family_mask, int(it.source)] for it in items], dtype=np.uint64).reshape(-1, 8)
    assert np.array_equal(got, wu)
    assert np.array_equal(np.array([[it.interval_t0, it.interval_t1] for it in items]).reshape(-1, 2), wi)
    assert np.array_equal(np.array([it.priority_score for it in items], dtype=np.float32), wp)
    names = ("raw_candidate_count", "proposal_output_count", "work_item_count", "fallback_count",
             "missing_proposal_fallback_count", "invalid_proposal_fallback_count", "ood_fallback_count",
             "high_uncertainty_fallback_count", "reordered_count", "monotonic_safe")
    assert [int(getattr(st, n)) for n in names] == G[f"p{k}_stats"].tolist()


@pytest.mark.parametrize("name", [str(n) for n in G["e_names"]])
def test_object_form_errors_match_reference(name):
    cands = G[f"e_{name}_cand"]
    with pytest.raises(ValueError) as e:
        cpp.schedule_runtime_exact_work_items(_cand_objects(cands), _row_objects(G[f"e_{name}_rows"], cands),
                                              _output_objects(G[f"e_{name}_outs"]), {1: 1, 2: 0, 3: 2},
                                              _cfg(f"e_{name}_cfg"))
    assert str(e.value) == str(G[f"e_{name}_msg"][0])


def _aligned(rng, n):
    cid = np.arange(1, n + 1, dtype=np.uint64) * 7
    qid = rng.integers(1, 5, n).astype(np.uint64)
    rt = rng.integers(1, 9, n).astype(np.uint32)
    feat = rng.normal(0, 1, (n, 32)).astype(np.float32)
    feat[:, 3] = rt
    feat[5, 7] = 2e5  # out of distribution
    feature_arrays = dict(schema_version=1, query_id=qid, candidate_id=cid, slab_id=np.zeros(n, np.uint32),
                          object_a_id=np.ones(n, np.uint32), patch_a_id=np.arange(n, dtype=np.uint32),
                          object_b_id=np.full(n, 2, np.uint32), patch_b_id=np.arange(n, dtype=np.uint32) + 100,
                          features=feat, interval_targets=np.zeros((n, 8), np.float32),
                          family_targets=np.zeros((n, 8), np.float32), priority_target=np.zeros(n, np.float32),
                          cost_target=rng.uniform(0.1, 1, n).astype(np.float32),
                          uncertainty_target=np.zeros(n, np.float32), target_mask=np.zeros(n, np.uint32))
    unc = rng.uniform(0, 1, n).astype(np.float32)
    pri = rng.uniform(0, 1, n).astype(np.float32)
    pri[3] = np.nan  # invalid proposal
    prediction_arrays = dict(interval_scores=rng.uniform(0, 1, (n, 8)).astype(np.float32),
                             family_scores=rng.uniform(0, 1, (n, 8)).astype(np.float32), priority_score=pri,
                             cost_score=rng.uniform(0, 1, n).astype(np.float32), uncertainty_score=unc)
    cands = [NS(schema_version=1, candidate_id=int(cid[i]), query_id=int(qid[i]), slab_id=0, object_a_id=1,
                patch_a_id=i, object_b_id=2, patch_b_id=i + 100, proxy_type_a=1, proxy_type_b=1, rt_hit_count=int(rt[i]),
                motion_bound=[0.0] * 4, proxy_features_offset=0, flags=3) for i in range(n)]
    return cands, feature_arrays, prediction_arrays


def test_array_forms_equal_object_form_on_aligned_inputs():
    cands, fa, pa = _aligned(np.random.default_rng(3), 64)
    cfg = cpp.ProposalSchedulingConfig()
    masks = {1: 1, 2: 0, 3: 2}
    a, sa = cpp.schedule_runtime_exact_work_items_from_arrays(cands, fa, pa, masks, cfg)
    b, sb = cpp.schedule_runtime_exact_work_items_from_proposal_arrays(fa, pa, masks, cfg)
    fields = ("work_item_id", "parent_candidate_id", "query_id", "slab_id", "patch_a_id", "patch_b_id", "interval_t0",
              "interval_t1", "feature_family_mask", "priority_score", "source")
    assert [[getattr(x, f) for f in fields] for x in a] == [[getattr(x, f) for f in fields] for x in b]
    for n in ("fallback_count", "invalid_proposal_fallback_count", "ood_fallback_count",
              "high_uncertainty_fallback_count", "reordered_count", "monotonic_safe"):
        assert getattr(sa, n) == getattr(sb, n), n
    assert sb.invalid_proposal_fallback_count == 1 and sb.ood_fallback_count >= 1 and sb.missing_proposal_fallback_count == 0
    pa2 = dict(pa, priority_score=pa["priority_score"][:-1])
    with pytest.raises(ValueError, match=r"priority_score must have shape \[64\]"):
        cpp.schedule_runtime_exact_work_items_from_proposal_arrays(fa, pa2, masks, cfg)
    zero_id = dict(fa, candidate_id=np.where(np.arange(64) == 9, 0, fa["candidate_id"]).astype(np.uint64))
    with pytest.raises(ValueError, match="preserve candidate_id order"):
        cpp.schedule_runtime_exact_work_items_from_proposal_arrays(zero_id, pa, masks, cfg)
    bad = cpp.ProposalSchedulingConfig()
    bad.first_work_item_id = 0
    with pytest.raises(ValueError, match="first_work_item_id must be non-zero"):
        cpp.schedule_runtime_exact_work_items_from_proposal_arrays(fa, pa, masks, bad)


@pytest.mark.parametrize("k", range(int(G["d_count"])))
def test_native_dense_group_equals_engine_on_converted_inputs(k):
    d = G[f"d{k}_in"]
    cfg = ast.literal_eval(str(G[f"d{k}_cfg"][0]))
    n = len(d)
    rng = np.random.default_rng(k)
    bins = rng.uniform(0, 1, (n, 8)).astype(np.float32)
    cost_target = rng.uniform(0.05, 1.0, n).astype(np.float32)
    cost_target[::5] = 0.0  # non-positive ratio: the narrow cost is the full cost
    trace = np.zeros((n, 8))
    trace[:, 0] = d["candidate_collides"]
    trace[:, 5] = d["full_exact_cost"]
    trace[:, 6], trace[:, 7] = d["contact_t0"], d["contact_t1"]
    fa = dict(query_id=d["query_id"], candidate_id=d["candidate_id"], cost_target=cost_target, oracle_trace=trace)
    pa = dict(interval_scores=bins, priority_score=d["priority_score"], cost_score=d["cost_score"],
              uncertainty_score=d["uncertainty_score"])
    got = cpp.run_native_dense_group_exact_early_stop(fa, pa, **cfg)
    # the conversion, restated: argmax bin (first of equals), [b/8, min(1, b/8 + 1/8)], narrow = max(1e-12, full * ratio)
    conv = d.copy()
    best = bins.argmax(1)
    conv["interval_t0"] = best * 0.125
    conv["interval_t1"] = np.minimum(1.0, conv["interval_t0"] + 0.125)
    ratio = cost_target.astype(np.float64)
    conv["narrow_exact_cost"] = np.where(ratio > 0, np.maximum(1e-12, d["full_exact_cost"] * ratio),
                                         d["full_exact_cost"])
    want = policy.run_dense_group_exact_early_stop(conv, policy.DenseGroupConfig(**cfg))
    for key in ("candidate_count", "group_count", "positive_group_count", "no_proposal_exact_calls",
                "learned_exact_calls", "learned_fallback_calls", "learned_interval_hit_count",
                "learned_interval_miss_count", "tp", "tn", "fp", "fn", "reordered_count",
                "high_uncertainty_group_count", "no_proposal_exact_work", "learned_exact_work"):
        assert got[key] == want[key], key
    pos = want["positive_group_count"]
    assert got["first_positive_rank_mean"] == (want["first_positive_rank_sum"] / pos if pos else 0.0)
    assert {"parse_ms", "native_total_ms", "exact_call_reduction", "exact_work_reduction"} <= set(got)
