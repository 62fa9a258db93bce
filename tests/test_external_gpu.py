"""External query batches on the GPU (p2cccd_cpp.generate_candidates_for_
external_batch over p2c_external_batch_host) against the reference's own
compiled GenerateCandidatesForExternalBatch (golden vectors, written by
tests/golden/make_external_golden.py): every candidate field, the runtime
query ids, counts, recall and the device-detected errors; the reference's KAT
(src/tests/cpp/test_rt_candidate.cpp:163-212); and a 100k-query batch
against the same rule evaluated on the host with the same slab shifts."""
from pathlib import Path
from types import SimpleNamespace as NS

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
cpp = pytest.importorskip("p2cccd_cpp")
G = np.load(Path(__file__).resolve().parent / "golden" / "external_golden.npz")


def _batch_from(q):
    items = []
    for r in q:
        items.append(NS(query_id=int(r["source_query_id"]), source_query_index=int(r["source_query_index"]),
                        family="vf" if r["family"] == 0 else "ee", vertices_t0=r["v0"].tolist(),
                        vertices_t1=r["v1"].tolist(),
                        ground_truth_collides=bool(r["ground_truth_collides"]) if r["has_ground_truth"] else None,
                        box_pair=tuple(int(v) for v in r["box_pair"]) if r["has_box_pair"] else None))
    return NS(queries=items)


@pytest.mark.parametrize("k", range(int(G["e_ok_count"])))
def test_external_batch_matches_reference(gpu, k):
    q = G[f"e{k}_q"]
    r = cpp.generate_candidates_for_external_batch(_batch_from(q), backend_name="cpu_reference")
    ref = G[f"e{k}_cand"]
    assert r.backend_name == "hip_exact"
    assert [r.primitive_count, r.raw_hit_count, r.compact_candidate_count] == G[f"e{k}_stats"].tolist()
    assert r.candidate_recall == float(G[f"e{k}_recall"])
    assert [m.runtime_query_id for m in r.runtime_query_ids] == G[f"e{k}_rid"].tolist()
    assert [m.source_query_id for m in r.runtime_query_ids] == q["source_query_id"].tolist()
    assert len(r.candidates) == len(ref)
    for c, e in zip(r.candidates, ref):
        got = (c.candidate_id, c.query_id, c.slab_id, c.object_a_id, c.patch_a_id, c.object_b_id, c.patch_b_id,
               int(c.proxy_type_a), int(c.proxy_type_b), c.rt_hit_count, c.flags, c.proxy_features_offset)
        want = tuple(int(e[f]) for f in ("candidate_id", "query_id", "slab_id", "object_a_id", "patch_a_id",
                                          "object_b_id", "patch_b_id", "proxy_type_a", "proxy_type_b",
                                          "rt_hit_count", "flags", "reserved"))
        assert got == want
        assert np.array_equal(np.array(c.motion_bound, np.float32), e["motion_bound"])


@pytest.mark.parametrize("k", [2, 3])
def test_external_batch_device_errors_match_reference(gpu, k):
    # shifted bounds that overflow, and a merged motion bound that overflows a
    # float: both found by the kernel, reported with the reference's text
    with pytest.raises(ValueError) as e:
        cpp.generate_candidates_for_external_batch(_batch_from(G[f"x{k}_q"]))
    assert str(e.value) == str(G[f"x{k}_msg"][0])


def test_external_batch_reference_kat(gpu):
    tri = [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]]
    q5 = NS(query_id=5, source_query_index=5, family="vf", vertices_t0=[[0.25, 0.25, 1.0]] + tri,
            vertices_t1=[[0.25, 0.25, -1.0]] + tri, ground_truth_collides=True, box_pair=None)
    q6 = NS(query_id=6, source_query_index=6, family="vf", vertices_t0=[[0.25, 0.25, 2.0]] + tri,
            vertices_t1=[[0.25, 0.25, 2.0]] + tri, ground_truth_collides=False, box_pair=None)
    r = cpp.generate_candidates_for_external_batch(NS(queries=[q5, q6]))
    assert (r.primitive_count, r.raw_hit_count, r.compact_candidate_count, r.candidate_recall) == (4, 1, 1, 1.0)
    assert len(r.runtime_query_ids) == 2 and len(r.candidates) == 1
    c = r.candidates[0]
    assert (c.query_id, c.patch_a_id, c.patch_b_id, c.object_a_id, c.object_b_id) == (5, 0, 6, 1, 2)


def test_large_batch_matches_host_rule(gpu):
    rng = np.random.default_rng(7)
    n = 100_000
    q = np.zeros(n, oracle.EXT_QUERY)
    q["source_query_id"] = np.arange(n) * 3 + 1
    q["source_query_index"] = np.arange(n)
    q["family"] = rng.integers(0, 2, n)
    q["v0"] = rng.uniform(-1, 1, (n, 1, 3)) + rng.uniform(-0.2, 0.2, (n, 4, 3))
    q["v1"] = q["v0"] + rng.normal(0, 0.05, (n, 4, 3))
    r = cpp.generate_candidates_for_external_batch(_batch_from(q))
    # the same rule on the host: per query boxes, the running x shift in
    # input order, inclusive overlap of the shifted boxes
    vf = q["family"] == 0
    pts = np.concatenate([q["v0"], q["v1"]], axis=1)  # (n, 8, 3)
    first = np.where(vf[:, None, None], pts[:, [0, 4, 0, 4]], pts[:, [0, 4, 1, 5]])
    second = np.where(vf[:, None, None], pts[:, [1, 5, 2, 6, 3, 7]], pts[:, [2, 6, 3, 7, 3, 7]])
    lo = np.stack([first.min(1), second.min(1)], 1)
    hi = np.stack([first.max(1), second.max(1)], 1)
    xlo = np.minimum(lo[:, 0, 0], lo[:, 1, 0])
    xhi = np.maximum(hi[:, 0, 0], hi[:, 1, 0])
    off = np.empty(n)
    cursor = 0.0
    for i in range(n):
        off[i] = cursor - xlo[i]
        cursor += (xhi[i] - xlo[i]) + 1.0e-2
    lo[:, :, 0] += off[:, None]
    hi[:, :, 0] += off[:, None]
    hit = ~((hi[:, 0] < lo[:, 1]) | (hi[:, 1] < lo[:, 0])).any(1)
    want = np.flatnonzero(hit)  # runtime ids ascend with the input here
    assert len(r.candidates) == len(want) > 0
    got = np.array([c.slab_id for c in r.candidates])
    assert np.array_equal(got, want)
    ids = np.array([c.candidate_id for c in r.candidates], dtype=np.uint64)
    assert np.array_equal(ids, (want * 3 + 1).astype(np.uint64) * np.uint64(1_000_003) + np.arange(len(want),
                                                                                                    dtype=np.uint64) + 1)
