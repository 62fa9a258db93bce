"""The reference package's drop-in certificate path (CS1:
execute_certificate_engine_for_external_batch, src/python/p2cccd/
certificate_engine.py:218-250 over rt_exact.py:356-1145) on the GPU:
- the reference's own first RT-exact case (src/tests/python/
  test_rt_exact_baseline.py:49-80): one candidate, one RAW work item, one
  collision, queue conserved, no false negative, no proposal stage;
- a queue of 32 or more items takes the CUDA batch twin's walk, below 32 the
  Python twin's, as the reference's rt_exact does; every certificate equals
  the oracle's restatement of that walk, the candidate set equals the
  oracle's literal restatement of the batched proxy scene, and the audit
  trail has one RT row per candidate and DEQUEUED + EXACT rows per item with
  running event ids and timestamps."""
import numpy as np
import pytest

import oracle
from p2cccd import (
    AuditStage,
    CertificateEngineWrapperConfig,
    CertificateStatus,
    ProposalSource,
    execute_certificate_engine_for_external_batch,
)
from p2cccd import workloads
from p2cccd.datasets import CCD_ADAPTER_SCHEMA_VERSION, CCDQueryFamily, DatasetQueryBatch, ExternalCCDQuery
from p2cccd.pure_exact import PureExactCPUConfig

pytestmark = pytest.mark.gpu


def _batch(x, family, labels=None, ids=None):
    fam = CCDQueryFamily.VERTEX_FACE if family == "vf" else CCDQueryFamily.EDGE_EDGE
    qs = []
    for i in range(len(x)):
        qid = int(ids[i]) if ids is not None else i
        qs.append(ExternalCCDQuery(schema_version=CCD_ADAPTER_SCHEMA_VERSION, source_name="unit", scene_name="s",
                                   batch_id="b", query_id=qid, source_query_index=i, family=fam,
                                   vertices_t0=tuple(map(tuple, x[i, :4])), vertices_t1=tuple(map(tuple, x[i, 4:])),
                                   ground_truth_collides=None if labels is None else bool(labels[i])))
    return DatasetQueryBatch(schema_version=CCD_ADAPTER_SCHEMA_VERSION, source_name="unit", scene_name="s",
                             batch_id="b", family=fam, queries=tuple(qs))


def test_reference_rt_exact_case(gpu):
    tri = ((0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    x = np.array([[(0.25, 0.25, 1.0), *tri, (0.25, 0.25, -1.0), *tri],
                  [(0.25, 0.25, 2.0), *tri, (0.25, 0.25, 2.0), *tri]], dtype=np.float64)
    cfg = CertificateEngineWrapperConfig(exact=PureExactCPUConfig(eps_time=1e-6, eps_space=1e-8,
                                                                  max_subdivision_depth=28))
    r = execute_certificate_engine_for_external_batch(_batch(x, "vf", labels=[True, False]), config=cfg)
    assert len(r.candidates) == len(r.work_items) == len(r.certificates) == 1
    assert r.certificates[0].status == CertificateStatus.COLLISION
    assert r.work_items[0].source == ProposalSource.RAW and r.work_items[0].feature_family_mask == 1
    assert r.queue_conserved and r.final_fn_zero and r.used_cpp_backend
    assert all(row.stage != AuditStage.PROPOSAL for row in r.audit_log)
    assert [q.status for q in r.query_results] == ["collision", "rt_candidate_separation"]


@pytest.mark.parametrize("n,family", [(20, "vf"), (20, "ee"), (400, "vf"), (400, "ee")])
def test_walks_and_audit_match_oracle(gpu, n, family):
    qs = workloads.make_query_set(n, n, seed=5 + n)
    x = qs.vf if family == "vf" else qs.ee
    ids = np.arange(n) * 2 + 3
    b = _batch(x, family, ids=ids)
    cfg = CertificateEngineWrapperConfig(first_work_item_id=7, first_event_id=100, first_timestamp_us=50)
    r = execute_certificate_engine_for_external_batch(b, config=cfg)
    # candidates: the literal batched-proxy-scene rule
    q = np.zeros(n, oracle.EXT_QUERY)
    q["source_query_id"], q["source_query_index"] = ids, np.arange(n)
    q["family"] = 0 if family == "vf" else 1
    q["v0"], q["v1"] = x[:, :4], x[:, 4:]
    want, _ = oracle.external_batch(q)
    assert [c.candidate_id for c in r.candidates] == [w["candidate_id"] for w in want]
    assert [c.slab_id for c in r.candidates] == [w["slab_id"] for w in want]
    m = len(r.work_items)
    assert m == len(want) > 0 and [w.work_item_id for w in r.work_items] == list(range(7, 7 + m))
    # certificates: the walk the reference takes for a queue this long
    rows = np.array([c.slab_id for c in r.candidates])
    fam = 1 if family == "vf" else 2
    ex = PureExactCPUConfig()
    prims = oracle.prims_from_positions(x[rows], ids=np.tile([0, 1], (m, 1)))
    if m >= 32:
        ref = oracle.ref_cert_twin(fam, prims, eps_time=ex.eps_time, eps_space=ex.eps_space,
                                   max_depth=ex.max_subdivision_depth)
        for c, e in zip(r.certificates, ref):
            assert int(c.status) == int(e["status"])
            assert (c.interval_t0, c.interval_t1, c.toi_upper) == (e["interval_t0"], e["interval_t1"], e["toi_upper"])
            assert c.safe_margin_lb == max(0.0, e["safe_margin_lb"])
    else:
        ref = oracle.ref_cert_py(fam, prims, eps_time=ex.eps_time, eps_space=ex.eps_space,
                                 max_depth=ex.max_subdivision_depth)
        for c, e, w in zip(r.certificates, ref, r.work_items):
            assert int(c.status) == int(e["status"])
            if c.status == CertificateStatus.COLLISION:
                assert c.toi_upper == min(1.0, max(0.0, e["toi_upper"]))
                assert (c.witness_id_a, c.witness_id_b) == (w.patch_a_id, w.patch_b_id)
            elif c.status == CertificateStatus.SEPARATION:
                assert c.safe_margin_lb == max(0.0, e["safe_margin_lb"]) and c.covered_feature_mask == fam
    # audit trail
    a = r.audit_log
    assert len(a) == m + 2 * m
    assert [row.event_id for row in a] == list(range(100, 100 + len(a)))
    assert [row.timestamp_us for row in a] == list(range(50, 50 + len(a)))
    assert all(row.stage == AuditStage.RT for row in a[:m])
    for k, (w, c) in enumerate(zip(r.work_items, r.certificates)):
        deq, done = a[m + 2 * k], a[m + 2 * k + 1]
        assert (deq.action, deq.work_item_id) == (1, w.work_item_id)
        assert done.action == {0: 2, 1: 3, 2: 4}[int(c.status)]
    assert r.queue_conserved and len(r.query_results) == n
