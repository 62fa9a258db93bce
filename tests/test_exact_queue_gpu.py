"""The exact work queue on the device against the reference's own compiled
ProcessExactWorkQueueCpu / GenerateConservativeRefinementWorkItems
(oracle/_ref) and the golden vectors it wrote:
- p2c_exact_queue_host (C-ABI) on 10,000 entries: certificates, audit rows
  and refinement children field for field;
- the reference's pybind name process_exact_work_queue_cpu on the same
  10,000 entries (ExactCertificateQuery objects), field for field;
- the first refusal of every refused queue, message for message;
- DeviceQueue.refine_until_settled (device-resident passes, children fed
  back as the next queue) against the oracle's pass-by-pass loop."""
from pathlib import Path

import numpy as np
import pytest

import oracle
import queue_cases

pytestmark = pytest.mark.gpu
G = np.load(Path(__file__).resolve().parent / "golden" / "queue_golden.npz")
REFINE = dict(first_child=1000, max_child_depth=5, min_interval_width=1.0e-3)


def _eq():
    from p2cccd import exact_queue

    return exact_queue


def _expected(e, pts, ees, **kw):
    if oracle.ref_available():
        return oracle.queue_process(e, pts, ees, compiled=True, **kw)
    return oracle.queue_process(e, pts, ees, **kw)


def test_queue_golden_on_device(gpu):
    eq = _eq()
    cfg = eq.QueueConfig(first_event_id=40, first_timestamp_us=9000)
    rc = eq.RefineConfig(first_child_work_item_id=REFINE["first_child"], max_child_depth=REFINE["max_child_depth"],
                         min_interval_width=REFINE["min_interval_width"])
    print("queue golden: 300 entries", flush=True)
    certs, rows, children = eq.process_queue_host(G["entries"], G["pts"], G["ees"], cfg, rc)
    assert queue_cases.first_difference(certs.view(oracle.CERTIFICATE), G["certs"]) == ""
    assert queue_cases.first_difference(rows.view(oracle.AUDIT_ROW), G["rows"]) == ""
    assert queue_cases.first_difference(children.view(oracle.QUEUE_ENTRY), G["children"]) == ""
    _, rows1, none = eq.process_queue_host(G["entries"], G["pts"], G["ees"],
                                          eq.QueueConfig(first_event_id=40, first_timestamp_us=9000,
                                                         emit_dequeue_events=False))
    assert none is None
    assert queue_cases.first_difference(rows1.view(oracle.AUDIT_ROW), G["rows_no_dequeue"]) == ""


def test_queue_10k_matches_compiled_reference(gpu):
    eq = _eq()
    e, pts, ees = queue_cases.make_queue(10_000, 5)
    print("queue 10k: seed 5", flush=True)
    want_c, want_r = _expected(e, pts, ees, first_event_id=3, first_timestamp_us=11)
    rc = eq.RefineConfig(first_child_work_item_id=77, max_child_depth=4, min_interval_width=1e-3)
    certs, rows, children = eq.process_queue_host(e, pts, ees, eq.QueueConfig(3, 11), rc)
    assert queue_cases.first_difference(certs.view(oracle.CERTIFICATE), want_c) == ""
    assert queue_cases.first_difference(rows.view(oracle.AUDIT_ROW), want_r) == ""
    want_k = oracle.queue_refine(e, want_c, first_child=77, max_child_depth=4, min_interval_width=1e-3,
                                 compiled=oracle.ref_available())
    assert len(want_k) > 1000
    assert queue_cases.first_difference(children.view(oracle.QUEUE_ENTRY), want_k) == ""
    assert set(np.unique(certs["status"])) == {0, 1, 2} and 3 in set(np.unique(certs["reason_code"]))


def _objects(cpp, e, pts, ees):
    def traj(fid, p0, p1):
        t = cpp.LinearVertexTrajectory()
        t.feature_id, t.position_t0, t.position_t1 = int(fid), [float(v) for v in p0], [float(v) for v in p1]
        return t

    def prim(p, pt):
        o = cpp.PointTriangleIntervalPrimitive() if pt else cpp.EdgeEdgeIntervalPrimitive()
        tr = [traj(p["v"][k]["feature_id"], p["v"][k]["p0"], p["v"][k]["p1"]) for k in range(4)]
        if pt:
            o.point_id, o.triangle_id = int(p["id_a"]), int(p["id_b"])
            o.point, o.triangle_v0, o.triangle_v1, o.triangle_v2 = tr
        else:
            o.edge_a_id, o.edge_b_id = int(p["id_a"]), int(p["id_b"])
            o.edge_a0, o.edge_a1, o.edge_b0, o.edge_b1 = tr
        return o

    pt_objs = [prim(p, True) for p in pts]
    ee_objs = [prim(p, False) for p in ees]
    out = []
    for r in e:
        q = cpp.ExactCertificateQuery()
        w = cpp.ExactWorkItem()
        w.work_item_id, w.parent_candidate_id, w.query_id = int(r["work_item_id"]), int(r["parent_candidate_id"]), \
            int(r["query_id"])
        w.slab_id, w.patch_a_id, w.patch_b_id = int(r["slab_id"]), int(r["patch_a_id"]), int(r["patch_b_id"])
        w.interval_t0, w.interval_t1 = float(r["interval_t0"]), float(r["interval_t1"])
        w.feature_family_mask, w.topk_feature_ids_offset = int(r["feature_family_mask"]), \
            int(r["topk_feature_ids_offset"])
        w.depth, w.priority_score, w.source = int(r["depth"]), float(r["priority_score"]), \
            cpp.ProposalSource(int(r["source"]))
        c = cpp.CertificateEngineConfig()
        c.eps_time, c.eps_space, c.max_subdivision_depth = float(r["eps_time"]), float(r["eps_space"]), \
            int(r["max_subdivision_depth"])
        q.work_item, q.config = w, c
        b, n = int(r["pt_begin"]), int(r["pt_count"])
        q.point_triangle_primitives = pt_objs[b:b + n]
        b, n = int(r["ee_begin"]), int(r["ee_count"])
        q.edge_edge_primitives = ee_objs[b:b + n]
        out.append(q)
    return out


def test_process_exact_work_queue_cpu_10k_matches_compiled_reference(gpu):
    cpp = pytest.importorskip("p2cccd_cpp")
    e, pts, ees = queue_cases.make_queue(10_000, 9)
    print("process_exact_work_queue_cpu 10k: seed 9", flush=True)
    want_c, want_r = _expected(e, pts, ees, first_event_id=1, first_timestamp_us=1)
    res = cpp.process_exact_work_queue_cpu(_objects(cpp, e, pts, ees), cpp.ExactWorkQueueConfig())
    assert res.processed_count == len(e) and len(res.audit_log) == 2 * len(e)
    got_c = np.zeros(len(e), oracle.CERTIFICATE)
    for i, c in enumerate(res.certificates):
        got_c[i] = (c.work_item_id, c.query_id, c.interval_t0, c.interval_t1, c.toi_upper, c.safe_margin_lb,
                    c.witness_id_a, c.witness_id_b, c.eps_time, c.eps_space, c.covered_feature_mask, c.reason_code,
                    int(c.status), c.witness_family, int(c.next_refinement_mode), [0] * 7)
    got_r = np.zeros(2 * len(e), oracle.AUDIT_ROW)
    for i, r in enumerate(res.audit_log):
        got_r[i] = (r.event_id, r.query_id, r.candidate_id, r.work_item_id, r.interval_t0, r.interval_t1,
                    r.timestamp_us, r.aux_value0, r.aux_value1, r.action, r.depth, int(r.stage), [0] * 3)
    assert queue_cases.first_difference(got_c, want_c) == ""
    assert queue_cases.first_difference(got_r, want_r) == ""
    assert cpp.validate_exact_work_queue_coverage(_objects(cpp, e[:50], pts, ees),
                                                  cpp.process_exact_work_queue_cpu(_objects(cpp, e[:50], pts, ees),
                                                                                   cpp.ExactWorkQueueConfig()))


@pytest.mark.parametrize("k", range(len(G["error_names"])))
def test_queue_refusals_on_device(gpu, k):
    eq = _eq()
    from p2cccd._native import NativeError

    cases = {c[0]: c for c in queue_cases.error_cases()}
    name, e, pts, ees, kw = cases[str(G["error_names"][k])]
    print(f"queue refusal: {name}", flush=True)
    cfg = eq.QueueConfig(first_event_id=kw.get("first_event_id", 1), first_timestamp_us=kw.get("first_timestamp_us", 1),
                         emit_dequeue_events=kw.get("emit", True))
    with pytest.raises(NativeError) as ex:
        eq.process_queue_host(e, pts, ees, cfg)
    assert str(ex.value) == str(G["error_messages"][k])


def test_device_queue_refines_until_settled(gpu):
    eq = _eq()
    e, pts, ees = queue_cases.make_queue(2_000, 13)
    e["max_subdivision_depth"] = np.minimum(e["max_subdivision_depth"], 5)
    print("device queue passes: 2000 entries, seed 13", flush=True)
    cfg = eq.QueueConfig(first_event_id=5, first_timestamp_us=50)
    rc = eq.RefineConfig(first_child_work_item_id=10**6, max_child_depth=8, min_interval_width=1e-3)
    dq = eq.DeviceQueue(pts, ees)
    certs, rows, entries = dq.refine_until_settled(e, cfg, rc, max_passes=32)
    # the oracle, pass by pass with running ids
    want_c, want_r, want_e = [], [], []
    cur, ev, ts, child = e, 5, 50, 10**6
    for _ in range(32):
        if len(cur) == 0:
            break
        c, r = oracle.queue_process(cur, pts, ees, first_event_id=ev, first_timestamp_us=ts)
        kids = oracle.queue_refine(cur, c, first_child=child, max_child_depth=8, min_interval_width=1e-3)
        want_c.append(c), want_r.append(r), want_e.append(cur)
        ev, ts, child = ev + 2 * len(cur), ts + 2 * len(cur), child + len(kids)
        cur = kids
    assert len(want_e) >= 3, "the workload should need several passes"
    assert queue_cases.first_difference(entries.view(oracle.QUEUE_ENTRY), np.concatenate(want_e)) == ""
    assert queue_cases.first_difference(certs.view(oracle.CERTIFICATE), np.concatenate(want_c)) == ""
    assert queue_cases.first_difference(rows.view(oracle.AUDIT_ROW), np.concatenate(want_r)) == ""
