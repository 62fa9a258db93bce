"""Pins the oracle's exact work queue (oracle/ref_queue.c) to the reference:
- against the golden vectors the compiled reference wrote
  (tests/golden/queue_golden.npz): certificates, audit rows with and without
  dequeue events, refinement children with a running first child id, and the
  message of every refused queue;
- against the compiled reference itself (oracle/_ref, where built) on a
  10,000-entry queue with every aggregation outcome.
Closed-form cases pin what the fixture leaves implicit: the event/timestamp
numbering and the children of a collision and of a bisection."""
from pathlib import Path

import numpy as np
import pytest

import oracle
import queue_cases

G = np.load(Path(__file__).resolve().parent / "golden" / "queue_golden.npz")
REFINE = dict(first_child=1000, max_child_depth=5, min_interval_width=1.0e-3)


def test_oracle_queue_matches_golden():
    certs, rows = oracle.queue_process(G["entries"], G["pts"], G["ees"], first_event_id=40, first_timestamp_us=9000)
    assert queue_cases.first_difference(certs, G["certs"]) == ""
    assert queue_cases.first_difference(rows, G["rows"]) == ""
    _, rows1 = oracle.queue_process(G["entries"], G["pts"], G["ees"], first_event_id=40, first_timestamp_us=9000,
                                    emit=False)
    assert queue_cases.first_difference(rows1, G["rows_no_dequeue"]) == ""
    children = oracle.queue_refine(G["entries"], certs, **REFINE)
    assert queue_cases.first_difference(children, G["children"]) == ""
    # the fixture covers every outcome
    assert set(np.unique(G["certs"]["status"])) == {0, 1, 2}
    assert set(np.unique(G["certs"]["reason_code"])) >= {0, 1, 2, 3}
    assert set(np.unique(G["rows"]["action"])) == {1, 2, 3, 4, 5}


@pytest.mark.parametrize("k", range(len(G["error_names"])))
def test_oracle_queue_refusals_match_golden(k):
    cases = {c[0]: c for c in queue_cases.error_cases()}
    name, e, pts, ees, kw = cases[str(G["error_names"][k])]
    with pytest.raises(ValueError) as ex:
        oracle.queue_process(e, pts, ees, **kw)
    assert str(ex.value) == str(G["error_messages"][k])


def test_oracle_queue_numbering_and_children_closed_form():
    e = np.zeros(3, oracle.QUEUE_ENTRY)
    e["work_item_id"], e["parent_candidate_id"], e["query_id"] = [5, 6, 7], 9, [1, 1, 2]
    e["interval_t0"], e["interval_t1"] = [0.0, 0.25, 0.5], [1.0, 0.75, 0.5]
    e["feature_family_mask"], e["eps_time"], e["eps_space"], e["max_subdivision_depth"] = 1, 1e-4, 1e-6, 32
    # entry 0: the KAT point (0.25, 0.25, 1 -> -3) through the unit triangle: contact at t = 0.25
    x = np.zeros((1, 8, 3))
    x[0, 0], x[0, 4] = [0.25, 0.25, 1.0], [0.25, 0.25, -3.0]
    x[0, 2], x[0, 6], x[0, 3], x[0, 7] = [1, 0, 0], [1, 0, 0], [0, 1, 0], [0, 1, 0]
    pts = oracle.prims_from_positions(x)
    e["pt_count"] = [1, 0, 0]  # entries 1, 2: no geometry
    certs, rows = oracle.queue_process(e, pts, np.zeros(0, oracle.PRIM), first_event_id=100, first_timestamp_us=7)
    assert list(rows["event_id"]) == list(range(100, 106)) and list(rows["timestamp_us"]) == list(range(7, 13))
    assert list(rows["stage"]) == [2, 4] * 3 and list(rows["action"]) == [1, 2, 1, 4, 1, 4]
    assert certs["status"][0] == 0 and abs(certs["toi_upper"][0] - 0.25) <= 1e-12
    assert list(certs["reason_code"][1:]) == [1, 1] and list(certs["next_refinement_mode"][1:]) == [2, 2]
    # undecided with a bisection request: two halves; collision: the prefix [t0, toi]
    certs["status"][1], certs["reason_code"][1], certs["next_refinement_mode"][1] = 2, 2, 1
    ch = oracle.queue_refine(e, certs, first_child=50, max_child_depth=8, min_interval_width=1e-3)
    assert list(ch["work_item_id"]) == [50, 51, 52]
    assert [(a, b) for a, b in zip(ch["interval_t0"], ch["interval_t1"])] == [
        (0.0, certs["toi_upper"][0]), (0.25, 0.5), (0.5, 0.75)]
    assert list(ch["depth"]) == [1, 1, 1] and list(ch["source"]) == [1, 1, 1]


@pytest.mark.skipif(not oracle.ref_available(), reason="compiled reference not built")
def test_oracle_queue_matches_compiled_reference_10k():
    e, pts, ees = queue_cases.make_queue(10_000, 5)
    want = oracle.queue_process(e, pts, ees, first_event_id=3, first_timestamp_us=11, compiled=True)
    got = oracle.queue_process(e, pts, ees, first_event_id=3, first_timestamp_us=11)
    assert queue_cases.first_difference(got[0], want[0]) == ""
    assert queue_cases.first_difference(got[1], want[1]) == ""
    kids = oracle.queue_refine(e, want[0], first_child=77, max_child_depth=4, min_interval_width=1e-3, compiled=True)
    assert queue_cases.first_difference(oracle.queue_refine(e, want[0], first_child=77, max_child_depth=4,
                                                            min_interval_width=1e-3), kids) == ""
    assert len(kids) > 1000
