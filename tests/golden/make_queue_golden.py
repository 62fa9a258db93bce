"""Writes tests/golden/queue_golden.npz from the reference's own compiled
ProcessExactWorkQueueCpu and GenerateConservativeRefinementWorkItems
(src/cpp/certificate/certificate_engine.cpp:593-720, built in place into
oracle/_ref by `make -C oracle ref`).

Run in the container that holds the reference:  python tests/golden/make_queue_golden.py
The fixture holds inputs and expected outputs only (no reference source):
a 300-entry queue (tests/queue_cases.make_queue, seed 21) with its
certificates, audit rows (dequeue events on) and refinement children, the
same queue's rows without dequeue events, and the message of every refused
queue of tests/queue_cases.error_cases.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "p2c-ccd_amd"))
sys.path.insert(0, str(ROOT / "tests"))

import oracle  # noqa: E402
import queue_cases  # noqa: E402

OUT = Path(__file__).resolve().parent / "queue_golden.npz"
REFINE = dict(first_child=1000, max_child_depth=5, min_interval_width=1.0e-3)


def main():
    e, pts, ees = queue_cases.make_queue(300, 21, pool=160)
    certs, rows = oracle.queue_process(e, pts, ees, first_event_id=40, first_timestamp_us=9000, compiled=True)
    _, rows1 = oracle.queue_process(e, pts, ees, first_event_id=40, first_timestamp_us=9000, emit=False,
                                    compiled=True)
    children = oracle.queue_refine(e, certs, compiled=True, **REFINE)
    out = dict(entries=e, pts=pts, ees=ees, certs=certs, rows=rows, rows_no_dequeue=rows1, children=children)
    names, msgs = [], []
    for name, ce, cp, cx, kw in queue_cases.error_cases():
        try:
            oracle.queue_process(ce, cp, cx, compiled=True, **kw)
            raise SystemExit(f"error case {name} was accepted")
        except ValueError as ex:
            names.append(name)
            msgs.append(str(ex))
    out["error_names"] = np.array(names)
    out["error_messages"] = np.array(msgs)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes): {len(e)} entries, {len(children)} children, "
          f"{len(names)} refused queues")


if __name__ == "__main__":
    main()
