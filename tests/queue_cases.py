"""Exact work queues for the queue tests: entries (the oracle's QUEUE_ENTRY
layout, the library's p2c_queue_entry) over point-triangle and edge-edge
primitive pools, with every outcome of the reference's aggregation in the mix:
collisions, separations, undecided leaves (small max depths), families with no
primitives (missing geometry), invalid primitives (negative ids, a NaN
coordinate: the invalid-input certificate), masks 1, 2 and 3, sub-intervals,
zero-width intervals and shared or overlapping primitive ranges."""
from __future__ import annotations

import numpy as np

import oracle
from p2cccd import workloads


def make_queue(n: int, seed: int, pool: int | None = None):
    rng = np.random.default_rng(seed)
    pool = pool or max(64, n // 2)
    qs = workloads.make_query_set(pool, pool, seed=seed, touch_frac=0.2, degenerate_frac=0.05)
    pts = oracle.prims_from_positions(qs.vf)
    ees = oracle.prims_from_positions(qs.ee)
    bad = rng.choice(pool, max(1, pool // 100), replace=False)
    pts["id_a"][bad[: len(bad) // 2 + 1]] = -1
    ees["v"][bad[len(bad) // 2:], 2]["p1"][:, 1] = np.nan
    e = np.zeros(n, oracle.QUEUE_ENTRY)
    e["work_item_id"] = np.arange(1, n + 1, dtype=np.uint64) * 3
    e["parent_candidate_id"] = rng.integers(1, 1 << 40, n).astype(np.uint64)
    e["query_id"] = rng.integers(1, max(2, n // 4), n).astype(np.uint64)
    e["slab_id"] = rng.integers(0, 8, n)
    e["patch_a_id"] = rng.integers(0, 100, n)
    e["patch_b_id"] = rng.integers(0, 100, n)
    e["feature_family_mask"] = rng.choice([1, 2, 3], n, p=[0.4, 0.4, 0.2])
    t = np.sort(rng.uniform(0.0, 1.0, (n, 2)), axis=1)
    full = rng.uniform(size=n) < 0.5
    t[full] = (0.0, 1.0)
    zero = rng.uniform(size=n) < 0.02
    t[zero, 1] = t[zero, 0]
    e["interval_t0"], e["interval_t1"] = t[:, 0], t[:, 1]
    e["topk_feature_ids_offset"] = rng.integers(0, 1000, n)
    e["depth"] = rng.integers(0, 6, n)
    e["source"] = rng.integers(0, 3, n)
    e["priority_score"] = rng.normal(0.0, 1.0, n).astype(np.float32)
    e["max_subdivision_depth"] = rng.choice([0, 2, 5, 12, 32], n, p=[0.05, 0.15, 0.2, 0.3, 0.3])
    e["eps_time"] = rng.choice([1.0e-3, 1.0e-4, 1.0e-6], n)
    e["eps_space"] = rng.choice([1.0e-4, 1.0e-6, 1.0e-8], n)
    for fam, src in (("pt", pts), ("ee", ees)):
        count = rng.choice([0, 1, 2, 3, 5], n, p=[0.1, 0.35, 0.25, 0.2, 0.1])
        e[f"{fam}_count"] = count
        e[f"{fam}_begin"] = rng.integers(0, len(src) - count + 1)
    return e, pts, ees


def first_difference(a: np.ndarray, b: np.ndarray) -> str:
    """'' when the records are equal field for field (NaN equal to NaN), else
    the first differing field and index."""
    if a.shape != b.shape:
        return f"shape {a.shape} != {b.shape}"
    for name in a.dtype.names:
        if name.startswith("pad"):
            continue
        x, y = a[name], b[name]
        same = (x == y) | ((x != x) & (y != y)) if x.dtype.kind == "f" else (x == y)
        if same.ndim > 1:
            same = same.all(axis=tuple(range(1, same.ndim)))
        if not same.all():
            i = int(np.flatnonzero(~same)[0])
            return f"{name}[{i}]: {x[i]!r} != {y[i]!r}"
    return ""


def error_cases():
    """(name, entries, pts, ees, queue kwargs) queues the reference refuses,
    each with an earlier valid entry and, where it matters, a later failing
    one (the first failure in queue order must be reported)."""
    base, pts, ees = make_queue(12, 77, pool=64)
    base["pt_count"] = np.minimum(base["pt_count"], 1)
    base["ee_count"] = np.minimum(base["ee_count"], 1)
    base["max_subdivision_depth"] = 8
    base["eps_time"], base["eps_space"] = 1.0e-4, 1.0e-6
    pts = pts.copy()
    ees = ees.copy()
    pts["id_a"] = np.abs(pts["id_a"])
    ees["v"]["p1"] = np.nan_to_num(ees["v"]["p1"])
    out = []

    def case(name, at, field, value, later=None, **kw):
        e = base.copy()
        for f, v in (field.items() if isinstance(field, dict) else [(field, value)]):
            e[f][at] = v
        if later is not None:
            e[later[0]][at + 3] = later[1]
        out.append((name, e, pts, ees, kw))

    case("work_item_id", 4, "work_item_id", 0, later=("query_id", 0))
    case("parent_candidate_id", 4, "parent_candidate_id", 0)
    case("query_id", 2, "query_id", 0, later=("work_item_id", 0))
    case("t0_nan", 5, "interval_t0", np.nan)
    case("t1_inf", 5, "interval_t1", np.inf)
    case("interval_order", 5, {"interval_t0": 0.9, "interval_t1": 0.5}, None)
    case("t1_above_one", 3, "interval_t1", 1.5)
    case("family_mask", 6, "feature_family_mask", 0)
    case("priority_nan", 6, "priority_score", np.nan)
    case("source", 7, "source", 3)
    case("eps_time_zero", 3, "eps_time", 0.0, later=("work_item_id", 0))
    case("eps_time_nan", 3, "eps_time", np.nan)
    case("eps_space_negative", 8, "eps_space", -1.0)
    case("eps_space_inf", 8, "eps_space", np.inf)
    dup = base.copy()
    dup["query_id"][9], dup["work_item_id"][9] = dup["query_id"][2], dup["work_item_id"][2]
    out.append(("duplicate", dup, pts, ees, {}))
    out.append(("first_event_zero", base.copy(), pts, ees, dict(first_event_id=0)))
    out.append(("first_timestamp_zero", base.copy(), pts, ees, dict(first_timestamp_us=0)))
    # event ids and timestamps wrap to 0 partway through the queue
    out.append(("event_wrap", base.copy(), pts, ees, dict(first_event_id=2**64 - 7)))
    out.append(("timestamp_wrap", base.copy(), pts, ees, dict(first_timestamp_us=2**64 - 4)))
    out.append(("event_wrap_no_dequeue", base.copy(), pts, ees, dict(first_event_id=2**64 - 5, emit=False)))
    return out
