"""Writes tests/golden/external_golden.npz from the reference's own compiled
GenerateCandidatesForExternalBatch and BuildProxyScene (oracle/_ref, built
from the reference sources in place by `make -C oracle ref`).

External batches (src/cpp/rt_candidate/external_batch_candidate.cpp:231-339):
seeded vertex-face and edge-edge queries, with
  - boxes that touch exactly, or miss by one double before the x shift of
    their slab, after earlier queries with large x extents make that shift
    coarse (the shift's rounding decides whether they touch),
  - source ids of 0 (runtime id index + 1) and explicit ids that collide with
    those (runtime ids bumped past taken ones), box pairs, ground truth,
and the reference's error cases (non-finite vertices, duplicate ids, bounds
that overflow, motion bounds that overflow a float).

Proxy-scene builds (src/cpp/rt_candidate/proxy_scene.cpp:156-213 with
geometry/proxy.cpp and geometry/motion_utils.cpp): swept-AABB and capsule
objects, several patches, motion segments and slabs, rotations on both the
slerp and the nearly-parallel lerp branch, and the reference's error cases.

    python tests/golden/make_external_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
import oracle  # noqa: E402

OUT = Path(__file__).resolve().parent / "external_golden.npz"


def _queries(rng, n, coarse=False):
    q = np.zeros(n, oracle.EXT_QUERY)
    q["family"] = rng.integers(0, 2, n)
    base = rng.uniform(-1, 1, (n, 1, 3))
    q["v0"] = base + rng.uniform(-0.3, 0.3, (n, 4, 3))
    q["v1"] = q["v0"] + rng.normal(0, 0.1, (n, 4, 3))
    if coarse:  # large x extents early in the batch make later slab shifts coarse
        q["v0"][:3, :, 0] *= 1e9
        q["v1"][:3, :, 0] *= 1e9
    for i in range(n):
        kind = rng.integers(0, 4)
        first = [0] if q["family"][i] == 0 else [0, 1]
        second = [1, 2, 3] if q["family"][i] == 0 else [2, 3]
        if kind in (1, 2):  # rebuild so the first side's x range ends where the second's begins
            lo = min(q["v0"][i][second][:, 0].min(), q["v1"][i][second][:, 0].min())
            top = lo if kind == 1 else np.nextafter(lo, -np.inf)
            hi_first = max(q["v0"][i][first][:, 0].max(), q["v1"][i][first][:, 0].max())
            shift = top - hi_first
            q["v0"][i, first, 0] += shift
            q["v1"][i, first, 0] += shift
            # keep the shifted maximum exactly at `top`
            m0 = q["v0"][i, first, 0]
            m1 = q["v1"][i, first, 0]
            j = int(np.argmax(np.concatenate([m0, m1])))
            if j < len(first):
                q["v0"][i, first[j], 0] = top
            else:
                q["v1"][i, first[j - len(first)], 0] = top
        elif kind == 3:  # separated in z
            q["v0"][i, first, 2] += 5.0
            q["v1"][i, first, 2] += 5.0
    ids = rng.integers(0, 3, n)
    q["source_query_id"] = np.where(ids == 0, 0, np.where(ids == 1, np.arange(n) + 2, rng.integers(10, 10**6, n)))
    _, first_seen = np.unique(q["source_query_id"], return_index=True)
    dup = np.ones(n, bool)
    dup[first_seen] = False
    q["source_query_id"][dup & (q["source_query_id"] != 0)] = 10**7 + np.arange(n)[dup & (q["source_query_id"] != 0)]
    zero = np.flatnonzero(q["source_query_id"] == 0)
    q["source_query_id"][zero[1:]] = 2 * 10**7 + zero[1:]  # only one id may be 0 (ids must be unique)
    q["source_query_index"] = rng.integers(0, 1000, n)
    q["has_ground_truth"] = rng.integers(0, 2, n)
    q["ground_truth_collides"] = rng.integers(0, 2, n)
    q["has_box_pair"] = rng.integers(0, 2, n)
    q["box_pair"] = rng.integers(1, 9, (n, 2))
    return q


def _external_cases(rng):
    ok, err = [], []
    for k, (n, coarse) in enumerate([(2, False), (40, False), (200, True), (300, True), (64, False)]):
        ok.append(_queries(rng, n, coarse))
    q = _queries(rng, 10)
    q["v0"][3, 1, 2] = np.nan
    q["v1"][2, 0, 0] = np.inf
    err.append(q)  # the first bad query in order reports its field
    q = _queries(rng, 10)
    q["source_query_id"][7] = q["source_query_id"][2] = 99
    err.append(q)
    q = _queries(rng, 6)
    q["v0"][0, :, 0] = [-1.7e308, 1.7e308, 0.0, 0.0]
    err.append(q)  # x extent overflows: every shifted bound after it is infinite
    q = _queries(rng, 6)
    q["family"][4] = 0
    q["v0"][4] = q["v1"][4] = [[0.1, 0.1, 0], [0, -1e39, 0], [1, 0, 0], [0, 1e39, 0]]
    err.append(q)  # overlapping boxes whose half extent overflows a float
    return ok, err


def _quat(rng):
    q = rng.normal(size=4)
    return q * rng.uniform(0.2, 3.0) / np.linalg.norm(q)


def _build(rng, n_objects, segs, slabs, eps_scale=0.05):
    objs = np.zeros(n_objects, oracle.BUILD_OBJECT)
    patches, segments = [], []
    bounds = np.linspace(0.0, 1.0, segs + 1)
    for o in range(n_objects):
        npat = int(rng.integers(1, 4))
        objs[o] = (10 * (o + 1), int(rng.integers(1, 3)), slabs, npat, segs, len(patches), len(segments),
                   float(rng.uniform(0, eps_scale)))
        for p in range(npat):
            patches.append((100 * o + p, rng.uniform(-0.5, 0.5, 3), float(rng.uniform(0.0, 0.3))))
        for s in range(segs):
            q0 = _quat(rng)
            branch = rng.integers(0, 3)
            if branch == 0:  # nearly parallel: the normalised lerp branch
                q1 = q0 + rng.normal(0, 1e-3, 4)
            elif branch == 1:  # opposite hemisphere: the sign flip
                q1 = -_quat(rng)
            else:
                q1 = _quat(rng)
            segments.append((bounds[s], bounds[s + 1], rng.uniform(-1, 1, 3), q0, rng.uniform(-1, 1, 3), q1))
    P = np.zeros(len(patches), oracle.BUILD_PATCH)
    for i, (pid, c, r) in enumerate(patches):
        P[i] = (pid, c, r)
    S = np.zeros(len(segments), oracle.BUILD_SEGMENT)
    for i, (t0, t1, a, qa, b, qb) in enumerate(segments):
        S[i] = (t0, t1, a, qa, b, qb)
    return objs, P, S


def _build_cases(rng):
    ok, err = [], []
    for n_objects, segs, slabs in ((2, 1, 2), (3, 2, 3), (4, 3, 1), (2, 2, 4)):
        ok.append(_build(rng, n_objects, segs, slabs))

    def bad(fn):
        o, p, s = _build(rng, 2, 2, 2)
        fn(o, p, s)
        err.append((o, p, s))

    def set_field(arr, field, index, value):
        arr[field][index] = value

    bad(lambda o, p, s: set_field(o, "proxy_type", 0, 0))                         # unknown proxy type
    bad(lambda o, p, s: set_field(o, "slabs_per_motion_segment", 1, 3))           # non-global slab grid
    bad(lambda o, p, s: set_field(o, "slabs_per_motion_segment", 0, 0))           # no slabs
    bad(lambda o, p, s: set_field(o, "eps_proxy", 1, -1.0))                       # negative eps
    bad(lambda o, p, s: set_field(s, "rotation0", 1, [0.0, 0.0, 0.0, 0.0]))      # zero quaternion
    bad(lambda o, p, s: set_field(s, "translation1", 0, [0.0, np.nan, 0.0]))     # non-finite pose
    bad(lambda o, p, s: set_field(p, "radius", 0, -0.5))                          # negative radius
    bad(lambda o, p, s: set_field(s, "t1", 2, s["t0"][2]))                        # empty segment
    return ok, err


def main() -> int:
    if not oracle.ref_available():
        raise SystemExit("the compiled reference (oracle/_ref) is required: make -C oracle ref")
    rng = np.random.default_rng(20261203)
    out = {}
    ok, err = _external_cases(rng)
    out["e_ok_count"] = np.array(len(ok))
    for k, q in enumerate(ok):
        cand, rid, st = oracle.ref_external_batch(q)
        out[f"e{k}_q"], out[f"e{k}_cand"], out[f"e{k}_rid"] = q, cand, rid
        out[f"e{k}_stats"] = np.array([st["primitive_count"], st["raw_hit_count"], st["compact_candidate_count"]],
                                      np.uint64)
        out[f"e{k}_recall"] = np.array(st["candidate_recall"])
        print(f"external batch {k}: {len(q)} queries -> {len(cand)} candidates")
    out["e_err_count"] = np.array(len(err))
    for k, q in enumerate(err):
        try:
            oracle.ref_external_batch(q)
            raise SystemExit(f"external error case {k} was accepted")
        except ValueError as e:
            out[f"x{k}_q"], out[f"x{k}_msg"] = q, np.array([str(e)])
            print(f"external error {k}: {e}")
    ok, err = _build_cases(rng)
    out["b_ok_count"] = np.array(len(ok))
    for k, (o, p, s) in enumerate(ok):
        prims = oracle.ref_build_proxy_scene(42 + k, o, p, s)
        out[f"b{k}_o"], out[f"b{k}_p"], out[f"b{k}_s"], out[f"b{k}_prims"] = o, p, s, prims
        print(f"proxy build {k}: {len(prims)} primitives")
    out["b_err_count"] = np.array(len(err))
    for k, (o, p, s) in enumerate(err):
        try:
            oracle.ref_build_proxy_scene(7, o, p, s)
            raise SystemExit(f"build error case {k} was accepted")
        except ValueError as e:
            out[f"y{k}_o"], out[f"y{k}_p"], out[f"y{k}_s"], out[f"y{k}_msg"] = o, p, s, np.array([str(e)])
            print(f"build error {k}: {e}")
    np.savez_compressed(OUT, **out)
    print("wrote", OUT)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
