"""Pins the proxy-scene candidate oracle (oracle/proxy.c) against the
reference's own compiled GenerateRawCandidatesCpu / CompactRawCandidates /
ComputeCandidateDensityStats (src/cpp/rt_candidate/candidate_buffer.cpp:
144-228, candidate_stats.cpp:128-170) through golden vectors
(tests/golden/proxy_golden.npz, written by tests/golden/make_proxy_golden.py),
the reference's KAT (src/tests/cpp/test_rt_candidate.cpp:97-150), and a
brute-force restatement of the raw-hit set definition."""
from pathlib import Path

import numpy as np
import pytest

import oracle

GOLDEN = Path(__file__).resolve().parent / "golden" / "proxy_golden.npz"


def golden_scenes():
    g = np.load(GOLDEN)
    for i in range(int(g["count"])):
        scene = {k: g[f"s{i}_{k}"] for k in oracle.PROXY_FIELDS}
        yield int(g[f"s{i}_query"]), scene, g[f"s{i}_raw"], g[f"s{i}_cand"], g[f"s{i}_density"][0]


def _same(a, b, dtype):
    assert len(a) == len(b)
    for f in dtype.names:
        assert np.array_equal(a[f], b[f]), f


def test_oracle_matches_golden_vectors():
    n = 0
    for q, scene, raw, cand, dens in golden_scenes():
        r, c, d = oracle.proxy_generate(q, scene)
        _same(r, raw, oracle.RAW_HIT)
        _same(c, cand, oracle.CANDIDATE)
        _same(np.array([d]), np.array([dens]), oracle.DENSITY)
        n += 1
    assert n == 7


def test_reference_kat():
    q, scene, raw, cand, dens = next(golden_scenes())
    assert q == 42 and len(scene["t0"]) == 4
    assert len(raw) == 2 and (raw["query_id"] == 42).all() and ((raw["flags"] & 3) == 3).all()
    assert (raw["motion_bound"][:, 3] > 0).all()
    assert len(cand) == 2 and (cand["candidate_id"] != 0).all()
    assert (cand["object_a_id"] == 10).all() and (cand["object_b_id"] == 20).all() and (cand["rt_hit_count"] == 1).all()
    assert (dens["proxy_count"], dens["object_count"], dens["slab_count"]) == (4, 2, 2)
    assert dens["cross_object_same_slab_pair_count"] == 2 and dens["aabb_overlap_ratio"] == 1.0
    # duplicate compaction: the same raw hit twice accumulates (test_rt_candidate.cpp:136-141)
    r, c, _ = oracle.proxy_generate(q, scene)
    assert len(c) == 2


def test_raw_hits_equal_brute_force_definition():
    rng = np.random.default_rng(4)
    for n, objs, slabs in ((50, 2, 1), (300, 4, 3)):
        s = oracle.random_proxy_scene(rng, n, objs, slabs, extent=0.5, size=0.3)
        raw, _, _ = oracle.proxy_generate(7, s)
        want = []
        b = s["bounds"]
        for i in range(n):
            for j in range(i + 1, n):
                if s["object_id"][i] == s["object_id"][j] or s["slab_id"][i] != s["slab_id"][j]:
                    continue
                if not (s["t0"][i] < s["t1"][j] and s["t0"][j] < s["t1"][i]):
                    continue
                if np.all(b[i, 3:] >= b[j, :3]) and np.all(b[j, 3:] >= b[i, :3]):
                    want.append((i, j))
        assert [(int(h["proxy_a_index"]), int(h["proxy_b_index"])) for h in raw] == want
        assert len(want) > 10


@pytest.mark.skipif(not oracle.ref_available(), reason="compiled reference not built here")
def test_oracle_matches_compiled_reference_random():
    rng = np.random.default_rng(9)
    for n, objs, slabs in ((100, 3, 2), (400, 5, 1)):
        s = oracle.random_proxy_scene(rng, n, objs, slabs, extent=0.5, size=0.3)
        s["patch_id"] = (s["patch_id"] % 2).astype(np.uint32)
        a = oracle.proxy_generate(11, s)
        b = oracle.ref_proxy_generate(11, s)
        _same(a[0], b[0], oracle.RAW_HIT)
        _same(a[1], b[1], oracle.CANDIDATE)
