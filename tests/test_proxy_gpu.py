"""gfx950 proxy-scene candidates against the reference's compiled CPU path
(golden vectors, tests/golden/proxy_golden.npz, including the KAT scene of
src/tests/cpp/test_rt_candidate.cpp:97-150) and against the oracle on a
larger seeded scene: raw hits identical field by field and in order,
compact candidates identical in key order, density identical."""
import numpy as np
import pytest

import oracle
from p2cccd import proxy

from test_oracle_proxy import golden_scenes

pytestmark = pytest.mark.gpu

RAW_FIELDS = ("query_id", "pair_key", "proxy_a_index", "proxy_b_index", "slab_id", "rt_hit_count", "flags",
              "motion_bound")
CAND_FIELDS = ("candidate_id", "query_id", "slab_id", "object_a_id", "patch_a_id", "object_b_id", "patch_b_id",
               "proxy_type_a", "proxy_type_b", "rt_hit_count", "motion_bound", "flags")


def _check(res, raw, cand, dens):
    r, c = res["raw_buffer"], res["candidates"]
    assert len(r) == len(raw) and len(c) == len(cand)
    for f in RAW_FIELDS:
        assert np.array_equal(r[f], raw[f]), f
    for f in CAND_FIELDS:
        assert np.array_equal(c[f].astype(np.uint64) if f != "motion_bound" else c[f],
                              cand[f].astype(np.uint64) if f != "motion_bound" else cand[f]), f
    for f in oracle.DENSITY.names:
        assert res["density"][f] == dens[f], f


def test_golden_scenes(gpu):
    n = 0
    for q, scene, raw, cand, dens in golden_scenes():
        _check(proxy.generate_candidates_for_proxy_scene(scene, q), raw, cand, dens)
        n += 1
    assert n == 7


def test_large_scene_against_oracle(gpu):
    rng = np.random.default_rng(31)
    s = oracle.random_proxy_scene(rng, 6000, 6, 3, extent=1.0, size=0.12)
    s["patch_id"] = (s["patch_id"] % 3).astype(np.uint32)
    res = proxy.generate_candidates_for_proxy_scene(s, 77)
    raw, cand, dens = oracle.proxy_generate(77, s)
    assert len(raw) > 1000 and len(cand) < len(raw)
    _check(res, raw, cand, dens)
