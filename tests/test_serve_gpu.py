"""Serving on the GPU: requests of mixed scene sizes share batched launches
under an HBM budget small enough to force evictions, and every request's
candidate pairs, hit flags and TOIs are bit-identical to serving it alone."""
import numpy as np
import pytest

from p2cccd import serve

pytestmark = pytest.mark.gpu


def test_each_result_is_the_same_alone_and_in_the_mix(gpu):
    sizes = [1200, 5000, 20000, 2500, 40000, 9000]
    table = {i: serve.make_scene(i, n, 11 + i) for i, n in enumerate(sizes)}
    reqs = [serve.step_request(k, k % len(sizes), table[k % len(sizes)], 500 + k) for k in range(14)]
    big = max(serve.ResidentScene.estimate_bytes(s) for s in table.values())
    srv = serve.Server(table, hbm_budget=int(1.3 * big), max_batch_requests=4)
    mixed = {c.rid: c for c in srv.run(reqs, realtime=False)}
    assert len(mixed) == len(reqs) and srv.batches >= 4 and srv.res.evictions > 0
    for r in reqs:
        alone = serve.Server({r.scene_id: table[r.scene_id]}, hbm_budget=8 * big).run(
            [serve.Request(r.rid, r.scene_id, r.x0, r.x1)], realtime=False)[0]
        m = mixed[r.rid]
        assert np.array_equal(m.vf, alone.vf) and np.array_equal(m.ee, alone.ee), r.rid
        assert np.array_equal(m.hit, alone.hit) and np.array_equal(m.toi, alone.toi), r.rid
    assert sum(c.hit.sum() for c in mixed.values()) > 0
    rep = serve.report(list(mixed.values()), 1.0)
    assert rep["p99_ms"] >= rep["p50_ms"] > 0
