"""Serving on the GPU: requests of mixed scene sizes share batched launches
under an HBM budget small enough to force evictions, and every request's
candidate pairs, hit flags and TOIs are bit-identical to serving it alone.
The contacts a completion carries (pairs and TOIs gathered on the device)
are exactly the hits among its full per-candidate results."""
import numpy as np
import pytest

from p2cccd import serve

pytestmark = pytest.mark.gpu


def _contacts(full, n_vf):
    hit = full["hit"] != 0
    return full["vf"][hit[:n_vf]], full["toi"][:n_vf][hit[:n_vf]], full["ee"][hit[n_vf:]], full["toi"][n_vf:][hit[n_vf:]]


def test_each_result_is_the_same_alone_and_in_the_mix(gpu):
    sizes = [1200, 5000, 20000, 2500, 40000, 9000]
    table = {i: serve.make_scene(i, n, 11 + i) for i, n in enumerate(sizes)}
    reqs = [serve.step_request(k, k % len(sizes), table[k % len(sizes)], 500 + k) for k in range(14)]
    big = max(serve.ResidentScene.estimate_bytes(s) for s in table.values())
    srv = serve.Server(table, hbm_budget=int(1.3 * big), max_batch_requests=4, full_results=True)
    mixed = {c.rid: c for c in srv.run(reqs, realtime=False)}
    assert len(mixed) == len(reqs) and srv.batches >= 4 and srv.res.evictions > 0
    for r in reqs:
        m = mixed[r.rid]
        # the device-side gather returns exactly the hits of the full results
        for a, b in zip((m.vf_hits, m.vf_toi, m.ee_hits, m.ee_toi), _contacts(m.full, m.n_vf)):
            assert np.array_equal(a, b), r.rid
        assert len(m.full["vf"]) == m.n_vf and len(m.full["ee"]) == m.n_ee
        alone = serve.Server({r.scene_id: table[r.scene_id]}, hbm_budget=8 * big, full_results=True).run(
            [serve.Request(r.rid, r.scene_id, r.x0, r.x1)], realtime=False)[0]
        assert (m.n_vf, m.n_ee) == (alone.n_vf, alone.n_ee), r.rid
        for k in ("vf", "ee", "hit", "toi"):
            assert np.array_equal(m.full[k], alone.full[k]), (r.rid, k)
        for a, b in zip((m.vf_hits, m.vf_toi, m.ee_hits, m.ee_toi),
                        (alone.vf_hits, alone.vf_toi, alone.ee_hits, alone.ee_toi)):
            assert np.array_equal(a, b), r.rid
    assert sum(len(c.vf_hits) + len(c.ee_hits) for c in mixed.values()) > 0
    rep = serve.report(list(mixed.values()), 1.0)
    assert rep["p99_ms"] >= rep["p50_ms"] > 0


def test_contacts_only_serving_matches_full_results(gpu):
    sizes = [3000, 15000]
    table = {i: serve.make_scene(i, n, 31 + i) for i, n in enumerate(sizes)}
    reqs = [serve.step_request(k, k % 2, table[k % 2], 900 + k) for k in range(6)]
    big = max(serve.ResidentScene.estimate_bytes(s) for s in table.values())
    lean = {c.rid: c for c in serve.Server(table, 4 * big).run(reqs, realtime=False)}
    full = {c.rid: c for c in serve.Server(table, 4 * big, full_results=True).run(reqs, realtime=False)}
    for r in reqs:
        a, b = lean[r.rid], full[r.rid]
        assert a.full is None and (a.n_vf, a.n_ee) == (b.n_vf, b.n_ee)
        for x, y in zip((a.vf_hits, a.vf_toi, a.ee_hits, a.ee_toi), _contacts(b.full, b.n_vf)):
            assert np.array_equal(x, y), r.rid
        assert a.earliest_toi == b.earliest_toi
