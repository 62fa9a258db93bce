"""Scheduled any-hit certification on the GPU against the reference's
scheduled rows (src/tools/scene_object_envelope_native_ti_walltime.cpp:
1283-1410) restated over oracle outcomes: per group, the top-K candidates by
score (descending; ties by index), then the rest by index, certified until the
first hit.  The first-hit position of every group, the reference's exact-call
count and the detected groups must match; detected groups must equal the
groups holding any hit (zero false negatives).  Spatial cells keep 64-4096
pairs and the group scorer runs on their features."""
import numpy as np
import pytest
import torch

import oracle
from p2cccd import scenes, schedule, stpf
from p2cccd.broadphase import BroadPhase
from p2cccd.tuv import TUV_RESULT_DTYPE, TuvCertifier

pytestmark = pytest.mark.gpu

WEIGHTS = oracle.HERE.parent / "tests" / "golden" / "frozen_tiny_stpf_weights.txt"


def _reference_rows(keys, scores, hits, top_k):
    """Per group (ascending key): sequence = top_k by (-score, index), then
    the rest by index; first hit position or -1; exact calls."""
    out = {}
    for k in np.unique(keys):
        members = np.flatnonzero(keys == k)
        by_score = members[np.lexsort((members, -scores[members]))]
        top = by_score[:top_k]
        rest = np.sort(by_score[top_k:])
        seq = np.concatenate([top, rest])
        h = np.flatnonzero(hits[seq])
        out[int(k)] = (int(h[0]) if h.size else -1, len(seq))
    return out


def _scene_candidates():
    s = scenes.two_icospheres(level=3, gap=-0.01)
    bp = BroadPhase.for_scene(s, "cuda")
    x0, x1 = torch.from_numpy(s.x0).cuda(), torch.from_numpy(s.x1).cuda()
    vf, ee = bp.candidates(x0, x1, s.thickness)
    return s, x0, x1, vf, ee


def _positions(scene, kind, pairs):
    t = scene.topo
    if kind == "vf":
        vid = np.concatenate([pairs[:, :1], t.faces[pairs[:, 1]]], axis=1)
    else:
        vid = np.concatenate([t.edges[pairs[:, 0]], t.edges[pairs[:, 1]]], axis=1)
    return np.concatenate([scene.x0[vid], scene.x1[vid]], axis=1)


@pytest.mark.parametrize("policy,top_k", [("learned", 32), ("proximity", 4096), ("random", 1)])
def test_object_pair_groups_match_reference_rows(gpu, policy, top_k):
    s, x0, x1, vf, ee = _scene_candidates()
    t = s.topo
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    w = torch.from_numpy(stpf.load_tiny_stpf_weights(WEIGHTS)).cuda()
    cert = TuvCertifier()
    for kind, pairs, topo, oa, ob in (("vf", vf, t.faces, t.vertex_object, t.face_object),
                                      ("ee", ee, t.edges, t.edge_object, t.edge_object)):
        n = pairs.shape[0]
        scores, keys = stpf.tiny_scores(kind, x0, x1, dev(topo), pairs, n, dev(oa), dev(ob), policy, w)
        g = schedule.order_groups(keys, scores, top_k)
        r = schedule.any_hit(cert, kind, x0, x1, dev(topo), pairs, g)
        p = pairs.cpu().numpy()
        hits = oracle.tuv_cert(kind == "ee", _positions(s, kind, p), threads=8)["hit"].astype(bool)
        ref = _reference_rows(keys.cpu().numpy(), scores.cpu().numpy(), hits, top_k)
        gk = g.keys.cpu().numpy()
        assert sorted(ref) == list(gk)
        for i, k in enumerate(gk):
            assert (int(r.first_hit[i]), int(r.sizes[i])) == ref[int(k)], (kind, k)
        assert r.exact_calls == sum(f + 1 if f >= 0 else z for f, z in ref.values())
        assert r.certified >= r.exact_calls
        # zero false negatives: a group is detected iff it holds a hit
        kk = keys.cpu().numpy()
        for i, k in enumerate(gk):
            assert (r.first_hit[i] >= 0) == bool(hits[kk == k].any())


def test_spatial_cells_and_group_priority(gpu):
    s, x0, x1, vf, ee = _scene_candidates()
    t = s.topo
    topo = torch.from_numpy(t.edges).cuda()
    rows = schedule.candidate_rows(x0, x1, topo, ee, "ee")
    cell = schedule.spatial_cells(rows, 64, 4096)
    sizes = torch.bincount(cell).cpu().numpy()
    assert sizes.sum() == ee.shape[0] and sizes.max() <= 4096
    assert (sizes[:-1] >= 64).all()
    idx = torch.arange(rows.shape[0], device="cuda")
    f = stpf.tiny_features("ee", rows.contiguous(), idx)
    gf = schedule.group_features(rows, f, cell, int(cell.max()) + 1)
    assert gf.shape == (len(sizes), 32) and torch.isfinite(gf).all()
    torch.manual_seed(0)
    m = stpf.StpfPriority().cuda()
    pr = stpf.stpf_priority(m.packed().cuda(), gf)
    with torch.no_grad():
        assert torch.allclose(pr, m(gf), rtol=1e-4, atol=2e-6)
    # scheduling by cells: every cell's decision is exact
    cert = TuvCertifier()
    scores = -f[:, 11]  # closest vertex gaps first
    g = schedule.order_groups(cell, scores, 4096)
    r = schedule.any_hit(cert, "ee", x0, x1, topo, ee, g)
    hits = oracle.tuv_cert(True, _positions(s, "ee", ee.cpu().numpy()), threads=8)["hit"].astype(bool)
    c = cell.cpu().numpy()
    for i, k in enumerate(g.keys.cpu().numpy()):
        assert (r.first_hit[i] >= 0) == bool(hits[c == k].any())


def test_dense_group_early_stop_on_gpu_certificate_outcomes(gpu):
    # RunDenseGroupExactEarlyStop driven by the GPU certificate's own outcomes
    # (contact flag, TOI, boxes as cost) instead of cost traces; with the
    # proposed interval the whole step, every contact lies inside it, so the
    # groups holding a contact are exactly the detected ones.
    from p2cccd import policy

    s, x0, x1, vf, ee = _scene_candidates()
    t = s.topo
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    cert = TuvCertifier()
    pairs, topo = ee, dev(t.edges)
    n = pairs.shape[0]
    cert.reserve(n)
    cert.setup_pairs(x0, x1, topo, pairs, True, n)
    cert.certify(n, "stream")
    torch.cuda.synchronize()
    res = cert.result_tensor(n).cpu().numpy().view(TUV_RESULT_DTYPE)
    scores, keys = stpf.tiny_scores("ee", x0, x1, topo, pairs, n, dev(t.edge_object), dev(t.edge_object), "proximity")
    cell = schedule.spatial_cells(schedule.candidate_rows(x0, x1, topo, pairs, "ee"))
    d = policy.dense_group_from_certificates(cell.cpu().numpy() + 1, res["hit"], res["toi"], res["boxes"],
                                             scores.cpu().numpy().astype(np.float32))
    st = policy.run_dense_group_exact_early_stop(d)
    positive = len(np.unique(d["query_id"][d["candidate_collides"]]))
    assert st["positive_group_count"] == positive > 0
    assert st["tp"] == positive and st["fn"] == 0 and st["fp"] == 0
    assert st["learned_exact_calls"] < st["no_proposal_exact_calls"]
