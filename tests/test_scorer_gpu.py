"""gfx950 scheduling scorers against the oracle: TinySTPF features, MLP and
every score policy (oracle/mlp.c; reference
src/tools/scene_object_envelope_native_ti_walltime.cpp:756-1131) on broad-phase
candidates of a small scene, and the FP32 STPF priority kernel against torch
(src/python/p2cccd/proposal/stpf_model.py:45-93).  Scores agree to 1e-12
relative (log1p and exp may differ from the host's in the last bit); group keys
and the hash policy are exact."""
import numpy as np
import pytest
import torch

import oracle
from p2cccd import scenes, stpf
from p2cccd.broadphase import BroadPhase

pytestmark = pytest.mark.gpu

WEIGHTS = oracle.HERE.parent / "tests" / "golden" / "frozen_tiny_stpf_weights.txt"


def _rows(scene, kind, pairs):
    t = scene.topo
    if kind == "vf":
        vid = np.concatenate([pairs[:, :1], t.faces[pairs[:, 1]]], axis=1)
    else:
        vid = np.concatenate([t.edges[pairs[:, 0]], t.edges[pairs[:, 1]]], axis=1)
    return np.concatenate([scene.x0[vid], scene.x1[vid]], axis=1)


def _keys(scene, kind, pairs):
    t = scene.topo
    if kind == "vf":
        a, b = t.vertex_object[pairs[:, 0]], t.face_object[pairs[:, 1]]
    else:
        a, b = t.edge_object[pairs[:, 0]], t.edge_object[pairs[:, 1]]
    lo, hi = np.minimum(a, b).astype(np.uint64), np.maximum(a, b).astype(np.uint64)
    return (lo << np.uint64(32)) | hi


def test_features_and_predict_match_oracle(gpu):
    rng = np.random.default_rng(21)
    rows = rng.uniform(-1, 1, (3000, 8, 3))
    rows[:, 4:] = rows[:, :4] + rng.normal(0, 0.2, (3000, 4, 3))
    idx = rng.integers(0, 1 << 40, 3000).astype(np.uint64)
    w = stpf.load_tiny_stpf_weights(WEIGHTS)
    wd = torch.from_numpy(w).cuda()
    for kind in ("vf", "ee"):
        got = stpf.tiny_features(kind, torch.from_numpy(rows).cuda(), torch.from_numpy(idx.view(np.int64)).cuda())
        ref = oracle.tiny_features(kind == "vf", rows, idx)
        got = got.cpu().numpy()
        exact = [k for k in range(32) if k not in (6, 30)]
        assert np.array_equal(got[:, exact], ref[:, exact]), kind
        assert np.allclose(got[:, [6, 30]], ref[:, [6, 30]], rtol=1e-15, atol=0)
        pred = stpf.tiny_predict(wd, torch.from_numpy(ref).cuda()).cpu().numpy()
        assert np.array_equal(pred, oracle.tiny_predict(w, ref)), kind


@pytest.mark.parametrize("policy", list(stpf.POLICIES))
def test_candidate_scores_and_group_keys(gpu, policy):
    s = scenes.two_icospheres(level=3, gap=-0.01)
    bp = BroadPhase.for_scene(s, "cuda")
    x0, x1 = torch.from_numpy(s.x0).cuda(), torch.from_numpy(s.x1).cuda()
    vf, ee = bp.candidates(x0, x1, s.thickness)
    w = stpf.load_tiny_stpf_weights(WEIGHTS)
    wd = torch.from_numpy(w).cuda()
    t = s.topo
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    for kind, pairs, topo, oa, ob in (("vf", vf, t.faces, t.vertex_object, t.face_object),
                                      ("ee", ee, t.edges, t.edge_object, t.edge_object)):
        n = pairs.shape[0]
        assert n > 100
        sc, keys = stpf.tiny_scores(kind, x0, x1, dev(topo), pairs, n, dev(oa), dev(ob), policy, wd)
        p = pairs.cpu().numpy()
        want_keys = _keys(s, kind, p)
        assert np.array_equal(keys.cpu().numpy().view(np.uint64), want_keys)
        ref = oracle.tiny_scores(stpf.POLICIES[policy], kind == "vf", _rows(s, kind, p), np.arange(n, dtype=np.uint64),
                                 want_keys, w)
        got = sc.cpu().numpy()
        if policy in ("random", "motion"):
            assert np.array_equal(got, ref), (policy, kind)
        else:
            assert np.allclose(got, ref, rtol=1e-12, atol=1e-15), (policy, kind)


def test_learned_policy_without_weights_is_rejected(gpu):
    from p2cccd import _native

    x = torch.zeros((4, 3), dtype=torch.float64, device="cuda")
    i = torch.zeros((1, 3), dtype=torch.int32, device="cuda")
    with pytest.raises(_native.NativeError, match="learned policy requested without STPF weights"):
        stpf.tiny_scores("vf", x, x, i, i, 1, i, i, "learned", None)


def test_stpf_priority_kernel_matches_torch_fp32(gpu):
    torch.manual_seed(5)
    m = stpf.StpfPriority().cuda()
    x = torch.randn(5000, 32, device="cuda") * 2.0
    got = stpf.stpf_priority(m.packed().cuda(), x)
    with torch.no_grad():
        ref = m(x)
    assert torch.allclose(got, ref, rtol=1e-4, atol=2e-6)
    assert float(ref.std()) > 0.01
