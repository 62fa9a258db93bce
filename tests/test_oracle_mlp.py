"""Pins the TinySTPF oracle (oracle/mlp.c), the restatement of
src/tools/scene_object_envelope_native_ti_walltime.cpp:773-782 (Mix64 /
UnitHash), :893-993 (features), :995-1066 (weights file), :1068-1131
(PredictTinyStpf, policies): hand-derived features on simple geometry, the
published SplitMix64 sequence, torch float64 as the library oracle of the
MLP on the reference's frozen weights (tests/golden/frozen_tiny_stpf_weights
.txt, a data file of the reference's benchmark run
src/benchmark/scene_object_envelope_strong_native_run_id/), planted
mistakes, and the product's weight loader."""
import math

import numpy as np
import pytest
import torch

import oracle
from p2cccd import stpf

WEIGHTS = oracle.HERE.parent / "tests" / "golden" / "frozen_tiny_stpf_weights.txt"


def _rows(a, b):
    return np.array([list(a) + list(b)], dtype=np.float64)


def test_splitmix64_known_answers():
    # SplitMix64 (Steele, Lea, Flood 2014) from state 0: the first outputs are
    # mix64(k * 0x9e3779b97f4a7c15) for k = 0, 1, 2 with the increment inside mix64
    L = oracle.lib()
    L.orc_mix64.argtypes = [oracle.ctypes.c_uint64]
    L.orc_mix64.restype = oracle.ctypes.c_uint64
    gamma = 0x9E3779B97F4A7C15
    outs = [L.orc_mix64((k * gamma) & (2**64 - 1)) for k in range(3)]
    assert outs == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    assert oracle.lib().orc_unit_hash(0) == (0xE220A8397B1DCDAF >> 11) / 2.0**53


def test_features_vertex_face_by_hand():
    tri = [(0, 0, 0), (1, 0, 0), (0, 1, 0)]
    f = oracle.tiny_features(True, _rows([(0, 0, 1)] + tri, [(0, 0, -1)] + tri), np.array([5]))[0]
    want = {0: 1.0, 1: 0.0, 2: 1.0, 3: 1.0, 4: 2.0, 5: math.sqrt(6), 6: math.log1p(2.0), 7: 1.0, 8: 2.0, 9: 1.0,
            10: 1.0, 11: 1.0, 12: 0.25, 13: 0.25, 14: 0.0, 15: 0.5, 16: 1.0, 17: math.sqrt(0.625), 18: 2.0,
            19: 0.0, 20: 0.0, 21: 0.0, 22: 1.0, 23: math.sqrt(2), 24: math.sqrt(2), 25: 0.5,
            26: 2.0 / (1.0 + math.sqrt(6)), 27: 0.0, 28: 0.0, 29: 0.5, 30: math.log1p(5.0), 31: 1.0}
    for k, v in want.items():
        assert f[k] == pytest.approx(v, rel=1e-15, abs=1e-15), k


def test_features_edge_edge_by_hand():
    # edge a along x moving +y by 1, edge b static along z at (0.5, 2, *)
    a0, a1, b0, b1 = (0, 0, 0), (1, 0, 0), (0.5, 2, -1), (0.5, 2, 1)
    f = oracle.tiny_features(False, _rows([a0, a1, b0, b1], [(0, 1, 0), (1, 1, 0), b0, b1]), np.array([0]))[0]
    assert (f[0], f[1]) == (0.0, 1.0)
    assert list(f[2:5]) == [1.0, 2.0, 2.0]
    assert f[7] == 1.0 and f[8] == 1.0          # relative motion of the means, largest vertex motion
    assert f[9] == 1.0 and f[10] == 1.0          # min vertex gap at t0 and t1 (the edge lengths)
    assert f[27] == 0.0 and f[28] == 0.0         # motion[0] . motion[3], |motion[0] x motion[3]| (b static)
    assert f[22] == 1.0 and f[23] == pytest.approx(math.sqrt(0.25 + 4 + 1), rel=1e-15)
    assert f[30] == 0.0 and f[25] == 0.5


def test_weight_loader_matches_file_and_reference_messages(tmp_path):
    w = stpf.load_tiny_stpf_weights(WEIGHTS)
    toks = WEIGHTS.read_text().split()
    assert w.size == 4289 and w[0] == float(toks[toks.index("mean") + 1]) and w[-1] == float(toks[-1])
    bad = tmp_path / "bad.txt"
    bad.write_text(WEIGHTS.read_text().replace("hidden0 64", "hidden0 65"))
    with pytest.raises(RuntimeError, match="unexpected TinySTPF dimensions"):
        stpf.load_tiny_stpf_weights(bad)
    bad.write_text(WEIGHTS.read_text().replace("version 1", "version 2"))
    with pytest.raises(RuntimeError, match="unsupported STPF weight version"):
        stpf.load_tiny_stpf_weights(bad)
    bad.write_text(WEIGHTS.read_text().replace("w1 ", "wx "))
    with pytest.raises(RuntimeError, match="bad STPF weight tag: expected w1, got wx"):
        stpf.load_tiny_stpf_weights(bad)
    with pytest.raises(RuntimeError, match="failed to open STPF weights"):
        stpf.load_tiny_stpf_weights(tmp_path / "missing.txt")


def _torch_predict(w, feats):
    t = torch.from_numpy
    mean, sd = t(w[:32]), t(w[32:64])
    w0 = t(w[64:64 + 2048]).reshape(64, 32)
    b0 = t(w[2112:2176])
    w1 = t(w[2176:2176 + 2048]).reshape(32, 64)
    b1 = t(w[4224:4256])
    w2 = t(w[4256:4288])
    b2 = float(w[4288])
    x = (t(feats) - mean) / torch.clamp(sd, min=1e-12)
    h0 = torch.relu(x @ w0.T + b0)
    h1 = torch.relu(h0 @ w1.T + b1)
    return (h1 @ w2 + b2).numpy()


def _features(n, seed):
    rng = np.random.default_rng(seed)
    rows = rng.uniform(-1, 1, (n, 8, 3))
    rows[:, 4:] = rows[:, :4] + rng.normal(0, 0.1, (n, 4, 3))
    return oracle.tiny_features(True, rows[: n // 2], np.arange(n // 2)), \
        oracle.tiny_features(False, rows[n // 2:], np.arange(n // 2, n))


def test_predict_matches_torch_float64_on_frozen_weights():
    w = stpf.load_tiny_stpf_weights(WEIGHTS)
    fv, fe = _features(400, 3)
    feats = np.concatenate([fv, fe])
    mine = oracle.tiny_predict(w, feats)
    ref = _torch_predict(w, feats)
    assert np.allclose(mine, ref, rtol=1e-12, atol=1e-12)
    assert np.std(mine) > 0.1  # the frozen model is not constant on these inputs


@pytest.mark.parametrize("mutation", [1, 2])
def test_planted_mistakes_are_caught(mutation):
    w = stpf.load_tiny_stpf_weights(WEIGHTS)
    fv, fe = _features(200, 4)
    feats = np.concatenate([fv, fe])
    ref = _torch_predict(w, feats)
    oracle.set_mlp_mutation(mutation)
    try:
        mine = oracle.tiny_predict(w, feats)
    finally:
        oracle.set_mlp_mutation(0)
    assert not np.allclose(mine, ref, rtol=1e-9, atol=1e-9)


def test_policies_compose_as_the_reference():
    w = stpf.load_tiny_stpf_weights(WEIGHTS)
    rng = np.random.default_rng(8)
    rows = rng.uniform(-1, 1, (50, 8, 3))
    idx = np.arange(50, dtype=np.uint64) * 7
    keys = (np.arange(50, dtype=np.uint64) % 5) << np.uint64(32) | np.uint64(3)
    f = oracle.tiny_features(True, rows, idx)
    learned = oracle.tiny_scores(0, True, rows, idx, keys, w)
    assert np.array_equal(learned, oracle.tiny_predict(w, f))
    resid = oracle.tiny_scores(1, True, rows, idx, keys, w)
    sig = 1.0 / (1.0 + np.exp(-learned))
    assert np.allclose(resid, f[:, 25] + 0.08 * sig + 0.004 * f[:, 26] - 0.00025 * f[:, 30], rtol=1e-14)
    assert np.array_equal(oracle.tiny_scores(2, True, rows, idx, keys, w), f[:, 25])
    assert np.array_equal(oracle.tiny_scores(3, True, rows, idx, keys, w), f[:, 8])
    rnd = oracle.tiny_scores(4, True, rows, idx, keys, None)
    L = oracle.lib()
    want = [L.orc_unit_hash(int(i) ^ 0x767666 ^ ((int(k) << 1) & (2**64 - 1))) for i, k in zip(idx, keys)]
    assert np.array_equal(rnd, np.array(want)) and np.all((rnd >= 0) & (rnd < 1))
