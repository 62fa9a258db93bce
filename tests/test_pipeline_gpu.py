"""End-to-end step on the GPU (LBVH candidates + (t,u,v) certificate on every
candidate) against the oracle (sort-and-sweep candidate sets + canonical
certificate per pair): identical candidate sets, identical hit decisions and
bit-identical TOI for every pair, in the plain, batched and graph-replayed
forms; zero false negatives follows from identity with the oracle."""
import numpy as np
import pytest

import oracle
from p2cccd import scenes
from p2cccd.pipeline import CcdPipeline

pytestmark = pytest.mark.gpu


def _positions(scene, kind, pairs):
    t = scene.topo
    if kind == "vf":
        vid = np.concatenate([pairs[:, :1], t.faces[pairs[:, 1]]], axis=1)
    else:
        vid = np.concatenate([t.edges[pairs[:, 0]], t.edges[pairs[:, 1]]], axis=1)
    return np.concatenate([scene.x0[vid], scene.x1[vid]], axis=1)


def _sorted(pairs, res):
    order = np.lexsort((pairs[:, 1], pairs[:, 0]))
    return pairs[order], res[order]


def _check_against_oracle(scene, r, threads=16):
    ref = oracle.candidates(scene)
    for kind in ("vf", "ee"):
        pairs, res = _sorted(r[kind], r[kind + "_res"])
        assert np.array_equal(pairs, ref[kind]), kind
        o = oracle.tuv_cert(kind == "ee", _positions(scene, kind, ref[kind]), threads=threads)
        assert np.array_equal(res["hit"], o["hit"].astype(res["hit"].dtype)), kind
        assert np.array_equal(res["toi"], o["toi"]), kind


@pytest.mark.parametrize("make", [
    lambda: scenes.two_icospheres(level=3, gap=-0.01),
    lambda: scenes.cloth_on_sphere(n=48, sphere_level=3, cloth_size=1.5),
])
def test_small_scene_all_forms_match_oracle(gpu, make):
    s = make()
    p = CcdPipeline(s)
    p.step_plain()
    r_plain = p.results()
    _check_against_oracle(s, r_plain)
    for use_graph in (False, True, True):
        p.step_batched(use_graph=use_graph)
        r = p.results()
        for kind in ("vf", "ee"):
            a = _sorted(r_plain[kind], r_plain[kind + "_res"])
            b = _sorted(r[kind], r[kind + "_res"])
            assert np.array_equal(a[0], b[0])
            assert np.array_equal(a[1]["hit"], b[1]["hit"]) and np.array_equal(a[1]["toi"], b[1]["toi"])


def test_two_icospheres_level6_matches_oracle(gpu):
    s = scenes.two_icospheres()
    p = CcdPipeline(s)
    p.step_batched()
    p.step_batched()  # a graph replay
    r = p.results()
    assert r["vf_res"]["hit"].sum() > 0
    _check_against_oracle(s, r)


@pytest.mark.parametrize("make", [
    lambda: scenes.two_icospheres(level=3, gap=-0.01),
    lambda: scenes.cloth_on_sphere(n=64, sphere_level=3, cloth_size=1.5),
])
@pytest.mark.parametrize("mode", ["stream", "plain"])
def test_compact_records_equal_full_records(gpu, make, mode):
    """64-byte pair records expanded from the vertex arrays give every pair
    the result of the 288-byte records, flags included, in the stream form
    with its wave pass and replay and in the plain form; in the plain form
    (one lane per pair, no shared pruning) the box count too."""
    s = make()
    out = {}
    for compact in (False, True):
        p = CcdPipeline(s)
        p.compact_records = compact
        p.batched_mode = mode
        p.step_batched(use_graph=False)
        p.step_batched(use_graph=True)
        r = p.results()
        out[compact] = r
    for kind in ("vf", "ee"):
        a = _sorted(out[False][kind], out[False][kind + "_res"])
        b = _sorted(out[True][kind], out[True][kind + "_res"])
        assert np.array_equal(a[0], b[0]), kind
        for f in ("hit", "toi", "flags") + (("boxes",) if mode == "plain" else ()):
            assert np.array_equal(a[1][f], b[1][f]), (kind, f)
    _check_against_oracle(s, out[True])
