"""gfx950 LBVH broad phase against the sort-and-sweep oracle: candidate sets
identical as sets (VF and EE), on small scenes and on the full-size two
icosphere (level 6) and folded-cloth (256 x 256) scenes."""
import numpy as np
import pytest
import torch

import oracle
from p2cccd import scenes
from p2cccd.broadphase import BroadPhase

pytestmark = pytest.mark.gpu


def _gpu_sets(scene):
    bp = BroadPhase.for_scene(scene)
    x0 = torch.from_numpy(scene.x0).cuda()
    x1 = torch.from_numpy(scene.x1).cuda()
    vf, ee = bp.candidates(x0, x1, scene.thickness)
    out = {}
    for k, t in (("vf", vf), ("ee", ee)):
        a = t.cpu().numpy()
        out[k] = a[np.lexsort((a[:, 1], a[:, 0]))] if len(a) else a.reshape(0, 2)
    return out, (vf, ee)


def _compare(scene):
    got, raw = _gpu_sets(scene)
    ref = oracle.candidates(scene)
    for k in ("vf", "ee"):
        assert got[k].shape == ref[k].shape, (k, got[k].shape, ref[k].shape)
        assert np.array_equal(got[k], ref[k]), k
    return raw


@pytest.mark.parametrize("make", [
    lambda: scenes.two_icospheres(level=2, gap=0.01),
    lambda: scenes.two_icospheres(level=3, gap=-0.02, thickness=1e-3),
    lambda: scenes.cloth_on_sphere(n=16, sphere_level=1, cloth_size=1.0),
    lambda: scenes.cloth_on_sphere(n=48, sphere_level=3, cloth_size=1.5, thickness=1e-4),
])
def test_small_scenes_match_oracle(gpu, make):
    _compare(make())


def test_single_triangle_and_tiny_meshes(gpu):
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float)
    s = scenes.Scene("tri", x, x + 0.1, scenes.prepare_topology(np.array([[0, 1, 2]], np.int32), 3), 0.0)
    got, _ = _gpu_sets(s)
    assert len(got["vf"]) == 0 and len(got["ee"]) == 0


def test_two_icospheres_level6_matches_oracle(gpu):
    vf, ee = _compare(scenes.two_icospheres())
    assert vf.shape[0] > 1_000_000


def test_folded_cloth_matches_oracle_and_is_deterministic(gpu):
    s = scenes.cloth_on_sphere()
    vf, ee = _compare(s)
    assert vf.shape[0] + ee.shape[0] > 10_000_000
    _, (vf2, ee2) = _gpu_sets(s)
    assert torch.equal(vf, vf2) and torch.equal(ee, ee2)


@pytest.mark.parametrize("slots", [0, 1, 16])
def test_hit_cache_size_does_not_change_the_output(gpu, slots):
    """The count pass's hit cache only saves the fill its traversal: any
    cache size (none, one partner, some) gives the same pairs in the same
    order as the default."""
    s = scenes.two_icospheres(level=4, gap=-0.01)
    x0, x1 = torch.from_numpy(s.x0).cuda(), torch.from_numpy(s.x1).cuda()
    ref = BroadPhase.for_scene(s, "cuda").candidates(x0, x1, s.thickness)
    got = BroadPhase.for_scene(s, "cuda", cache_slots=slots).candidates(x0, x1, s.thickness)
    assert ref[0].shape[0] > 1000 and ref[1].shape[0] > 1000
    assert torch.equal(ref[0], got[0]) and torch.equal(ref[1], got[1])
