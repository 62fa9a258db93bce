"""Batched broad phase (p2c_bp_batch_count / p2c_bp_batch_emit: one launch
per stage for a batch of resident scenes).  Every scene's pairs, in order,
equal the one-scene broad phase's on that scene alone, whatever else shares
the batch and whatever the hit-cache size; the sets equal the sort-and-sweep
oracle's; and compact certificate records set up for the whole batch's pair
list (p2c_tuv_setup_pairs_lite_batch) certify every pair exactly as the
one-scene records do."""
import numpy as np
import pytest
import torch

import oracle
from p2cccd import scenes, serve
from p2cccd.broadphase import BatchBroadPhase, BroadPhase, ResidentMesh
from p2cccd.tuv import TUV_RESULT_DTYPE, TUV_SCENE_DTYPE, TuvCertifier
from test_broadphase_gpu import _many_objects

pytestmark = pytest.mark.gpu


def _tri_scene():
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [3, 3, 3]], float)  # vertex 3 is in no face
    return scenes.Scene("tri", x, x + 0.1, scenes.prepare_topology(np.array([[0, 1, 2]], np.int32), 4), 0.0)


def _two_tris():
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 0.01], [1, 0, 0.01], [0, 1, 0.01]], float)
    x1 = x.copy()
    x1[3:, 2] -= 0.02
    f = np.array([[0, 1, 2], [3, 4, 5]], np.int32)
    return scenes.Scene("two_tris", x, x1, scenes.prepare_topology(f, 6), 0.0)


def _served(sid, tri, seed):
    s = serve.make_scene(sid, tri, seed)
    r = serve.step_request(0, sid, s, seed + 1)
    return scenes.Scene(s.name, r.x0, r.x1, s.topo, s.thickness)


def _scene_list():
    return [
        scenes.two_icospheres(level=2, gap=0.01),
        _tri_scene(),
        scenes.cloth_on_sphere(n=48, sphere_level=3, cloth_size=1.5, thickness=1e-4),
        _many_objects(65, 4),
        _two_tris(),
        scenes.two_icospheres(level=3, gap=-0.02, thickness=1e-3),
        _served(0, 3000, 7),
        _served(1, 40000, 8),
        _many_objects(3, 1),
    ]


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()


def _single(scene):
    return BroadPhase.for_scene(scene).candidates(_dev(scene.x0), _dev(scene.x1), scene.thickness)


def _items(meshes, pos, scs):
    return [(m, p0.data_ptr(), p1.data_ptr(), s.thickness) for m, (p0, p1), s in zip(meshes, pos, scs)]


@pytest.fixture(scope="module")
def batch_setup():
    scs = _scene_list()
    meshes = [ResidentMesh(s.topo, s.n_vertices) for s in scs]
    pos = [(_dev(s.x0), _dev(s.x1)) for s in scs]
    ref = [_single(s) for s in scs]
    return scs, meshes, pos, ref


def _check(offs, pairs, ref, order):
    for k, j in enumerate(order):
        vf, ee = ref[j]
        a, b, c = (int(v) for v in offs[2 * k: 2 * k + 3])
        assert torch.equal(pairs[a:b], vf), (k, j, "vf")
        assert torch.equal(pairs[b:c], ee), (k, j, "ee")


@pytest.mark.parametrize("slots", [16, 0, 1, 64])
def test_batch_equals_each_scene_alone(gpu, batch_setup, slots):
    scs, meshes, pos, ref = batch_setup
    assert sum(r[0].shape[0] + r[1].shape[0] for r in ref) > 100_000
    offs, pairs = BatchBroadPhase("cuda", slots).candidates(_items(meshes, pos, scs))
    assert offs[2 * len(scs) + 1] == 0
    _check(offs, pairs, ref, range(len(scs)))


def test_batch_composition_does_not_matter(gpu, batch_setup):
    """Sub-batches in other orders, with the same resident meshes and one
    batch workspace reused: the same pairs per scene."""
    scs, meshes, pos, ref = batch_setup
    bb = BatchBroadPhase("cuda", 8)
    rng = np.random.default_rng(3)
    for n in (1, 3, len(scs)):
        order = rng.permutation(len(scs))[:n]
        offs, pairs = bb.candidates(_items([meshes[j] for j in order], [pos[j] for j in order],
                                           [scs[j] for j in order]))
        _check(offs, pairs, ref, order)
    order = [5, 0, 7]  # again after the other batches
    offs, pairs = bb.candidates(_items([meshes[j] for j in order], [pos[j] for j in order], [scs[j] for j in order]))
    _check(offs, pairs, ref, order)


def test_batch_sets_match_oracle(gpu, batch_setup):
    scs, meshes, pos, ref = batch_setup
    pick = [0, 3, 4, 6]
    offs, pairs = BatchBroadPhase("cuda").candidates(_items([meshes[j] for j in pick], [pos[j] for j in pick],
                                                            [scs[j] for j in pick]))
    host = pairs.cpu().numpy()
    for k, j in enumerate(pick):
        want = oracle.candidates(scs[j])
        a, b, c = (int(v) for v in offs[2 * k: 2 * k + 3])
        for got, key in ((host[a:b], "vf"), (host[b:c], "ee")):
            got = got[np.lexsort((got[:, 1], got[:, 0]))] if len(got) else got.reshape(0, 2)
            assert np.array_equal(got, want[key]), (j, key)


def test_batch_lite_records_certify_like_each_scene_alone(gpu, batch_setup):
    scs, meshes, pos, ref = batch_setup
    pick = [0, 2, 4, 5, 6, 7]
    offs, pairs = BatchBroadPhase("cuda").candidates(_items([meshes[j] for j in pick], [pos[j] for j in pick],
                                                            [scs[j] for j in pick]))
    n = int(offs[2 * len(pick)])
    base = np.concatenate([[0], np.cumsum([scs[j].n_vertices for j in pick])])
    xb0 = torch.cat([pos[j][0] for j in pick]).contiguous()
    xb1 = torch.cat([pos[j][1] for j in pick]).contiguous()
    tab = np.zeros(len(pick), TUV_SCENE_DTYPE)
    for k, j in enumerate(pick):
        tab[k] = (meshes[j].faces.data_ptr(), meshes[j].edges.data_ptr(), base[k])
    dtab = torch.from_numpy(tab.view(np.uint8)).cuda()
    cert = TuvCertifier()
    cert.setup_batch_lite(xb0, xb1, dtab, len(pick), pairs, torch.from_numpy(offs).cuda(), n)
    cert.certify(n, "stream")
    got = cert.result_tensor(n).cpu().numpy().view(TUV_RESULT_DTYPE)
    assert got["hit"].sum() > 0
    one = TuvCertifier()
    for k, j in enumerate(pick):
        vf, ee = ref[j]
        nvf, nee = vf.shape[0], ee.shape[0]
        if nvf + nee == 0:
            continue
        one.reserve(nvf + nee)
        p0, p1 = pos[j]
        if nvf:
            one.setup_pairs(p0, p1, meshes[j].faces, vf, False, nvf, lite=True)
        if nee:
            one.setup_pairs(p0, p1, meshes[j].edges, ee, True, nee, out_offset=nvf, lite=True)
        one.certify(nvf + nee, "stream")
        want = one.result_tensor(nvf + nee).cpu().numpy().view(TUV_RESULT_DTYPE)
        a = int(offs[2 * k])
        mine = got[a:a + nvf + nee]
        # the results; `boxes` counts work, which depends on the schedule: the
        # pairs a draining stream wave hands to the wave pass, where lanes and
        # pieces prune with a shared best, depend on what else is in flight
        for f in ("hit", "toi", "flags"):
            assert np.array_equal(mine[f], want[f]), (j, f)
