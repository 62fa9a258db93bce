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


def _many_objects(n_obj, seed, spacing=0.3, level=0):
    """n_obj small icospheres on a jittered grid, each moving on its own:
    neighbours' envelopes overlap or not depending on the jitter."""
    rng = np.random.default_rng(seed)
    v, f = scenes.icosphere(level, 0.1)
    side = int(np.ceil(n_obj ** (1 / 3)))
    xs, fs = [], []
    for k in range(n_obj):
        c = np.array([k % side, (k // side) % side, k // (side * side)], float) * spacing
        c += rng.uniform(-0.08, 0.08, 3)
        xs.append(v + c)
        fs.append(f + k * len(v))
    x0 = np.concatenate(xs)
    x1 = x0 + np.repeat(rng.normal(0.0, 0.04, (n_obj, 3)), len(v), axis=0)
    faces = np.concatenate(fs).astype(np.int32)
    return scenes.Scene(f"objects{n_obj}", x0, x1, scenes.prepare_topology(faces, len(x0)), 1e-3)


@pytest.mark.parametrize("n_obj,seed", [(3, 1), (63, 2), (64, 3), (65, 4), (700, 5), (2000, 6)])
def test_many_object_scenes_match_oracle(gpu, n_obj, seed):
    s = _many_objects(n_obj, seed)
    print(f"objects: {n_obj} seed {seed}", flush=True)
    assert s.topo.object_count == n_obj
    _compare(s)


def test_disjoint_envelopes_exclude_object_pairs(gpu, tmp_path):
    """Three objects read from PLY frames; the envelopes of objects 0 and 2
    do not overlap (the reference's object_pair_allowed, :528-550), so no
    candidate pairs them, though 0-1 and 1-2 pairs exist."""
    import struct

    from p2cccd.datasets.ply import scene_from_ply_frames

    v, f = scenes.icosphere(1, 0.5)
    x0 = np.concatenate([v + (0, 0, 0), v + (0.95, 0, 0), v + (1.9, 0, 0)])
    x1 = x0 + 0.01
    faces = np.concatenate([f, f + len(v), f + 2 * len(v)])

    def write(path, x):
        head = (f"ply\nformat binary_little_endian 1.0\nelement vertex {len(x)}\nproperty double x\n"
                f"property double y\nproperty double z\nelement face {len(faces)}\n"
                f"property list uchar int vertex_indices\nend_header\n").encode()
        body = b"".join(struct.pack("<3d", *p) for p in x) + b"".join(struct.pack("<B3i", 3, *t) for t in faces)
        path.write_bytes(head + body)
        return path

    s = scene_from_ply_frames(write(tmp_path / "f0.ply", x0), write(tmp_path / "f1.ply", x1), thickness=1e-3)
    assert s.topo.object_count == 3
    got, _ = _gpu_sets(s)
    ref = oracle.candidates(s)
    for k in ("vf", "ee"):
        assert np.array_equal(got[k], ref[k]), k
    vo, fo, eo = s.topo.vertex_object, s.topo.face_object, s.topo.edge_object
    pairs_vf = {tuple(sorted((int(vo[a]), int(fo[b])))) for a, b in got["vf"]}
    pairs_ee = {tuple(sorted((int(eo[a]), int(eo[b])))) for a, b in got["ee"]}
    assert (0, 1) in pairs_vf and (1, 2) in pairs_vf and (0, 2) not in pairs_vf
    assert (0, 1) in pairs_ee and (0, 2) not in pairs_ee


def test_traversal_visit_replay_matches_a_python_walk(gpu):
    """The host replay of the count and emit walks (used for the broad-phase
    roofline) against a plain Python walk of the same exported trees: the
    count walk of each query from its start node, and the emit walk of the
    queries with more candidates than the hit cache from their resume node."""
    from p2cccd.broadphase import BroadPhase, _LAYOUT_FIELDS
    from p2cccd import _native as nat

    s = scenes.two_icospheres(level=2, gap=-0.02, thickness=1e-3)
    for slots in (0, 2):
        bp = BroadPhase.for_scene(s, cache_slots=slots)
        x0 = torch.from_numpy(s.x0).cuda()
        x1 = torch.from_numpy(s.x1).cuda()
        bp.candidates(x0, x1, s.thickness)
        got = bp.traversal_visits()
        lay = np.zeros(len(_LAYOUT_FIELDS), np.uint64)
        nat.call("p2c_bp_workspace_layout", bp.n_vertices, bp.n_faces, bp.n_edges, bp.n_objects, lay.ctypes.data)
        off = dict(zip(_LAYOUT_FIELDS, (int(v) for v in lay)))
        ws = bp.workspace.cpu().numpy()
        nv, nf, ne = bp.n_vertices, bp.n_faces, bp.n_edges
        node_t = np.dtype([("mn", "<f8", 3), ("mx", "<f8", 3), ("left", "<i4"), ("right", "<i4"), ("escape", "<i4"),
                           ("prim", "<i4")])
        F = ws[off["fnodes"]:off["fnodes"] + 64 * (2 * nf - 1)].view(node_t)
        E = ws[off["enodes"]:off["enodes"] + 64 * (2 * ne - 1)].view(node_t)
        VB = ws[off["vbox"]:off["vbox"] + 48 * nv].view("<f8").reshape(nv, 6)
        EB = ws[off["ebox"]:off["ebox"] + 48 * ne].view("<f8").reshape(ne, 6)
        QO = ws[off["qorder"]:off["qorder"] + 4 * nv].view("<u4")
        EO = ws[off["eorder"]:off["eorder"] + 4 * ne].view("<u4")
        C = ws[off["counts"]:off["counts"] + 4 * (nv + ne)].view("<u4")
        R = ws[off["resume"]:off["resume"] + 4 * (nv + ne)].view("<i4")

        def walk(T, start, q):
            n, node = 0, int(start)
            while node >= 0:
                t = T[node]
                n += 1
                hit = (q[:3] <= t["mx"]).all() and (t["mn"] <= q[3:]).all()
                node = int(t["left"]) if hit and t["left"] >= 0 else int(t["escape"])
            return n

        want = dict(vf_count_visits=0, vf_emit_visits=0, ee_count_visits=0, ee_emit_visits=0)
        for q in range(nv):
            v = QO[q]
            if s.topo.vertex_object[v] < 0:
                continue
            want["vf_count_visits"] += walk(F, 0, VB[v])
            if C[q] > slots:
                want["vf_emit_visits"] += walk(F, R[q] if slots else 0, VB[v])
        for e in range(ne):
            start = E[ne - 1 + e]["escape"]
            want["ee_count_visits"] += walk(E, start, EB[EO[e]])
            if C[nv + e] > slots:
                want["ee_emit_visits"] += walk(E, R[nv + e] if slots else start, EB[EO[e]])
        assert got == want, (slots, got, want)
        assert got["vf_count_visits"] > nv and got["vf_emit_visits"] <= got["vf_count_visits"]
