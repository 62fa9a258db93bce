"""Broad-phase oracle (oracle/broadphase.c) pinned against brute force with
the same set rules, hand-worked known answers for the adjacency and object
rules, and the reference's topology conventions
(src/tools/scene_object_envelope_native_ti_walltime.cpp:497-696)."""
import numpy as np
import pytest

import oracle
from p2cccd import scenes


def _small_scenes():
    yield scenes.two_icospheres(level=2, gap=0.01)
    yield scenes.two_icospheres(level=3, gap=-0.02, thickness=1e-3)
    yield scenes.cloth_on_sphere(n=16, sphere_level=1, cloth_size=1.0)
    yield scenes.cloth_on_sphere(n=24, sphere_level=2, cloth_size=1.5, thickness=1e-4)


@pytest.mark.parametrize("idx", range(4))
def test_sweep_equals_brute_force(idx):
    s = list(_small_scenes())[idx]
    a = oracle.candidates(s, "sweep")
    b = oracle.candidates(s, "brute")
    for k in ("vf", "ee"):
        assert len(a[k]) > 0
        assert np.array_equal(a[k], b[k]), k


def _two_triangle_scene(offset, thickness=0.0):
    # object 0: triangles (0,1,2), (1,3,2) sharing edge (1,2); object 1: one
    # triangle (4,5,6) hovering `offset` above vertex 0's corner
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0],
                  [0.1, 0.1, offset], [0.3, 0.1, offset], [0.1, 0.3, offset]], float)
    faces = np.array([[0, 1, 2], [1, 3, 2], [4, 5, 6]], np.int32)
    topo = scenes.prepare_topology(faces, 7)
    return scenes.Scene("kat", x, x.copy(), topo, thickness)


def test_known_answers_adjacency_and_objects():
    s = _two_triangle_scene(0.5)
    t = s.topo
    assert t.object_count == 2 and list(t.face_object) == [0, 0, 1] and list(t.vertex_object) == [0] * 4 + [1] * 3
    # edges in first-appearance order over (face, corner)
    assert t.edges.tolist() == [[0, 1], [1, 2], [0, 2], [1, 3], [2, 3], [4, 5], [5, 6], [4, 6]]
    c = oracle.candidates(s)
    # The objects' envelopes do not overlap (z = 0 vs 0.5), so only self pairs
    # remain.  Both faces of object 0 have the box [0,1]^2 x {0}: vertex 3
    # lies in face 0's box and vertex 0 in face 1's, and neither face holds
    # that vertex; vertices 1 and 2 belong to both faces.  Object 1's vertices
    # all belong to its single face.
    assert c["vf"].tolist() == [[0, 1], [3, 0]]
    # EE: the only non-adjacent edge pairs of object 0 are (0,1)-(2,3) (y = 0
    # vs y = 1) and (0,2)-(1,3) (x = 0 vs x = 1): disjoint boxes.  Object 1's
    # three edges are pairwise adjacent.
    assert c["ee"].tolist() == []
    # lower object 1 into the plane: cross-object pairs appear
    s2 = _two_triangle_scene(0.0)
    c2 = oracle.candidates(s2)
    assert [0, 1] in c2["vf"].tolist() and any(p[0] >= 4 for p in c2["vf"].tolist())
    assert len(c2["ee"]) > 0


def test_topology_sizes():
    v, f = scenes.icosphere(6)
    t = scenes.prepare_topology(f, len(v))
    assert (len(v), len(f), len(t.edges)) == (40962, 81920, 122880)
    v, f = scenes.grid_cloth(256)
    t = scenes.prepare_topology(f, len(v))
    assert (len(v), len(f), len(t.edges)) == (65536, 130050, 195585)
