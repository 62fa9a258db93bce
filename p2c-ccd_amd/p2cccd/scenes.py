"""Synthetic scenes for the broad phase and the full pipeline, plus the
reference's per-mesh topology preparation.

Topology follows PrepareScene (src/tools/scene_object_envelope_native_ti_walltime
.cpp:497-586): objects are connected components of the face graph, numbered
in face order of first appearance; vertices take their faces' object (-1 when
no face uses them); edges are the unique (min, max) vertex pairs in order of
first appearance over faces f, corners k = 0, 1, 2 (edge (v[k], v[k+1])),
each tagged with its face's object.  Self-object pairs are allowed, as the
reference's default (`exclude_self_object_pairs = false`, :48).

Scene sizes:
  two_icospheres   level-6 spheres (40,962 V, 81,920 F each), one static, one
                   translated 0.05 toward it with a rotation <= 10 degrees
  cloth_on_sphere  256 x 256 cloth (65,536 V, 130,050 F) folded over itself,
                   N(0, 1e-3) noise, per-vertex velocities N(0, 0.02), over a
                   noisy level-5 sphere (10,242 V, 20,480 F)
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

from .workloads import BASE_SEED


@dataclass
class Topology:
    faces: np.ndarray          # (F, 3) int32
    edges: np.ndarray          # (E, 2) int32, (min, max), reference order
    vertex_object: np.ndarray  # (V,) int32, -1 for unused vertices
    face_object: np.ndarray    # (F,) int32
    edge_object: np.ndarray    # (E,) int32
    object_count: int


@dataclass
class Scene:
    name: str
    x0: np.ndarray             # (V, 3) float64 positions at t = 0
    x1: np.ndarray             # (V, 3) float64 positions at t = 1
    topo: Topology
    thickness: float = 0.0
    info: dict = field(default_factory=dict)

    @property
    def n_vertices(self) -> int:
        return self.x0.shape[0]


def prepare_topology(faces: np.ndarray, n_vertices: int) -> Topology:
    faces = np.ascontiguousarray(faces, dtype=np.int32)
    nf = faces.shape[0]
    # connected components over the face graph (edges v0-v1, v1-v2)
    rows = np.concatenate([faces[:, 0], faces[:, 1]])
    cols = np.concatenate([faces[:, 1], faces[:, 2]])
    g = coo_matrix((np.ones(rows.size, np.int8), (rows, cols)), shape=(n_vertices, n_vertices))
    _, label = connected_components(g, directed=False)
    face_label = label[faces[:, 0]]
    uniq, first = np.unique(face_label, return_index=True)
    order = np.argsort(first, kind="stable")
    obj_of_label = np.full(label.max() + 1, -1, np.int64)
    obj_of_label[uniq[order]] = np.arange(uniq.size)
    face_object = obj_of_label[face_label].astype(np.int32)
    vertex_object = np.full(n_vertices, -1, np.int32)
    vertex_object[faces.reshape(-1)] = np.repeat(face_object, 3)
    # edges in order of first appearance over (face, corner)
    a = faces.reshape(-1)
    b = faces[:, [1, 2, 0]].reshape(-1)
    lo, hi = np.minimum(a, b).astype(np.int64), np.maximum(a, b).astype(np.int64)
    key = lo * n_vertices + hi
    _, first_pos = np.unique(key, return_index=True)
    first_pos = np.sort(first_pos)
    edges = np.stack([lo[first_pos], hi[first_pos]], axis=1).astype(np.int32)
    edge_object = face_object[first_pos // 3]
    return Topology(faces=faces, edges=np.ascontiguousarray(edges), vertex_object=vertex_object,
                    face_object=face_object, edge_object=np.ascontiguousarray(edge_object, np.int32),
                    object_count=int(uniq.size))


def icosphere(level: int, radius: float = 1.0) -> tuple[np.ndarray, np.ndarray]:
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    verts = np.array(v, np.float64)
    verts /= np.linalg.norm(verts, axis=1, keepdims=True)
    faces = np.array(f, np.int64)
    for _ in range(level):
        nv = verts.shape[0]
        e = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
        e = np.sort(e, axis=1)
        key = e[:, 0] * (nv + 4 * faces.shape[0]) + e[:, 1]
        uk, inv = np.unique(key, return_inverse=True)
        ue = np.stack([uk // (nv + 4 * faces.shape[0]), uk % (nv + 4 * faces.shape[0])], axis=1)
        mids = verts[ue[:, 0]] + verts[ue[:, 1]]
        mids /= np.linalg.norm(mids, axis=1, keepdims=True)
        m = nv + inv.reshape(3, -1)
        a, b, c = faces[:, 0], faces[:, 1], faces[:, 2]
        mab, mbc, mca = m[0], m[1], m[2]
        faces = np.concatenate([np.stack([a, mab, mca], 1), np.stack([b, mbc, mab], 1), np.stack([c, mca, mbc], 1),
                                np.stack([mab, mbc, mca], 1)])
        verts = np.concatenate([verts, mids])
    return verts * radius, faces.astype(np.int32)


def _rotation(axis: np.ndarray, angle: float) -> np.ndarray:
    axis = axis / np.linalg.norm(axis)
    k = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + np.sin(angle) * k + (1 - np.cos(angle)) * (k @ k)


def two_icospheres(level: int = 6, gap: float = 0.02, step: float = 0.05, max_rot_deg: float = 10.0,
                   thickness: float = 1.0e-6, seed: int = BASE_SEED) -> Scene:
    """Sphere A static at the origin; sphere B starts `gap` away (surface to
    surface) on +x and translates `step` toward A while rotating about its
    centre by a random angle <= max_rot_deg; vertices move linearly."""
    rng = np.random.Generator(np.random.PCG64(seed + 2))
    v, f = icosphere(level)
    nv = v.shape[0]
    c0 = np.array([2.0 + gap, 0.0, 0.0])
    c1 = c0 - np.array([step, 0.0, 0.0])
    rot = _rotation(rng.normal(size=3), np.deg2rad(rng.uniform(0, max_rot_deg)))
    x0 = np.concatenate([v, v + c0])
    x1 = np.concatenate([v, v @ rot.T + c1])
    faces = np.concatenate([f, f + nv]).astype(np.int32)
    topo = prepare_topology(faces, 2 * nv)
    return Scene("two_icospheres_l%d" % level, x0, x1, topo, thickness,
                 dict(level=level, gap=gap, step=step, seed=seed))


def grid_cloth(n: int = 256, size: float = 2.0) -> tuple[np.ndarray, np.ndarray]:
    """n x n vertex grid in the z = 0 plane, two triangles per cell."""
    s = np.linspace(-size / 2, size / 2, n)
    X, Y = np.meshgrid(s, s, indexing="ij")
    v = np.stack([X.ravel(), Y.ravel(), np.zeros(n * n)], axis=1)
    i, j = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    a = (i * n + j).ravel()
    b, c, d = a + n, a + 1, a + n + 1
    faces = np.concatenate([np.stack([a, b, c], 1), np.stack([c, b, d], 1)]).astype(np.int32)
    # interleave the two triangles of each cell
    faces = faces.reshape(2, -1, 3).transpose(1, 0, 2).reshape(-1, 3)
    return v, faces


def cloth_on_sphere(n: int = 256, sphere_level: int = 5, cloth_size: float = 4.0, crease_radius: float = 0.004,
                    noise: float = 1.0e-3, velocity_sigma: float = 0.02, fall: float = 0.02,
                    thickness: float = 0.0, seed: int = BASE_SEED) -> Scene:
    """A cloth folded over itself around a crease of radius `crease_radius`
    (the two layers 2 r apart), lying on top of a noisy unit sphere, moving
    down by `fall` plus N(0, velocity_sigma) per vertex over the step."""
    rng = np.random.Generator(np.random.PCG64(seed + 3))
    v, f = grid_cloth(n, cloth_size)
    # fold: arc length along x; first half flat, crease semicircle, second half back on top
    half = cloth_size / 2
    sx = v[:, 0] + half  # 0 .. size
    L = cloth_size
    flat = (L - np.pi * crease_radius) / 2
    xf = np.empty_like(sx)
    zf = np.zeros_like(sx)
    m1 = sx <= flat
    xf[m1], zf[m1] = sx[m1], 0.0
    m2 = (sx > flat) & (sx < flat + np.pi * crease_radius)
    th = (sx[m2] - flat) / crease_radius
    xf[m2] = flat + crease_radius * np.sin(th)
    zf[m2] = crease_radius * (1 - np.cos(th))
    m3 = sx >= flat + np.pi * crease_radius
    xf[m3] = flat - (sx[m3] - flat - np.pi * crease_radius)
    zf[m3] = 2 * crease_radius
    cloth = np.stack([xf - flat / 2, v[:, 1], zf], axis=1)
    cloth += rng.normal(0.0, noise, cloth.shape)
    sv, sf = icosphere(sphere_level)
    sv = sv * (1.0 + rng.normal(0.0, noise, (sv.shape[0], 1)))
    top = cloth[:, 2].min()
    cloth[:, 2] += 1.0 + noise - top  # cloth rests just above the sphere's top
    x0 = np.concatenate([cloth, sv])
    vel = rng.normal(0.0, velocity_sigma, cloth.shape)
    vel[:, 2] -= fall
    x1 = np.concatenate([cloth + vel, sv])
    faces = np.concatenate([f, sf + cloth.shape[0]]).astype(np.int32)
    topo = prepare_topology(faces, x0.shape[0])
    return Scene("cloth%d_on_sphere_l%d" % (n, sphere_level), x0, x1, topo, thickness,
                 dict(n=n, sphere_level=sphere_level, crease_radius=crease_radius, noise=noise,
                      velocity_sigma=velocity_sigma, fall=fall, seed=seed))
