"""Binary PLY frames -> broad-phase scenes, the reference scene tool's input
(src/tools/scene_object_envelope_native_ti_walltime.cpp:396-586).

Two frames of one mesh (t = 0 and t = 1) are read natively
(p2cccd_cpp.load_binary_ply: x, y, z of each vertex, polygons fanned into
triangles, triangles with a repeated corner dropped); they must agree in
vertex and triangle counts.  Objects, per-vertex objects and the edge list
follow PrepareScene (scenes.prepare_topology)."""
from __future__ import annotations

from pathlib import Path

import numpy as np

from ..scenes import Scene, prepare_topology


def load_binary_ply(path) -> tuple[np.ndarray, np.ndarray]:
    """(vertices float64[V, 3], triangles int32[F, 3])."""
    import p2cccd_cpp

    return p2cccd_cpp.load_binary_ply(str(path))


def scene_from_ply_frames(frame0, frame1, *, thickness: float = 0.0, name: str | None = None) -> Scene:
    x0, f0 = load_binary_ply(frame0)
    x1, f1 = load_binary_ply(frame1)
    if x0.shape[0] != x1.shape[0]:
        raise ValueError("frame vertex counts differ")
    if f0.shape[0] != f1.shape[0]:
        raise ValueError("frame face counts differ")
    if f0.size and (f0.min() < 0 or f0.max() >= x0.shape[0]):
        raise ValueError("PLY face index outside the vertex list")
    return Scene(name=name or Path(frame0).stem, x0=np.ascontiguousarray(x0), x1=np.ascontiguousarray(x1),
                 topo=prepare_topology(f0, x0.shape[0]), thickness=float(thickness),
                 info=dict(frame0=str(frame0), frame1=str(frame1)))
