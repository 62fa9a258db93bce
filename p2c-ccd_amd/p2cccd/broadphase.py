"""LBVH broad phase on the GPU: host driver over the p2c_bp_* C-ABI.

A `BroadPhase` holds one mesh's topology on the device (faces, edges, object
ids -- prepared once, as the reference's PrepareScene does) and a reusable
workspace; each step builds both BVHs from that step's positions and emits
the VF and EE candidate pairs.

For many scenes at once (serving), a `ResidentMesh` keeps a scene's topology
and its trees on the device, and a `BatchBroadPhase` builds, counts and emits
a whole batch of them with one launch per stage (p2c_bp_batch_count /
p2c_bp_batch_emit): every scene's pairs, in the same order, as `BroadPhase`
gives for that scene alone, in one pair list.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native as nat
from .scenes import Scene, Topology


class _Mesh(ctypes.Structure):
    _fields_ = [
        ("faces", ctypes.c_void_p),
        ("edges", ctypes.c_void_p),
        ("vertex_object", ctypes.c_void_p),
        ("face_object", ctypes.c_void_p),
        ("edge_object", ctypes.c_void_p),
        ("n_vertices", ctypes.c_int64),
        ("n_faces", ctypes.c_int64),
        ("n_edges", ctypes.c_int64),
        ("n_objects", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("hit_cache", ctypes.c_void_p),
        ("hit_cache_slots", ctypes.c_int64),
    ]


_vp, _i64, _i32 = nat._vp, nat._i64, nat._i32
nat.lib.p2c_bp_workspace_bytes.argtypes = [_i64, _i64, _i64, _i32]
nat.lib.p2c_bp_workspace_bytes.restype = ctypes.c_size_t
nat._sig("p2c_bp_build", [_vp, _vp, _vp, nat._f64, _vp, ctypes.c_size_t, _vp])
nat._sig("p2c_bp_count", [_vp, _vp, _vp, _vp])
nat._sig("p2c_bp_emit", [_vp, _vp, _vp, _i64, _vp, _i64, _vp])
nat._sig("p2c_bp_workspace_layout", [_i64, _i64, _i64, _i32, _vp])
_LAYOUT_FIELDS = ("fnodes", "enodes", "vbox", "ebox", "qorder", "eorder", "counts", "resume")


class _Scene(ctypes.Structure):
    _fields_ = [
        ("mesh", _Mesh),
        ("x0", ctypes.c_void_p),
        ("x1", ctypes.c_void_p),
        ("thickness", ctypes.c_double),
        ("workspace", ctypes.c_void_p),
        ("workspace_bytes", ctypes.c_size_t),
    ]


nat.lib.p2c_bp_scene_workspace_bytes.argtypes = [_i64, _i64, _i64, _i32]
nat.lib.p2c_bp_scene_workspace_bytes.restype = ctypes.c_size_t
nat.lib.p2c_bp_batch_workspace_bytes.argtypes = [_vp, _i32, _i64]
nat.lib.p2c_bp_batch_workspace_bytes.restype = ctypes.c_size_t
nat._sig("p2c_bp_batch_count", [_vp, _i32, _i64, _vp, ctypes.c_size_t, _vp, _vp])
nat._sig("p2c_bp_batch_emit", [_vp, _i32, _i64, _vp, ctypes.c_size_t, _vp, _i64, _vp])


DEFAULT_CACHE_SLOTS = 256  # partners per query kept by the count pass (covers nearly all queries of dense scenes)


class BroadPhase:
    def __init__(self, topo: Topology, n_vertices: int, device: str | torch.device = "cuda",
                 cache_slots: int = DEFAULT_CACHE_SLOTS):
        self.device = torch.device(device)
        dev = self.device

        def up(a):
            return torch.from_numpy(np.ascontiguousarray(a, np.int32)).to(dev)

        self.faces, self.edges = up(topo.faces), up(topo.edges)
        self.vobj, self.fobj, self.eobj = up(topo.vertex_object), up(topo.face_object), up(topo.edge_object)
        self.n_vertices, self.n_faces, self.n_edges = n_vertices, topo.faces.shape[0], topo.edges.shape[0]
        self.cache_slots = int(cache_slots)
        self.cache = torch.empty(max((self.n_vertices + self.n_edges) * self.cache_slots, 1), dtype=torch.int32,
                                 device=dev)
        self.mesh = _Mesh(self.faces.data_ptr(), self.edges.data_ptr(), self.vobj.data_ptr(), self.fobj.data_ptr(),
                          self.eobj.data_ptr(), self.n_vertices, self.n_faces, self.n_edges, topo.object_count, 0,
                          self.cache.data_ptr() if self.cache_slots else None, self.cache_slots)
        self.ws_bytes = int(nat.lib.p2c_bp_workspace_bytes(self.n_vertices, self.n_faces, self.n_edges,
                                                           topo.object_count))
        self.workspace = torch.empty(self.ws_bytes, dtype=torch.uint8, device=dev)
        self.totals = torch.zeros(3, dtype=torch.int64, device=dev)
        self.n_objects = topo.object_count
        self.vertex_object = np.ascontiguousarray(topo.vertex_object, np.int32)

    @staticmethod
    def workspace_bytes(n_vertices: int, n_faces: int, n_edges: int, n_objects: int,
                        cache_slots: int = DEFAULT_CACHE_SLOTS) -> int:
        """Workspace plus the hit cache."""
        return int(nat.lib.p2c_bp_workspace_bytes(n_vertices, n_faces, n_edges, n_objects)) + \
            4 * (n_vertices + n_edges) * cache_slots

    @classmethod
    def for_scene(cls, scene: Scene, device="cuda", cache_slots: int = DEFAULT_CACHE_SLOTS) -> "BroadPhase":
        return cls(scene.topo, scene.n_vertices, device, cache_slots)

    def build(self, x0: torch.Tensor, x1: torch.Tensor, thickness: float, stream=None) -> None:
        assert x0.dtype == torch.float64 and x0.is_cuda and x0.shape == (self.n_vertices, 3) and x0.is_contiguous()
        assert x1.dtype == torch.float64 and x1.is_cuda and x1.shape == (self.n_vertices, 3) and x1.is_contiguous()
        nat.call("p2c_bp_build", ctypes.addressof(self.mesh), x0.data_ptr(), x1.data_ptr(), float(thickness),
                 self.workspace.data_ptr(), self.ws_bytes, nat.stream_handle(stream))

    def count(self, stream=None) -> None:
        nat.call("p2c_bp_count", ctypes.addressof(self.mesh), self.workspace.data_ptr(), self.totals.data_ptr(),
                 nat.stream_handle(stream))

    def emit(self, vf: torch.Tensor, ee: torch.Tensor, stream=None) -> None:
        nat.call("p2c_bp_emit", ctypes.addressof(self.mesh), self.workspace.data_ptr(), vf.data_ptr(), vf.shape[0],
                 ee.data_ptr(), ee.shape[0], nat.stream_handle(stream))

    def candidates(self, x0: torch.Tensor, x1: torch.Tensor, thickness: float) -> tuple[torch.Tensor, torch.Tensor]:
        """Build, count, read the totals (one host sync), allocate and emit."""
        self.build(x0, x1, thickness)
        self.count()
        n_vf, n_ee, bad = (int(v) for v in self.totals.cpu())
        if bad:
            raise RuntimeError(f"{bad} BVH traversals exceeded their step bound")
        vf = torch.empty((n_vf, 2), dtype=torch.int32, device=self.device)
        ee = torch.empty((n_ee, 2), dtype=torch.int32, device=self.device)
        self.emit(vf, ee)
        return vf, ee

    # -- measurement ----------------------------------------------------------
    def traversal_visits(self) -> dict:
        """Nodes the last count and emit passes visited, replayed on the host
        (p2cccd_cpp.lbvh_traversal_visits) from the trees, query data, counts
        and resume nodes in the workspace (p2c_bp_workspace_layout)."""
        import p2cccd_cpp

        lay = np.zeros(len(_LAYOUT_FIELDS), np.uint64)
        nat.call("p2c_bp_workspace_layout", self.n_vertices, self.n_faces, self.n_edges, self.n_objects,
                 lay.ctypes.data)
        off = dict(zip(_LAYOUT_FIELDS, (int(v) for v in lay)))
        nv, nf, ne = self.n_vertices, self.n_faces, self.n_edges
        sizes = dict(fnodes=64 * max(2 * nf - 1, 0), enodes=64 * max(2 * ne - 1, 0), vbox=48 * nv, ebox=48 * ne,
                     qorder=4 * nv, eorder=4 * ne, counts=4 * (nv + ne), resume=4 * (nv + ne))
        torch.cuda.synchronize(self.device)
        host = {k: self.workspace[off[k]:off[k] + sizes[k]].cpu().numpy() for k in _LAYOUT_FIELDS}
        return p2cccd_cpp.lbvh_traversal_visits(host["fnodes"], nf, host["enodes"], ne, host["vbox"], host["ebox"],
                                                host["qorder"], host["eorder"], self.vertex_object, host["counts"],
                                                host["resume"], nv, self.cache_slots)



class ResidentMesh:
    """A scene's topology and broad-phase state (boxes, envelopes, object-pair
    matrix, both trees) kept on the device for batched builds."""

    TOPOLOGY = ("faces", "edges", "vertex_object", "face_object", "edge_object")

    def __init__(self, topo: Topology, n_vertices: int, device: str | torch.device = "cuda",
                 host: dict | None = None):
        """host: pinned copies of the topology (pinned_topology); the uploads
        are then asynchronous on the current stream."""
        self.device = torch.device(device)

        def up(name):
            if host is not None:
                return host[name].to(self.device, non_blocking=True)
            return torch.from_numpy(np.ascontiguousarray(getattr(topo, name), np.int32)).to(self.device)

        self.faces, self.edges = up("faces"), up("edges")
        self.vobj, self.fobj, self.eobj = up("vertex_object"), up("face_object"), up("edge_object")
        self.n_vertices, self.n_faces, self.n_edges = n_vertices, topo.faces.shape[0], topo.edges.shape[0]
        self.n_objects = topo.object_count
        self.ws_bytes = self.workspace_bytes(n_vertices, self.n_faces, self.n_edges, self.n_objects)
        self.workspace = torch.empty(max(self.ws_bytes, 1), dtype=torch.uint8, device=self.device)
        self.mesh = _Mesh(self.faces.data_ptr(), self.edges.data_ptr(), self.vobj.data_ptr(), self.fobj.data_ptr(),
                          self.eobj.data_ptr(), self.n_vertices, self.n_faces, self.n_edges, self.n_objects, 0,
                          None, 0)

    @classmethod
    def pinned_topology(cls, topo: Topology) -> dict:
        return {k: torch.from_numpy(np.ascontiguousarray(getattr(topo, k), np.int32)).pin_memory()
                for k in cls.TOPOLOGY}

    @staticmethod
    def workspace_bytes(n_vertices: int, n_faces: int, n_edges: int, n_objects: int) -> int:
        return int(nat.lib.p2c_bp_scene_workspace_bytes(n_vertices, n_faces, n_edges, n_objects))

    @classmethod
    def footprint(cls, topo: Topology, n_vertices: int) -> int:
        """Device bytes of a resident mesh: its workspace and its topology."""
        nf, ne = topo.faces.shape[0], topo.edges.shape[0]
        return cls.workspace_bytes(n_vertices, nf, ne, topo.object_count) + 4 * (3 * nf + 2 * ne + n_vertices + nf + ne)

    def nbytes(self) -> int:
        return self.workspace.numel() + 4 * sum(t.numel() for t in (self.faces, self.edges, self.vobj, self.fobj,
                                                                     self.eobj))


class BatchBroadPhase:
    """Broad phase of a batch of resident meshes, one launch per stage: count()
    builds every mesh's trees and counts its pairs, emit() writes the batch's
    pair list (scene after scene: its VF pairs, then its EE pairs, in
    scene-local ids).  The batch workspace (sort keys, per-query counts, the
    hit cache of `cache_slots` partners per query) grows to the largest batch
    seen; allocate and call on one stream."""

    def __init__(self, device: str | torch.device = "cuda", cache_slots: int = 16):
        self.device = torch.device(device)
        self.cache_slots = int(cache_slots)
        self.workspace = torch.empty(0, dtype=torch.uint8, device=self.device)
        self.offsets = torch.zeros(2, dtype=torch.int64, device=self.device)
        self.scenes = None
        self.n_scenes = 0
        self.slots = self.cache_slots

    def _describe(self, items) -> None:
        arr = (_Scene * len(items))()
        for k, (m, x0, x1, th) in enumerate(items):
            arr[k] = _Scene(m.mesh, x0, x1, float(th), m.workspace.data_ptr(), m.ws_bytes)
        self.scenes, self.n_scenes = arr, len(items)

    def count(self, items, stream=None, cache_slots: int | None = None) -> torch.Tensor:
        """items: (ResidentMesh, x0 address, x1 address, thickness) per scene,
        positions device (V, 3) float64; cache_slots overrides the default
        for this batch (and its emit).  Returns the device offsets
        (int64[2 S + 2], p2c_bp_batch_count)."""
        if not items:
            raise ValueError("a batch holds at least one scene")
        self._describe(items)
        self.slots = int(self.cache_slots if cache_slots is None else cache_slots)
        need = int(nat.lib.p2c_bp_batch_workspace_bytes(self.scenes, self.n_scenes, self.slots))
        if need == 0:
            raise ValueError("invalid batch")
        if self.workspace.numel() < need:
            self.workspace = torch.empty(int(need * 1.25) + (1 << 20), dtype=torch.uint8, device=self.device)
        if self.offsets.numel() < 2 * self.n_scenes + 2:
            self.offsets = torch.zeros(2 * self.n_scenes + 2, dtype=torch.int64, device=self.device)
        nat.call("p2c_bp_batch_count", self.scenes, self.n_scenes, self.slots, self.workspace.data_ptr(),
                 self.workspace.numel(), self.offsets.data_ptr(), nat.stream_handle(stream))
        return self.offsets[: 2 * self.n_scenes + 2]

    def emit(self, pairs: torch.Tensor, stream=None) -> None:
        """Fills `pairs` (int32 (cap, 2)) for the batch of the last count()."""
        assert pairs.dtype == torch.int32 and pairs.is_contiguous()
        nat.call("p2c_bp_batch_emit", self.scenes, self.n_scenes, self.slots, self.workspace.data_ptr(),
                 self.workspace.numel(), pairs.data_ptr(), pairs.shape[0], nat.stream_handle(stream))

    def candidates(self, items) -> tuple[np.ndarray, torch.Tensor]:
        """Count, read the offsets (one host sync), emit; returns the host
        offsets and the device pair list."""
        offs = self.count(items).cpu().numpy().astype(np.int64)
        s = self.n_scenes
        if offs[2 * s + 1]:
            raise RuntimeError(f"{offs[2 * s + 1]} BVH traversals exceeded their step bound")
        pairs = torch.empty((max(int(offs[2 * s]), 1), 2), dtype=torch.int32, device=self.device)
        self.emit(pairs)
        return offs, pairs[: int(offs[2 * s])]
