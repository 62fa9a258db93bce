"""Proxy-scene candidates on the GPU: raw hits, compaction, density stats.

The reference path (src/cpp/rt_candidate/candidate_buffer.cpp:144-228,
candidate_stats.cpp:128-170, proxy_scene.cpp:18-154):
  raw hits   every pair i < j of proxy primitives with different objects, the
             same slab, strictly overlapping times and inclusively overlapping
             AABBs, in (i, j) order (p2c_proxy_raw_count, an exclusive scan,
             p2c_proxy_raw_fill);
  compaction one candidate per key (query, slab, first object, first patch,
             second object, second patch, first type, second type) with
             first = the smaller (object, patch, proxy id); rt_hit_count
             summed, motion bound maxed, flags OR'ed; candidate_id = 1, 2, ...
             in key order (std::map order) -- p2c_proxy_compact, a stable
             radix sort over the key fields and a merge per key on the device;
  density    proxy / object / slab counts, cross-object same-slab pairs and
             the ratios of ComputeCandidateDensityStats (host bookkeeping).
The scene is given as arrays (ProxyPrimitive fields, proxy_id = index);
validation raises ValueError with the reference's messages.  The same device
entry points serve p2cccd_cpp.generate_candidates_for_proxy_scene.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np
import torch

from . import _native as nat

RAW_HIT_DTYPE = np.dtype([("query_id", "<u8"), ("pair_key", "<u8"), ("proxy_a_index", "<u4"),
                          ("proxy_b_index", "<u4"), ("slab_id", "<u4"), ("rt_hit_count", "<u4"), ("flags", "<u4"),
                          ("reserved0", "<u4"), ("motion_bound", "<f4", (4,)), ("reserved1", "<u4"),
                          ("reserved2", "<u4")])
assert RAW_HIT_DTYPE.itemsize == 64
CANDIDATE_DTYPE = np.dtype([("schema_version", "<u4"), ("candidate_id", "<u8"), ("query_id", "<u8"),
                            ("slab_id", "<u4"), ("object_a_id", "<u4"), ("patch_a_id", "<u4"),
                            ("object_b_id", "<u4"), ("patch_b_id", "<u4"), ("proxy_type_a", "u1"),
                            ("proxy_type_b", "u1"), ("rt_hit_count", "<u4"), ("motion_bound", "<f4", (4,)),
                            ("proxy_features_offset", "<u4"), ("flags", "<u4")])
FIELDS = ("object_id", "patch_id", "slab_id", "proxy_type", "t0", "t1", "bounds", "motion")
BACKEND_NAME = "hip_exact"


class ProxySet(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in FIELDS] + [("n", ctypes.c_int64), ("query_id", ctypes.c_uint64)]


# p2c_candidate: the C++ CandidateRecord with its natural alignment (80 bytes)
CANDIDATE_C_DTYPE = np.dtype({
    "names": ["schema_version", "candidate_id", "query_id", "slab_id", "object_a_id", "patch_a_id", "object_b_id",
              "patch_b_id", "proxy_type_a", "proxy_type_b", "rt_hit_count", "motion_bound", "proxy_features_offset",
              "flags"],
    "formats": ["<u4", "<u8", "<u8", "<u4", "<u4", "<u4", "<u4", "<u4", "u1", "u1", "<u4", ("<f4", (4,)), "<u4", "<u4"],
    "offsets": [0, 8, 16, 24, 28, 32, 36, 40, 44, 45, 48, 52, 68, 72],
    "itemsize": 80})

_vp, _i64 = nat._vp, nat._i64
nat._sig("p2c_proxy_raw_count", [_vp, _vp, _vp, ctypes.c_size_t, _vp])
nat._sig("p2c_proxy_raw_fill", [_vp, _vp, _vp, ctypes.c_size_t, _vp, _i64, _vp])
nat._sig("p2c_exclusive_scan_u32", [_vp, _i64, _vp, _vp, _vp])
nat._sig("p2c_proxy_compact", [_vp, _vp, _i64, _vp, ctypes.c_size_t, _vp, _i64, _vp, _vp])
nat.lib.p2c_proxy_compact_workspace_bytes.argtypes = [_i64]
nat.lib.p2c_proxy_compact_workspace_bytes.restype = ctypes.c_size_t
nat.lib.p2c_proxy_raw_workspace_bytes.restype = ctypes.c_size_t
nat.lib.p2c_proxy_raw_workspace_bytes.argtypes = [_i64]
nat.lib.p2c_proxy_status_message.argtypes = [ctypes.c_uint64]
nat.lib.p2c_proxy_status_message.restype = ctypes.c_char_p


def _time_tol(t0, t1):
    return 1.0e-12 * max(1.0, abs(t0), abs(t1), abs(t1 - t0))


def validate_proxy_scene(query_id: int, scene: dict) -> None:
    """ValidateProxyScene (src/cpp/rt_candidate/proxy_scene.cpp:18-154)."""
    if query_id == 0:
        raise ValueError("ProxyScene.query_id is required")
    t0, t1 = np.asarray(scene["t0"], np.float64), np.asarray(scene["t1"], np.float64)
    ptype = np.asarray(scene["proxy_type"])
    b = np.asarray(scene["bounds"], np.float64).reshape(-1, 6)
    for i in range(len(t0)):
        if ptype[i] not in (1, 2):
            raise ValueError("proxy primitive has an unsupported proxy type")
        if not (np.isfinite(t0[i]) and np.isfinite(t1[i])) or t0[i] >= t1[i]:
            raise ValueError("proxy primitive must satisfy finite t0 < t1")
        if not np.isfinite(b[i]).all():
            raise ValueError("proxy bounds contain a non-finite value")
        if (b[i, :3] > b[i, 3:]).any():
            raise ValueError("proxy bounds min cannot exceed max")
    slabs: dict = {}
    for s, a, c in zip(np.asarray(scene["slab_id"]).tolist(), t0.tolist(), t1.tolist()):
        if s not in slabs:
            slabs[s] = (a, c)
            continue
        f0, f1 = slabs[s]
        tol = _time_tol(f0, f1)
        if abs(f0 - a) > tol or abs(f1 - c) > tol:
            raise ValueError("proxy scene slab_id must map to one global time interval")
    ordered = sorted(slabs.items(), key=lambda kv: (kv[1][0], kv[0]))
    for (_, (a0, a1)), (_, (b0, b1)) in zip(ordered, ordered[1:]):
        tol = max(_time_tol(a0, a1), _time_tol(b0, b1))
        if a0 < b1 - tol and b0 < a1 - tol:
            raise ValueError("proxy scene global slab intervals must not overlap")


def _device_arrays(scene: dict, device):
    out = {k: torch.from_numpy(np.ascontiguousarray(scene[k], dtype=np.uint32).view(np.int32)).to(device)
           for k in FIELDS[:4]}
    out.update({k: torch.from_numpy(np.ascontiguousarray(scene[k], dtype=np.float64)).to(device) for k in FIELDS[4:]})
    return out


def _set(query_id: int, arr: dict) -> ProxySet:
    return ProxySet(*[arr[k].data_ptr() for k in FIELDS], arr["t0"].numel(), query_id)


def raw_hits(query_id: int, arr: dict, stream=None) -> torch.Tensor:
    """Raw hits as a device uint8 tensor of n_raw * 64 bytes (RawCandidateHit layout)."""
    n = arr["t0"].numel()
    s = _set(query_id, arr)
    dev = arr["t0"].device
    counts = torch.zeros(max(n, 1), dtype=torch.int32, device=dev)
    offsets = torch.zeros(n + 1, dtype=torch.int64, device=dev)
    part = torch.empty(n // 1024 + 2, dtype=torch.int64, device=dev)
    ws_bytes = int(nat.lib.p2c_proxy_raw_workspace_bytes(n))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    h = nat.stream_handle(stream)
    if n:
        nat.call("p2c_proxy_raw_count", ctypes.addressof(s), counts.data_ptr(), ws.data_ptr(), ws_bytes, h)
        nat.call("p2c_exclusive_scan_u32", counts.data_ptr(), n, offsets.data_ptr(), part.data_ptr(), h)
    total = int(offsets[n]) if n else 0
    out = torch.empty(max(total, 1) * 64, dtype=torch.uint8, device=dev)
    if total:
        nat.call("p2c_proxy_raw_fill", ctypes.addressof(s), offsets.data_ptr(), ws.data_ptr(), ws_bytes,
                 out.data_ptr(), total, h)
    return out[: total * 64]


def compact(query_id: int, arr: dict, raw: torch.Tensor, stream=None) -> np.ndarray:
    """CompactRawCandidates on the device; returns a CANDIDATE_DTYPE array.
    Raises ValueError with the reference's message for an invalid raw hit or
    an rt_hit_count overflow."""
    m = raw.numel() // 64
    dev = arr["t0"].device
    s = _set(query_id, arr)
    ws_bytes = int(nat.lib.p2c_proxy_compact_workspace_bytes(m))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    out = torch.empty(max(m, 1) * CANDIDATE_C_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    status = torch.zeros(2, dtype=torch.int64, device=dev)
    raw_ptr = raw.data_ptr() if m else out.data_ptr()
    nat.call("p2c_proxy_compact", ctypes.addressof(s), raw_ptr, m, ws.data_ptr(), ws_bytes, out.data_ptr(), m,
             status.data_ptr(), nat.stream_handle(stream))
    count, err = (int(v) & 0xFFFFFFFFFFFFFFFF for v in status.cpu())
    if err:
        raise ValueError(nat.lib.p2c_proxy_status_message(err).decode())
    c = out[: count * CANDIDATE_C_DTYPE.itemsize].cpu().numpy().view(CANDIDATE_C_DTYPE)
    res = np.zeros(count, CANDIDATE_DTYPE)
    for name in CANDIDATE_DTYPE.names:
        res[name] = c[name]
    return res


def density(scene: dict, n_raw: int, cand: np.ndarray) -> dict:
    obj = np.asarray(scene["object_id"], np.int64)
    slab = np.asarray(scene["slab_id"], np.int64)
    n = len(obj)
    cross = 0
    for s in np.unique(slab):
        members = obj[slab == s]
        k = len(members)
        _, per = np.unique(members, return_counts=True)
        cross += k * (k - 1) // 2 - int((per * (per - 1) // 2).sum())
    nobj, nslab = len(np.unique(obj)), len(np.unique(slab))
    c = len(cand)
    ratio = lambda a, b: 0.0 if b == 0 else a / b  # noqa: E731
    return dict(query_id=None, proxy_count=n, object_count=nobj, slab_count=nslab,
                cross_object_same_slab_pair_count=cross, raw_hit_count=n_raw, compact_candidate_count=c,
                raw_hits_per_proxy=ratio(n_raw, n), candidates_per_proxy=ratio(c, n), candidates_per_slab=ratio(c, nslab),
                aabb_overlap_ratio=ratio(n_raw, cross),
                avg_rt_hits_per_candidate=ratio(int(cand["rt_hit_count"].sum()) if c else 0, c))


def generate_candidates_for_proxy_scene(scene: dict, query_id: int, device="cuda") -> dict:
    """The reference's GenerateCandidatesForProxyScene result for the
    `hip_exact` backend: raw_buffer (RAW_HIT_DTYPE), candidates
    (CANDIDATE_DTYPE), density and timing (ms)."""
    t_start = time.perf_counter()
    validate_proxy_scene(query_id, scene)
    dev = torch.device(device)
    arr = _device_arrays(scene, dev)
    torch.cuda.synchronize(dev)
    t_build = time.perf_counter()
    raw = raw_hits(query_id, arr)
    torch.cuda.synchronize(dev)
    t_trace = time.perf_counter()
    cand = compact(query_id, arr, raw)
    t_compact = time.perf_counter()
    raw_np = raw.cpu().numpy().view(RAW_HIT_DTYPE)
    dens = density(scene, len(raw_np), cand)
    dens["query_id"] = query_id
    t_end = time.perf_counter()
    ms = lambda a, b: (b - a) * 1e3  # noqa: E731
    timing = dict(build_ms=ms(t_start, t_build), update_ms=0.0, trace_ms=ms(t_build, t_trace),
                  compact_ms=ms(t_trace, t_compact), stats_ms=ms(t_compact, t_end), total_ms=ms(t_start, t_end))
    dens.update(timing=timing, backend_name=BACKEND_NAME)
    return dict(backend_name=BACKEND_NAME, raw_buffer=raw_np, candidates=cand, density=dens, timing=timing)
