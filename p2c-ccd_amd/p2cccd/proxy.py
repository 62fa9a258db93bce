"""Proxy-scene candidates on the GPU: raw hits, compaction, density stats.

The reference path (src/cpp/rt_candidate/candidate_buffer.cpp:144-228,
candidate_stats.cpp:128-170, proxy_scene.cpp:18-154):
  raw hits   every pair i < j of proxy primitives with different objects, the
             same slab, strictly overlapping times and inclusively overlapping
             AABBs, in (i, j) order (p2c_proxy_raw_count / _fill);
  compaction one candidate per key (query, slab, first object, first patch,
             second object, second patch, first type, second type) with
             first = the smaller (object, patch, proxy id); rt_hit_count
             summed, motion bound maxed, flags OR'ed; candidate_id = 1, 2, ...
             in key order (std::map order);
  density    proxy / object / slab counts, cross-object same-slab pairs and
             the ratios of ComputeCandidateDensityStats.
Compaction is a segmented reduction after a stable lexicographic sort, done
with device tensors.  The scene is given as arrays (ProxyPrimitive fields,
proxy_id = index); validation raises ValueError with the reference's messages.
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


_vp, _i64 = nat._vp, nat._i64
nat._sig("p2c_proxy_raw_count", [_vp, _vp, _vp])
nat._sig("p2c_proxy_raw_fill", [_vp, _vp, _vp, _i64, _vp])


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


def raw_hits(query_id: int, arr: dict, stream=None) -> torch.Tensor:
    """Raw hits as a device uint8 tensor of n_raw * 64 bytes (RawCandidateHit layout)."""
    n = arr["t0"].numel()
    s = ProxySet(*[arr[k].data_ptr() for k in FIELDS], n, query_id)
    dev = arr["t0"].device
    counts = torch.zeros(max(n, 1), dtype=torch.int32, device=dev)
    nat.call("p2c_proxy_raw_count", ctypes.addressof(s), counts.data_ptr(), nat.stream_handle(stream))
    c64 = counts[:n].long()
    offsets = torch.cumsum(c64, 0) - c64
    total = int(c64.sum()) if n else 0
    out = torch.empty(max(total, 1) * 64, dtype=torch.uint8, device=dev)
    if total:
        nat.call("p2c_proxy_raw_fill", ctypes.addressof(s), offsets.data_ptr(), out.data_ptr(), total,
                 nat.stream_handle(stream))
    return out[: total * 64]


def compact(query_id: int, arr: dict, raw: torch.Tensor):
    """CompactRawCandidates on device tensors; returns a CANDIDATE_DTYPE array."""
    m = raw.numel() // 64
    if m == 0:
        return np.zeros(0, CANDIDATE_DTYPE)
    words = raw.view(torch.int32).view(m, 16)
    a = words[:, 4].long() & 0xFFFFFFFF
    b = words[:, 5].long() & 0xFFFFFFFF
    slab = words[:, 6].long() & 0xFFFFFFFF
    rt = words[:, 7].long() & 0xFFFFFFFF
    flags = words[:, 8].long() & 0xFFFFFFFF
    motion = raw.view(torch.float32).view(m, 16)[:, 10:14]
    obj = arr["object_id"].long() & 0xFFFFFFFF
    patch = arr["patch_id"].long() & 0xFFFFFFFF
    ptype = arr["proxy_type"].long() & 0xFFFFFFFF
    oa, ob, pa, pb = obj[a], obj[b], patch[a], patch[b]
    swap = (oa > ob) | ((oa == ob) & ((pa > pb) | ((pa == pb) & (a > b))))
    first = torch.where(swap, b, a)
    second = torch.where(swap, a, b)
    fields = [slab, obj[first], patch[first], obj[second], patch[second], ptype[first], ptype[second]]
    order = torch.arange(m, device=raw.device)
    for f in reversed(fields):  # stable LSD passes: lexicographic by the key
        order = order[torch.argsort(f[order], stable=True)]
    keys = torch.stack([f[order] for f in fields], 1)
    new = torch.ones(m, dtype=torch.bool, device=raw.device)
    new[1:] = (keys[1:] != keys[:-1]).any(1)
    seg = torch.cumsum(new.long(), 0) - 1
    C = int(seg[-1]) + 1
    rt_sum = torch.zeros(C, dtype=torch.int64, device=raw.device).index_add_(0, seg, rt[order])
    mb = torch.full((C, 4), float("-inf"), dtype=torch.float32, device=raw.device)
    mb.scatter_reduce_(0, seg[:, None].expand(-1, 4), motion[order], reduce="amax")
    fl = torch.zeros(C, dtype=torch.int64, device=raw.device)
    for bit in range(int(flags.max()).bit_length()):  # OR as a per-bit maximum
        hit = (flags[order] >> bit) & 1
        fl |= torch.zeros(C, dtype=torch.int64, device=raw.device).scatter_reduce_(0, seg, hit, reduce="amax") << bit
    k = keys[new].cpu().numpy()
    out = np.zeros(C, CANDIDATE_DTYPE)
    out["schema_version"] = 1
    out["candidate_id"] = np.arange(1, C + 1, dtype=np.uint64)
    out["query_id"] = query_id
    out["slab_id"], out["object_a_id"], out["patch_a_id"] = k[:, 0], k[:, 1], k[:, 2]
    out["object_b_id"], out["patch_b_id"] = k[:, 3], k[:, 4]
    out["proxy_type_a"], out["proxy_type_b"] = k[:, 5], k[:, 6]
    if int(rt_sum.max()) > 0xFFFFFFFF:
        raise ValueError("candidate rt_hit_count overflow during compaction")
    out["rt_hit_count"] = rt_sum.cpu().numpy()
    out["motion_bound"] = mb.cpu().numpy()
    out["flags"] = fl.cpu().numpy()
    return out


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
