"""The exact work queue on the device: ProcessExactWorkQueueCpu and
GenerateConservativeRefinementWorkItems (src/cpp/certificate/
certificate_engine.cpp:593-720) over whole queues held in HBM.

A queue is a structured array (or device tensor of its bytes) of
QUEUE_ENTRY_DTYPE records: an ExactWorkItem, its query's
CertificateEngineConfig and its primitive ranges into a point-triangle and an
edge-edge PRIM array.  One pass (p2c_exact_queue_run) returns every entry's
certificate, its audit rows and, with a refinement config, its children as
entries with the parent's config and ranges, ready to be the next queue.
Children of the whole queue are numbered in queue order: entry i's first
child id is first_child_work_item_id plus the children of the entries before
it, what calling the reference's per-parent function with a running first id
gives.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _native
from ._native import PRIM_DTYPE, _i64, _vp, call, ptr

QUEUE_ENTRY_DTYPE = np.dtype([
    ("work_item_id", "<u8"), ("parent_candidate_id", "<u8"), ("query_id", "<u8"), ("slab_id", "<u4"),
    ("patch_a_id", "<u4"), ("patch_b_id", "<u4"), ("feature_family_mask", "<u4"), ("interval_t0", "<f8"),
    ("interval_t1", "<f8"), ("topk_feature_ids_offset", "<u4"), ("depth", "<u2"), ("source", "u1"),
    ("pad0", "u1"), ("priority_score", "<f4"), ("max_subdivision_depth", "<u4"), ("eps_time", "<f8"),
    ("eps_space", "<f8"), ("pt_begin", "<i8"), ("pt_count", "<i8"), ("ee_begin", "<i8"), ("ee_count", "<i8"),
])
CERTIFICATE_DTYPE = np.dtype([
    ("work_item_id", "<u8"), ("query_id", "<u8"), ("interval_t0", "<f8"), ("interval_t1", "<f8"),
    ("toi_upper", "<f8"), ("safe_margin_lb", "<f8"), ("witness_id_a", "<i8"), ("witness_id_b", "<i8"),
    ("eps_time", "<f8"), ("eps_space", "<f8"), ("covered_feature_mask", "<u4"), ("reason_code", "<u2"),
    ("status", "u1"), ("witness_family", "u1"), ("next_refinement_mode", "u1"), ("pad", "u1", (7,)),
])
AUDIT_ROW_DTYPE = np.dtype([
    ("event_id", "<u8"), ("query_id", "<u8"), ("candidate_id", "<u8"), ("work_item_id", "<u8"),
    ("interval_t0", "<f8"), ("interval_t1", "<f8"), ("timestamp_us", "<u8"), ("aux_value0", "<f8"),
    ("aux_value1", "<f8"), ("action", "<u2"), ("depth", "<u2"), ("stage", "u1"), ("pad", "u1", (3,)),
])
QUEUE_CONFIG_DTYPE = np.dtype([("first_event_id", "<u8"), ("first_timestamp_us", "<u8"),
                               ("emit_dequeue_events", "<u4"), ("walk_flags", "<u4")])
REFINE_CONFIG_DTYPE = np.dtype([("first_child_work_item_id", "<u8"), ("min_interval_width", "<f8"),
                                ("max_child_depth", "<u4"), ("pad", "<u4")])
assert QUEUE_ENTRY_DTYPE.itemsize == 120 and CERTIFICATE_DTYPE.itemsize == 96 and AUDIT_ROW_DTYPE.itemsize == 80

STATUS_WORDS = 6

_size = ctypes.c_size_t
_native._sig("p2c_exact_queue_run", [_vp, _i64, _vp, _i64, _vp, _i64, _i64, _i64, _vp, _vp, _vp, _size, _vp, _vp, _vp,
                                     _i64, _vp, _vp])
_native._sig("p2c_exact_queue_error", [_vp])
_native._sig("p2c_exact_queue_host", [_vp, _i64, _vp, _i64, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _i64, _vp])
_native.lib.p2c_exact_queue_workspace_bytes.argtypes = [_i64, _i64, _i64]
_native.lib.p2c_exact_queue_workspace_bytes.restype = _size


@dataclass(frozen=True)
class QueueConfig:
    """ExactWorkQueueConfig (certificate_engine.h:83-87) and the walk flags
    (0: the CPU engine's walk)."""
    first_event_id: int = 1
    first_timestamp_us: int = 1
    emit_dequeue_events: bool = True
    walk_flags: int = 0

    def record(self) -> np.ndarray:
        r = np.zeros(1, QUEUE_CONFIG_DTYPE)
        r["first_event_id"], r["first_timestamp_us"] = self.first_event_id, self.first_timestamp_us
        r["emit_dequeue_events"], r["walk_flags"] = int(self.emit_dequeue_events), self.walk_flags
        return r


@dataclass(frozen=True)
class RefineConfig:
    """ExactRefinementConfig (certificate_engine.h:95-99)."""
    first_child_work_item_id: int = 1
    max_child_depth: int = 64
    min_interval_width: float = 1.0e-4

    def record(self) -> np.ndarray:
        r = np.zeros(1, REFINE_CONFIG_DTYPE)
        r["first_child_work_item_id"], r["min_interval_width"] = self.first_child_work_item_id, self.min_interval_width
        r["max_child_depth"] = self.max_child_depth
        return r


def _pair_bounds(entries: np.ndarray) -> tuple[int, int]:
    mask = entries["feature_family_mask"]
    pt = int(np.where((mask & 1) != 0, np.maximum(entries["pt_count"], 0), 0).sum())
    ee = int(np.where((mask & 2) != 0, np.maximum(entries["ee_count"], 0), 0).sum())
    return pt, ee


def process_queue_host(entries: np.ndarray, pts: np.ndarray, ees: np.ndarray, config: QueueConfig = QueueConfig(),
                       refine: RefineConfig | None = None):
    """Host arrays in and out (p2c_exact_queue_host).  Returns (certificates,
    audit rows, children); children is None without a refinement config."""
    entries = np.ascontiguousarray(entries, QUEUE_ENTRY_DTYPE)
    pts = np.ascontiguousarray(pts, PRIM_DTYPE)
    ees = np.ascontiguousarray(ees, PRIM_DTYPE)
    n = entries.shape[0]
    certs = np.zeros(n, CERTIFICATE_DTYPE)
    rows = np.zeros(n * (2 if config.emit_dequeue_events else 1), AUDIT_ROW_DTYPE)
    cap = 2 * n if refine is not None else 0
    children = np.zeros(cap, QUEUE_ENTRY_DTYPE)
    got = np.zeros(1, np.int64)
    rc = refine.record() if refine is not None else None
    call("p2c_exact_queue_host", ptr(entries), n, ptr(pts), pts.shape[0], ptr(ees), ees.shape[0],
         ptr(config.record()), ptr(rc), ptr(certs), ptr(rows), ptr(children), cap, ptr(got))
    return certs, rows, (children[: int(got[0])] if refine is not None else None)


class DeviceQueue:
    """A queue's primitives resident on the device, certified pass after
    pass: each pass takes a queue of entries (a uint8 tensor of
    QUEUE_ENTRY_DTYPE records on the device) and returns device tensors of
    its certificates, rows and children, reading back only the status words.
    Workspace is kept between passes and grown when a pass needs more."""

    def __init__(self, pts: np.ndarray, ees: np.ndarray, device="cuda"):
        import torch

        self.torch = torch
        self.device = torch.device(device)
        self.pts = self._upload(np.ascontiguousarray(pts, PRIM_DTYPE))
        self.ees = self._upload(np.ascontiguousarray(ees, PRIM_DTYPE))
        self.n_pt, self.n_ee = int(len(pts)), int(len(ees))
        self.status = torch.zeros(STATUS_WORDS, dtype=torch.int64, device=self.device)
        self.workspace = torch.empty(0, dtype=torch.uint8, device=self.device)

    def _upload(self, a: np.ndarray):
        t = self.torch.from_numpy(a.view(np.uint8).reshape(-1).copy())
        return t.to(self.device) if t.numel() else self.torch.empty(0, dtype=self.torch.uint8, device=self.device)

    def upload_entries(self, entries: np.ndarray):
        return self._upload(np.ascontiguousarray(entries, QUEUE_ENTRY_DTYPE))

    def run(self, entries, pt_pairs: int, ee_pairs: int, config: QueueConfig = QueueConfig(),
            refine: RefineConfig | None = None):
        """One pass over `entries` (device tensor from upload_entries or a
        previous pass's children).  pt_pairs / ee_pairs bound the pairs the
        entries certify (their PT and EE counts).  Returns (certificates,
        rows, children, child count) as device uint8 tensors; child_pairs
        then holds the children's pair counts, the next pass's bounds."""
        torch = self.torch
        n = entries.numel() // QUEUE_ENTRY_DTYPE.itemsize
        need = int(_native.lib.p2c_exact_queue_workspace_bytes(n, pt_pairs, ee_pairs))
        if self.workspace.numel() < need:
            self.workspace = torch.empty(need, dtype=torch.uint8, device=self.device)
        certs = torch.empty(n * CERTIFICATE_DTYPE.itemsize, dtype=torch.uint8, device=self.device)
        rows = torch.empty(n * (2 if config.emit_dequeue_events else 1) * AUDIT_ROW_DTYPE.itemsize, dtype=torch.uint8,
                           device=self.device)
        cap = 2 * n if refine is not None else 0
        children = torch.empty(max(cap, 1) * QUEUE_ENTRY_DTYPE.itemsize, dtype=torch.uint8, device=self.device)
        rc = refine.record() if refine is not None else None
        call("p2c_exact_queue_run", ptr(entries), n, ptr(self.pts), self.n_pt, ptr(self.ees), self.n_ee, pt_pairs,
             ee_pairs, ptr(config.record()), ptr(rc), ptr(self.workspace), self.workspace.numel(), ptr(certs),
             ptr(rows), ptr(children), cap, ptr(self.status), _native.stream_handle())
        st = self.status.cpu().numpy().view(np.uint64)
        call("p2c_exact_queue_error", ptr(st))
        got = int(st[1])
        self.child_pairs = (int(st[4]), int(st[5]))
        return certs, rows, children[: got * QUEUE_ENTRY_DTYPE.itemsize], got

    def refine_until_settled(self, entries: np.ndarray, config: QueueConfig = QueueConfig(),
                             refine: RefineConfig = RefineConfig(), max_passes: int = 64):
        """Certifies the queue, then its children, then theirs, until a pass
        makes none (or max_passes).  Event ids, timestamps and child ids run
        on across passes.  Returns host arrays of every pass's certificates,
        rows and entries, concatenated in pass order."""
        cur = self.upload_entries(entries)
        host_entries = np.ascontiguousarray(entries, QUEUE_ENTRY_DTYPE)
        pt_pairs, ee_pairs = _pair_bounds(host_entries)
        out_c, out_r, out_e = [], [], []
        first_event, first_ts, first_child = config.first_event_id, config.first_timestamp_us, \
            refine.first_child_work_item_id
        for _ in range(max_passes):
            n = cur.numel() // QUEUE_ENTRY_DTYPE.itemsize
            if n == 0:
                break
            cfg = QueueConfig(first_event, first_ts, config.emit_dequeue_events, config.walk_flags)
            rcfg = RefineConfig(first_child, refine.max_child_depth, refine.min_interval_width)
            certs, rows, children, got = self.run(cur, pt_pairs, ee_pairs, cfg, rcfg)
            per = 2 if config.emit_dequeue_events else 1
            first_event += per * n
            first_ts += per * n
            first_child += got
            out_e.append(cur.cpu().numpy().view(QUEUE_ENTRY_DTYPE))
            out_c.append(certs.cpu().numpy().view(CERTIFICATE_DTYPE))
            out_r.append(rows.cpu().numpy().view(AUDIT_ROW_DTYPE))
            pt_pairs, ee_pairs = self.child_pairs
            cur = children
        cat = np.concatenate
        return (cat(out_c) if out_c else np.zeros(0, CERTIFICATE_DTYPE),
                cat(out_r) if out_r else np.zeros(0, AUDIT_ROW_DTYPE),
                cat(out_e) if out_e else np.zeros(0, QUEUE_ENTRY_DTYPE))
