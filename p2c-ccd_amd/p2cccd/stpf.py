"""Scheduling scorers over the C-ABI: TinySTPF candidate scores (FP64) and the
STPF priority trunk (FP32, 32-128-128-1).

TinySTPF follows src/tools/scene_object_envelope_native_ti_walltime.cpp
(weights file format and checks :995-1066, features :893-993, predict
:1068-1091, policies :1102-1131, group key :756-771).  The STPF trunk is the
feature_dim 32, hidden 128, two-layer GELU trunk with the sigmoid priority
head of src/python/p2cccd/proposal/stpf_model.py:45-93.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import torch

from . import _native as nat

TINY_STPF_WEIGHTS = 4289
STPF_WEIGHTS = 20865
POLICIES = {"learned": 0, "learned_residual": 1, "proximity": 2, "motion": 3, "random": 4}

_vp, _i64, _i32 = nat._vp, nat._i64, nat._i32
nat._sig("p2c_tiny_stpf_score", [_i32, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _vp])
nat._sig("p2c_tiny_stpf_features", [_i32, _vp, _vp, _i64, _vp, _vp])
nat._sig("p2c_tiny_stpf_predict", [_vp, _vp, _i64, _vp, _vp])
nat._sig("p2c_stpf_priority", [_vp, _vp, _i64, _vp, _vp])


def load_tiny_stpf_weights(path: str | Path) -> np.ndarray:
    """Reads a TinySTPF weights text file into the packed layout of
    p2c_tiny_stpf_score (mean, std, w0, b0, w1, b1, w2, b2).  Raises
    RuntimeError with the reference loader's messages."""
    path = Path(path)
    try:
        tokens = path.read_text().split()
    except OSError:
        raise RuntimeError(f"failed to open STPF weights {path}") from None
    pos = 0

    def take():
        nonlocal pos
        tok = tokens[pos] if pos < len(tokens) else ""
        pos += 1
        return tok

    def expect(tag):
        got = take()
        if got != tag:
            raise RuntimeError(f"bad STPF weight tag: expected {tag}, got {got}")

    expect("version")
    if int(take()) != 1:
        raise RuntimeError("unsupported STPF weight version")
    dims = []
    for tag in ("input_dim", "hidden0", "hidden1"):
        expect(tag)
        dims.append(int(take()))
    if dims != [32, 64, 32]:
        raise RuntimeError("unexpected TinySTPF dimensions")
    out = []
    for tag, count in (("mean", 32), ("std", 32), ("w0", 64 * 32), ("b0", 64), ("w1", 32 * 64), ("b1", 32),
                       ("w2", 32), ("b2", 1)):
        expect(tag)
        vals = tokens[pos:pos + count]
        if len(vals) != count:
            raise RuntimeError(f"truncated STPF weights in {tag}")
        out.extend(float(v) for v in vals)
        pos += count
    w = np.asarray(out, dtype=np.float64)
    assert w.size == TINY_STPF_WEIGHTS
    return w


def tiny_scores(kind: str, x0: torch.Tensor, x1: torch.Tensor, topo: torch.Tensor, pairs: torch.Tensor, n: int,
                object_a: torch.Tensor, object_b: torch.Tensor, policy: str = "learned",
                weights: torch.Tensor | None = None, totals: torch.Tensor | None = None, stream=None):
    """Scores and group keys of n VF (kind "vf") or EE ("ee") candidates,
    indexed 0..n-1 in list order: (scores float64[n], keys uint64-as-int64[n])."""
    if policy not in POLICIES:
        raise RuntimeError(f"unknown score policy: {policy}")
    dev = x0.device
    scores = torch.empty(max(n, 1), dtype=torch.float64, device=dev)
    keys = torch.empty(max(n, 1), dtype=torch.int64, device=dev)
    if weights is not None:
        assert weights.dtype == torch.float64 and weights.numel() == TINY_STPF_WEIGHTS and weights.is_cuda
    nat.call("p2c_tiny_stpf_score", 1 if kind == "vf" else 2, x0.data_ptr(), x1.data_ptr(), topo.data_ptr(),
             pairs.data_ptr(), n, 0 if totals is None else totals.data_ptr(), object_a.data_ptr(),
             object_b.data_ptr(), POLICIES[policy], 0 if weights is None else weights.data_ptr(), scores.data_ptr(),
             keys.data_ptr(), nat.stream_handle(stream))
    return scores[:n], keys[:n]


def tiny_features(kind: str, rows: torch.Tensor, index: torch.Tensor, stream=None) -> torch.Tensor:
    n = rows.shape[0]
    out = torch.empty((max(n, 1), 32), dtype=torch.float64, device=rows.device)
    nat.call("p2c_tiny_stpf_features", 1 if kind == "vf" else 2, rows.data_ptr(), index.data_ptr(), n,
             out.data_ptr(), nat.stream_handle(stream))
    return out[:n]


def tiny_predict(weights: torch.Tensor, feats: torch.Tensor, stream=None) -> torch.Tensor:
    n = feats.shape[0]
    out = torch.empty(max(n, 1), dtype=torch.float64, device=feats.device)
    nat.call("p2c_tiny_stpf_predict", weights.data_ptr(), feats.data_ptr(), n, out.data_ptr(),
             nat.stream_handle(stream))
    return out[:n]


# ---- STPF priority trunk (FP32) ---------------------------------------------
class StpfPriority(torch.nn.Module):
    """Trunk (32 -> 128 GELU -> 128 GELU) and priority head of the STPF model;
    the training-side twin of the hand-written inference kernel."""

    def __init__(self):
        super().__init__()
        self.l1 = torch.nn.Linear(32, 128)
        self.l2 = torch.nn.Linear(128, 128)
        self.head = torch.nn.Linear(128, 1)

    def logits(self, x: torch.Tensor) -> torch.Tensor:
        g = torch.nn.functional.gelu
        return self.head(g(self.l2(g(self.l1(x))))).squeeze(-1)

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        return torch.sigmoid(self.logits(x))

    def packed(self) -> torch.Tensor:
        """Weights in the kernel's layout (w1, b1, w2, b2, wp, bp), float32."""
        parts = [self.l1.weight, self.l1.bias, self.l2.weight, self.l2.bias, self.head.weight, self.head.bias]
        return torch.cat([p.detach().reshape(-1).float() for p in parts]).contiguous()


def stpf_priority(packed: torch.Tensor, feats: torch.Tensor, stream=None) -> torch.Tensor:
    """Priority in (0, 1) per row of feats (float32[n][32]) through the HIP kernel."""
    assert packed.dtype == torch.float32 and packed.numel() == STPF_WEIGHTS and packed.is_cuda
    feats = feats.contiguous().float()
    n = feats.shape[0]
    out = torch.empty(max(n, 1), dtype=torch.float32, device=feats.device)
    nat.call("p2c_stpf_priority", packed.data_ptr(), feats.data_ptr(), n, out.data_ptr(), nat.stream_handle(stream))
    return out[:n]


def save_stpf(path: str | Path, model: StpfPriority, meta: dict | None = None) -> None:
    np.savez(path, packed=model.packed().cpu().numpy(), **{k: np.asarray(v) for k, v in (meta or {}).items()})


DEFAULT_PRIORITY_WEIGHTS = Path(__file__).resolve().parent / "data" / "stpf_priority.npz"


def load_stpf(path: str | Path, device) -> torch.Tensor:
    with np.load(path, allow_pickle=False) as z:
        packed = torch.from_numpy(np.ascontiguousarray(z["packed"], dtype=np.float32))
    if packed.numel() != STPF_WEIGHTS:
        raise RuntimeError("unexpected STPF priority weight count")
    return packed.to(device)
