"""The reference's Python certificate twin on the GPU.

Mirrors p2cccd.bench.pure_exact_cpu (src/python/p2cccd/bench/pure_exact_cpu.py):
`PureExactCPUConfig` (:21-26), `PureExactQueryResult` (:29-39),
`evaluate_external_ccd_query` (:247-308), with the same validation messages
(:54-62).  Queries are certified in one batched launch of the reference
interval certificate with the Python twin's rules (P2C_CERT_PYTHON_TWIN: EE
guard abs(denom) > 1e-14, motion radius from the summed largest
displacements, contact at the first smallest sample) over [0, 1].  A query
carries `query_id`, `family` ("vf"/"ee", or an enum whose value is one of
them), `vertices_t0`, `vertices_t1` (four points each) and optionally
`ground_truth_collides`.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Sequence

import numpy as np
import torch

from . import _native as nat

_STATUS = {0: "collision", 1: "separation", 2: "undecided"}
_WITNESS = {"vf": "point_triangle", "ee": "edge_edge"}


@dataclass(frozen=True, slots=True)
class PureExactCPUConfig:
    eps_time: float = 1.0e-5
    eps_space: float = 1.0e-8
    max_subdivision_depth: int = 32
    conservative_undecided_as_collision: bool = True


@dataclass(frozen=True, slots=True)
class PureExactQueryResult:
    query_id: int
    family: str
    predicted_collision: bool
    ground_truth_collision: bool | None
    status: str
    toi_upper: float
    safe_margin_lb: float
    exact_evals: int
    max_depth: int


def _validate_config(config: PureExactCPUConfig) -> PureExactCPUConfig:
    if not math.isfinite(config.eps_time) or config.eps_time <= 0.0:
        raise ValueError("PureExactCPUConfig.eps_time must be finite and positive")
    if not math.isfinite(config.eps_space) or config.eps_space <= 0.0:
        raise ValueError("PureExactCPUConfig.eps_space must be finite and positive")
    if config.max_subdivision_depth < 0:
        raise ValueError("PureExactCPUConfig.max_subdivision_depth must be non-negative")
    return config


def _family(query) -> str:
    f = str(getattr(query.family, "value", query.family)).lower()
    if f not in _WITNESS:
        raise ValueError(f"unsupported external CCD query family: {query.family}")
    return f


def evaluate_external_ccd_queries(queries: Sequence, config: PureExactCPUConfig | None = None,
                                  device: str | torch.device = "cuda") -> list[PureExactQueryResult]:
    """Certifies every query (one launch per family); results in query order."""
    cfg = _validate_config(config or PureExactCPUConfig())
    out: list[PureExactQueryResult | None] = [None] * len(queries)
    fams = [_family(q) for q in queries]
    for fam, code in (("vf", 1), ("ee", 2)):
        idx = [i for i, f in enumerate(fams) if f == fam]
        if not idx:
            continue
        x = np.empty((len(idx), 8, 3), np.float64)
        for r, i in enumerate(idx):
            x[r, :4] = np.asarray(queries[i].vertices_t0, np.float64)
            x[r, 4:] = np.asarray(queries[i].vertices_t1, np.float64)
        if not np.isfinite(x).all():
            raise ValueError("external CCD query vertices must be finite")
        ids = np.stack([np.arange(len(idx)), np.arange(len(idx))], 1)
        prims = nat.prims_from_positions(x, ids=ids)
        c = nat.cert_config(cfg.eps_time, cfg.eps_space, cfg.max_subdivision_depth)
        c["flags"] = 2  # P2C_CERT_PYTHON_TWIN
        d_in = torch.from_numpy(prims.view(np.uint8).copy()).to(device)
        d_out = torch.zeros(len(idx) * nat.PRIM_RESULT_DTYPE.itemsize, dtype=torch.uint8, device=device)
        nat.call("p2c_cert_ref_batch", code, d_in.data_ptr(), len(idx), 0.0, 1.0, nat.ptr(c), d_out.data_ptr(),
                 0, nat.stream_handle())
        res = d_out.cpu().numpy().view(nat.PRIM_RESULT_DTYPE)
        for r, i in enumerate(idx):
            status = _STATUS[int(res["status"][r])]
            q = queries[i]
            out[i] = PureExactQueryResult(
                query_id=q.query_id, family=_WITNESS[fam],
                predicted_collision=status == "collision" or (status == "undecided"
                                                               and cfg.conservative_undecided_as_collision),
                ground_truth_collision=getattr(q, "ground_truth_collides", None), status=status,
                toi_upper=float(res["toi_upper"][r]), safe_margin_lb=float(res["safe_margin_lb"][r]),
                exact_evals=3 * int(res["nodes"][r]), max_depth=int(res["pad"][r][0]))
    return out


def evaluate_external_ccd_query(query, config: PureExactCPUConfig | None = None) -> PureExactQueryResult:
    return evaluate_external_ccd_queries([query], config)[0]
