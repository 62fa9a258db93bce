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


WALK_PYTHON_TWIN = 2  # P2C_CERT_PYTHON_TWIN
WALK_CUDA_TWIN = 1    # P2C_CERT_CUDA_TWIN


def certify_queries(queries: Sequence, config: PureExactCPUConfig, walk: int,
                    device: str | torch.device = "cuda") -> np.ndarray:
    """Raw certificate results (PRIM_RESULT_DTYPE, query order) of external
    queries over [0, 1] with one of the reference's walks, one device launch
    per family.  Primitive ids are 0 and 1, as the reference's twin path
    numbers them (rt_exact.py:223-281)."""
    cfg = _validate_config(config)
    out = np.zeros(len(queries), nat.PRIM_RESULT_DTYPE)
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
        prims = nat.prims_from_positions(x, ids=np.tile([0, 1], (len(idx), 1)))
        c = nat.cert_config(cfg.eps_time, cfg.eps_space, cfg.max_subdivision_depth)
        c["flags"] = walk
        d_in = torch.from_numpy(prims.view(np.uint8).copy()).to(device)
        d_out = torch.zeros(len(idx) * nat.PRIM_RESULT_DTYPE.itemsize, dtype=torch.uint8, device=device)
        nat.call("p2c_cert_ref_batch", code, d_in.data_ptr(), len(idx), 0.0, 1.0, nat.ptr(c), d_out.data_ptr(),
                 0, nat.stream_handle())
        out[idx] = d_out.cpu().numpy().view(nat.PRIM_RESULT_DTYPE)
    return out


def evaluate_external_ccd_queries(queries: Sequence, config: PureExactCPUConfig | None = None,
                                  device: str | torch.device = "cuda") -> list[PureExactQueryResult]:
    """Certifies every query with the Python twin's walk (one launch per
    family); results in query order."""
    cfg = _validate_config(config or PureExactCPUConfig())
    res = certify_queries(queries, cfg, WALK_PYTHON_TWIN, device)
    out = []
    for i, q in enumerate(queries):
        status = _STATUS[int(res["status"][i])]
        out.append(PureExactQueryResult(
            query_id=q.query_id, family=_WITNESS[_family(q)],
            predicted_collision=status == "collision" or (status == "undecided"
                                                           and cfg.conservative_undecided_as_collision),
            ground_truth_collision=getattr(q, "ground_truth_collides", None), status=status,
            toi_upper=float(res["toi_upper"][i]), safe_margin_lb=float(res["safe_margin_lb"][i]),
            exact_evals=3 * int(res["nodes"][i]), max_depth=int(res["pad"][i][0])))
    return out


def evaluate_external_ccd_query(query, config: PureExactCPUConfig | None = None) -> PureExactQueryResult:
    return evaluate_external_ccd_queries([query], config)[0]
