"""Candidate generation for external CCD query batches: the reference package's
generate_candidates_for_external_batch (src/python/p2cccd/candidate_generation.py:
31-153, rt_exact.py:470-679) over the device path.

Each query's two swept boxes are built and tested on the GPU
(p2cccd_cpp.generate_candidates_for_external_batch); candidates come back as
contract records, validated, with the runtime query ids and each runtime
query's conservative family mask.  There is no Python broad-phase fallback:
the native module is required."""
from __future__ import annotations

import importlib
from dataclasses import dataclass
from typing import Any

from .contracts import CandidateRecord, ProxyType
from .validators import validate_candidate_record

FEATURE_FAMILY_POINT_TRIANGLE = 1 << 0
FEATURE_FAMILY_EDGE_EDGE = 1 << 1
FEATURE_FAMILY_CONSERVATIVE = FEATURE_FAMILY_POINT_TRIANGLE | FEATURE_FAMILY_EDGE_EDGE

CPP_CANDIDATE_ENTRYPOINTS = (
    "generate_candidates_for_internal_samples",
    "generate_candidates_for_external_batch",
    "generate_candidates",
)
BACKEND_NAMES = {"cpu_reference_rt", "optix_compatible", "optix_rt", "hip_exact"}


@dataclass(frozen=True, slots=True)
class RtCandidateTiming:
    build_ms: float = 0.0
    update_ms: float = 0.0
    trace_ms: float = 0.0
    compact_ms: float = 0.0
    total_ms: float = 0.0


@dataclass(frozen=True, slots=True)
class RtCandidateStats:
    backend_name: str
    primitive_count: int
    raw_hit_count: int
    compact_candidate_count: int
    candidate_recall: float
    timing: RtCandidateTiming


@dataclass(frozen=True, slots=True)
class CandidateGenerationWrapperConfig:
    backend_name: str = "cpu_reference_rt"
    same_query_only: bool = True
    prefer_cpp_backend: bool = False
    allow_python_fallback: bool = True


@dataclass(frozen=True, slots=True)
class CandidateGenerationWrapperResult:
    candidates: tuple[CandidateRecord, ...]
    candidate_stats: RtCandidateStats
    runtime_query_ids: dict[int, int]
    family_by_runtime_query_id: dict[int, int]
    source_name: str
    scene_name: str
    batch_id: str
    used_cpp_backend: bool
    fallback_reason: str

    @property
    def candidate_count(self) -> int:
        return len(self.candidates)


def _load_cpp_module() -> Any | None:
    try:
        return importlib.import_module("p2cccd_cpp")
    except ImportError:
        return None


def is_cpp_candidate_generation_available() -> bool:
    m = _load_cpp_module()
    return m is not None and any(callable(getattr(m, n, None)) for n in CPP_CANDIDATE_ENTRYPOINTS)


def family_mask_for_external(query) -> int:
    f = str(getattr(query.family, "value", query.family))
    if f == "vf":
        return FEATURE_FAMILY_POINT_TRIANGLE
    if f == "ee":
        return FEATURE_FAMILY_EDGE_EDGE
    raise ValueError(f"unsupported external CCD query family: {query.family}")


def _record(c) -> CandidateRecord:
    return validate_candidate_record(CandidateRecord(
        schema_version=int(c.schema_version), candidate_id=int(c.candidate_id), query_id=int(c.query_id),
        slab_id=int(c.slab_id), object_a_id=int(c.object_a_id), patch_a_id=int(c.patch_a_id),
        object_b_id=int(c.object_b_id), patch_b_id=int(c.patch_b_id), proxy_type_a=ProxyType(int(c.proxy_type_a)),
        proxy_type_b=ProxyType(int(c.proxy_type_b)), rt_hit_count=int(c.rt_hit_count),
        motion_bound=[float(v) for v in c.motion_bound], proxy_features_offset=int(c.proxy_features_offset),
        flags=int(c.flags)))


def generate_candidates_for_external_batch(batch, config: CandidateGenerationWrapperConfig | None = None, *,
                                           backend=None) -> CandidateGenerationWrapperResult:
    if not batch.queries:
        raise ValueError("candidate generation requires at least one external query")
    cfg = config or CandidateGenerationWrapperConfig()
    if cfg.backend_name not in BACKEND_NAMES:
        raise ValueError("CandidateGenerationWrapperConfig.backend_name is unsupported")
    if backend is not None:
        raise ValueError("external candidates come from the device path; a Python broad-phase backend is not used")
    cpp = _load_cpp_module()
    if cpp is None:
        raise RuntimeError("p2cccd_cpp is required: external candidates are generated on the device")
    r = cpp.generate_candidates_for_external_batch(batch, backend_name="hip_exact",
                                                   allow_optix_cpu_fallback=cfg.backend_name == "optix_compatible")
    runtime_ids = {int(m.source_query_id): int(m.runtime_query_id) for m in r.runtime_query_ids}
    t = r.timing
    stats = RtCandidateStats(backend_name=str(r.backend_name), primitive_count=int(r.primitive_count),
                             raw_hit_count=int(r.raw_hit_count), compact_candidate_count=int(r.compact_candidate_count),
                             candidate_recall=float(r.candidate_recall),
                             timing=RtCandidateTiming(build_ms=t.build_ms, update_ms=t.update_ms, trace_ms=t.trace_ms,
                                                      compact_ms=t.compact_ms, total_ms=t.total_ms))
    return CandidateGenerationWrapperResult(
        candidates=tuple(_record(c) for c in r.candidates), candidate_stats=stats, runtime_query_ids=runtime_ids,
        family_by_runtime_query_id={runtime_ids[q.query_id]: family_mask_for_external(q) for q in batch.queries},
        source_name=str(getattr(batch, "source_name", "")), scene_name=str(getattr(batch, "scene_name", "")),
        batch_id=str(getattr(batch, "batch_id", "")), used_cpp_backend=True, fallback_reason="")
