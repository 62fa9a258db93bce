"""Runtime contracts of the certificate path: the reference package's record
types, enum codes and stable schemas (src/python/p2cccd/contracts.py:8-130,
222-290), which the native layer mirrors field for field
(src/cpp/common/runtime_contracts.h:8-114).  Benchmark-report rows are not
part of the path and are not kept."""
from __future__ import annotations

from dataclasses import dataclass, field, fields, is_dataclass
from enum import IntEnum
from typing import List

UINT8_MAX = 2**8 - 1
UINT16_MAX = 2**16 - 1
UINT32_MAX = 2**32 - 1
UINT64_MAX = 2**64 - 1
INT64_MIN = -(2**63)
INT64_MAX = 2**63 - 1

CONTRACT_SCHEMA_VERSION = 1
MAX_INTERVAL_SCORES = 8
MAX_FAMILY_SCORES = 8
MOTION_BOUND_SIZE = 4


class ProxyType(IntEnum):
    UNKNOWN = 0
    SWEPT_AABB = 1
    CAPSULE = 2


class ProposalSource(IntEnum):
    RAW = 0
    REFINED = 1
    FALLBACK = 2


class CertificateStatus(IntEnum):
    COLLISION = 0
    SEPARATION = 1
    UNDECIDED = 2


class CertificateRefinementMode(IntEnum):
    NONE = 0
    BISECT_INTERVAL = 1
    REQUEST_GEOMETRY = 2
    ESCALATE_PRECISION = 3


class AuditStage(IntEnum):
    RT = 0
    PROPOSAL = 1
    EXACT = 2
    REFINE = 3
    CERTIFY = 4


@dataclass(slots=True)
class CandidateRecord:
    schema_version: int = CONTRACT_SCHEMA_VERSION
    candidate_id: int = 0
    query_id: int = 0
    slab_id: int = 0
    object_a_id: int = 0
    patch_a_id: int = 0
    object_b_id: int = 0
    patch_b_id: int = 0
    proxy_type_a: ProxyType = ProxyType.UNKNOWN
    proxy_type_b: ProxyType = ProxyType.UNKNOWN
    rt_hit_count: int = 0
    motion_bound: List[float] = field(default_factory=lambda: [0.0] * MOTION_BOUND_SIZE)
    proxy_features_offset: int = 0
    flags: int = 0


@dataclass(slots=True)
class ProposalOutput:
    candidate_id: int = 0
    interval_scores: List[float] = field(default_factory=lambda: [0.0] * MAX_INTERVAL_SCORES)
    family_scores: List[float] = field(default_factory=lambda: [0.0] * MAX_FAMILY_SCORES)
    priority_score: float = 0.0
    cost_score: float = 0.0
    uncertainty_score: float = 0.0


@dataclass(slots=True)
class ExactWorkItem:
    work_item_id: int = 0
    parent_candidate_id: int = 0
    query_id: int = 0
    slab_id: int = 0
    patch_a_id: int = 0
    patch_b_id: int = 0
    interval_t0: float = 0.0
    interval_t1: float = 1.0
    feature_family_mask: int = 0
    topk_feature_ids_offset: int = 0
    depth: int = 0
    priority_score: float = 0.0
    source: ProposalSource = ProposalSource.RAW


@dataclass(slots=True)
class CertificateResult:
    work_item_id: int = 0
    query_id: int = 0
    status: CertificateStatus = CertificateStatus.UNDECIDED
    interval_t0: float = 0.0
    interval_t1: float = 1.0
    toi_upper: float = 1.0
    safe_margin_lb: float = 0.0
    witness_family: int = 0
    witness_id_a: int = -1
    witness_id_b: int = -1
    covered_feature_mask: int = 0
    eps_time: float = 1.0e-4
    eps_space: float = 1.0e-6
    reason_code: int = 0
    next_refinement_mode: CertificateRefinementMode = CertificateRefinementMode.NONE


@dataclass(slots=True)
class AuditLogRow:
    event_id: int = 0
    query_id: int = 0
    candidate_id: int = 0
    work_item_id: int = 0
    stage: AuditStage = AuditStage.RT
    action: int = 0
    depth: int = 0
    interval_t0: float = 0.0
    interval_t1: float = 1.0
    timestamp_us: int = 0
    aux_value0: float = 0.0
    aux_value1: float = 0.0


def _schema(cls, types: dict) -> tuple:
    return tuple((f.name, types[f.name]) for f in fields(cls))


# field name -> stable wire type, in declaration order
CONTRACT_SCHEMAS: dict[str, tuple[tuple[str, str], ...]] = {
    "CandidateRecord": _schema(CandidateRecord, dict(
        schema_version="uint32", candidate_id="uint64", query_id="uint64", slab_id="uint32", object_a_id="uint32",
        patch_a_id="uint32", object_b_id="uint32", patch_b_id="uint32", proxy_type_a="ProxyType:uint8",
        proxy_type_b="ProxyType:uint8", rt_hit_count="uint32", motion_bound="float32[4]",
        proxy_features_offset="uint32", flags="uint32")),
    "ProposalOutput": _schema(ProposalOutput, dict(
        candidate_id="uint64", interval_scores="float32[8]", family_scores="float32[8]", priority_score="float32",
        cost_score="float32", uncertainty_score="float32")),
    "ExactWorkItem": _schema(ExactWorkItem, dict(
        work_item_id="uint64", parent_candidate_id="uint64", query_id="uint64", slab_id="uint32",
        patch_a_id="uint32", patch_b_id="uint32", interval_t0="float64", interval_t1="float64",
        feature_family_mask="uint32", topk_feature_ids_offset="uint32", depth="uint16", priority_score="float32",
        source="ProposalSource:uint8")),
    "CertificateResult": _schema(CertificateResult, dict(
        work_item_id="uint64", query_id="uint64", status="CertificateStatus:uint8", interval_t0="float64",
        interval_t1="float64", toi_upper="float64", safe_margin_lb="float64", witness_family="uint8",
        witness_id_a="int64", witness_id_b="int64", covered_feature_mask="uint32", eps_time="float64",
        eps_space="float64", reason_code="uint16", next_refinement_mode="CertificateRefinementMode:uint8")),
    "AuditLogRow": _schema(AuditLogRow, dict(
        event_id="uint64", query_id="uint64", candidate_id="uint64", work_item_id="uint64",
        stage="AuditStage:uint8", action="uint16", depth="uint16", interval_t0="float64", interval_t1="float64",
        timestamp_us="uint64", aux_value0="float64", aux_value1="float64")),
}

CONTRACT_TYPES: dict[str, type] = {
    "CandidateRecord": CandidateRecord,
    "ProposalOutput": ProposalOutput,
    "ExactWorkItem": ExactWorkItem,
    "CertificateResult": CertificateResult,
    "AuditLogRow": AuditLogRow,
}


def schema_field_names(contract_type: str | type | object) -> tuple[str, ...]:
    """Field names of a contract given by name, class or instance (KeyError if unknown)."""
    if isinstance(contract_type, str):
        name = contract_type
    elif isinstance(contract_type, type):
        name = contract_type.__name__
    elif is_dataclass(contract_type):
        name = type(contract_type).__name__
    else:
        raise KeyError(contract_type)
    return tuple(n for n, _ in CONTRACT_SCHEMAS[name])
