"""Validators of the runtime contracts: the reference package's rules and
messages (src/python/p2cccd/validators.py:34-300), so a record accepted or
refused there is accepted or refused here, with the same ValidationError
text.  Each validator returns its argument."""
from __future__ import annotations

import math
from dataclasses import asdict, is_dataclass
from enum import IntEnum
from typing import Any

from .contracts import (
    CONTRACT_SCHEMA_VERSION,
    CONTRACT_TYPES,
    INT64_MAX,
    INT64_MIN,
    MAX_FAMILY_SCORES,
    MAX_INTERVAL_SCORES,
    MOTION_BOUND_SIZE,
    UINT8_MAX,
    UINT16_MAX,
    UINT32_MAX,
    UINT64_MAX,
    AuditLogRow,
    AuditStage,
    CandidateRecord,
    CertificateRefinementMode,
    CertificateResult,
    CertificateStatus,
    ExactWorkItem,
    ProposalOutput,
    ProposalSource,
    ProxyType,
    schema_field_names,
)


class ValidationError(ValueError):
    """A runtime contract violates its stable schema."""


def _fail(message: str):
    raise ValidationError(message)


def _uint(name: str, v: Any, hi: int, *, lo: int = 0, required: bool = False) -> int:
    if isinstance(v, bool) or not isinstance(v, int):
        _fail(f"{name} must be an integer")
    if v < lo:
        _fail(f"{name} must be >= {lo}")
    if v > hi:
        _fail(f"{name} must be <= {hi}")
    if required and v == 0:
        _fail(f"{name} is required")
    return v


def _int64(name: str, v: Any) -> int:
    return _uint(name, v, INT64_MAX, lo=INT64_MIN)


def _finite(name: str, v: Any) -> float:
    if isinstance(v, bool) or not isinstance(v, (int, float)):
        _fail(f"{name} must be numeric")
    x = float(v)
    if not math.isfinite(x):
        _fail(f"{name} must be finite")
    return x


def _nonneg(name: str, v: Any) -> float:
    x = _finite(name, v)
    if x < 0.0:
        _fail(f"{name} must be non-negative")
    return x


def _positive(name: str, v: Any) -> float:
    x = _finite(name, v)
    if x <= 0.0:
        _fail(f"{name} must be positive")
    return x


def _enum(name: str, v: Any, kind: type[IntEnum], *, concrete: bool = False) -> IntEnum:
    if isinstance(v, bool):
        _fail(f"{name} must be an integer enum value, not bool")
    try:
        e = kind(v)
    except (TypeError, ValueError) as exc:
        raise ValidationError(f"{name} has invalid enum value {v!r}") from exc
    if concrete and e.value == 0:
        _fail(f"{name} must be a concrete enum value")
    return e


def _numbers(name: str, v: Any, n: int, *, nonneg: bool = False) -> None:
    if not isinstance(v, list):
        _fail(f"{name} must be a list")
    if len(v) != n:
        _fail(f"{name} must have length {n}")
    check = _nonneg if nonneg else _finite
    for i, x in enumerate(v):
        check(f"{name}[{i}]", x)


def _unit_interval(t0: Any, t1: Any) -> None:
    a, b = _finite("interval_t0", t0), _finite("interval_t1", t1)
    if a < 0.0 or b > 1.0 or a > b:
        _fail("interval must satisfy 0 <= interval_t0 <= interval_t1 <= 1")


def validate_dict_schema(data: dict, contract_type: str | type | object) -> None:
    if not isinstance(data, dict):
        _fail("contract payload must be a dict")
    try:
        required = set(schema_field_names(contract_type))
    except KeyError as exc:
        raise ValidationError(f"unknown contract type {contract_type!r}") from exc
    missing, extra = sorted(required - set(data)), sorted(set(data) - required)
    if missing:
        _fail(f"missing required field(s): {', '.join(missing)}")
    if extra:
        _fail(f"unknown field(s): {', '.join(extra)}")


def validate_candidate_record(r: CandidateRecord) -> CandidateRecord:
    p = "CandidateRecord."
    _uint(p + "schema_version", r.schema_version, UINT32_MAX, lo=1)
    if r.schema_version != CONTRACT_SCHEMA_VERSION:
        _fail("CandidateRecord.schema_version is unsupported")
    _uint(p + "candidate_id", r.candidate_id, UINT64_MAX, required=True)
    _uint(p + "query_id", r.query_id, UINT64_MAX, required=True)
    for f in ("slab_id", "object_a_id", "patch_a_id", "object_b_id", "patch_b_id"):
        _uint(p + f, getattr(r, f), UINT32_MAX)
    _enum(p + "proxy_type_a", r.proxy_type_a, ProxyType, concrete=True)
    _enum(p + "proxy_type_b", r.proxy_type_b, ProxyType, concrete=True)
    _uint(p + "rt_hit_count", r.rt_hit_count, UINT32_MAX, required=True)
    _numbers(p + "motion_bound", r.motion_bound, MOTION_BOUND_SIZE, nonneg=True)
    _uint(p + "proxy_features_offset", r.proxy_features_offset, UINT32_MAX)
    _uint(p + "flags", r.flags, UINT32_MAX)
    return r


def validate_proposal_output(o: ProposalOutput) -> ProposalOutput:
    p = "ProposalOutput."
    _uint(p + "candidate_id", o.candidate_id, UINT64_MAX, required=True)
    _numbers(p + "interval_scores", o.interval_scores, MAX_INTERVAL_SCORES)
    _numbers(p + "family_scores", o.family_scores, MAX_FAMILY_SCORES)
    _finite(p + "priority_score", o.priority_score)
    _nonneg(p + "cost_score", o.cost_score)
    _nonneg(p + "uncertainty_score", o.uncertainty_score)
    return o


def validate_exact_work_item(w: ExactWorkItem) -> ExactWorkItem:
    p = "ExactWorkItem."
    _uint(p + "work_item_id", w.work_item_id, UINT64_MAX, required=True)
    _uint(p + "parent_candidate_id", w.parent_candidate_id, UINT64_MAX, required=True)
    _uint(p + "query_id", w.query_id, UINT64_MAX, required=True)
    for f in ("slab_id", "patch_a_id", "patch_b_id"):
        _uint(p + f, getattr(w, f), UINT32_MAX)
    _unit_interval(w.interval_t0, w.interval_t1)
    _uint(p + "feature_family_mask", w.feature_family_mask, UINT32_MAX, required=True)
    _uint(p + "topk_feature_ids_offset", w.topk_feature_ids_offset, UINT32_MAX)
    _uint(p + "depth", w.depth, UINT16_MAX)
    _finite(p + "priority_score", w.priority_score)
    _enum(p + "source", w.source, ProposalSource)
    return w


def validate_certificate_result(c: CertificateResult) -> CertificateResult:
    p = "CertificateResult."
    _uint(p + "work_item_id", c.work_item_id, UINT64_MAX, required=True)
    _uint(p + "query_id", c.query_id, UINT64_MAX, required=True)
    status = _enum(p + "status", c.status, CertificateStatus)
    _unit_interval(c.interval_t0, c.interval_t1)
    toi = _finite(p + "toi_upper", c.toi_upper)
    margin = _finite(p + "safe_margin_lb", c.safe_margin_lb)
    _uint(p + "witness_family", c.witness_family, UINT8_MAX)
    _int64(p + "witness_id_a", c.witness_id_a)
    _int64(p + "witness_id_b", c.witness_id_b)
    _uint(p + "covered_feature_mask", c.covered_feature_mask, UINT32_MAX)
    _positive(p + "eps_time", c.eps_time)
    _positive(p + "eps_space", c.eps_space)
    _uint(p + "reason_code", c.reason_code, UINT16_MAX)
    mode = _enum(p + "next_refinement_mode", c.next_refinement_mode, CertificateRefinementMode)
    if status == CertificateStatus.COLLISION:
        if toi < c.interval_t0 or toi > c.interval_t1:
            _fail("CertificateResult.toi_upper must lie inside the certified interval")
        if c.witness_id_a < 0 or c.witness_id_b < 0:
            _fail("CertificateResult collision witnesses are required")
        if mode != CertificateRefinementMode.NONE:
            _fail("CertificateResult collision cannot request refinement")
    if status == CertificateStatus.SEPARATION:
        if margin < 0.0:
            _fail("CertificateResult.safe_margin_lb must be non-negative")
        if c.covered_feature_mask == 0:
            _fail("CertificateResult.covered_feature_mask is required for separation")
        if mode != CertificateRefinementMode.NONE:
            _fail("CertificateResult separation cannot request refinement")
    if status == CertificateStatus.UNDECIDED and c.reason_code == 0:
        _fail("CertificateResult.reason_code is required for undecided results")
    if status == CertificateStatus.UNDECIDED and mode == CertificateRefinementMode.NONE:
        _fail("CertificateResult.next_refinement_mode is required for undecided results")
    return c


def validate_audit_log_row(a: AuditLogRow) -> AuditLogRow:
    p = "AuditLogRow."
    _uint(p + "event_id", a.event_id, UINT64_MAX, required=True)
    _uint(p + "query_id", a.query_id, UINT64_MAX, required=True)
    _uint(p + "candidate_id", a.candidate_id, UINT64_MAX)
    _uint(p + "work_item_id", a.work_item_id, UINT64_MAX)
    _enum(p + "stage", a.stage, AuditStage)
    _uint(p + "action", a.action, UINT16_MAX)
    _uint(p + "depth", a.depth, UINT16_MAX)
    _unit_interval(a.interval_t0, a.interval_t1)
    _uint(p + "timestamp_us", a.timestamp_us, UINT64_MAX, required=True)
    _finite(p + "aux_value0", a.aux_value0)
    _finite(p + "aux_value1", a.aux_value1)
    return a


_VALIDATORS = {
    CandidateRecord: validate_candidate_record,
    ProposalOutput: validate_proposal_output,
    ExactWorkItem: validate_exact_work_item,
    CertificateResult: validate_certificate_result,
    AuditLogRow: validate_audit_log_row,
}


def validate_contract(value: Any) -> Any:
    """Validates any contract record (its schema's fields, then its rules)."""
    kind = type(value)
    if kind not in _VALIDATORS or not is_dataclass(value):
        raise ValidationError(f"unknown contract type {kind.__name__!r}")
    validate_dict_schema(asdict(value), kind)
    return _VALIDATORS[kind](value)


__all__ = [
    "CONTRACT_TYPES",
    "ValidationError",
    "validate_audit_log_row",
    "validate_candidate_record",
    "validate_certificate_result",
    "validate_contract",
    "validate_dict_schema",
    "validate_exact_work_item",
    "validate_proposal_output",
]
