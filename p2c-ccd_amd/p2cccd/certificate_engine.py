"""The certificate engine over external CCD query batches: the reference
package's execute_certificate_engine_for_external_batch
(src/python/p2cccd/certificate_engine.py:44-250, rt_exact.py:356-474,
741-1012, 1104-1145) with every certificate computed on the GPU.

Candidates -> one exact work item per candidate (interval [0, 1], priority =
rt_hit_count) -> one certificate per work item -> an audit trail of RT,
DEQUEUED and EXACT rows with running event ids and timestamps -> coverage
and contract validation.  The certificate walk is the one the reference's
Python layer takes when its native module is present: its CUDA batch twin's
(UNDECIDED at the first unsplittable interval) for 32 work items or more,
its Python twin's (pure_exact_cpu.py) below that; both run as one device
launch per family (p2cccd.pure_exact.certify_queries)."""
from __future__ import annotations

import importlib
import time
from dataclasses import dataclass, field
from typing import Any, Sequence

from .candidate_generation import (
    FEATURE_FAMILY_CONSERVATIVE,
    CandidateGenerationWrapperConfig,
    CandidateGenerationWrapperResult,
    generate_candidates_for_external_batch,
)
from .contracts import (
    AuditLogRow,
    AuditStage,
    CandidateRecord,
    CertificateRefinementMode,
    CertificateResult,
    CertificateStatus,
    ExactWorkItem,
    ProposalSource,
)
from .pure_exact import (
    WALK_CUDA_TWIN,
    WALK_PYTHON_TWIN,
    PureExactCPUConfig,
    PureExactQueryResult,
    certify_queries,
)
from .validators import (
    validate_audit_log_row,
    validate_candidate_record,
    validate_certificate_result,
    validate_exact_work_item,
)

CPP_CERTIFICATE_ENTRYPOINTS = ("execute_certificate_engine", "process_exact_work_queue", "evaluate_certificate_query")
RT_AUDIT_CANDIDATE = 1
EXACT_AUDIT_DEQUEUED = 1
EXACT_AUDIT_COLLISION = 2
EXACT_AUDIT_SEPARATION = 3
EXACT_AUDIT_UNDECIDED = 4
CERTIFICATE_REASON_NONE = 0
CERTIFICATE_REASON_MAX_SUBDIVISION = 2
TWIN_BATCH_MIN_ITEMS = 32  # the reference sends a queue this long to its CUDA batch (rt_exact.py:363)


@dataclass(frozen=True, slots=True)
class CertificateEngineWrapperConfig:
    exact: PureExactCPUConfig = field(default_factory=PureExactCPUConfig)
    first_work_item_id: int = 1
    first_event_id: int = 1
    first_timestamp_us: int = 1
    prefer_cpp_backend: bool = False
    allow_python_fallback: bool = True


@dataclass(frozen=True, slots=True)
class CertificateEngineWrapperResult:
    query_results: tuple[PureExactQueryResult, ...]
    candidates: tuple[CandidateRecord, ...]
    work_items: tuple[ExactWorkItem, ...]
    certificates: tuple[CertificateResult, ...]
    audit_log: tuple[AuditLogRow, ...]
    exact_ms: float
    source_name: str
    scene_name: str
    batch_id: str
    used_cpp_backend: bool
    fallback_reason: str

    @property
    def final_fn_zero(self) -> bool:
        return all(r.ground_truth_collision is not True or r.predicted_collision for r in self.query_results)

    @property
    def queue_conserved(self) -> bool:
        return len(self.candidates) == len(self.work_items) == len(self.certificates)


def is_cpp_certificate_engine_available() -> bool:
    try:
        m: Any = importlib.import_module("p2cccd_cpp")
    except ImportError:
        return False
    return any(callable(getattr(m, n, None)) for n in CPP_CERTIFICATE_ENTRYPOINTS) or hasattr(m, "CertificateEngine")


def _check_config(c: CertificateEngineWrapperConfig) -> CertificateEngineWrapperConfig:
    for name in ("first_work_item_id", "first_event_id", "first_timestamp_us"):
        if getattr(c, name) <= 0:
            raise ValueError(f"CertificateEngineWrapperConfig.{name} must be positive")
    return c


def schedule_exact_work_items_without_stpf(candidates: Sequence[CandidateRecord], *,
                                           family_by_runtime_query_id: dict[int, int],
                                           first_work_item_id: int = 1) -> tuple[ExactWorkItem, ...]:
    if first_work_item_id <= 0:
        raise ValueError("first_work_item_id must be positive")
    items = []
    for k, c in enumerate(candidates):
        validate_candidate_record(c)
        items.append(validate_exact_work_item(ExactWorkItem(
            work_item_id=first_work_item_id + k, parent_candidate_id=c.candidate_id, query_id=c.query_id,
            slab_id=c.slab_id, patch_a_id=c.patch_a_id, patch_b_id=c.patch_b_id, interval_t0=0.0, interval_t1=1.0,
            feature_family_mask=family_by_runtime_query_id.get(c.query_id, FEATURE_FAMILY_CONSERVATIVE),
            topk_feature_ids_offset=0, depth=0, priority_score=float(c.rt_hit_count), source=ProposalSource.RAW)))
    return tuple(items)


_STATUS = {0: "collision", 1: "separation", 2: "undecided"}
_MODE = {1: CertificateRefinementMode.BISECT_INTERVAL, 2: CertificateRefinementMode.REQUEST_GEOMETRY}


def _twin_certificate(r, item: ExactWorkItem, cfg: PureExactCPUConfig) -> CertificateResult:
    """A certificate from the batch twin's primitive result (rt_exact.py:300-330)."""
    return CertificateResult(
        work_item_id=item.work_item_id, query_id=item.query_id, status=CertificateStatus(int(r["status"])),
        interval_t0=float(r["interval_t0"]), interval_t1=float(r["interval_t1"]), toi_upper=float(r["toi_upper"]),
        safe_margin_lb=max(0.0, float(r["safe_margin_lb"])), witness_family=int(r["witness_family"]),
        witness_id_a=int(r["witness_id_a"]), witness_id_b=int(r["witness_id_b"]),
        covered_feature_mask=int(r["covered_feature_mask"]), eps_time=cfg.eps_time, eps_space=cfg.eps_space,
        reason_code=int(r["reason_code"]),
        next_refinement_mode=_MODE.get(int(r["next_refinement_mode"]), CertificateRefinementMode.NONE))


def _twin_query_result(q, r, cfg: PureExactCPUConfig) -> PureExactQueryResult:
    status = _STATUS[int(r["status"])]
    return PureExactQueryResult(
        query_id=q.query_id, family="point_triangle" if str(q.family) == "vf" else "edge_edge",
        predicted_collision=status == "collision" or (status == "undecided" and cfg.conservative_undecided_as_collision),
        ground_truth_collision=q.ground_truth_collides, status=status, toi_upper=float(r["toi_upper"]),
        safe_margin_lb=max(0.0, float(r["safe_margin_lb"])), exact_evals=1,
        max_depth=cfg.max_subdivision_depth if status == "undecided" else 0)


def _python_query_result(q, r, cfg: PureExactCPUConfig) -> PureExactQueryResult:
    status = _STATUS[int(r["status"])]
    return PureExactQueryResult(
        query_id=q.query_id, family="point_triangle" if str(q.family) == "vf" else "edge_edge",
        predicted_collision=status == "collision" or (status == "undecided" and cfg.conservative_undecided_as_collision),
        ground_truth_collision=q.ground_truth_collides, status=status, toi_upper=float(r["toi_upper"]),
        safe_margin_lb=float(r["safe_margin_lb"]), exact_evals=3 * int(r["nodes"]), max_depth=int(r["pad"][0]))


def _python_certificate(res: PureExactQueryResult, item: ExactWorkItem, cfg: PureExactCPUConfig) -> CertificateResult:
    """A certificate from the Python twin's query result (rt_exact.py:774-833)."""
    base = dict(work_item_id=item.work_item_id, query_id=item.query_id, interval_t0=item.interval_t0,
                interval_t1=item.interval_t1, eps_time=cfg.eps_time, eps_space=cfg.eps_space)
    if res.status == "collision":
        return CertificateResult(**base, status=CertificateStatus.COLLISION,
                                 toi_upper=min(item.interval_t1, max(item.interval_t0, res.toi_upper)),
                                 safe_margin_lb=0.0, witness_family=1 if res.family == "point_triangle" else 2,
                                 witness_id_a=int(item.patch_a_id), witness_id_b=int(item.patch_b_id),
                                 covered_feature_mask=0, reason_code=CERTIFICATE_REASON_NONE,
                                 next_refinement_mode=CertificateRefinementMode.NONE)
    if res.status == "separation":
        return CertificateResult(**base, status=CertificateStatus.SEPARATION, toi_upper=item.interval_t1,
                                 safe_margin_lb=max(0.0, res.safe_margin_lb), witness_family=0, witness_id_a=-1,
                                 witness_id_b=-1, covered_feature_mask=item.feature_family_mask,
                                 reason_code=CERTIFICATE_REASON_NONE,
                                 next_refinement_mode=CertificateRefinementMode.NONE)
    return CertificateResult(**base, status=CertificateStatus.UNDECIDED, toi_upper=item.interval_t1,
                             safe_margin_lb=0.0, witness_family=0, witness_id_a=-1, witness_id_b=-1,
                             covered_feature_mask=0, reason_code=CERTIFICATE_REASON_MAX_SUBDIVISION,
                             next_refinement_mode=CertificateRefinementMode.BISECT_INTERVAL)


def _audit(event: int, query: int, candidate: int, item: int, stage: AuditStage, action: int, depth: int, ts: int,
           aux0: float = 0.0, aux1: float = 0.0) -> AuditLogRow:
    return validate_audit_log_row(AuditLogRow(event_id=event, query_id=query, candidate_id=candidate, work_item_id=item,
                                              stage=stage, action=action, depth=depth, interval_t0=0.0,
                                              interval_t1=1.0, timestamp_us=ts, aux_value0=aux0, aux_value1=aux1))


def validate_rt_exact_coverage(candidates, work_items, certificates) -> None:
    if len(candidates) != len(work_items):
        raise ValueError("RTExact must schedule exactly one exact work item per candidate")
    if len(work_items) != len(certificates):
        raise ValueError("RTExact must emit exactly one certificate per exact work item")
    ids: set[int] = set()
    for c in candidates:
        validate_candidate_record(c)
        if c.candidate_id in ids:
            raise ValueError("RTExact candidate queue contains duplicate candidate_id")
        ids.add(c.candidate_id)
    parents: set[int] = set()
    items: set[int] = set()
    for w in work_items:
        validate_exact_work_item(w)
        if w.parent_candidate_id not in ids:
            raise ValueError("RTExact work item references an unknown candidate")
        if w.parent_candidate_id in parents:
            raise ValueError("RTExact work queue duplicates a parent candidate")
        parents.add(w.parent_candidate_id)
        if w.work_item_id in items:
            raise ValueError("RTExact work queue contains duplicate work_item_id")
        items.add(w.work_item_id)
    covered: set[int] = set()
    for c in certificates:
        validate_certificate_result(c)
        if c.work_item_id not in items:
            raise ValueError("RTExact certificate references an unknown work item")
        if c.work_item_id in covered:
            raise ValueError("RTExact certificates duplicate a work item")
        covered.add(c.work_item_id)
    if covered != items:
        raise ValueError("RTExact work item disappeared without certificate coverage")


def execute_certificate_engine_for_external_batch(batch, candidate_result: CandidateGenerationWrapperResult | None = None,
                                                  config: CertificateEngineWrapperConfig | None = None, *,
                                                  candidate_config: CandidateGenerationWrapperConfig | None = None,
                                                  backend=None) -> CertificateEngineWrapperResult:
    if not batch.queries:
        raise ValueError("certificate engine wrapper requires at least one external query")
    cfg = _check_config(config or CertificateEngineWrapperConfig())
    cands = candidate_result or generate_candidates_for_external_batch(batch, candidate_config, backend=backend)
    items = schedule_exact_work_items_without_stpf(cands.candidates,
                                                   family_by_runtime_query_id=cands.family_by_runtime_query_id,
                                                   first_work_item_id=cfg.first_work_item_id)
    by_runtime = {cands.runtime_query_ids[q.query_id]: q for q in batch.queries}
    by_candidate = {c.candidate_id: c for c in cands.candidates}
    queries = []
    for w in items:
        q = by_runtime.get(w.query_id)
        if q is None:
            raise ValueError("exact work item references an unknown external query")
        queries.append(q)
    twin = len(items) >= TWIN_BATCH_MIN_ITEMS
    start = time.perf_counter()
    raw = certify_queries(queries, cfg.exact, WALK_CUDA_TWIN if twin else WALK_PYTHON_TWIN) if queries else []
    exact_ms = (time.perf_counter() - start) * 1000.0

    audit: list[AuditLogRow] = []
    event, ts = cfg.first_event_id, cfg.first_timestamp_us
    for c in cands.candidates:
        audit.append(_audit(event, c.query_id, c.candidate_id, 0, AuditStage.RT, RT_AUDIT_CANDIDATE, 0, ts,
                            float(c.rt_hit_count)))
        event, ts = event + 1, ts + 1
    certificates: list[CertificateResult] = []
    exact_by_runtime: dict[int, PureExactQueryResult] = {}
    for k, w in enumerate(items):
        parent = by_candidate[w.parent_candidate_id]
        audit.append(_audit(event, w.query_id, parent.candidate_id, w.work_item_id, AuditStage.EXACT,
                            EXACT_AUDIT_DEQUEUED, w.depth, ts))
        event, ts = event + 1, ts + 1
        if twin:
            res = _twin_query_result(queries[k], raw[k], cfg.exact)
            cert = validate_certificate_result(_twin_certificate(raw[k], w, cfg.exact))
        else:
            res = _python_query_result(queries[k], raw[k], cfg.exact)
            cert = validate_certificate_result(_python_certificate(res, w, cfg.exact))
        exact_by_runtime[w.query_id] = res
        certificates.append(cert)
        action = {CertificateStatus.COLLISION: EXACT_AUDIT_COLLISION,
                  CertificateStatus.SEPARATION: EXACT_AUDIT_SEPARATION}.get(cert.status, EXACT_AUDIT_UNDECIDED)
        aux0 = cert.safe_margin_lb if cert.status == CertificateStatus.SEPARATION else cert.toi_upper
        audit.append(_audit(event, w.query_id, parent.candidate_id, w.work_item_id, AuditStage.EXACT, action, w.depth,
                            ts, aux0, float(cert.next_refinement_mode)))
        event, ts = event + 1, ts + 1
    results = []
    for q in batch.queries:
        rid = cands.runtime_query_ids[q.query_id]
        if rid in exact_by_runtime:
            results.append(exact_by_runtime[rid])
        else:
            results.append(PureExactQueryResult(
                query_id=q.query_id, family="point_triangle" if str(q.family) == "vf" else "edge_edge",
                predicted_collision=False, ground_truth_collision=q.ground_truth_collides,
                status="rt_candidate_separation", toi_upper=1.0, safe_margin_lb=0.0, exact_evals=0, max_depth=0))
    validate_rt_exact_coverage(cands.candidates, items, certificates)
    return CertificateEngineWrapperResult(
        query_results=tuple(results), candidates=cands.candidates, work_items=items,
        certificates=tuple(certificates), audit_log=tuple(audit), exact_ms=exact_ms, source_name=cands.source_name,
        scene_name=cands.scene_name, batch_id=cands.batch_id, used_cpp_backend=True, fallback_reason="")
