"""Runtime contracts and validators of the p2cccd package against the
reference's own expectations (src/tests/python/test_contracts.py:100-230):
field order and wire types equal the C++ structs the native layer mirrors
(bindings/certificate_hip.h, bindings/proxy_hip.h, bindings/policy_hip.h),
and every accepted and refused record of those tests is accepted and refused
here; plus the external-query batch rules (datasets/ccd/contracts.py:69-151)
and the package's exported names (src/python/p2cccd/__init__.py:1-143)."""
import re
from dataclasses import replace
from pathlib import Path

import pytest

import p2cccd
from p2cccd import (
    CONTRACT_SCHEMAS,
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
    ValidationError,
    schema_field_names,
    validate_audit_log_row,
    validate_candidate_record,
    validate_certificate_result,
    validate_contract,
    validate_dict_schema,
    validate_exact_work_item,
    validate_proposal_output,
)
from p2cccd.datasets import CCD_ADAPTER_SCHEMA_VERSION, CCDQueryFamily, DatasetQueryBatch, ExternalCCDQuery

BIND = Path(p2cccd.__file__).resolve().parent.parent / "bindings"
TYPES = {"std::uint32_t": "uint32", "std::uint64_t": "uint64", "std::uint16_t": "uint16", "std::uint8_t": "uint8",
         "std::int64_t": "int64", "double": "float64", "float": "float32", "std::array<float, 4>": "float32[4]",
         "std::array<float, 8>": "float32[8]", "std::array<float, kProposalIntervalBinCount>": "float32[8]",
         "std::array<float, kProposalFamilyCount>": "float32[8]", "ProxyType": "ProxyType:uint8",
         "ProposalSource": "ProposalSource:uint8", "CertificateStatus": "CertificateStatus:uint8",
         "CertificateRefinementMode": "CertificateRefinementMode:uint8", "AuditStage": "AuditStage:uint8"}


def _cpp_schema(name):
    for header in ("certificate_hip.h", "proxy_hip.h", "policy_hip.h"):
        m = re.search(rf"struct\s+(?:alignas\(\d+\)\s+)?{name}\s*\{{(?P<body>.*?)\n\}};", (BIND / header).read_text(),
                      re.S)
        if m:
            out = []
            for line in m.group("body").splitlines():
                d = line.strip().rstrip(";").split("=", 1)[0].split("{", 1)[0].strip()
                if d and not d.startswith("//"):
                    t, n = re.match(r"(.+?)\s+([A-Za-z_]\w*)$", d).groups()
                    out.append((n, TYPES[re.sub(r"\s+", " ", t)]))
            return tuple(out)
    raise KeyError(name)


@pytest.mark.parametrize("name", ["CandidateRecord", "ProposalOutput", "ExactWorkItem", "CertificateResult",
                                  "AuditLogRow"])
def test_python_contracts_match_the_native_structs(name):
    assert CONTRACT_SCHEMAS[name] == _cpp_schema(name)
    assert schema_field_names(name) == tuple(n for n, _ in _cpp_schema(name))


def test_candidate_record_rules():
    c = CandidateRecord(candidate_id=1, query_id=10, proxy_type_a=ProxyType.SWEPT_AABB,
                        proxy_type_b=ProxyType.CAPSULE, rt_hit_count=1, motion_bound=[0.1, 0.2, 0.3, 0.4])
    assert validate_candidate_record(c) is c and validate_contract(c) is c
    payload = {f: getattr(c, f) for f in schema_field_names(CandidateRecord)}
    validate_dict_schema(payload, CandidateRecord)
    missing = dict(payload)
    missing.pop("query_id")
    with pytest.raises(ValidationError, match="missing required field"):
        validate_dict_schema(missing, CandidateRecord)
    with pytest.raises(ValidationError, match="unknown contract type"):
        validate_dict_schema(payload, "NotAContract")
    for bad, msg in ((replace(c, proxy_type_a=ProxyType.UNKNOWN), "concrete enum value"),
                     (replace(c, proxy_type_a=True), "not bool"),
                     (replace(c, rt_hit_count=0), "rt_hit_count is required"),
                     (replace(c, candidate_id=2**64), "must be <= "),
                     (replace(c, slab_id=2**32), "must be <= "),
                     (replace(c, motion_bound=[0.0, -1.0, 0.0, 0.0]), r"motion_bound\[1\] must be non-negative"),
                     (replace(c, schema_version=2), "schema_version is unsupported")):
        with pytest.raises(ValidationError, match=msg):
            validate_candidate_record(bad)


def test_proposal_work_item_certificate_and_audit_rules():
    p = ProposalOutput(candidate_id=1, interval_scores=[0.1] * 8, family_scores=[0.2] * 8, priority_score=0.5,
                       cost_score=1.0, uncertainty_score=0.05)
    assert validate_proposal_output(p) is p
    with pytest.raises(ValidationError, match="cost_score must be non-negative"):
        validate_proposal_output(replace(p, cost_score=-1.0))
    w = ExactWorkItem(work_item_id=2, parent_candidate_id=1, query_id=10, interval_t0=0.25, interval_t1=0.5,
                      feature_family_mask=1, priority_score=0.4, source=ProposalSource.RAW)
    assert validate_exact_work_item(w) is w
    for bad in (replace(w, interval_t0=0.75, interval_t1=0.5), replace(w, source=True), replace(w, depth=2**16)):
        with pytest.raises(ValidationError):
            validate_exact_work_item(bad)
    r = CertificateResult(work_item_id=2, query_id=10, status=CertificateStatus.SEPARATION, interval_t0=0.25,
                          interval_t1=0.5, toi_upper=0.5, safe_margin_lb=1.0e-3, covered_feature_mask=1, eps_time=1e-4,
                          eps_space=1e-6, next_refinement_mode=CertificateRefinementMode.NONE)
    assert validate_certificate_result(r) is r
    bad_results = (
        (replace(r, witness_id_a=-(2**63) - 1), "must be >="),
        (replace(r, status=CertificateStatus.UNDECIDED, reason_code=0,
                 next_refinement_mode=CertificateRefinementMode.BISECT_INTERVAL), "reason_code is required"),
        (replace(r, witness_family=2**8), "must be <= 255"),
        (replace(r, status=CertificateStatus.UNDECIDED, reason_code=1,
                 next_refinement_mode=CertificateRefinementMode.NONE), "next_refinement_mode is required"),
        (replace(r, covered_feature_mask=0), "covered_feature_mask is required for separation"),
        (replace(r, status=CertificateStatus.COLLISION, toi_upper=0.75, witness_id_a=1, witness_id_b=2,
                 covered_feature_mask=0), "toi_upper must lie inside"),
        (replace(r, status=CertificateStatus.COLLISION, toi_upper=0.3), "collision witnesses are required"))
    for bad, msg in bad_results:
        with pytest.raises(ValidationError, match=msg):
            validate_certificate_result(bad)
    a = AuditLogRow(event_id=1, query_id=10, candidate_id=1, work_item_id=2, stage=AuditStage.RT, interval_t0=0.0,
                    interval_t1=1.0, timestamp_us=123)
    assert validate_audit_log_row(a) is a
    with pytest.raises(ValidationError, match="action must be <= 65535"):
        validate_audit_log_row(replace(a, action=2**16))
    with pytest.raises(ValidationError, match="timestamp_us is required"):
        validate_audit_log_row(replace(a, timestamp_us=0))


def _q(qid, family=CCDQueryFamily.VERTEX_FACE, **kw):
    v = ((0.25, 0.25, 1.0), (0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0))
    base = dict(schema_version=CCD_ADAPTER_SCHEMA_VERSION, source_name="unit", scene_name="pt", batch_id="b",
                query_id=qid, source_query_index=qid, family=family, vertices_t0=v, vertices_t1=v)
    base.update(kw)
    return ExternalCCDQuery(**base)


def test_external_query_batch_rules():
    b = DatasetQueryBatch(schema_version=1, source_name="unit", scene_name="pt", batch_id="b",
                          family=CCDQueryFamily.VERTEX_FACE, queries=(_q(0), _q(1)))
    assert b.query_count == 2 and str(b.queries[0].family) == "vf"
    assert CCDQueryFamily.EDGE_EDGE.p2cccd_witness_family == "edge_edge"
    with pytest.raises(ValueError, match="query_id values must be unique"):
        DatasetQueryBatch(schema_version=1, source_name="unit", scene_name="pt", batch_id="b",
                          family=CCDQueryFamily.VERTEX_FACE, queries=(_q(0), _q(0)))
    with pytest.raises(ValueError, match="query family must match batch family"):
        DatasetQueryBatch(schema_version=1, source_name="unit", scene_name="pt", batch_id="b",
                          family=CCDQueryFamily.EDGE_EDGE, queries=(_q(0),))
    with pytest.raises(ValueError, match="four vertices"):
        _q(0, vertices_t1=((0, 0, 0),))
    with pytest.raises(ValueError, match="query_id must be non-negative"):
        _q(-1)


def test_package_exports_the_reference_entry_points():
    for name in ("generate_candidates_for_external_batch", "execute_certificate_engine_for_external_batch",
                 "evaluate_point_triangle_batch_cuda", "evaluate_edge_edge_batch_cuda", "cross_check_cpu_cuda_exact",
                 "get_cuda_binding_status", "is_cuda_exact_binding_available", "CertificateResult", "ExactWorkItem",
                 "CandidateRecord", "AuditLogRow", "ProposalOutput", "validate_candidate_record",
                 "validate_certificate_result", "validate_exact_work_item", "validate_audit_log_row",
                 "validate_proposal_output", "validate_contract", "validate_dict_schema", "ValidationError",
                 "CONTRACT_SCHEMAS", "schema_field_names", "CandidateGenerationWrapperConfig",
                 "CertificateEngineWrapperConfig", "is_cpp_candidate_generation_available",
                 "is_cpp_certificate_engine_available"):
        assert hasattr(p2cccd, name), name
    assert p2cccd.is_cpp_candidate_generation_available() and p2cccd.is_cpp_certificate_engine_available()
    import p2cccd_cpp
    for name in ("generate_candidates_for_external_batch", "build_proxy_scene", "schedule_runtime_exact_work_items",
                 "schedule_runtime_exact_work_items_from_arrays",
                 "schedule_runtime_exact_work_items_from_proposal_arrays", "run_native_dense_group_exact_early_stop",
                 "generate_candidates_for_proxy_scene", "ExternalBatchCandidateResult", "RuntimeQueryIdMapping",
                 "ProposalFeatureRow", "ProposalSchedulingConfig", "ProxySceneBuildInput"):
        assert hasattr(p2cccd_cpp, name), name
