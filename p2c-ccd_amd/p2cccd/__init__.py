"""p2cccd -- MI355X (gfx950) proposal-to-certificate path.

Keeps the reference package's entry-point names for the certificate path
(src/python/p2cccd/__init__.py:1-143): the runtime contracts and their
validators, external-batch candidate generation and the certificate engine
over it, and the CUDA-binding probes.  The work runs in libp2c_hip.so,
through the p2cccd_cpp module and the device-pointer C-ABI.  The reference's
benchmark-report rows, run configuration files, serialization helpers and
internal sample generators are not part of the path and are not kept.
"""
from .candidate_generation import (
    CandidateGenerationWrapperConfig,
    CandidateGenerationWrapperResult,
    generate_candidates_for_external_batch,
    is_cpp_candidate_generation_available,
)
from .certificate_engine import (
    CertificateEngineWrapperConfig,
    CertificateEngineWrapperResult,
    execute_certificate_engine_for_external_batch,
    is_cpp_certificate_engine_available,
)
from .contracts import (
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
    schema_field_names,
)
from .cuda_bindings import (
    CUDA_BINDING_ENTRYPOINTS,
    CudaBindingStatus,
    cross_check_cpu_cuda_exact,
    evaluate_edge_edge_batch_cuda,
    evaluate_point_triangle_batch_cuda,
    get_cuda_binding_status,
    is_cuda_exact_binding_available,
)
from .validators import (
    ValidationError,
    validate_audit_log_row,
    validate_candidate_record,
    validate_certificate_result,
    validate_contract,
    validate_dict_schema,
    validate_exact_work_item,
    validate_proposal_output,
)

__all__ = [
    "AuditLogRow",
    "AuditStage",
    "CONTRACT_SCHEMAS",
    "CUDA_BINDING_ENTRYPOINTS",
    "CandidateGenerationWrapperConfig",
    "CandidateGenerationWrapperResult",
    "CandidateRecord",
    "CertificateEngineWrapperConfig",
    "CertificateEngineWrapperResult",
    "CertificateRefinementMode",
    "CertificateResult",
    "CertificateStatus",
    "CudaBindingStatus",
    "ExactWorkItem",
    "ProposalOutput",
    "ProposalSource",
    "ProxyType",
    "ValidationError",
    "cross_check_cpu_cuda_exact",
    "evaluate_edge_edge_batch_cuda",
    "evaluate_point_triangle_batch_cuda",
    "execute_certificate_engine_for_external_batch",
    "generate_candidates_for_external_batch",
    "get_cuda_binding_status",
    "is_cpp_candidate_generation_available",
    "is_cpp_certificate_engine_available",
    "is_cuda_exact_binding_available",
    "schema_field_names",
    "validate_audit_log_row",
    "validate_candidate_record",
    "validate_certificate_result",
    "validate_contract",
    "validate_dict_schema",
    "validate_exact_work_item",
    "validate_proposal_output",
]
