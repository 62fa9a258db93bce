"""p2cccd -- MI355X (gfx950) proposal-to-certificate path.

Keeps the reference package's entry-point names (src/python/p2cccd/__init__.py:
27-50) for the certificate path; the work runs in libp2c_hip.so through the
p2cccd_cpp module and the device-pointer C-ABI.
"""
from .cuda_bindings import (
    CUDA_BINDING_ENTRYPOINTS,
    CudaBindingStatus,
    cross_check_cpu_cuda_exact,
    evaluate_edge_edge_batch_cuda,
    evaluate_point_triangle_batch_cuda,
    get_cuda_binding_status,
    is_cuda_exact_binding_available,
)

__all__ = [
    "CUDA_BINDING_ENTRYPOINTS",
    "CudaBindingStatus",
    "cross_check_cpu_cuda_exact",
    "evaluate_edge_edge_batch_cuda",
    "evaluate_point_triangle_batch_cuda",
    "get_cuda_binding_status",
    "is_cuda_exact_binding_available",
]
