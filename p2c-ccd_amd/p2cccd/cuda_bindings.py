"""Python wrappers of the native batch certificate entry points.

Same public names, arguments and error behaviour as the reference's
src/python/p2cccd/cuda_bindings.py:8-156: the native module is found by name
(``p2cccd_cpp``) and attribute probing; a missing module or entry point
raises RuntimeError carrying the fallback reason.  The module found here is
the gfx950 one, so ``backend_name`` reports the HIP backend.
"""
from __future__ import annotations

import importlib
from dataclasses import dataclass
from typing import Any, Sequence

CUDA_BINDING_ENTRYPOINTS = (
    "is_cuda_exact_built",
    "cuda_binding_status",
    "evaluate_point_triangle_batch_cuda",
    "evaluate_edge_edge_batch_cuda",
    "cross_check_cpu_cuda_exact",
)


@dataclass(frozen=True, slots=True)
class CudaBindingStatus:
    cpp_module_available: bool
    cuda_exact_built: bool
    host_batch_exact_api: bool
    device_pointer_abi: bool
    backend_name: str
    fallback_reason: str

    @property
    def ready_for_host_batches(self) -> bool:
        return self.cpp_module_available and self.host_batch_exact_api

    @property
    def ready_for_cuda_execution(self) -> bool:
        return self.ready_for_host_batches and self.cuda_exact_built


def _load_cpp_module() -> Any | None:
    try:
        return importlib.import_module("p2cccd_cpp")
    except ImportError:
        return None


def _has_entrypoints(module: Any) -> bool:
    return all(callable(getattr(module, name, None)) for name in CUDA_BINDING_ENTRYPOINTS)


def get_cuda_binding_status(cpp_module: Any | None = None) -> CudaBindingStatus:
    module = cpp_module if cpp_module is not None else _load_cpp_module()
    if module is None:
        return CudaBindingStatus(False, False, False, False, "unavailable", "p2cccd_cpp is not importable")
    if not _has_entrypoints(module):
        return CudaBindingStatus(True, False, False, False, "missing_entrypoints",
                                 "p2cccd_cpp is missing CUDA-aware binding entrypoints")
    raw = dict(module.cuda_binding_status())
    built = bool(raw.get("cuda_exact_built", module.is_cuda_exact_built()))
    return CudaBindingStatus(
        cpp_module_available=True,
        cuda_exact_built=built,
        host_batch_exact_api=bool(raw.get("host_batch_exact_api", True)),
        device_pointer_abi=bool(raw.get("device_pointer_abi", False)),
        backend_name=str(raw.get("backend_name", "cuda_exact" if built else "cuda_exact_stub")),
        fallback_reason="" if built else "CUDA exact backend is not built; use CPU exact fallback",
    )


def is_cuda_exact_binding_available(*, require_built: bool = True) -> bool:
    status = get_cuda_binding_status()
    return status.ready_for_cuda_execution if require_built else status.ready_for_host_batches


def _require_module(cpp_module: Any | None, *, require_built: bool) -> Any:
    module = cpp_module if cpp_module is not None else _load_cpp_module()
    status = get_cuda_binding_status(module)
    if not status.ready_for_host_batches or (require_built and not status.cuda_exact_built):
        raise RuntimeError(status.fallback_reason)
    return module


def evaluate_point_triangle_batch_cuda(primitives: Sequence[Any], interval_t0: float, interval_t1: float,
                                       config: Any, *, cpp_module: Any | None = None,
                                       require_built: bool = True) -> tuple[Any, ...]:
    module = _require_module(cpp_module, require_built=require_built)
    return tuple(module.evaluate_point_triangle_batch_cuda(list(primitives), float(interval_t0),
                                                            float(interval_t1), config))


def evaluate_edge_edge_batch_cuda(primitives: Sequence[Any], interval_t0: float, interval_t1: float,
                                  config: Any, *, cpp_module: Any | None = None,
                                  require_built: bool = True) -> tuple[Any, ...]:
    module = _require_module(cpp_module, require_built=require_built)
    return tuple(module.evaluate_edge_edge_batch_cuda(list(primitives), float(interval_t0),
                                                       float(interval_t1), config))


def cross_check_cpu_cuda_exact(point_triangles: Sequence[Any], edge_edges: Sequence[Any], interval_t0: float,
                               interval_t1: float, config: Any, eps_cert: float, *,
                               cpp_module: Any | None = None, require_built: bool = True) -> bool:
    module = _require_module(cpp_module, require_built=require_built)
    return bool(module.cross_check_cpu_cuda_exact(list(point_triangles), list(edge_edges), float(interval_t0),
                                                  float(interval_t1), config, float(eps_cert)))
