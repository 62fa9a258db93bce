"""ctypes view of libp2c_hip.so (include/p2c_hip.h) and the numpy record
layouts that cross the C-ABI.

The library is required: importing this module on a machine where the
library was not built raises immediately, and every call checks the C return
code and raises with the library's own message.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

try:  # one HIP runtime per process: bind the library to the one torch loads
    import torch  # noqa: F401
except ImportError:
    pass

_PKG_ROOT = Path(__file__).resolve().parent.parent
LIB_PATH = _PKG_ROOT / "lib" / "libp2c_hip.so"

# ---- record layouts (must match include/p2c_hip.h) -------------------------
TRAJ_DTYPE = np.dtype([("feature_id", "<i8"), ("position_t0", "<f8", (3,)), ("position_t1", "<f8", (3,))])
PRIM_DTYPE = np.dtype([("id_a", "<i8"), ("id_b", "<i8"), ("v", TRAJ_DTYPE, (4,))])
CERT_CONFIG_DTYPE = np.dtype(
    [("eps_time", "<f8"), ("eps_space", "<f8"), ("max_subdivision_depth", "<u4"), ("flags", "<u4")]
)
PRIM_RESULT_DTYPE = np.dtype(
    [
        ("status", "u1"),
        ("witness_family", "u1"),
        ("reason_code", "<u2"),
        ("covered_feature_mask", "<u4"),
        ("interval_t0", "<f8"),
        ("interval_t1", "<f8"),
        ("toi_upper", "<f8"),
        ("safe_margin_lb", "<f8"),
        ("witness_id_a", "<i8"),
        ("witness_id_b", "<i8"),
        ("next_refinement_mode", "u1"),
        ("pad", "u1", (3,)),
        ("nodes", "<u4"),
    ]
)
assert PRIM_DTYPE.itemsize == 240 and PRIM_RESULT_DTYPE.itemsize == 64

P2C_OK = 0
FAMILY_POINT_TRIANGLE = 1
FAMILY_EDGE_EDGE = 2
STATUS_COLLISION, STATUS_SEPARATION, STATUS_UNDECIDED, STATUS_INVALID = 0, 1, 2, 255


class NativeError(RuntimeError):
    """A libp2c_hip entry point returned a non-zero code."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


def _load() -> ctypes.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C p2c-ccd_amd). "
            "The HIP library is required: there is no CPU fallback."
        )
    return ctypes.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | ctypes.RTLD_GLOBAL)


lib = _load()
_vp, _i64, _i32, _u32, _f64 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_uint32, ctypes.c_double
lib.p2c_last_error.restype = ctypes.c_char_p
lib.p2c_version.restype = _i32
lib.p2c_device_available.restype = _i32
lib.p2c_stream_synchronize.argtypes = [_vp]


def _sig(name, args):
    fn = getattr(lib, name)
    fn.argtypes = args
    fn.restype = _i32
    return fn


_sig("p2c_cert_ref_batch", [_i32, _vp, _i64, _f64, _f64, _vp, _vp, _vp, _vp])
_sig("p2c_cert_ref_distance", [_i32, _vp, _vp, _i64, _vp, _vp])
_sig("p2c_cert_ref_batch_host", [_i32, _vp, _i64, _f64, _f64, _vp, _vp])
_sig("p2c_cert_ref_eval_cpu", [_i32, _vp, _f64, _f64, _vp, _vp])


def check(code: int) -> None:
    if code != P2C_OK:
        raise NativeError(code, lib.p2c_last_error().decode("utf-8", "replace"))


def call(name: str, *args) -> None:
    check(getattr(lib, name)(*args))


def ptr(x) -> int:
    """Address of a numpy array's data or a torch tensor's storage."""
    if isinstance(x, np.ndarray):
        return x.ctypes.data
    if x is None:
        return 0
    return int(x.data_ptr())


def stream_handle(stream=None) -> int:
    """hipStream_t of a torch stream (the current stream when None)."""
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def device_available() -> bool:
    return bool(lib.p2c_device_available())


def cert_config(eps_time=1.0e-4, eps_space=1.0e-6, max_subdivision_depth=32, cuda_twin=False) -> np.ndarray:
    """p2c_cert_config; cuda_twin selects the CUDA batch twin's walk
    (certificate_cuda.cu:286-346) over the CPU engine's."""
    c = np.zeros(1, CERT_CONFIG_DTYPE)
    c["eps_time"], c["eps_space"], c["max_subdivision_depth"] = eps_time, eps_space, max_subdivision_depth
    c["flags"] = 1 if cuda_twin else 0
    return c


def prims_from_positions(x: np.ndarray, ids: np.ndarray | None = None) -> np.ndarray:
    """Pack (n, 8, 3) float64 positions (4 vertices at t0, then the same 4 at
    t1) into PRIM records; ids (n, 2) default to (2i, 2i+1)."""
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[0]
    out = np.zeros(n, PRIM_DTYPE)
    if ids is None:
        ids = np.stack([np.arange(n) * 2, np.arange(n) * 2 + 1], axis=1)
    out["id_a"], out["id_b"] = ids[:, 0], ids[:, 1]
    for k in range(4):
        out["v"][:, k]["feature_id"] = np.arange(n) * 4 + k
        out["v"][:, k]["position_t0"] = x[:, k]
        out["v"][:, k]["position_t1"] = x[:, 4 + k]
    return out


def cert_ref_batch_host(family: int, prims: np.ndarray, t0: float, t1: float, cfg: np.ndarray) -> np.ndarray:
    prims = np.ascontiguousarray(prims, dtype=PRIM_DTYPE)
    out = np.zeros(prims.shape[0], PRIM_RESULT_DTYPE)
    call("p2c_cert_ref_batch_host", family, ptr(prims), prims.shape[0], t0, t1, ptr(cfg), ptr(out))
    return out


def cert_ref_eval_cpu(family: int, prims: np.ndarray, t0: float, t1: float, cfg: np.ndarray) -> np.ndarray:
    prims = np.ascontiguousarray(prims, dtype=PRIM_DTYPE)
    out = np.zeros(prims.shape[0], PRIM_RESULT_DTYPE)
    for i in range(prims.shape[0]):
        call("p2c_cert_ref_eval_cpu", family, ptr(prims[i:i + 1]), t0, t1, ptr(cfg), ptr(out[i:i + 1]))
    return out
