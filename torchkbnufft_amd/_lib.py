"""ctypes binding of ``libtkbn_hip.so`` (the C ABI declared in ``include/tkbn_hip.h``).

The library is opened after ``import torch`` so its ``libamdhip64.so.7``
dependency resolves to the HIP runtime torch has already loaded.  There is no
fallback: if the library is missing every operator raises.
"""
from __future__ import annotations

import ctypes
import os
import threading
from typing import Optional

import torch

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libtkbn_hip.so")

TKBN_OK = 0
TKBN_ERR_INVALID = 1
TKBN_ERR_UNSUPPORTED = 2
TKBN_ERR_HIP = 3
TKBN_ERR_WORKSPACE = 4

TKBN_F32 = 0
TKBN_F64 = 1

TKBN_ADJ_ATOMIC = 0
TKBN_ADJ_DETERMINISTIC = 1

TKBN_PLAN_REUSE = 1

_c_i32 = ctypes.c_int32
_c_i64 = ctypes.c_int64
_c_dbl = ctypes.c_double
_c_vp = ctypes.c_void_p


class InterpDesc(ctypes.Structure):
    """Mirror of ``tkbn_interp_desc``."""

    _fields_ = [
        ("dtype", _c_i32),
        ("ndim", _c_i32),
        ("grid_size", _c_i64 * 3),
        ("numpoints", _c_i64 * 3),
        ("table_oversamp", _c_i64 * 3),
        ("n_shift", _c_dbl * 3),
        ("tables", _c_vp * 3),
        ("batch", _c_i64),
        ("coils", _c_i64),
        ("klength", _c_i64),
        ("omega_batched", _c_i32),
        ("flags", _c_i32),
        ("tile", _c_i32 * 3),
        ("work_target", _c_i32),
    ]


_SIGNATURES = {
    "tkbn_abi_version": (ctypes.c_int, []),
    "tkbn_last_error": (ctypes.c_char_p, []),
    "tkbn_interp_fwd_workspace_bytes": (_c_i64, [ctypes.POINTER(InterpDesc)]),
    "tkbn_interp_fwd": (ctypes.c_int, [ctypes.POINTER(InterpDesc), _c_vp, _c_vp, _c_vp, _c_vp, _c_i64, _c_vp]),
    "tkbn_interp_adj_workspace_bytes": (_c_i64, [ctypes.POINTER(InterpDesc), _c_i32]),
    "tkbn_interp_adj": (
        ctypes.c_int,
        [ctypes.POINTER(InterpDesc), _c_vp, _c_vp, _c_vp, _c_i32, _c_vp, _c_i64, _c_vp],
    ),
    "tkbn_interp_adj_noalloc_max_bytes": (_c_i64, []),
    "tkbn_interp_adj_noalloc": (
        ctypes.c_int,
        [ctypes.POINTER(InterpDesc), _c_vp, _c_vp, _c_vp, _c_i32, _c_vp],
    ),
    "tkbn_batch_create": (
        ctypes.c_int,
        [ctypes.POINTER(InterpDesc), ctypes.POINTER(InterpDesc), _c_i32, _c_i32, ctypes.POINTER(_c_vp),
         ctypes.POINTER(_c_vp),
         ctypes.POINTER(_c_vp), ctypes.POINTER(_c_vp), ctypes.POINTER(_c_vp), ctypes.POINTER(_c_vp),
         ctypes.POINTER(_c_vp), ctypes.POINTER(_c_vp)],
    ),
    "tkbn_batch_destroy": (ctypes.c_int, [_c_vp]),
    "tkbn_batch_table_bytes": (_c_i64, [_c_vp]),
    "tkbn_batch_write_table": (ctypes.c_int, [_c_vp, _c_vp]),
    "tkbn_batch_workspace_bytes": (_c_i64, [_c_vp]),
    "tkbn_batch_interp_fwd": (ctypes.c_int, [_c_vp, _c_vp, _c_vp, _c_i64, _c_vp]),
    "tkbn_batch_interp_adj": (ctypes.c_int, [_c_vp, _c_vp, _c_vp, _c_i64, _c_vp]),
    "tkbn_interp_domega": (ctypes.c_int, [ctypes.POINTER(InterpDesc), _c_vp, _c_vp, _c_vp, _c_vp, _c_vp]),
    "tkbn_profile_next_kernel": (ctypes.c_int, [_c_vp, _c_vp]),
    "tkbn_last_interp_kernel": (ctypes.c_char_p, []),
    "tkbn_apod_pad": (
        ctypes.c_int,
        [_c_i32, _c_i32, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_i64), _c_i64, _c_i64, _c_i64,
         _c_vp, _c_vp, _c_vp, _c_dbl, _c_vp, _c_vp],
    ),
    "tkbn_fft_supported": (ctypes.c_int, [_c_i32, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_i64)]),
    "tkbn_fft_workspace_bytes": (_c_i64, [_c_i32, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_i64), _c_i64, _c_i64]),
    "tkbn_fft_grid_fwd": (
        ctypes.c_int,
        [_c_i32, _c_i32, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_i64), _c_i64, _c_i64, _c_i64,
         _c_vp, _c_vp, _c_vp, _c_dbl, _c_vp, _c_vp, _c_i64, _c_vp],
    ),
    "tkbn_fft_grid_adj": (
        ctypes.c_int,
        [_c_i32, _c_i32, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_i64), _c_i64, _c_i64,
         _c_vp, _c_vp, _c_vp, _c_dbl, _c_vp, _c_vp, _c_i64, _c_vp],
    ),
    "tkbn_fft_filter": (
        ctypes.c_int,
        [_c_i32, _c_i32, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_i64), _c_i64, _c_i64, _c_i64,
         _c_vp, _c_vp, _c_vp, _c_i32, _c_dbl, _c_dbl, _c_vp, _c_vp, _c_i64, _c_vp],
    ),
    "tkbn_spmv_block_rows": (ctypes.c_int, []),
    "tkbn_spmv_workspace_bytes": (_c_i64, [_c_i32, _c_i64, _c_i64, _c_i64, _c_i64]),
    "tkbn_spmv": (
        ctypes.c_int,
        [_c_i32, _c_i64, _c_i64, _c_i64, _c_i64, _c_vp, _c_i64, _c_vp, _c_i64, _c_vp, _c_vp, _c_vp, _c_vp,
         _c_vp, _c_i64, _c_vp],
    ),
    "tkbn_crop_deapod": (
        ctypes.c_int,
        [_c_i32, _c_i32, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_i64), _c_i64, _c_i64,
         _c_vp, _c_vp, _c_vp, _c_dbl, _c_vp, _c_vp],
    ),
}

ABI_VERSION = 4

# bytes at the head of an interpolation workspace that hold the plan-cache header
PLAN_CACHE_HEADER_BYTES = 256

_lock = threading.Lock()
_lib: Optional[ctypes.CDLL] = None


class TkbnError(RuntimeError):
    """Error status returned by the HIP library."""


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"HIP extension {LIB_PATH} is not built; run `python -m torchkbnufft_amd._build` "
                "(torchkbnufft_amd has no CPU fallback)"
            )
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if lib.tkbn_abi_version() != ABI_VERSION:
            raise RuntimeError(
                f"libtkbn_hip ABI {lib.tkbn_abi_version()} != expected {ABI_VERSION}; rebuild"
            )
        _lib = lib
        return lib


def check(status: int, what: str) -> None:
    if status == TKBN_OK:
        return
    msg = load().tkbn_last_error().decode(errors="replace")
    if status in (TKBN_ERR_INVALID, TKBN_ERR_WORKSPACE):
        raise ValueError(f"{what}: {msg}")
    if status == TKBN_ERR_UNSUPPORTED:
        raise NotImplementedError(f"{what}: {msg}")
    raise TkbnError(f"{what}: status {status}: {msg}")


def i64_array(values) -> ctypes.Array:
    vals = [int(v) for v in values]
    return (_c_i64 * max(len(vals), 1))(*vals)


def stream_of(t: torch.Tensor) -> int:
    return torch.cuda.current_stream(t.device).cuda_stream


def dtype_code(t: torch.Tensor) -> int:
    if t.dtype in (torch.complex64, torch.float32):
        return TKBN_F32
    if t.dtype in (torch.complex128, torch.float64):
        return TKBN_F64
    raise TypeError(f"unsupported dtype {t.dtype}; expected complex64 or complex128")
