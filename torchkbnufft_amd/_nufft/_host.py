"""Host helpers: small-buffer caching (no per-call device syncs) and argument checks."""
from __future__ import annotations

import weakref
from typing import Dict, Tuple

import torch
from torch import Tensor

# (id, version, data_ptr) -> python values of a small plan buffer (n_shift, numpoints, ...)
_HOST_CACHE: Dict[Tuple[int, int, int], tuple] = {}


def host_values(t) -> tuple:
    """Python values of a small plan tensor, read from the device once per version.

    Plan buffers (``n_shift``, ``numpoints``, ``table_oversamp``, ``grid_size``,
    ``im_size``) live on the device next to the data; reading them on every call
    would stall the stream, so their values are cached per tensor version.
    """
    if not isinstance(t, Tensor):
        return tuple(t)
    if t.device.type == "cpu":
        return tuple(t.reshape(-1).tolist())
    key = (id(t), t._version, t.data_ptr())
    hit = _HOST_CACHE.get(key)
    if hit is not None:
        return hit
    vals = tuple(t.detach().reshape(-1).cpu().tolist())
    _HOST_CACHE[key] = vals
    try:
        weakref.finalize(t, _HOST_CACHE.pop, key, None)
    except TypeError:
        pass
    return vals


def require_gpu(*tensors: Tensor) -> torch.device:
    dev = None
    for t in tensors:
        if t is None:
            continue
        if t.device.type != "cuda":
            raise RuntimeError(
                "torchkbnufft_amd runs on the GPU (HIP) only; got a tensor on "
                f"{t.device}. Move data, trajectory and module to a cuda device."
            )
        if dev is None:
            dev = t.device
        elif t.device != dev:
            raise RuntimeError(f"tensors on different devices: {dev} and {t.device}")
    return dev


def real_dtype_of(t: Tensor) -> torch.dtype:
    return {torch.complex64: torch.float32, torch.complex128: torch.float64}[t.dtype]


def as_complex_input(t: Tensor, name: str) -> Tensor:
    if t.dtype not in (torch.complex64, torch.complex128):
        raise TypeError(f"{name} must be complex64 or complex128, got {t.dtype}")
    return t.contiguous()
