"""Oversampled FFT stages: fused HIP scale/pad and crop/deapodise around torch.fft.

``fft_and_scale`` / ``ifft_and_scale`` keep the signatures of
``torchkbnufft/_nufft/fft.py:36-118``.  The element-wise work (scaling, zero-pad,
crop, conjugate scaling, the SENSE coil expansion and coil combine of
``modules/kbnufft.py:182-183,404-405``) is one HIP pass on each side; the FFT is
``torch.fft`` (rocFFT) called unnormalised, with the ``"ortho"`` factor folded into
the fused pass.
"""
from __future__ import annotations

import ctypes
import math
from typing import Optional

import torch
from torch import Tensor

from .. import _lib
from .._autograd._apply import run as _run
from ._host import host_values, require_gpu


def _norm_scale(norm: Optional[str], grid_size) -> float:
    if norm is None:
        return 1.0
    if norm == "ortho":
        return 1.0 / math.sqrt(float(math.prod(int(v) for v in grid_size)))
    raise ValueError("Only option for norm is 'ortho'.")


def apod_pad(image: Tensor, smaps: Optional[Tensor], scaling_coef: Tensor, im_size, grid_size, scale: float) -> Tensor:
    """grid[b, c] = pad_end(image[b, c or 0] * smaps[b, c] * scaling_coef) * scale."""
    require_gpu(image, smaps, scaling_coef)
    image = image.contiguous()
    im = tuple(int(v) for v in im_size)
    gs = tuple(int(v) for v in grid_size)
    d = len(im)
    B = image.shape[0]
    C = smaps.shape[1] if smaps is not None else image.shape[1]
    if tuple(image.shape[2:]) != im:
        raise ValueError(f"image spatial shape {tuple(image.shape[2:])} != im_size {im}")
    if smaps is not None:
        smaps = smaps.to(image.dtype).contiguous()
        if tuple(smaps.shape) != (B, C) + im and not (smaps.shape[0] == 1 and tuple(smaps.shape[1:]) == (C,) + im):
            raise ValueError(f"smaps shape {tuple(smaps.shape)} does not match image {tuple(image.shape)}")
        if smaps.shape[0] != B:
            smaps = smaps.expand(B, *smaps.shape[1:]).contiguous()
        if image.shape[1] not in (1, C):
            raise ValueError("image coil dimension must be 1 or match smaps")
    sc = scaling_coef.to(device=image.device, dtype=image.dtype).contiguous()
    if sc.numel() != math.prod(im):
        raise ValueError("scaling_coef does not match im_size")
    grid = torch.empty((B, C) + gs, dtype=image.dtype, device=image.device)
    with torch.cuda.device(image.device):
        st = _lib.load().tkbn_apod_pad(
            _lib.dtype_code(image),
            d,
            _lib.i64_array(im),
            _lib.i64_array(gs),
            B,
            C,
            image.shape[1],
            image.data_ptr(),
            smaps.data_ptr() if smaps is not None else None,
            sc.data_ptr(),
            float(scale),
            grid.data_ptr(),
            _lib.stream_of(image),
        )
    _lib.check(st, "tkbn_apod_pad")
    return grid


def crop_deapod(grid: Tensor, smaps: Optional[Tensor], scaling_coef: Tensor, im_size, scale: float) -> Tensor:
    """image[b, c] = crop(grid[b, c]) * conj(scaling_coef) * scale; with smaps summed
    over coils against conj(smaps) into image[b, 0]."""
    require_gpu(grid, smaps, scaling_coef)
    grid = grid.contiguous()
    im = tuple(int(v) for v in im_size)
    d = len(im)
    gs = tuple(grid.shape[2:])
    B, C = grid.shape[0], grid.shape[1]
    if smaps is not None:
        smaps = smaps.to(grid.dtype).contiguous()
        if smaps.shape[0] != B:
            smaps = smaps.expand(B, *smaps.shape[1:]).contiguous()
        if tuple(smaps.shape) != (B, C) + im:
            raise ValueError(f"smaps shape {tuple(smaps.shape)} does not match data ({B}, {C}) + {im}")
    sc = scaling_coef.to(device=grid.device, dtype=grid.dtype).contiguous()
    out = torch.empty((B, 1 if smaps is not None else C) + im, dtype=grid.dtype, device=grid.device)
    with torch.cuda.device(grid.device):
        st = _lib.load().tkbn_crop_deapod(
            _lib.dtype_code(grid),
            d,
            _lib.i64_array(im),
            _lib.i64_array(gs),
            B,
            C,
            grid.data_ptr(),
            smaps.data_ptr() if smaps is not None else None,
            sc.data_ptr(),
            float(scale),
            out.data_ptr(),
            _lib.stream_of(grid),
        )
    _lib.check(st, "tkbn_crop_deapod")
    return out


def fft_and_scale(image: Tensor, scaling_coef: Tensor, im_size: Tensor, grid_size: Tensor, norm: Optional[str] = None) -> Tensor:
    """Oversampled FFT of ``image * scaling_coef`` (``fft.py:36-76``)."""
    from .._autograd.fft import ApodPad, GridForward

    gs = host_values(grid_size)
    im = host_values(im_size)
    scale = _norm_scale(norm, gs)
    if fused_supported(image.dtype, im, gs):
        return _run(GridForward, image, None, scaling_coef, im, gs, scale)
    grid = _run(ApodPad, image, None, scaling_coef, im, gs, scale)
    return torch.fft.fftn(grid, dim=tuple(range(-len(gs), 0)), norm="backward")


def ifft_and_scale(image: Tensor, scaling_coef: Tensor, im_size: Tensor, grid_size: Tensor, norm: Optional[str] = None) -> Tensor:
    """Unnormalised (or ortho) inverse FFT, crop, conjugate scaling (``fft.py:80-118``)."""
    from .._autograd.fft import CropDeapod, GridAdjoint

    gs = host_values(grid_size)
    im = host_values(im_size)
    scale = _norm_scale(norm, gs)
    if fused_supported(image.dtype, im, gs):
        return _run(GridAdjoint, image, None, scaling_coef, im, scale)
    x = torch.fft.ifftn(image, dim=tuple(range(-len(gs), 0)), norm="forward")
    return _run(CropDeapod, x, None, scaling_coef, im, scale)


# ---------------------------------------------------------------------------
# fused FFT path (complex64, grid sizes with prime factors 2, 3, 5)
# ---------------------------------------------------------------------------
_FUSED_OK = {}
_FUSED_DISABLED = False


def set_fused_fft(enabled: bool) -> None:
    """Use the fused HIP FFT where the geometry allows (default) or always torch.fft."""
    global _FUSED_DISABLED
    _FUSED_DISABLED = not enabled


def fused_supported(dtype: torch.dtype, im_size, grid_size) -> bool:
    if _FUSED_DISABLED or dtype != torch.complex64:
        return False
    key = (tuple(int(v) for v in im_size), tuple(int(v) for v in grid_size))
    ok = _FUSED_OK.get(key)
    if ok is None:
        d = len(key[0])
        ok = _lib.load().tkbn_fft_supported(d, _lib.i64_array(key[0]), _lib.i64_array(key[1])) == _lib.TKBN_OK
        _FUSED_OK[key] = ok
    return ok


def _out_or_new(out: Optional[Tensor], shape: tuple, like: Tensor) -> Tensor:
    if out is None:
        return torch.empty(shape, dtype=like.dtype, device=like.device)
    if tuple(out.shape) != tuple(shape) or out.dtype != like.dtype or out.device != like.device or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous {tuple(shape)} {like.dtype} tensor on {like.device}")
    return out


def _fft_workspace(dev, d, im, gs, B, C):
    n = _lib.load().tkbn_fft_workspace_bytes(d, _lib.i64_array(im), _lib.i64_array(gs), B, C)
    if n < 0:
        _lib.check(_lib.TKBN_ERR_INVALID, "tkbn_fft_workspace_bytes")
    return torch.empty(max(int(n), 1), dtype=torch.uint8, device=dev), int(n)


def fft_grid_fwd(image: Tensor, smaps: Optional[Tensor], scaling_coef: Tensor, im_size, grid_size, scale: float,
                 out: Optional[Tensor] = None) -> Tensor:
    """FFT(pad_end(image * smaps * scaling_coef)) * scale in one fused HIP pipeline."""
    require_gpu(image, smaps, scaling_coef)
    image = image.contiguous()
    im = tuple(int(v) for v in im_size)
    gs = tuple(int(v) for v in grid_size)
    d = len(im)
    B = image.shape[0]
    C = smaps.shape[1] if smaps is not None else image.shape[1]
    if tuple(image.shape[2:]) != im:
        raise ValueError(f"image spatial shape {tuple(image.shape[2:])} != im_size {im}")
    if smaps is not None:
        smaps = smaps.to(image.dtype)
        if smaps.shape[0] != B:
            smaps = smaps.expand(B, *smaps.shape[1:])
        smaps = smaps.contiguous()
        if tuple(smaps.shape) != (B, C) + im:
            raise ValueError(f"smaps shape {tuple(smaps.shape)} does not match image {tuple(image.shape)}")
        if image.shape[1] not in (1, C):
            raise ValueError("image coil dimension must be 1 or match smaps")
    sc = scaling_coef.to(device=image.device, dtype=image.dtype).contiguous()
    grid = _out_or_new(out, (B, C) + gs, image)
    ws, nws = _fft_workspace(image.device, d, im, gs, B, C)
    with torch.cuda.device(image.device):
        st = _lib.load().tkbn_fft_grid_fwd(
            _lib.dtype_code(image), d, _lib.i64_array(im), _lib.i64_array(gs), B, C, image.shape[1],
            image.data_ptr(), smaps.data_ptr() if smaps is not None else None, sc.data_ptr(), float(scale),
            grid.data_ptr(), ws.data_ptr(), nws, _lib.stream_of(image),
        )
    _lib.check(st, "tkbn_fft_grid_fwd")
    return grid


def fft_grid_adj(grid: Tensor, smaps: Optional[Tensor], scaling_coef: Tensor, im_size, scale: float,
                 out: Optional[Tensor] = None) -> Tensor:
    """sum_c crop(IFFT(grid)) * conj(scaling_coef) * conj(smaps_c) * scale (per coil without smaps)."""
    require_gpu(grid, smaps, scaling_coef)
    grid = grid.contiguous()
    im = tuple(int(v) for v in im_size)
    gs = tuple(grid.shape[2:])
    d = len(im)
    B, C = grid.shape[0], grid.shape[1]
    if smaps is not None:
        smaps = smaps.to(grid.dtype)
        if smaps.shape[0] != B:
            smaps = smaps.expand(B, *smaps.shape[1:])
        smaps = smaps.contiguous()
        if tuple(smaps.shape) != (B, C) + im:
            raise ValueError(f"smaps shape {tuple(smaps.shape)} does not match data ({B}, {C}) + {im}")
    sc = scaling_coef.to(device=grid.device, dtype=grid.dtype).contiguous()
    out = _out_or_new(out, (B, 1 if smaps is not None else C) + im, grid)
    ws, nws = _fft_workspace(grid.device, d, im, gs, B, C)
    with torch.cuda.device(grid.device):
        st = _lib.load().tkbn_fft_grid_adj(
            _lib.dtype_code(grid), d, _lib.i64_array(im), _lib.i64_array(gs), B, C, grid.data_ptr(),
            smaps.data_ptr() if smaps is not None else None, sc.data_ptr(), float(scale), out.data_ptr(),
            ws.data_ptr(), nws, _lib.stream_of(grid),
        )
    _lib.check(st, "tkbn_fft_grid_adj")
    return out


# ---------------------------------------------------------------------------
# FFT filtering on an oversampled grid (Toeplitz embedding)
# ---------------------------------------------------------------------------
_ONES = {}


def _ones_scaling(im, dtype, device) -> Tensor:
    key = (tuple(im), dtype, device)
    t = _ONES.get(key)
    if t is None:
        t = torch.ones(tuple(im), dtype=dtype, device=device)
        _ONES[key] = t
    return t


def fft_filter_fused(image: Tensor, smaps: Optional[Tensor], kernel: Tensor, kernel_batched: bool,
                     scale_fwd: float, scale_adj: float) -> Tensor:
    """One tkbn_fft_filter call: image (B, 1|C, *N) [, smaps (B|1, C, *N)], kernel
    (*K) or (B, *K) with kernel_batched; returns (B, 1|C, *N)."""
    require_gpu(image, smaps, kernel)
    image = image.contiguous()
    d = image.ndim - 2
    im = tuple(int(v) for v in image.shape[2:])
    gs = tuple(int(v) for v in kernel.shape[-d:])
    B = image.shape[0]
    C = smaps.shape[1] if smaps is not None else image.shape[1]
    if smaps is not None:
        smaps = smaps.to(image.dtype)
        if smaps.shape[0] != B:
            smaps = smaps.expand(B, *smaps.shape[1:])
        smaps = smaps.contiguous()
        if tuple(smaps.shape) != (B, C) + im:
            raise ValueError(f"smaps shape {tuple(smaps.shape)} does not match image {tuple(image.shape)}")
    kernel = kernel.to(image.dtype).contiguous()
    if kernel.numel() != (B if kernel_batched else 1) * math.prod(gs):
        raise ValueError(f"kernel of shape {tuple(kernel.shape)} does not match batch {B} and grid {gs}")
    out = torch.empty((B, 1 if smaps is not None else C) + im, dtype=image.dtype, device=image.device)
    ws, nws = _fft_workspace(image.device, d, im, gs, B, C)
    with torch.cuda.device(image.device):
        st = _lib.load().tkbn_fft_filter(
            _lib.dtype_code(image), d, _lib.i64_array(im), _lib.i64_array(gs), B, C, image.shape[1],
            image.data_ptr(), smaps.data_ptr() if smaps is not None else None, kernel.data_ptr(),
            1 if kernel_batched else 0, float(scale_fwd), float(scale_adj), out.data_ptr(), ws.data_ptr(), nws,
            _lib.stream_of(image),
        )
    _lib.check(st, "tkbn_fft_filter")
    return out


def toeplitz_filter(image: Tensor, smaps: Optional[Tensor], kernel: Tensor, norm: Optional[str]) -> Tensor:
    """crop(iFFT(kernel * FFT(pad_end(image [* smaps])))) [summed against conj(smaps)].

    The grid size is the kernel's trailing extents.  Where the fused FFT covers the
    geometry (complex64, 2/3/5-smooth sizes, two or more dimensions) the whole
    operator is one tkbn_fft_filter pipeline: the padded grid is never written
    (dimension 0 is transformed, multiplied by the kernel and transformed back in one
    pass).  One-dimensional problems, or a kernel or smaps that needs a gradient, use
    the fused pad + FFT and iFFT + crop stages with the product in between; other
    geometries use rocFFT.
    """
    from .._autograd.fft import FftFilter, GridAdjoint, GridForward

    if norm is not None and norm != "ortho":
        raise ValueError("Only option for norm is 'ortho'.")
    require_gpu(image, smaps, kernel)
    d = image.ndim - 2
    im = tuple(int(v) for v in image.shape[2:])
    gs = tuple(int(v) for v in kernel.shape[-d:])
    scale = _norm_scale(norm, gs)
    fused = fused_supported(image.dtype, im, gs) and all(g >= n for g, n in zip(gs, im))
    lead = tuple(kernel.shape[:-d])
    one_pass = (
        fused
        and d >= 2
        and kernel.dtype == image.dtype
        and (lead == () or (smaps is not None and lead == (image.shape[0], 1)))
        and not (torch.is_grad_enabled() and (kernel.requires_grad or (smaps is not None and smaps.requires_grad)))
    )
    if one_pass:
        return _run(FftFilter, image, smaps, kernel, lead != (), scale, scale)
    if fused:
        ones = _ones_scaling(im, image.dtype, image.device)
        grid = _run(GridForward, image, smaps, ones, im, gs, scale)
        return _run(GridAdjoint, grid * kernel, smaps, ones, im, scale)
    x = image * smaps if smaps is not None else image
    pad = []
    for g, n in zip(reversed(gs), reversed(im)):
        pad += [0, g - n]
    dims = tuple(range(-d, 0))
    grid = torch.fft.fftn(torch.nn.functional.pad(x, pad), dim=dims, norm="ortho" if norm == "ortho" else None)
    x = torch.fft.ifftn(grid * kernel, dim=dims, norm="ortho" if norm == "ortho" else "forward")
    for ax, n in enumerate(im):
        x = x.narrow(2 + ax, 0, n)
    if smaps is not None:
        x = torch.sum(x * smaps.conj(), dim=1, keepdim=True)
    return x


def fft_filter(image: Tensor, kernel: Tensor, norm: Optional[str] = "ortho") -> Tensor:
    """``iFFT(kernel * FFT(image))`` on an oversampled grid (``fft.py:121-173``)."""
    return toeplitz_filter(image, None, kernel, norm)
