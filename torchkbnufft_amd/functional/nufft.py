"""Functional NUFFT (``torchkbnufft/functional/nufft.py:15-260``): table and sparse matrix."""
from __future__ import annotations

from typing import List, Optional, Tuple

import torch
from torch import Tensor

from .._autograd._apply import run as _run
from .._autograd.fft import ApodPad, CropDeapod, GridAdjoint, GridForward
from .._nufft._host import host_values
from .._nufft.fft import _norm_scale, fused_supported
from .interp import kb_spmat_interp, kb_spmat_interp_adjoint, kb_table_interp, kb_table_interp_adjoint


def _to_grid(image, smaps, scaling_coef, im_size, grid_size, norm):
    gs = host_values(grid_size)
    im = host_values(im_size)
    if fused_supported(image.dtype, im, gs):
        return _run(GridForward, image, smaps, scaling_coef, im, gs, _norm_scale(norm, gs))
    grid = _run(ApodPad, image, smaps, scaling_coef, im, gs, _norm_scale(norm, gs))
    return torch.fft.fftn(grid, dim=tuple(range(-len(gs), 0)), norm="backward")


def _from_grid(grid, smaps, scaling_coef, im_size, grid_size, norm):
    gs = host_values(grid_size)
    im = host_values(im_size)
    if fused_supported(grid.dtype, im, gs):
        return _run(GridAdjoint, grid, smaps, scaling_coef, im, _norm_scale(norm, gs))
    grid = torch.fft.ifftn(grid, dim=tuple(range(-len(gs), 0)), norm="forward")
    return _run(CropDeapod, grid, smaps, scaling_coef, im, _norm_scale(norm, gs))


def _nufft_forward(image, smaps, scaling_coef, im_size, grid_size, omega, tables, n_shift, numpoints,
                   table_oversamp, offsets, norm):
    grid = _to_grid(image, smaps, scaling_coef, im_size, grid_size, norm)
    return kb_table_interp(grid, omega, tables, n_shift, numpoints, table_oversamp, offsets)


def _nufft_adjoint(data, smaps, scaling_coef, im_size, grid_size, omega, tables, n_shift, numpoints,
                   table_oversamp, offsets, norm):
    gs = host_values(grid_size)
    grid = kb_table_interp_adjoint(data, omega, tables, n_shift, numpoints, table_oversamp, offsets, gs)
    return _from_grid(grid, smaps, scaling_coef, im_size, grid_size, norm)


def _spmat_nufft_forward(image, smaps, scaling_coef, im_size, grid_size, interp_mats, norm):
    return kb_spmat_interp(_to_grid(image, smaps, scaling_coef, im_size, grid_size, norm), interp_mats)


def _spmat_nufft_adjoint(data, smaps, scaling_coef, im_size, grid_size, interp_mats, norm):
    grid = kb_spmat_interp_adjoint(data, interp_mats, host_values(grid_size))
    return _from_grid(grid, smaps, scaling_coef, im_size, grid_size, norm)


def kb_spmat_nufft(
    image: Tensor,
    scaling_coef: Tensor,
    im_size: Tensor,
    grid_size: Tensor,
    interp_mats: Tuple[Tensor, Tensor],
    norm: Optional[str] = None,
) -> Tensor:
    """Kaiser-Bessel NUFFT with sparse-matrix interpolation (image (B, C, *N) -> (B, C, M))."""
    is_complex = True
    if not image.is_complex():
        if not image.shape[-1] == 2:
            raise ValueError("For real inputs, last dimension must be size 2.")
        is_complex = False
        image = torch.view_as_complex(image.contiguous())
    data = _spmat_nufft_forward(image, None, scaling_coef, im_size, grid_size, interp_mats, norm)
    if not is_complex:
        data = torch.view_as_real(data)
    return data


def kb_spmat_nufft_adjoint(
    data: Tensor,
    scaling_coef: Tensor,
    im_size: Tensor,
    grid_size: Tensor,
    interp_mats: Tuple[Tensor, Tensor],
    norm: Optional[str] = None,
) -> Tensor:
    """Kaiser-Bessel NUFFT adjoint with sparse-matrix interpolation ((B, C, M) -> (B, C, *N))."""
    is_complex = True
    if not data.is_complex():
        if not data.shape[-1] == 2:
            raise ValueError("For real inputs, last dimension must be size 2.")
        is_complex = False
        data = torch.view_as_complex(data.contiguous())
    image = _spmat_nufft_adjoint(data, None, scaling_coef, im_size, grid_size, interp_mats, norm)
    if not is_complex:
        image = torch.view_as_real(image)
    return image


def kb_table_nufft(
    image: Tensor,
    scaling_coef: Tensor,
    im_size: Tensor,
    grid_size: Tensor,
    omega: Tensor,
    tables: List[Tensor],
    n_shift: Tensor,
    numpoints: Tensor,
    table_oversamp: Tensor,
    offsets: Tensor,
    norm: Optional[str] = None,
) -> Tensor:
    """Kaiser-Bessel NUFFT with table interpolation (image (B, C, *N) -> (B, C, M))."""
    is_complex = True
    if not image.is_complex():
        if not image.shape[-1] == 2:
            raise ValueError("For real inputs, last dimension must be size 2.")
        is_complex = False
        image = torch.view_as_complex(image.contiguous())
    data = _nufft_forward(image, None, scaling_coef, im_size, grid_size, omega, tables, n_shift, numpoints,
                          table_oversamp, offsets, norm)
    if not is_complex:
        data = torch.view_as_real(data)
    return data


def kb_table_nufft_adjoint(
    data: Tensor,
    scaling_coef: Tensor,
    im_size: Tensor,
    grid_size: Tensor,
    omega: Tensor,
    tables: List[Tensor],
    n_shift: Tensor,
    numpoints: Tensor,
    table_oversamp: Tensor,
    offsets: Tensor,
    norm: Optional[str] = None,
) -> Tensor:
    """Kaiser-Bessel NUFFT adjoint with table interpolation ((B, C, M) -> (B, C, *N))."""
    is_complex = True
    if not data.is_complex():
        if not data.shape[-1] == 2:
            raise ValueError("For real inputs, last dimension must be size 2.")
        is_complex = False
        data = torch.view_as_complex(data.contiguous())
    image = _nufft_adjoint(data, None, scaling_coef, im_size, grid_size, omega, tables, n_shift, numpoints,
                           table_oversamp, offsets, norm)
    if not is_complex:
        image = torch.view_as_real(image)
    return image
