"""Functional interpolation (``torchkbnufft/functional/interp.py:14-172``): table and sparse matrix."""
from __future__ import annotations

from typing import List, Tuple

import torch
from torch import Tensor

from .._autograd._apply import run as _run
from .._autograd.interp import (
    KbSpmatInterpAdjoint,
    KbSpmatInterpForward,
    KbTableInterpAdjoint,
    KbTableInterpForward,
)


def kb_table_interp(
    image: Tensor,
    omega: Tensor,
    tables: List[Tensor],
    n_shift: Tensor,
    numpoints: Tensor,
    table_oversamp: Tensor,
    offsets: Tensor,
) -> Tensor:
    """Kaiser-Bessel table interpolation of gridded ``image`` (B, C, *K) to ``omega``.

    Real inputs must carry a trailing dimension of size 2 (``view_as_real`` layout).
    """
    is_complex = True
    if not image.is_complex():
        if not image.shape[-1] == 2:
            raise ValueError("For real inputs, last dimension must be size 2.")
        is_complex = False
        image = torch.view_as_complex(image.contiguous())
    data = _run(KbTableInterpForward, image, omega, tables, n_shift, numpoints, table_oversamp, offsets)
    if not is_complex:
        data = torch.view_as_real(data)
    return data


def kb_table_interp_adjoint(
    data: Tensor,
    omega: Tensor,
    tables: List[Tensor],
    n_shift: Tensor,
    numpoints: Tensor,
    table_oversamp: Tensor,
    offsets: Tensor,
    grid_size: Tensor,
) -> Tensor:
    """Kaiser-Bessel table interpolation adjoint of ``data`` (B, C, M) onto ``grid_size``."""
    is_complex = True
    if not data.is_complex():
        if not data.shape[-1] == 2:
            raise ValueError("For real inputs, last dimension must be size 2.")
        is_complex = False
        data = torch.view_as_complex(data.contiguous())
    image = _run(KbTableInterpAdjoint, data, omega, tables, n_shift, numpoints, table_oversamp, offsets, grid_size)
    if not is_complex:
        image = torch.view_as_real(image)
    return image


def kb_spmat_interp(image: Tensor, interp_mats: Tuple[Tensor, Tensor]) -> Tensor:
    """Sparse-matrix interpolation of gridded ``image`` (B, C, *K) to the matrix's samples.

    ``interp_mats`` is the (real, imaginary) pair from ``calc_tensor_spmatrix``.
    """
    is_complex = True
    if not image.is_complex():
        if not image.shape[-1] == 2:
            raise ValueError("For real inputs, last dimension must be size 2.")
        is_complex = False
        image = torch.view_as_complex(image.contiguous())
    data = _run(KbSpmatInterpForward, image, interp_mats)
    if not is_complex:
        data = torch.view_as_real(data)
    return data


def kb_spmat_interp_adjoint(data: Tensor, interp_mats: Tuple[Tensor, Tensor], grid_size: Tensor) -> Tensor:
    """Sparse-matrix interpolation adjoint of ``data`` (B, C, M) onto ``grid_size``."""
    is_complex = True
    if not data.is_complex():
        if not data.shape[-1] == 2:
            raise ValueError("For real inputs, last dimension must be size 2.")
        is_complex = False
        data = torch.view_as_complex(data.contiguous())
    image = _run(KbSpmatInterpAdjoint, data, interp_mats, grid_size)
    if not is_complex:
        image = torch.view_as_real(image)
    return image
