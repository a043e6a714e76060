"""Density compensation by Pipe's fixed-point iteration on the HIP interpolators.

Same function, arguments, defaults and errors as
``torchkbnufft/_nufft/dcomp.py:10-119``: starting from ``w = 1`` it repeats
``w <- w / |G G^H w|`` ``num_iterations`` times, where ``G^H`` spreads samples onto
the oversampled grid and ``G`` interpolates back (no FFT, no apodisation).  Here
each half of an iteration is one HIP launch of the library's spread / gather, and
the trajectory plan built by the first iteration is reused by the other nine (the
plan cache compares the trajectory on the device).
"""
from __future__ import annotations

from typing import Optional, Sequence, Union

import torch
from torch import Tensor

from . import interp as _interp
from .utils import init_fn


def calc_density_compensation_function(
    ktraj: Tensor,
    im_size: Sequence[int],
    num_iterations: int = 10,
    grid_size: Optional[Sequence[int]] = None,
    numpoints: Union[int, Sequence[int]] = 6,
    n_shift: Optional[Sequence[int]] = None,
    table_oversamp: Union[int, Sequence[int]] = 2**10,
    kbwidth: float = 2.34,
    order: Union[float, Sequence[float]] = 0.0,
) -> Tensor:
    """Density compensation weights for ``ktraj`` (``(d, M)`` or ``(B, d, M)``).

    Returns a ``(B, 1, M)`` complex tensor (``B = 1`` for an unbatched trajectory)
    of real, positive values in the complex dtype matching ``ktraj``'s precision.
    """
    if ktraj.ndim not in (2, 3):
        raise ValueError("ktraj must have 2 or 3 dimensions")
    batch = 1
    if ktraj.ndim == 3:
        if ktraj.shape[0] == 1:
            ktraj = ktraj[0]
        else:
            batch = ktraj.shape[0]
    tables, _, grid_t, n_shift_t, numpoints_t, offsets_t, oversamp_t, _, _ = init_fn(
        im_size=im_size,
        grid_size=grid_size,
        numpoints=numpoints,
        n_shift=n_shift,
        table_oversamp=table_oversamp,
        kbwidth=kbwidth,
        order=order,
        dtype=ktraj.dtype,
        device=ktraj.device,
    )
    weights = torch.ones((batch, 1, ktraj.shape[-1]), dtype=tables[0].dtype, device=ktraj.device)
    for _ in range(num_iterations):
        grid = _interp.table_interp_adjoint(
            weights, ktraj, tables, n_shift_t, numpoints_t, oversamp_t, offsets_t, grid_t
        )
        back = _interp.table_interp(grid, ktraj, tables, n_shift_t, numpoints_t, oversamp_t, offsets_t)
        weights = weights / back.abs()
    return weights
