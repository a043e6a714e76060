"""Toeplitz embedding of the normal operator A^H W A.

``calc_toeplitz_kernel`` keeps the signature, defaults and errors of
``torchkbnufft/_nufft/toep.py:11-182``.  The kernel is the FFT of the point-spread
function of A^H W A sampled on a grid of twice the image size: its values at
offsets ``k`` in ``[-N+1, N-1]^d`` (wrapped modulo ``2N``) are
``sum_m w_m exp(i omega_m . k)``, evaluated by adjoint NUFFTs with ``n_shift = 0``.
One adjoint gives the offsets with non-negative coordinates; each sign pattern of
the dimensions after the first needs one more (``2^(d-1)`` adjoints), and the
first dimension's negative half follows from Hermitian symmetry
``K(-k) = conj(K(k))``.  The entries at offset ``N`` (the middle of each doubled
dimension) are zero, and the kernel is symmetrised as ``(K + conj(K(-k))) / 2``
before its FFT, as the reference does.  The adjoints run on the HIP library; the
assembly is a few element-wise passes over a ``(2N)^d`` array.
"""
from __future__ import annotations

import itertools
import math
from typing import Optional, Sequence, Union

import torch
from torch import Tensor

from . import fft as _fft


def _circular_reverse(x: Tensor, dims) -> Tensor:
    """x[..., (-m) mod n, ...] along each of ``dims``."""
    dims = tuple(dims)
    return torch.roll(torch.flip(x, dims), shifts=(1,) * len(dims), dims=dims)


def _one_kernel(
    omega: Tensor,
    im_size: Sequence[int],
    weights: Optional[Tensor],
    norm: Optional[str],
    grid_size: Optional[Sequence[int]],
    numpoints,
    table_oversamp,
    kbwidth: float,
    order,
) -> Tensor:
    from ..modules import KbNufftAdjoint

    d = omega.shape[0]
    im = tuple(int(v) for v in im_size)
    adj = KbNufftAdjoint(
        im_size=im,
        grid_size=grid_size,
        numpoints=numpoints,
        n_shift=[0] * d,
        table_oversamp=table_oversamp,
        kbwidth=kbwidth,
        order=order,
        dtype=omega.dtype,
        device=omega.device,
    )
    cdtype = adj.table_0.dtype
    if weights is None:
        w = torch.ones((1, 1, omega.shape[-1]), dtype=cdtype, device=omega.device)
    else:
        w = weights.to(adj.table_0).unsqueeze(0)

    # offsets with k_0 >= 0: one adjoint per sign pattern of dimensions 1 .. d-1
    half = torch.zeros((im[0],) + tuple(2 * n for n in im[1:]), dtype=cdtype, device=omega.device)
    for signs in itertools.product((1, -1), repeat=d - 1):
        flip = torch.tensor((1,) + signs, dtype=omega.dtype, device=omega.device).unsqueeze(-1)
        part = adj(w, omega * flip, norm=norm)[0, 0]
        dest = [slice(0, im[0])]
        for j, s in enumerate(signs, start=1):
            if s > 0:
                dest.append(slice(0, im[j]))
            else:  # offset -n (n >= 1) wraps to 2N - n; n = 0 belongs to the + pattern
                part = part.narrow(j, 1, im[j] - 1).flip(j)
                dest.append(slice(im[j] + 1, 2 * im[j]))
        half[tuple(dest)] = part

    # negative k_0 by Hermitian symmetry; the middle row stays zero
    full = torch.zeros(tuple(2 * n for n in im), dtype=cdtype, device=omega.device)
    full[: im[0]] = half
    if im[0] > 1:
        mirrored = torch.flip(half[1:], (0,))
        if d > 1:
            mirrored = _circular_reverse(mirrored, range(1, d))
        full[im[0] + 1 :] = mirrored.conj()
    full = (full + _circular_reverse(full, range(d)).conj()) / 2

    kgrid = adj.grid_size.tolist()
    if norm == "ortho":
        scale = math.sqrt(math.prod(2 * n for n in im) / math.prod(int(v) for v in kgrid))
    else:
        scale = 1.0 / math.prod(2 * n for n in im)
    shape = tuple(2 * n for n in im)
    if _fft.fused_supported(cdtype, shape, shape):
        # the fused HIP FFT (no padding, unit apodisation) instead of rocFFT, which
        # compiles kernels at run time for each new length
        if norm == "ortho":
            scale = scale / math.sqrt(math.prod(shape))
        ones = _fft._ones_scaling(shape, cdtype, omega.device)
        return _fft.fft_grid_fwd(full.reshape((1, 1) + shape), None, ones, shape, shape, scale)[0, 0]
    return torch.fft.fftn(full, dim=tuple(range(d)), norm="ortho" if norm == "ortho" else None) * scale


def calc_toeplitz_kernel(
    omega: Tensor,
    im_size: Sequence[int],
    weights: Optional[Tensor] = None,
    norm: Optional[str] = None,
    grid_size: Optional[Sequence[int]] = None,
    numpoints: Union[int, Sequence[int]] = 6,
    table_oversamp: Union[int, Sequence[int]] = 2**10,
    kbwidth: float = 2.34,
    order: Union[float, Sequence[float]] = 0.0,
) -> Tensor:
    """FFT-domain kernel D with F^H D F ~= A^H W A on a ``2 * im_size`` grid.

    ``omega``: ``(d, M)`` or ``(B, d, M)``; ``weights``: ``(1, M)`` (or ``(B, M)`` /
    ``(B, 1, M)`` for a batched trajectory), default ones.  Returns ``(2N_0, ...)``
    or ``(B, 2N_0, ...)``; use with :class:`ToepNufft`.
    """
    if omega.ndim not in (2, 3):
        raise ValueError("Unrecognized k-space shape.")
    if weights is not None:
        if weights.ndim not in (2, 3):
            raise ValueError("Unrecognized weights dimension.")
        if omega.ndim == 3 and weights.ndim == 2:
            if weights.shape[0] == 1:
                weights = weights.repeat(omega.shape[0], 1)
            if not weights.shape[0] == omega.shape[0]:
                raise ValueError("weights and omega do not have same batch size")
    args = (norm, grid_size, numpoints, table_oversamp, kbwidth, order)
    if omega.ndim == 2:
        return _one_kernel(omega, im_size, weights, *args)
    return torch.stack(
        [_one_kernel(om, im_size, None if weights is None else weights[i], *args) for i, om in enumerate(omega)]
    )
