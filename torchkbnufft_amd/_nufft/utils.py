"""Host-side FP64 plan: defaults, Kaiser-Bessel tables, apodisation, offsets.

Mirrors ``torchkbnufft/_nufft/utils.py`` (``init_fn`` :257-358, ``validate_args``
:361-427, ``build_table`` :122-179, ``kaiser_bessel_ft`` :182-209,
``compute_scaling_coefs`` :212-254) so module buffers are byte-compatible with the
reference's ``state_dict``.

The table is evaluated directly (no sparse matrix): for the ``L`` sub-sample
positions ``t = J/2 - 1 + l/L`` and the ``J`` neighbours ``j``, the reference's
distance ``d = om/gam - floor(om/gam - J/2) - j`` with ``om = t*2*pi/K`` and
``gam = 2*pi/K`` is computed with the same FP64 operations (so table entries
round to the same complex64/complex128 values), weighted by
``I_order(alpha*sqrt(1-(d/(J/2))^2))/I_order(alpha)`` and the phase
``conj(exp(i*gam*(N-1)/2*d))``, and laid out neighbour-major from ``j = J`` down
to ``j = 1`` followed by a trailing zero.
"""
from __future__ import annotations

import itertools
from typing import List, Optional, Sequence, Tuple, Union

import numpy as np
import torch
from scipy import special
from torch import Tensor

DTYPE_MAP = [
    (torch.complex128, torch.float64),
    (torch.complex64, torch.float32),
]


def _table_1d(im_size: int, grid_size: int, numpoints: int, table_oversamp: int, order: float, alpha: float) -> np.ndarray:
    t1 = numpoints / 2 - 1 + np.array(range(table_oversamp)) / table_oversamp
    om = t1 * 2 * np.pi / grid_size
    gam = 2 * np.pi / grid_size
    dist = om / gam - np.floor(om / gam - numpoints / 2)  # (L,)
    kern = -1 * np.arange(1, numpoints + 1)[None, :] + dist[:, None]  # (L, J)
    weight = np.zeros(kern.shape, dtype=np.complex128)
    inside = np.absolute(kern) < numpoints / 2
    arg = np.sqrt(1 - (kern[inside] / (numpoints / 2)) ** 2)
    weight[inside] = special.iv(order, alpha * arg) / special.iv(order, alpha)
    weight = np.real(weight)
    phase = np.exp(1j * gam * (im_size - 1) / 2 * kern)
    coef = np.conj(phase * weight) * np.exp(1j * 0.0)
    # Tap j of row k is stored in grid column (j + koff[k]) mod K, as the reference's
    # sparse matrix does (utils.py:83-85), and only columns 0 .. J-1 form the table
    # (utils.py:169-172; duplicates add).  koff is -1 except on rows where om / gam
    # rounds below J/2 - 1: there every tap moves one column to the right.
    koff = np.floor(om / gam - numpoints / 2)
    cols = np.mod(np.arange(1, numpoints + 1)[None, :] + koff[:, None], grid_size).astype(np.int64)
    tab = np.zeros((table_oversamp, numpoints), dtype=np.complex128)
    for j in range(numpoints):
        rows = np.nonzero(cols[:, j] < numpoints)[0]
        np.add.at(tab, (rows, cols[rows, j]), coef[rows, j])
    # column J-1 first, down to column 0, then the closing zero
    return np.concatenate([tab[:, ::-1].T.reshape(-1), np.array([0])])


def build_table(
    im_size: Sequence[int],
    grid_size: Sequence[int],
    numpoints: Sequence[int],
    table_oversamp: Sequence[int],
    order: Sequence[float],
    alpha: Sequence[float],
) -> List[Tensor]:
    """One complex128 table of ``J*L + 1`` entries per dimension."""
    return [
        torch.tensor(_table_1d(n, k, j, l, o, a))
        for n, k, j, l, o, a in zip(im_size, grid_size, numpoints, table_oversamp, order, alpha)
    ]


def kaiser_bessel_ft(omega: np.ndarray, numpoints: int, alpha: float, order: float, d: int) -> np.ndarray:
    """Fourier transform of the KB kernel (``utils.py:182-209``)."""
    z = np.sqrt((2 * np.pi * (numpoints / 2) * omega) ** 2 - alpha**2 + 0j)
    nu = d / 2 + order
    val = (
        (2 * np.pi) ** (d / 2)
        * ((numpoints / 2) ** d)
        * (alpha**order)
        / special.iv(order, alpha)
        * special.jv(nu, z)
        / (z**nu)
    )
    return np.real(val)


def compute_scaling_coefs(
    im_size: Sequence[int],
    grid_size: Sequence[int],
    numpoints: Sequence[int],
    alpha: Sequence[float],
    order: Sequence[float],
) -> Tensor:
    """Separable apodisation ``prod_d 1/KBFT((n_d - (N_d-1)/2)/K_d)`` (1 where J_d == 1)."""
    out = None
    for i in range(len(im_size)):
        w = np.array(range(im_size[i])) - (im_size[i] - 1) / 2
        f = 1 / kaiser_bessel_ft(w / grid_size[i], numpoints[i], alpha[i], order[i], 1)
        if numpoints[i] == 1:
            f = np.ones(f.shape)
        if out is None:
            out = f
        else:
            out = np.expand_dims(out, axis=-1) * f.reshape((1,) * i + (-1,))
    return torch.tensor(out)


def validate_args(
    im_size: Sequence[int],
    grid_size: Optional[Sequence[int]] = None,
    numpoints: Union[int, Sequence[int]] = 6,
    n_shift: Optional[Sequence[int]] = None,
    table_oversamp: Union[int, Sequence[int]] = 2**10,
    kbwidth: float = 2.34,
    order: Union[float, Sequence[float]] = 0.0,
    dtype: Optional[torch.dtype] = None,
    device: Optional[torch.device] = None,
):
    """Fill defaults (``utils.py:361-427``)."""
    im_size = tuple(im_size)
    grid_size = tuple(d * 2 for d in im_size) if grid_size is None else tuple(grid_size)
    numpoints = tuple(numpoints for _ in grid_size) if isinstance(numpoints, int) else tuple(numpoints)
    n_shift = tuple(d // 2 for d in im_size) if n_shift is None else tuple(n_shift)
    table_oversamp = (
        tuple(table_oversamp for _ in grid_size) if isinstance(table_oversamp, int) else tuple(table_oversamp)
    )
    alpha = tuple(kbwidth * j for j in numpoints)
    order = tuple(order for _ in grid_size) if isinstance(order, float) else tuple(order)
    dtype = torch.get_default_dtype() if dtype is None else dtype
    device = torch.device("cpu") if device is None else device
    for seq in (grid_size, n_shift, numpoints, alpha, order, table_oversamp):
        assert len(seq) == len(im_size)
    return im_size, grid_size, numpoints, n_shift, table_oversamp, order, alpha, dtype, device


def complex_and_real(dtype: torch.dtype) -> Tuple[torch.dtype, torch.dtype]:
    if dtype.is_floating_point:
        for c, r in DTYPE_MAP:
            if r == dtype:
                return c, r
    elif dtype.is_complex:
        for c, r in DTYPE_MAP:
            if c == dtype:
                return c, r
    raise TypeError("Unrecognized dtype.")


def init_fn(
    im_size: Sequence[int],
    grid_size: Optional[Sequence[int]] = None,
    numpoints: Union[int, Sequence[int]] = 6,
    n_shift: Optional[Sequence[int]] = None,
    table_oversamp: Union[int, Sequence[int]] = 2**10,
    kbwidth: float = 2.34,
    order: Union[float, Sequence[float]] = 0.0,
    dtype: Optional[torch.dtype] = None,
    device: Optional[torch.device] = None,
):
    """Returns ``(tables, im_size, grid_size, n_shift, numpoints, offsets,
    table_oversamp, order, alpha)`` as tensors, as ``utils.py:257-358`` does."""
    (im_size, grid_size, numpoints, n_shift, table_oversamp, order, alpha, dtype, device) = validate_args(
        im_size, grid_size, numpoints, n_shift, table_oversamp, kbwidth, order, dtype, device
    )
    tables = build_table(im_size, grid_size, numpoints, table_oversamp, order, alpha)
    offsets = list(itertools.product(*[range(j) for j in numpoints]))
    complex_dtype, real_dtype = complex_and_real(dtype)
    tables = [t.to(dtype=complex_dtype, device=device) for t in tables]
    return (
        tables,
        torch.tensor(im_size, dtype=torch.long, device=device),
        torch.tensor(grid_size, dtype=torch.long, device=device),
        torch.tensor(n_shift, dtype=real_dtype, device=device),
        torch.tensor(numpoints, dtype=torch.long, device=device),
        torch.tensor(offsets, dtype=torch.long, device=device),
        torch.tensor(table_oversamp, dtype=torch.long, device=device),
        torch.tensor(order, dtype=real_dtype, device=device),
        torch.tensor(alpha, dtype=real_dtype, device=device),
    )
