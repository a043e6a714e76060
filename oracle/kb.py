"""FP64 closed forms of the Kaiser-Bessel pieces (TEST INFRASTRUCTURE ONLY).

These are independent formulations used to pin the product's host plan and the
oracle itself; they do not retrace the reference's sparse-matrix construction.

* table (``torchkbnufft/_nufft/utils.py:122-179``, which samples
  ``build_numpy_spmatrix`` ``:16-119`` in 1-D with ``n_shift = 0``):
  ``table[m] = KB(u) * exp(-i*pi*(N-1)*u/K)`` with ``u = m/L - J/2`` for
  ``0 <= m < J*L`` and ``table[J*L] = 0``; ``KB(u) = I0(alpha*sqrt(1-(2u/J)^2))/I0(alpha)``
  for ``|u| < J/2`` else 0 (order 0; other orders use ``I_order`` as the
  reference does at ``utils.py:53-55``).
* apodisation (``utils.py:182-254``): ``s = prod_d 1/KBFT((n_d-(N_d-1)/2)/K_d)``,
  ``KBFT(w) = J*sin(z)/(z*I0(alpha))``, ``z = sqrt((pi*J*w)^2 - alpha^2)``
  (complex branch gives ``sinh``); for other orders the Bessel-J form.
* derivative ``dKB/du = I1(alpha*s)*alpha*(-4u/(J^2 s))/I0(alpha)``, ``s = sqrt(1-(2u/J)^2)``.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np
from scipy import special


def kb_value(u: np.ndarray, numpoints: int, alpha: float, order: float = 0.0) -> np.ndarray:
    u = np.asarray(u, dtype=np.float64)
    half = numpoints / 2.0
    out = np.zeros_like(u)
    inside = np.abs(u) < half
    arg = alpha * np.sqrt(1.0 - (u[inside] / half) ** 2)
    if order == 0.0:
        out[inside] = np.i0(arg) / np.i0(alpha)
    else:
        out[inside] = special.iv(order, arg) / special.iv(order, alpha)
    return out


def kb_deriv(u: np.ndarray, numpoints: int, alpha: float) -> np.ndarray:
    """d/du of ``kb_value`` (order 0)."""
    u = np.asarray(u, dtype=np.float64)
    half = numpoints / 2.0
    out = np.zeros_like(u)
    inside = np.abs(u) < half
    s = np.sqrt(1.0 - (u[inside] / half) ** 2)
    # I1(a s)/s is finite as s -> 0 (limit a/2)
    safe = np.where(s > 1e-300, s, 1.0)
    ratio = np.where(s > 1e-300, special.i1(alpha * s) / safe, alpha / 2.0)
    out[inside] = ratio * alpha * (-u[inside] / half**2) / np.i0(alpha)
    return out


def kb_table(
    im_size: int, grid_size: int, numpoints: int, table_oversamp: int, alpha: float, order: float = 0.0
) -> np.ndarray:
    m = np.arange(numpoints * table_oversamp, dtype=np.float64)
    u = m / table_oversamp - numpoints / 2.0
    val = kb_value(u, numpoints, alpha, order) * np.exp(-1j * np.pi * (im_size - 1) * u / grid_size)
    return np.concatenate([val, np.zeros(1, dtype=np.complex128)])


def kb_table_deriv(
    im_size: int, grid_size: int, numpoints: int, table_oversamp: int, alpha: float
) -> np.ndarray:
    """d/du of the table function ``KB(u) exp(-i pi (N-1) u / K)`` on the same grid."""
    m = np.arange(numpoints * table_oversamp, dtype=np.float64)
    u = m / table_oversamp - numpoints / 2.0
    ph = np.exp(-1j * np.pi * (im_size - 1) * u / grid_size)
    val = (kb_deriv(u, numpoints, alpha) + kb_value(u, numpoints, alpha) * (-1j * np.pi * (im_size - 1) / grid_size)) * ph
    return np.concatenate([val, np.zeros(1, dtype=np.complex128)])


def kb_ft(w: np.ndarray, numpoints: int, alpha: float, order: float = 0.0) -> np.ndarray:
    w = np.asarray(w, dtype=np.float64)
    z = np.sqrt((np.pi * numpoints * w) ** 2 - alpha**2 + 0j)
    if order == 0.0:
        # sin(z)/z for complex z; z == 0 cannot occur for alpha > 0 on these grids
        return np.real(numpoints * np.sin(z) / (z * np.i0(alpha)))
    nu = 0.5 + order
    val = (
        np.sqrt(2 * np.pi)
        * (numpoints / 2.0)
        * alpha**order
        / special.iv(order, alpha)
        * special.jv(nu, z)
        / z**nu
    )
    return np.real(val)


def apodization(
    im_size: Sequence[int],
    grid_size: Sequence[int],
    numpoints: Sequence[int],
    alpha: Sequence[float],
    order: Sequence[float],
) -> np.ndarray:
    out = np.ones((), dtype=np.float64)
    for n, k, j, a, o in zip(im_size, grid_size, numpoints, alpha, order):
        w = (np.arange(n) - (n - 1) / 2.0) / k
        f = np.ones(n) if j == 1 else 1.0 / kb_ft(w, j, a, o)
        out = np.multiply.outer(out, f)
    return out
