"""FP64 direct non-uniform DFT and an exact-KB operator (TEST INFRASTRUCTURE ONLY).

The table NUFFT approximates ``y_m = sum_n x_n exp(-i omega_m . (n - n_shift))``
with ``n`` running over the image grid and ``n_shift = N // 2`` by default
(the operator documented at ``torchkbnufft/modules/kbnufft.py:55-123``).
``exact_kb_interp`` is the table interpolation of ``interp.py:154-203`` with the
kernel evaluated in closed form at the exact distance (no table quantisation):
it is smooth in ``omega`` and is the operator finite differences are taken on.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np

from . import kb


def _grid_coords(im_size: Sequence[int], n_shift: Sequence[int]) -> np.ndarray:
    axes = [np.arange(n, dtype=np.float64) - s for n, s in zip(im_size, n_shift)]
    mesh = np.meshgrid(*axes, indexing="ij")
    return np.stack([m.reshape(-1) for m in mesh])  # (d, P)


def nudft(x: np.ndarray, omega: np.ndarray, n_shift: Sequence[int], chunk: int = 4096) -> np.ndarray:
    """x (..., *N) complex, omega (d, M) -> (..., M) complex128."""
    d = omega.shape[0]
    im_size = x.shape[-d:]
    lead = x.shape[:-d]
    xf = x.reshape(-1, int(np.prod(im_size))).astype(np.complex128)
    r = _grid_coords(im_size, n_shift)
    out = np.empty((xf.shape[0], omega.shape[1]), dtype=np.complex128)
    for s in range(0, omega.shape[1], chunk):
        om = omega[:, s : s + chunk].astype(np.float64)
        e = np.exp(-1j * (om.T @ r))  # (m, P)
        out[:, s : s + chunk] = xf @ e.T
    return out.reshape(*lead, omega.shape[1])


def nudft_adjoint(y: np.ndarray, omega: np.ndarray, im_size, n_shift, chunk: int = 4096) -> np.ndarray:
    lead = y.shape[:-1]
    yf = y.reshape(-1, y.shape[-1]).astype(np.complex128)
    r = _grid_coords(im_size, n_shift)
    out = np.zeros((yf.shape[0], r.shape[1]), dtype=np.complex128)
    for s in range(0, omega.shape[1], chunk):
        om = omega[:, s : s + chunk].astype(np.float64)
        e = np.exp(1j * (om.T @ r))  # (m, P)
        out += yf[:, s : s + chunk] @ e
    return out.reshape(*lead, *im_size)


def nudft_domega(x: np.ndarray, omega: np.ndarray, n_shift, chunk: int = 2048) -> np.ndarray:
    """dy_m/domega_{d,m} = sum_n x_n (-i (n_d - s_d)) exp(-i omega_m.(n - s)): (d, ..., M)."""
    d = omega.shape[0]
    im_size = x.shape[-d:]
    lead = x.shape[:-d]
    xf = x.reshape(-1, int(np.prod(im_size))).astype(np.complex128)
    r = _grid_coords(im_size, n_shift)
    out = np.empty((d, xf.shape[0], omega.shape[1]), dtype=np.complex128)
    for s in range(0, omega.shape[1], chunk):
        om = omega[:, s : s + chunk].astype(np.float64)
        e = np.exp(-1j * (om.T @ r))  # (m, P)
        for k in range(d):
            out[k, :, s : s + chunk] = (xf * (-1j * r[k])[None, :]) @ e.T
    return out.reshape(d, *lead, omega.shape[1])


def exact_kb_interp(
    grid: np.ndarray,
    omega: np.ndarray,
    im_size: Sequence[int],
    numpoints: Sequence[int],
    alpha: Sequence[float],
    n_shift: Sequence[float],
) -> np.ndarray:
    """Table interpolation with the kernel in closed form (FP64, smooth in omega).

    grid (..., *K) complex, omega (d, M) -> (..., M) complex128.
    """
    d = omega.shape[0]
    grid_size = grid.shape[-d:]
    lead = grid.shape[:-d]
    gf = grid.reshape(-1, int(np.prod(grid_size))).astype(np.complex128)
    om = omega.astype(np.float64)
    tm = om * (np.asarray(grid_size, dtype=np.float64)[:, None] / (2 * np.pi))
    base = 1 + np.floor(tm - np.asarray(numpoints, dtype=np.float64)[:, None] / 2.0)
    out = np.zeros((gf.shape[0], om.shape[1]), dtype=np.complex128)
    strides = [int(np.prod(grid_size[k + 1 :])) for k in range(d)]
    for offs in np.ndindex(*numpoints):
        coef = np.ones(om.shape[1], dtype=np.complex128)
        idx = np.zeros(om.shape[1], dtype=np.int64)
        for k in range(d):
            g = base[k] + offs[k]
            u = tm[k] - g
            coef = coef * kb.kb_value(u, numpoints[k], alpha[k]) * np.exp(
                -1j * np.pi * (im_size[k] - 1) * u / grid_size[k]
            )
            idx = idx + np.mod(g.astype(np.int64), grid_size[k]) * strides[k]
        out += coef[None, :] * gf[:, idx]
    phase = np.exp(1j * np.sum(om * np.asarray(n_shift, dtype=np.float64)[:, None], axis=0))
    return (out * phase[None, :]).reshape(*lead, om.shape[1])


def exact_kb_interp_domega(
    grid: np.ndarray,
    omega: np.ndarray,
    im_size: Sequence[int],
    numpoints: Sequence[int],
    alpha: Sequence[float],
    n_shift: Sequence[float],
) -> np.ndarray:
    """d/d omega_{k,m} of ``exact_kb_interp`` (away from footprint edges, where it is
    smooth): (d, ..., M) complex128.  Per tap the kernel factor of dimension k is
    replaced by its derivative in omega_k (``kb.kb_deriv`` and the table phase, times
    K_k / 2 pi), and the shift phase contributes i n_shift_k y."""
    d = omega.shape[0]
    grid_size = grid.shape[-d:]
    lead = grid.shape[:-d]
    gf = grid.reshape(-1, int(np.prod(grid_size))).astype(np.complex128)
    om = omega.astype(np.float64)
    kk = np.asarray(grid_size, dtype=np.float64)
    tm = om * (kk[:, None] / (2 * np.pi))
    base = 1 + np.floor(tm - np.asarray(numpoints, dtype=np.float64)[:, None] / 2.0)
    val = np.zeros((gf.shape[0], om.shape[1]), dtype=np.complex128)
    der = np.zeros((d, gf.shape[0], om.shape[1]), dtype=np.complex128)
    strides = [int(np.prod(grid_size[k + 1 :])) for k in range(d)]
    for offs in np.ndindex(*numpoints):
        f, df = [], []
        idx = np.zeros(om.shape[1], dtype=np.int64)
        for k in range(d):
            g = base[k] + offs[k]
            u = tm[k] - g
            ph = np.exp(-1j * np.pi * (im_size[k] - 1) * u / grid_size[k])
            v = kb.kb_value(u, numpoints[k], alpha[k])
            f.append(v * ph)
            dv = kb.kb_deriv(u, numpoints[k], alpha[k]) + v * (-1j * np.pi * (im_size[k] - 1) / grid_size[k])
            df.append(dv * ph * (kk[k] / (2 * np.pi)))
            idx = idx + np.mod(g.astype(np.int64), grid_size[k]) * strides[k]
        x = gf[:, idx]
        prod_all = np.prod(f, axis=0)
        val += prod_all[None, :] * x
        for k in range(d):
            others = np.prod([f[j] for j in range(d) if j != k], axis=0) if d > 1 else 1.0
            der[k] += (df[k] * others)[None, :] * x
    shift = np.asarray(n_shift, dtype=np.float64)
    phase = np.exp(1j * np.sum(om * shift[:, None], axis=0))
    out = (der + 1j * shift[:, None, None] * val[None]) * phase[None, None, :]
    return out.reshape(d, *lead, om.shape[1])
