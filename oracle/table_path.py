"""CPU restatement of torchkbnufft's table-mode NUFFT path (TEST INFRASTRUCTURE ONLY).

Every function follows the reference's op sequence in torch CPU ops so that
complex64 results reproduce the reference's float32 index arithmetic:

* ``tm = omega / (2*pi / K)`` is evaluated as ``omega / (reciprocal(K) * 2pi)``
  (``interp.py:171,663``; the scalar-over-tensor division lowers to
  ``reciprocal`` then ``mul``), then a correctly rounded divide;
* ``base = 1 + floor(tm - J/2)`` (``interp.py:177,670``);
* per offset ``o`` (``itertools.product`` order, last dim fastest,
  ``utils.py:329``): ``g = base + o``,
  ``i = round((tm - g) * L) + floor(J*L/2)`` (half-to-even),
  ``coef = ((1 * t_0[i_0]) * t_1[i_1]) ...``,
  ``idx = sum_d mod(g_d, K_d) * prod(K[d+1:])`` (``interp.py:129-148``);
* forward ``y = sum_o coef * X[idx]`` then ``y *= exp(i * sum_d omega_d * n_shift_d)``
  (``interp.py:185-203``); adjoint is the conjugate transpose with the
  reference's sort-by-``tm`` accumulation order (``interp.py:552-584,680-724``).

Also restated: density compensation (``dcomp.py``), the Toeplitz kernel and filter
(``toep.py``, ``fft.py:121-173``) and the sparse-matrix products (``interp.py:13-85``).
"""
from __future__ import annotations

import itertools
from typing import List, Optional, Sequence

import numpy as np
import torch
import torch.nn.functional as F

from . import kb

_REAL_OF = {torch.complex64: torch.float32, torch.complex128: torch.float64}
_COMPLEX_OF = {torch.float32: torch.complex64, torch.float64: torch.complex128}


# ----------------------------------------------------------------------------
# plan (restates utils.py:257-427 through the FP64 closed forms in ``kb``)
# ----------------------------------------------------------------------------
class Plan:
    """Buffers of a reference ``KbNufft`` module, built from ``kb`` closed forms."""

    def __init__(
        self,
        im_size: Sequence[int],
        grid_size: Optional[Sequence[int]] = None,
        numpoints=6,
        n_shift: Optional[Sequence[int]] = None,
        table_oversamp=2**10,
        kbwidth: float = 2.34,
        order: float = 0.0,
        dtype: torch.dtype = torch.complex64,
    ):
        d = len(im_size)
        self.im_size = tuple(int(v) for v in im_size)
        self.grid_size = (
            tuple(2 * v for v in self.im_size) if grid_size is None else tuple(int(v) for v in grid_size)
        )
        self.numpoints = (numpoints,) * d if isinstance(numpoints, int) else tuple(numpoints)
        self.n_shift = tuple(v // 2 for v in self.im_size) if n_shift is None else tuple(n_shift)
        self.table_oversamp = (
            (table_oversamp,) * d if isinstance(table_oversamp, int) else tuple(table_oversamp)
        )
        self.alpha = tuple(kbwidth * j for j in self.numpoints)
        self.order = (order,) * d if isinstance(order, float) else tuple(order)
        self.complex_dtype = dtype if dtype.is_complex else _COMPLEX_OF[dtype]
        self.real_dtype = _REAL_OF[self.complex_dtype]
        self.tables = [
            torch.from_numpy(
                kb.kb_table(n, k, j, l, a, o)
            ).to(self.complex_dtype)
            for n, k, j, l, a, o in zip(
                self.im_size, self.grid_size, self.numpoints, self.table_oversamp, self.alpha, self.order
            )
        ]
        self.scaling_coef = torch.from_numpy(
            kb.apodization(self.im_size, self.grid_size, self.numpoints, self.alpha, self.order)
        ).to(self.complex_dtype)
        self.offsets = torch.tensor(
            list(itertools.product(*[range(j) for j in self.numpoints])), dtype=torch.long
        )
        self.t_im_size = torch.tensor(self.im_size, dtype=torch.long)
        self.t_grid_size = torch.tensor(self.grid_size, dtype=torch.long)
        self.t_n_shift = torch.tensor(self.n_shift, dtype=self.real_dtype)
        self.t_numpoints = torch.tensor(self.numpoints, dtype=torch.long)
        self.t_table_oversamp = torch.tensor(self.table_oversamp, dtype=torch.long)


# ----------------------------------------------------------------------------
# interpolation (interp.py:88-203, 406-726)
# ----------------------------------------------------------------------------
def _tm(omega: torch.Tensor, grid_size: torch.Tensor) -> torch.Tensor:
    # interp.py:171 / :663 -- scalar / tensor lowers to reciprocal(K) * 2pi.
    gam = torch.reciprocal(grid_size.to(omega).unsqueeze(-1)) * (2 * np.pi)
    return omega / gam


def _coef_and_index(tm, base_offset, offset, tables, centers, table_oversamp, grid_size, conj):
    # interp.py:89-150
    gridind = base_offset + offset.unsqueeze(1)
    distind = torch.round((tm - gridind.to(tm)) * table_oversamp.unsqueeze(1)).to(torch.long)
    coef = torch.ones(tm.shape[1], dtype=tables[0].dtype)
    arr_ind = torch.zeros(tm.shape[1], dtype=torch.long)
    for d in range(len(tables)):
        t = tables[d][distind[d] + centers[d]]
        coef = coef * (t.conj() if conj else t)
        arr_ind = arr_ind + torch.remainder(gridind[d], grid_size[d]).view(-1) * torch.prod(
            grid_size[d + 1 :]
        )
    return coef, arr_ind


def _phase(omega: torch.Tensor, n_shift: torch.Tensor) -> torch.Tensor:
    # interp.py:200-203 with _math.py:120-135
    theta = torch.sum(omega * n_shift.unsqueeze(-1), dim=-2, keepdim=True)
    return torch.view_as_complex(torch.stack((torch.cos(theta), torch.sin(theta)), -1))


def _interp_one(image, omega, tables, n_shift, numpoints, table_oversamp, offsets):
    # interp.py:154-203
    grid_size = torch.tensor(image.shape[2:], dtype=torch.long)
    tm = _tm(omega, grid_size)
    centers = torch.floor(numpoints * table_oversamp / 2).to(dtype=torch.long)
    base_offset = 1 + torch.floor(tm - numpoints.unsqueeze(-1) / 2.0).to(dtype=torch.long)
    image = image.reshape(image.shape[0], image.shape[1], -1)
    kdat = torch.zeros(image.shape[0], image.shape[1], tm.shape[-1], dtype=image.dtype)
    for offset in offsets:
        coef, arr_ind = _coef_and_index(
            tm, base_offset, offset, tables, centers, table_oversamp, grid_size, False
        )
        kdat += coef * image[:, :, arr_ind]
    return kdat * _phase(omega, n_shift)


def table_interp(image, omega, tables, n_shift, numpoints, table_oversamp, offsets):
    """Forward table interpolation (``interp.py:315-403``); batched omega loops."""
    if omega.ndim == 3 and omega.shape[0] == 1:
        omega = omega[0]
    if omega.ndim == 2:
        return _interp_one(image, omega, tables, n_shift, numpoints, table_oversamp, offsets)
    return torch.cat(
        [
            _interp_one(im.unsqueeze(0), om, tables, n_shift, numpoints, table_oversamp, offsets)
            for im, om in zip(image, omega)
        ]
    )


def _sort_one(tm, omega, data, grid_size):
    # interp.py:552-562
    key = torch.zeros(omega.shape[1], dtype=omega.dtype)
    for d in range(grid_size.numel()):
        key = key + torch.remainder(tm[d], grid_size[d]) * torch.prod(grid_size[d + 1 :])
    _, idx = torch.sort(key)
    return tm[:, idx], omega[:, idx], data[:, :, idx]


def table_interp_adjoint(data, omega, tables, n_shift, numpoints, table_oversamp, offsets, grid_size):
    """Adjoint table interpolation (``interp.py:587-726``), serial index_add per row."""
    if omega.ndim == 3 and omega.shape[0] == 1:
        omega = omega[0]
    grid_size = torch.as_tensor(grid_size, dtype=torch.long)
    out_size = [data.shape[0], data.shape[1]] + [int(v) for v in grid_size]
    prod = int(torch.prod(grid_size))
    if omega.ndim == 3:
        return torch.cat(
            [
                table_interp_adjoint(
                    dt.unsqueeze(0), om, tables, n_shift, numpoints, table_oversamp, offsets, grid_size
                )
                for dt, om in zip(data, omega)
            ]
        )
    tm = _tm(omega, grid_size)
    tm, omega, data = _sort_one(tm, omega, data, grid_size)
    centers = torch.floor(numpoints * table_oversamp / 2).to(dtype=torch.long)
    base_offset = 1 + torch.floor(tm - numpoints.unsqueeze(-1) / 2.0).to(dtype=torch.long)
    image = torch.zeros(data.shape[0] * data.shape[1], prod, dtype=data.dtype)
    data = data * _phase(omega, n_shift).conj()
    rows = data.reshape(data.shape[0] * data.shape[1], -1)
    for offset in offsets:
        coef, arr_ind = _coef_and_index(
            tm, base_offset, offset, tables, centers, table_oversamp, grid_size, True
        )
        contrib = coef.unsqueeze(0) * rows
        for r in range(rows.shape[0]):
            image[r].index_add_(0, arr_ind, contrib[r])
    return image.view(out_size)


# ----------------------------------------------------------------------------
# FFT stages (fft.py:9-118) and SENSE (kbnufft.py:182-183, 404-405)
# ----------------------------------------------------------------------------
def fft_and_scale(image, scaling_coef, im_size, grid_size, norm=None):
    pad = []
    for gd, im in zip(reversed(list(grid_size)), reversed(list(im_size))):
        pad += [0, int(gd) - int(im)]
    dims = list(range(-len(im_size), 0))
    return torch.fft.fftn(F.pad(image * scaling_coef, pad), dim=dims, norm="ortho" if norm == "ortho" else None)


def ifft_and_scale(image, scaling_coef, im_size, grid_size, norm=None):
    dims = list(range(-len(im_size), 0))
    x = torch.fft.ifftn(image, dim=dims, norm="ortho" if norm == "ortho" else "forward")
    for ax, n in enumerate(im_size):
        x = torch.narrow(x, 2 + ax, 0, int(n))
    return x * scaling_coef.conj()


def kb_table_nufft(plan: Plan, image, omega, smaps=None, norm=None):
    """Forward NUFFT (``kbnufft.py:182-228`` -> ``functional/nufft.py:126-191``)."""
    if smaps is not None:
        image = image * smaps
    grid = fft_and_scale(image, plan.scaling_coef, plan.im_size, plan.grid_size, norm)
    return table_interp(
        grid, omega, plan.tables, plan.t_n_shift, plan.t_numpoints, plan.t_table_oversamp, plan.offsets
    )


def kb_table_nufft_adjoint(plan: Plan, data, omega, smaps=None, norm=None):
    """Adjoint NUFFT (``functional/nufft.py:194-260`` + ``kbnufft.py:404-405``)."""
    grid = table_interp_adjoint(
        data,
        omega,
        plan.tables,
        plan.t_n_shift,
        plan.t_numpoints,
        plan.t_table_oversamp,
        plan.offsets,
        plan.t_grid_size,
    )
    image = ifft_and_scale(grid, plan.scaling_coef, plan.im_size, plan.grid_size, norm)
    if smaps is not None:
        image = torch.sum(image * smaps.conj(), dim=1, keepdim=True)
    return image


def kb_table_interp(plan: Plan, grid, omega):
    return table_interp(
        grid, omega, plan.tables, plan.t_n_shift, plan.t_numpoints, plan.t_table_oversamp, plan.offsets
    )


def kb_table_interp_adjoint(plan: Plan, data, omega, grid_size=None):
    gs = plan.t_grid_size if grid_size is None else torch.as_tensor(grid_size, dtype=torch.long)
    return table_interp_adjoint(
        data, omega, plan.tables, plan.t_n_shift, plan.t_numpoints, plan.t_table_oversamp, plan.offsets, gs
    )


def density_compensation(plan: Plan, omega, num_iterations: int = 10):
    """Pipe's iteration (``dcomp.py:97-117``): ``w <- w / |G G^H w|`` from ``w = 1``.

    ``omega`` is ``(d, M)`` or ``(B, d, M)``; returns ``(B, 1, M)`` in the plan's
    complex dtype.  ``G^H`` / ``G`` are this module's adjoint / forward
    interpolation on the plan's grid.
    """
    if omega.ndim == 3 and omega.shape[0] == 1:
        omega = omega[0]
    batch = omega.shape[0] if omega.ndim == 3 else 1
    w = torch.ones((batch, 1, omega.shape[-1]), dtype=plan.tables[0].dtype)
    for _ in range(num_iterations):
        back = kb_table_interp(plan, kb_table_interp_adjoint(plan, w, omega), omega)
        w = w / torch.abs(back)
    return w


# ----------------------------------------------------------------------------
# Toeplitz embedding (toep.py:11-300, fft.py:121-173, kbnufft.py:413-547)
# ----------------------------------------------------------------------------
def _adjoint_flip_concat(dim, omega, weights, adjoint):
    """toep.py:185-236: adjoints over the sign flips of dims >= ``dim``, concatenated."""
    ax = dim + 2
    flip = torch.ones(omega.shape[0], dtype=omega.dtype).unsqueeze(-1)
    flip[dim] = -1
    if dim < omega.shape[0] - 1:
        k1 = _adjoint_flip_concat(dim + 1, omega, weights, adjoint)
        k2 = _adjoint_flip_concat(dim + 1, flip * omega, weights, adjoint)
    else:
        k1 = adjoint(weights, omega)
        k2 = adjoint(weights, flip * omega)
    zshape = list(k1.shape)
    zshape[ax] = 1
    k2 = k2.narrow(ax, 1, k2.shape[ax] - 1)
    return torch.cat((k1, torch.zeros(zshape, dtype=k1.dtype), k2.flip(ax)), ax)


def _reverse_index(x, d):
    return x.index_select(d, torch.remainder(-1 * torch.arange(x.shape[d]), x.shape[d]))


def _reflect_conj_concat(kernel, dim):
    """toep.py:239-274"""
    tmp = kernel.conj()
    for d in range(dim, kernel.ndim):
        tmp = _reverse_index(tmp, d)
    zshape = list(kernel.shape)
    zshape[dim] = 1
    tmp = torch.cat((torch.zeros(zshape, dtype=kernel.dtype), tmp.narrow(dim, 1, tmp.shape[dim] - 1)), dim)
    return torch.cat((kernel, tmp), dim)


def _hermitify(kernel, dim):
    """toep.py:277-300"""
    rev = kernel
    for d in range(dim, kernel.ndim):
        rev = _reverse_index(rev, d)
    return (kernel + rev.conj()) / 2


def toeplitz_kernel(
    omega, im_size, weights=None, norm=None, grid_size=None, numpoints=6, dtype=torch.complex64, adjoint=None
):
    """calc_toeplitz_kernel (``toep.py:11-182``) for a ``(d, M)`` or ``(B, d, M)`` omega.

    ``adjoint(w, omega)`` replaces the table NUFFT adjoint (``n_shift = 0``) when
    given, e.g. by an exact NUDFT adjoint for the exactness check.
    """
    if omega.ndim == 3:
        return torch.stack(
            [
                toeplitz_kernel(
                    om, im_size, None if weights is None else weights[i], norm, grid_size, numpoints, dtype, adjoint
                )
                for i, om in enumerate(omega)
            ]
        )
    d = omega.shape[0]
    plan = Plan(im_size, grid_size, numpoints=numpoints, n_shift=[0] * d, dtype=dtype)
    if adjoint is None:
        adjoint = lambda w_, om_: kb_table_nufft_adjoint(plan, w_, om_, norm=norm)  # noqa: E731
    if weights is None:
        w = torch.ones(omega.shape[-1], dtype=dtype).unsqueeze(0).unsqueeze(0)
    else:
        w = weights.to(dtype).unsqueeze(0)
    if d > 1:
        kernel = _adjoint_flip_concat(1, omega, w, adjoint)
    else:
        kernel = adjoint(w, omega)
    kernel = _hermitify(_reflect_conj_concat(kernel, 2), 2)
    im = torch.tensor(plan.im_size, dtype=torch.float64)
    if norm == "ortho":
        scale = torch.sqrt(torch.prod(2 * im) / torch.prod(torch.tensor(plan.grid_size, dtype=torch.float64)))
    else:
        scale = 1 / torch.prod(2 * im)
    dims = list(range(-d, 0))
    return torch.fft.fftn(kernel, dim=dims, norm="ortho" if norm == "ortho" else None)[0, 0] * scale.item()


def fft_filter(image, kernel, norm="ortho"):
    """fft.py:121-173: crop(iFFT(kernel * FFT(pad_end(image))))."""
    d = image.ndim - 2
    im = image.shape[2:]
    gs = kernel.shape[-d:]
    pad = []
    for g, n in zip(reversed(list(gs)), reversed(list(im))):
        pad += [0, int(g) - int(n)]
    dims = list(range(-d, 0))
    x = torch.fft.fftn(F.pad(image, pad), dim=dims, norm="ortho" if norm == "ortho" else None) * kernel
    x = torch.fft.ifftn(x, dim=dims, norm="ortho" if norm == "ortho" else "forward")
    for ax, n in enumerate(im):
        x = torch.narrow(x, 2 + ax, 0, int(n))
    return x


def toep_nufft(image, kernel, smaps=None, norm=None):
    """ToepNufft.forward (``kbnufft.py:491-547``) on complex tensors."""
    d = image.ndim - 2
    if kernel.ndim > d and kernel.shape[0] == 1:
        kernel = kernel[0]
    if smaps is None:
        return fft_filter(image, kernel, norm)
    out = []
    for b in range(image.shape[0]):
        kb = kernel[b] if kernel.ndim > d else kernel
        sm = smaps[b if smaps.shape[0] > 1 else 0].unsqueeze(0)
        x = fft_filter(image[b].unsqueeze(0) * sm, kb, norm)
        out.append(torch.sum(x * sm.conj(), dim=1, keepdim=True)[0])
    return torch.stack(out)


# ----------------------------------------------------------------------------
# sparse-matrix interpolation (interp.py:13-85): four real products per direction
# ----------------------------------------------------------------------------
def spmat_interp(image, interp_mats):
    """Grid (B, C, *K) -> samples (B, C, M) with the (real, imaginary) matrix pair
    (``interp.py:13-47``: ``(R x_re - I x_im) + i (R x_im + I x_re)``)."""
    re, im = interp_mats
    b, c = image.shape[:2]
    x = torch.view_as_real(image).reshape(b * c, -1, 2)
    xr, xi = x[..., 0].t().contiguous(), x[..., 1].t().contiguous()
    out = torch.stack([(torch.mm(re, xr) - torch.mm(im, xi)).t(), (torch.mm(re, xi) + torch.mm(im, xr)).t()], dim=-1)
    return torch.view_as_complex(out.contiguous()).reshape(b, c, -1)


def spmat_interp_adjoint(data, interp_mats, grid_size):
    """Samples (B, C, M) -> grid (B, C, *grid_size) with the transposed pair and the
    conjugate (``interp.py:50-85``: ``(R^T y_re + I^T y_im) + i (R^T y_im - I^T y_re)``)."""
    re, im = interp_mats
    b, c = data.shape[:2]
    y = torch.view_as_real(data).reshape(b * c, -1, 2)
    yr, yi = y[..., 0].t().contiguous(), y[..., 1].t().contiguous()
    rt, it = re.t(), im.t()
    out = torch.stack([(torch.mm(rt, yr) + torch.mm(it, yi)).t(), (torch.mm(rt, yi) - torch.mm(it, yr)).t()], dim=-1)
    return torch.view_as_complex(out.contiguous()).reshape([b, c] + [int(v) for v in grid_size])


def inner_product(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """``sum(b * conj(a))``: ``_math.py:138-166`` evaluates ``conj_complex_mult(val2, val1)``."""
    return torch.sum(b * a.conj())


def rel_l2(a, b) -> float:
    a = torch.as_tensor(a).to(torch.complex128)
    b = torch.as_tensor(b).to(torch.complex128)
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


__all__: List[str] = [
    "Plan",
    "density_compensation",
    "toeplitz_kernel",
    "fft_filter",
    "toep_nufft",
    "table_interp",
    "table_interp_adjoint",
    "fft_and_scale",
    "ifft_and_scale",
    "kb_table_nufft",
    "kb_table_nufft_adjoint",
    "kb_table_interp",
    "kb_table_interp_adjoint",
    "spmat_interp",
    "spmat_interp_adjoint",
    "inner_product",
    "rel_l2",
]
