"""Sparse-matrix Kaiser-Bessel interpolation.

``calc_tensor_spmatrix`` keeps the signature, defaults, errors and output of
``torchkbnufft/_nufft/spmat.py:10-105`` (with ``build_numpy_spmatrix``,
``torchkbnufft/_nufft/utils.py:16-119``): a pair of real sparse COO tensors (the real
and imaginary parts of the interpolation matrix, shape ``(klength, prod(grid_size))``)
whose row ``m`` holds the ``prod(numpoints)`` exact Kaiser-Bessel weights of sample
``m``.  The weights are evaluated in FP64 on the host; as in the reference, the grid
coordinate ``omega / (2 pi / K)`` and its floor are taken in omega's own precision
(float32 for a float32 trajectory) before the distances go to FP64, so the matrix
matches the reference's entry for entry.

``spmat_interp`` / ``spmat_interp_adjoint`` replace the reference's ``torch.mm``
products (``_nufft/interp.py:13-85``) with the HIP kernel ``tkbn_spmv``.  The two real
matrices are merged once into one complex matrix A = A_re + i A_im (and its conjugate
transpose for the adjoint), stored in blocks of ``_BLOCK_ROWS`` rows that share their
union of columns (neighbouring rows of an interpolation matrix share most columns, so
each input value is read once per block), cut into segments for load balance, and
cached per matrix pair.
"""
from __future__ import annotations

import math
import weakref
from typing import Dict, Optional, Sequence, Tuple, Union

import numpy as np
import torch
from scipy import special
from torch import Tensor

from .. import _lib
from ._host import host_values, require_gpu
from .utils import validate_args

_BLOCK_ROWS = 8  # rows per block of the sparse format (the library's tkbn_spmv_block_rows)
_SEG_COLS = 128  # union columns per segment (one wave each)


def _coo_entries(
    omega: np.ndarray,
    numpoints: Sequence[int],
    im_size: Sequence[int],
    grid_size: Sequence[int],
    n_shift: Sequence[int],
    order: Sequence[float],
    alpha: Sequence[float],
) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(row, col, value) of every entry, in the reference's entry order.

    Per dimension k with J taps on a grid of K points: ``t = omega_k / (2 pi / K)``
    and ``f = floor(t - J / 2)`` in omega's precision; tap ``j = 1 .. J`` sits at
    distance ``u = (t - f) - j`` from the sample (FP64 from here on), has the weight
    ``I_order(alpha sqrt(1 - (2u/J)^2)) / I_order(alpha)`` for ``|u| < J/2`` (else 0)
    times ``exp(i pi (N - 1) u / K)``, and the grid column ``(f + j) mod K``.  An entry
    is the conjugate of the product of its taps' factors over the dimensions (taken in
    dimension order) times ``exp(i omega . n_shift)``; the taps of the last dimension
    vary slowest within a row (``utils.py:88-104``).
    """
    d, m = omega.shape
    coef = None
    cols = None
    for k in range(d):
        J, K, N = int(numpoints[k]), int(grid_size[k]), int(im_size[k])
        gam = 2 * np.pi / K
        t = omega[k] / omega.dtype.type(gam)
        f = np.floor(t - J / 2)
        u = (t - f).astype(np.float64)[:, None] - np.arange(1, J + 1)[None, :]
        w = np.zeros(u.shape, dtype=np.float64)
        inside = np.absolute(u) < J / 2
        w[inside] = special.iv(order[k], alpha[k] * np.sqrt(1 - (u[inside] / (J / 2)) ** 2)) / special.iv(
            order[k], alpha[k]
        )
        c_k = np.exp((1j * gam * (N - 1) / 2) * u) * w
        col_k = np.mod(f.astype(np.float64)[:, None] + np.arange(1, J + 1)[None, :], K).astype(np.int64)
        col_k = col_k * int(np.prod(grid_size[k + 1 :]))
        if coef is None:
            coef, cols = c_k, col_k
        else:
            taps = coef.shape[1] * J
            coef = (coef[:, None, :] * c_k[:, :, None]).reshape(m, taps)
            cols = (cols[:, None, :] + col_k[:, :, None]).reshape(m, taps)
    shift = np.exp(1j * np.dot(np.transpose(omega), np.expand_dims(np.asarray(n_shift), 1)))
    vals = np.conj(coef) * shift
    rows = np.repeat(np.arange(m, dtype=np.int64), coef.shape[1])
    return rows, cols.reshape(-1), vals.reshape(-1)


def calc_tensor_spmatrix(
    omega: Tensor,
    im_size: Sequence[int],
    grid_size: Optional[Sequence[int]] = None,
    numpoints: Union[int, Sequence[int]] = 6,
    n_shift: Optional[Sequence[int]] = None,
    table_oversamp: Union[int, Sequence[int]] = 2**10,
    kbwidth: float = 2.34,
    order: Union[float, Sequence[float]] = 0.0,
) -> Tuple[Tensor, Tensor]:
    """(real, imaginary) sparse COO interpolation matrices for ``omega`` ``(d, klength)``.

    Pass the pair as ``interp_mats`` to ``KbInterp``, ``KbInterpAdjoint``, ``KbNufft``
    or ``KbNufftAdjoint`` (or the ``kb_spmat_*`` functions) instead of table
    interpolation.  ``table_oversamp`` is accepted for signature compatibility.
    """
    if not omega.ndim == 2:
        raise ValueError("Sparse matrix calculation not implemented for batched omega.")
    (im_size, grid_size, numpoints, n_shift, table_oversamp, order, alpha, dtype, device) = validate_args(
        im_size, grid_size, numpoints, n_shift, table_oversamp, kbwidth, order, omega.dtype, omega.device
    )
    rows, cols, vals = _coo_entries(
        omega.detach().cpu().numpy(), numpoints, im_size, grid_size, n_shift, order, alpha
    )
    inds = torch.tensor(np.stack((rows, cols)), dtype=torch.long, device=device)
    shape = torch.Size((omega.shape[1], int(np.prod(grid_size))))
    real_vals = torch.tensor(np.real(vals), dtype=dtype, device=device)
    imag_vals = torch.tensor(np.imag(vals), dtype=dtype, device=device)
    return (
        torch.sparse_coo_tensor(inds, real_vals, shape),
        torch.sparse_coo_tensor(inds, imag_vals, shape),
    )


class SpmvPlan:
    """A complex sparse matrix in row blocks for ``tkbn_spmv``.

    Rows are grouped in blocks of ``_BLOCK_ROWS``; a block keeps the union of its rows'
    columns (``ucol``) and, per union column, the block's ``_BLOCK_ROWS`` values
    (``uval``, zero where a row has no entry).  The union columns of each block are cut
    into segments of at most ``_SEG_COLS`` (one wave each); blocks with several
    segments add partial sums in a fix-up pass (``fix``).
    """

    def __init__(self, rows: Tensor, cols: Tensor, vals: Tensor, nrows: int, ncols: int):
        dev = vals.device
        R = _BLOCK_ROWS
        self.nrows, self.ncols = int(nrows), int(ncols)
        if ncols >= 2**31:
            raise ValueError("sparse interpolation supports at most 2^31 - 1 grid points")
        if rows.numel():
            lo = torch.stack([rows.min(), cols.min()]).tolist()
            hi = torch.stack([rows.max(), cols.max()]).tolist()
            if lo[0] < 0 or lo[1] < 0 or hi[0] >= nrows or hi[1] >= ncols:
                raise ValueError(f"sparse matrix indices outside its shape ({nrows}, {ncols})")
        # sum duplicate (row, col) entries
        key, perm = torch.sort(rows * ncols + cols, stable=True)
        vals = vals[perm]
        uniq, inverse = torch.unique_consecutive(key, return_inverse=True)
        if uniq.numel() != key.numel():
            summed = torch.zeros((uniq.numel(), 2), dtype=vals.real.dtype, device=dev)
            vals = torch.view_as_complex(summed.index_add_(0, inverse, torch.view_as_real(vals)))
        r = torch.div(uniq, ncols, rounding_mode="floor")
        c = uniq - r * ncols
        # union columns per row block, with the block's values per union column
        nblk = -(-self.nrows // R)
        bkey, bperm = torch.sort(torch.div(r, R, rounding_mode="floor") * ncols + c, stable=True)
        ukey, uinv = torch.unique_consecutive(bkey, return_inverse=True)
        ublk = torch.div(ukey, ncols, rounding_mode="floor")
        self.ucol = (ukey - ublk * ncols).to(torch.int32).contiguous()
        uval = torch.zeros((ukey.numel(), R), dtype=vals.dtype, device=dev)
        uval[uinv, r[bperm] % R] = vals[bperm]
        self.uval = uval.contiguous()
        ustart = torch.zeros(nblk + 1, dtype=torch.long, device=dev)
        ustart[1:] = torch.cumsum(torch.bincount(ublk, minlength=nblk), 0)
        # segments of at most _SEG_COLS union columns (an empty block keeps one, empty)
        n = ustart[1:] - ustart[:-1]
        nseg_b = torch.clamp(torch.div(n + _SEG_COLS - 1, _SEG_COLS, rounding_mode="floor"), min=1)
        ends = torch.cumsum(nseg_b, 0)
        self.nseg = int(ends[-1].item()) if nblk else 0
        first = ends - nseg_b
        seg_blk = torch.repeat_interleave(torch.arange(nblk, device=dev), nseg_b, output_size=self.nseg)
        within = torch.arange(self.nseg, device=dev) - first[seg_blk]
        u0 = ustart[seg_blk] + within * _SEG_COLS
        u1 = torch.minimum(u0 + _SEG_COLS, ustart[seg_blk + 1])
        split = (nseg_b > 1)[seg_blk]
        slot = torch.where(split, torch.cumsum(split.long(), 0) - 1, torch.full_like(seg_blk, -1))
        self.seg = torch.stack([seg_blk, u0, u1, slot], 1).contiguous()
        self.nslots = int(split.sum().item()) if self.nseg else 0
        fix_blk = torch.nonzero(nseg_b > 1).reshape(-1)
        self.fix = torch.stack([fix_blk, slot[first[fix_blk]], nseg_b[fix_blk]], 1).contiguous()
        self.nfix = int(fix_blk.numel())
        self.nnz = int(uniq.numel())

    def apply(self, x: Tensor) -> Tensor:
        """y (batch, nrows) = A x for x (batch, ncols), complex, on the plan's device."""
        x = x.contiguous()
        batch = x.shape[0]
        y = torch.empty((batch, self.nrows), dtype=x.dtype, device=x.device)
        if batch == 0 or self.nrows == 0:
            return y
        lib = _lib.load()
        if lib.tkbn_spmv_block_rows() != _BLOCK_ROWS:
            raise RuntimeError("libtkbn_hip was built with another sparse row-block size; rebuild")
        code = _lib.dtype_code(x)
        nbytes = lib.tkbn_spmv_workspace_bytes(code, self.nrows, self.ncols, batch, self.nslots)
        if nbytes < 0:
            _lib.check(_lib.TKBN_ERR_INVALID, "tkbn_spmv_workspace_bytes")
        ws = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=x.device)
        st = lib.tkbn_spmv(
            code, self.nrows, self.ncols, batch, self.nseg, self.seg.data_ptr(), self.nfix,
            self.fix.data_ptr() if self.nfix else None, self.nslots, self.ucol.data_ptr(),
            self.uval.data_ptr(), x.data_ptr(), y.data_ptr(), ws.data_ptr(), nbytes, _lib.stream_of(x),
        )
        _lib.check(st, "tkbn_spmv")
        return y


# (id(real), id(imag), adjoint, dtype, device) -> (weakref real, weakref imag, versions, plan)
_PLANS: Dict[tuple, tuple] = {}


def _coo_parts(mat: Tensor):
    if not isinstance(mat, Tensor):
        raise TypeError("interp_mats must be 2-tuple of (real_mat, imag_mat.")
    if mat.layout != torch.sparse_coo:
        mat = mat.to_sparse_coo() if mat.layout != torch.strided else mat.to_sparse()
    return mat._indices(), mat._values(), mat.shape


def _plan(interp_mats, adjoint: bool, like: Tensor) -> SpmvPlan:
    if not isinstance(interp_mats, tuple) or len(interp_mats) != 2:
        raise TypeError("interp_mats must be 2-tuple of (real_mat, imag_mat.")
    re, im = interp_mats
    ver = (getattr(re, "_version", 0), getattr(im, "_version", 0))
    key = (id(re), id(im), adjoint, like.dtype, like.device)
    hit = _PLANS.get(key)
    if hit is not None and hit[0]() is re and hit[1]() is im and hit[2] == ver:
        return hit[3]
    (ri, rv, rshape), (ii, iv, ishape) = _coo_parts(re), _coo_parts(im)
    if tuple(rshape) != tuple(ishape) or len(rshape) != 2:
        raise ValueError("interp_mats must be two 2-D sparse matrices of the same shape")
    real = rv.real.dtype if rv.is_complex() else rv.dtype
    cd = like.dtype
    rt = {torch.complex64: torch.float32, torch.complex128: torch.float64}[cd]
    if real not in (torch.float32, torch.float64):
        raise TypeError(f"interp_mats values must be float32 or float64, got {real}")
    dev = like.device
    ri, ii = ri.to(dev), ii.to(dev)
    if ri.shape == ii.shape and torch.equal(ri, ii):  # calc_tensor_spmatrix's pairs share indices
        rows, cols = ri[0], ri[1]
        vals = torch.complex(rv.to(dev, rt), iv.to(dev, rt))
    else:  # A = A_re + i A_im, entries summed when the block form is built
        rows, cols = torch.cat([ri[0], ii[0]]), torch.cat([ri[1], ii[1]])
        z_r = torch.zeros_like(rv, dtype=rt, device=dev)
        z_i = torch.zeros_like(iv, dtype=rt, device=dev)
        vals = torch.cat([torch.complex(rv.to(dev, rt), z_r), torch.complex(z_i, iv.to(dev, rt))])
    nrows, ncols = int(rshape[0]), int(rshape[1])
    if adjoint:  # conjugate transpose
        rows, cols, vals, nrows, ncols = cols, rows, vals.conj(), ncols, nrows
    plan = SpmvPlan(rows.contiguous(), cols.contiguous(), vals.resolve_conj().contiguous(), nrows, ncols)
    fresh = key not in _PLANS
    _PLANS[key] = (weakref.ref(re), weakref.ref(im), ver, plan)
    if fresh:
        weakref.finalize(re, _PLANS.pop, key, None)
    return plan


def clear_spmat_cache() -> None:
    """Drop every cached block form of interpolation matrices."""
    _PLANS.clear()


def spmat_interp(image: Tensor, interp_mats: Tuple[Tensor, Tensor]) -> Tensor:
    """Sparse-matrix interpolation: grid (B, C, *K) -> samples (B, C, klength)."""
    if not isinstance(interp_mats, tuple):
        raise TypeError("interp_mats must be 2-tuple of (real_mat, imag_mat.")
    require_gpu(image)
    if not image.is_complex():
        raise TypeError(f"image must be complex64 or complex128, got {image.dtype}")
    b, c = image.shape[:2]
    plan = _plan(interp_mats, False, image)
    x = image.reshape(b * c, math.prod(image.shape[2:]))
    if x.shape[1] != plan.ncols:
        raise RuntimeError(
            f"interpolation matrix has {plan.ncols} grid columns but the grid has {x.shape[1]} points"
        )
    return plan.apply(x).reshape(b, c, plan.nrows)


def spmat_interp_adjoint(data: Tensor, interp_mats: Tuple[Tensor, Tensor], grid_size) -> Tensor:
    """Sparse-matrix interpolation adjoint: samples (B, C, klength) -> grid (B, C, *grid_size)."""
    if not isinstance(interp_mats, tuple):
        raise TypeError("interp_mats must be 2-tuple of (real_mat, imag_mat.")
    require_gpu(data)
    if not data.is_complex():
        raise TypeError(f"data must be complex64 or complex128, got {data.dtype}")
    gs = [int(v) for v in host_values(grid_size)]
    b, c = data.shape[:2]
    plan = _plan(interp_mats, True, data)
    x = data.reshape(b * c, data.shape[-1])
    if x.shape[1] != plan.ncols:
        raise RuntimeError(f"interpolation matrix has {plan.ncols} samples but the data has {x.shape[1]}")
    if math.prod(gs) != plan.nrows:
        raise RuntimeError(f"grid_size {tuple(gs)} does not match the matrix's {plan.nrows} grid columns")
    return plan.apply(x).reshape([b, c] + gs)
