"""Autograd Functions of the table interpolation (``torchkbnufft/_autograd/interp.py:78-178``).

Image/data gradients are the reference's: backward(forward) = adjoint and
backward(adjoint) = forward, in PyTorch's complex convention.  Unlike the
reference (which returns ``None`` for ``omega``), the trajectory gets a gradient
when it requires one.
"""
from __future__ import annotations

import torch
from torch.autograd import Function

from .._nufft import interp as _interp
from .._nufft import spmat as _spmat


class KbTableInterpForward(Function):
    @staticmethod
    def forward(ctx, image, omega, tables, n_shift, numpoints, table_oversamp, offsets):
        output = _interp.table_interp(image, omega, tables, n_shift, numpoints, table_oversamp, offsets)
        ctx.tables = tables
        ctx.grid_size = tuple(image.shape[2:])
        ctx.save_for_backward(
            omega, n_shift, numpoints, table_oversamp, offsets, image if omega.requires_grad else None
        )
        return output

    @staticmethod
    def backward(ctx, grad):
        omega, n_shift, numpoints, table_oversamp, offsets, image = ctx.saved_tensors
        grad_image = grad_omega = None
        if ctx.needs_input_grad[0]:
            grad_image = _interp.table_interp_adjoint(
                grad, omega, ctx.tables, n_shift, numpoints, table_oversamp, offsets, ctx.grid_size
            )
        if ctx.needs_input_grad[1]:
            grad_omega = _interp.table_interp_domega(
                image, grad, omega, ctx.tables, n_shift, numpoints, table_oversamp
            )
        return grad_image, grad_omega, None, None, None, None, None


class KbTableInterpAdjoint(Function):
    @staticmethod
    def forward(ctx, data, omega, tables, n_shift, numpoints, table_oversamp, offsets, grid_size):
        image = _interp.table_interp_adjoint(data, omega, tables, n_shift, numpoints, table_oversamp, offsets, grid_size)
        ctx.tables = tables
        ctx.save_for_backward(
            omega, n_shift, numpoints, table_oversamp, offsets, data if omega.requires_grad else None
        )
        return image

    @staticmethod
    def backward(ctx, grad):
        omega, n_shift, numpoints, table_oversamp, offsets, data = ctx.saved_tensors
        grad_data = grad_omega = None
        if ctx.needs_input_grad[0]:
            grad_data = _interp.table_interp(grad, omega, ctx.tables, n_shift, numpoints, table_oversamp, offsets)
        if ctx.needs_input_grad[1]:
            # d/domega Re<grad, A^H data> = d/domega Re<A grad, data>: the forward's
            # trajectory gradient with the roles of grid and samples swapped
            grad_omega = _interp.table_interp_domega(
                grad.contiguous(), data, omega, ctx.tables, n_shift, numpoints, table_oversamp
            )
        return grad_data, grad_omega, None, None, None, None, None, None


class KbSpmatInterpForward(Function):
    """Sparse-matrix interpolation (``torchkbnufft/_autograd/interp.py:12-43``)."""

    @staticmethod
    def forward(ctx, image, interp_mats):
        ctx.interp_mats = interp_mats
        ctx.grid_size = tuple(image.shape[2:])
        return _spmat.spmat_interp(image, interp_mats)

    @staticmethod
    def backward(ctx, data):
        return _spmat.spmat_interp_adjoint(data, ctx.interp_mats, ctx.grid_size), None


class KbSpmatInterpAdjoint(Function):
    """Sparse-matrix interpolation adjoint (``torchkbnufft/_autograd/interp.py:46-75``)."""

    @staticmethod
    def forward(ctx, data, interp_mats, grid_size):
        ctx.interp_mats = interp_mats
        return _spmat.spmat_interp_adjoint(data, interp_mats, grid_size)

    @staticmethod
    def backward(ctx, image):
        return _spmat.spmat_interp(image, ctx.interp_mats), None, None
