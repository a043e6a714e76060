"""Autograd for the fused scale/pad and crop/deapodise stages.

Each stage's backward is the other stage (they are adjoint to each other), so
``torch.autograd`` gradients of the NUFFT match the reference's composition of
ATen multiply / ``F.pad`` / index-select gradients (``fft.py:36-118``,
``kbnufft.py:182-183,404-405``).
"""
from __future__ import annotations

import torch
from torch.autograd import Function

from .._nufft.fft import apod_pad, crop_deapod


def _sum_to(t: torch.Tensor, shape) -> torch.Tensor:
    if tuple(t.shape) == tuple(shape):
        return t
    dims = [i for i, (a, b) in enumerate(zip(t.shape, shape)) if b == 1 and a != 1]
    return t.sum(dim=dims, keepdim=True)


class ApodPad(Function):
    @staticmethod
    def forward(ctx, image, smaps, scaling_coef, im_size, grid_size, scale):
        ctx.im_size = tuple(im_size)
        ctx.scale = scale
        ctx.image_coils = image.shape[1]
        ctx.smaps_shape = None if smaps is None else tuple(smaps.shape)
        ctx.save_for_backward(image if (smaps is not None and smaps.requires_grad) else None, smaps, scaling_coef)
        return apod_pad(image, smaps, scaling_coef, im_size, grid_size, scale)

    @staticmethod
    def backward(ctx, grad):
        image, smaps, scaling_coef = ctx.saved_tensors
        grad_image = grad_smaps = None
        per_coil = None
        if smaps is None:
            if ctx.needs_input_grad[0]:
                grad_image = crop_deapod(grad, None, scaling_coef, ctx.im_size, ctx.scale)
        else:
            if ctx.needs_input_grad[0] and ctx.image_coils == 1:
                grad_image = crop_deapod(grad, smaps, scaling_coef, ctx.im_size, ctx.scale)
            if ctx.needs_input_grad[1] or (ctx.needs_input_grad[0] and ctx.image_coils != 1):
                per_coil = crop_deapod(grad, None, scaling_coef, ctx.im_size, ctx.scale)
            if ctx.needs_input_grad[0] and ctx.image_coils != 1:
                grad_image = per_coil * smaps.conj()
            if ctx.needs_input_grad[1]:
                grad_smaps = _sum_to(per_coil * image.conj(), ctx.smaps_shape)
        return grad_image, grad_smaps, None, None, None, None


class CropDeapod(Function):
    @staticmethod
    def forward(ctx, grid, smaps, scaling_coef, im_size, scale):
        ctx.im_size = tuple(im_size)
        ctx.grid_size = tuple(grid.shape[2:])
        ctx.scale = scale
        ctx.smaps_shape = None if smaps is None else tuple(smaps.shape)
        ctx.save_for_backward(grid if (smaps is not None and smaps.requires_grad) else None, smaps, scaling_coef)
        return crop_deapod(grid, smaps, scaling_coef, im_size, scale)

    @staticmethod
    def backward(ctx, grad):
        grid, smaps, scaling_coef = ctx.saved_tensors
        grad_grid = grad_smaps = None
        if ctx.needs_input_grad[0]:
            grad_grid = apod_pad(grad, smaps, scaling_coef, ctx.im_size, ctx.grid_size, ctx.scale)
        if smaps is not None and ctx.needs_input_grad[1]:
            per_coil = crop_deapod(grid, None, scaling_coef, ctx.im_size, ctx.scale)
            grad_smaps = _sum_to(per_coil * grad.conj(), ctx.smaps_shape)
        return grad_grid, grad_smaps, None, None, None


class GridForward(Function):
    """image (B, 1|C, *N) [, smaps] -> FFT(pad(image * smaps * scaling)) * scale, fused."""

    @staticmethod
    def forward(ctx, image, smaps, scaling_coef, im_size, grid_size, scale):
        from .._nufft.fft import fft_grid_fwd

        ctx.im_size = tuple(im_size)
        ctx.scale = scale
        ctx.image_coils = image.shape[1]
        ctx.smaps_shape = None if smaps is None else tuple(smaps.shape)
        ctx.save_for_backward(image if (smaps is not None and smaps.requires_grad) else None, smaps, scaling_coef)
        return fft_grid_fwd(image, smaps, scaling_coef, im_size, grid_size, scale)

    @staticmethod
    def backward(ctx, grad):
        from .._nufft.fft import fft_grid_adj

        image, smaps, scaling_coef = ctx.saved_tensors
        grad_image = grad_smaps = per_coil = None
        if smaps is None:
            if ctx.needs_input_grad[0]:
                grad_image = fft_grid_adj(grad, None, scaling_coef, ctx.im_size, ctx.scale)
        else:
            if ctx.needs_input_grad[0] and ctx.image_coils == 1:
                grad_image = fft_grid_adj(grad, smaps, scaling_coef, ctx.im_size, ctx.scale)
            if ctx.needs_input_grad[1] or (ctx.needs_input_grad[0] and ctx.image_coils != 1):
                per_coil = fft_grid_adj(grad, None, scaling_coef, ctx.im_size, ctx.scale)
            if ctx.needs_input_grad[0] and ctx.image_coils != 1:
                grad_image = per_coil * smaps.conj()
            if ctx.needs_input_grad[1]:
                grad_smaps = _sum_to(per_coil * image.conj(), ctx.smaps_shape)
        return grad_image, grad_smaps, None, None, None, None


class GridAdjoint(Function):
    """grid (B, C, *K) [, smaps] -> sum_c crop(IFFT(grid)) * conj(scaling) * conj(smaps) * scale."""

    @staticmethod
    def forward(ctx, grid, smaps, scaling_coef, im_size, scale):
        from .._nufft.fft import fft_grid_adj

        ctx.im_size = tuple(im_size)
        ctx.grid_size = tuple(grid.shape[2:])
        ctx.scale = scale
        ctx.smaps_shape = None if smaps is None else tuple(smaps.shape)
        ctx.save_for_backward(grid if (smaps is not None and smaps.requires_grad) else None, smaps, scaling_coef)
        return fft_grid_adj(grid, smaps, scaling_coef, im_size, scale)

    @staticmethod
    def backward(ctx, grad):
        from .._nufft.fft import fft_grid_adj, fft_grid_fwd

        grid, smaps, scaling_coef = ctx.saved_tensors
        grad_grid = grad_smaps = None
        if ctx.needs_input_grad[0]:
            grad_grid = fft_grid_fwd(grad, smaps, scaling_coef, ctx.im_size, ctx.grid_size, ctx.scale)
        if smaps is not None and ctx.needs_input_grad[1]:
            per_coil = fft_grid_adj(grid, None, scaling_coef, ctx.im_size, ctx.scale)
            grad_smaps = _sum_to(per_coil * grad.conj(), ctx.smaps_shape)
        return grad_grid, grad_smaps, None, None, None


class FftFilter(Function):
    """image [, smaps] -> crop(IFFT(kernel * FFT(pad(image * smaps)))) [combined], fused.

    The operator is P^T F^H D F P (with the SENSE expansion and combination around
    it), so its adjoint is the same filter with conj(kernel): that gives the image
    gradient.  Gradients for smaps or the kernel go through the unfused path
    (the caller only uses this Function when neither needs one).
    """

    @staticmethod
    def forward(ctx, image, smaps, kernel, kernel_batched, scale_fwd, scale_adj):
        from .._nufft.fft import fft_filter_fused

        ctx.args = (kernel_batched, scale_fwd, scale_adj)
        ctx.save_for_backward(smaps, kernel)
        return fft_filter_fused(image, smaps, kernel, kernel_batched, scale_fwd, scale_adj)

    @staticmethod
    def backward(ctx, grad):
        from .._nufft.fft import fft_filter_fused

        smaps, kernel = ctx.saved_tensors
        kernel_batched, scale_fwd, scale_adj = ctx.args
        grad_image = None
        if ctx.needs_input_grad[0]:
            grad_image = fft_filter_fused(grad, smaps, kernel.conj(), kernel_batched, scale_adj, scale_fwd)
        return grad_image, None, None, None, None, None
