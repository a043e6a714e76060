"""Toeplitz-embedded normal operator (``torchkbnufft/modules/kbnufft.py:413-547``)."""
from __future__ import annotations

from typing import Optional

import torch
from torch import Tensor

from .._nufft.fft import toeplitz_filter


class ToepNufft(torch.nn.Module):
    r"""Applies :math:`T x \approx A^H W A x` with an FFT kernel from
    :func:`calc_toeplitz_kernel` (no interpolation).

    Same arguments, checks and errors as the reference module.  With ``smaps`` the
    SENSE expansion, zero-padded FFT, kernel product, inverse FFT, crop and coil
    combination of all batch items run together (the fused HIP FFT pipelines where
    the geometry allows) instead of a Python loop over the batch; a batched kernel
    ``(B, 2N_0, ...)`` pairs kernel ``b`` with batch item ``b``.
    """

    def __init__(self):
        super().__init__()

    def forward(
        self,
        image: Tensor,
        kernel: Tensor,
        smaps: Optional[Tensor] = None,
        norm: Optional[str] = None,
    ) -> Tensor:
        if not kernel.dtype == image.dtype:
            raise TypeError("kernel and image must have same dtype.")
        if smaps is not None and not smaps.dtype == image.dtype:
            raise TypeError("image dtype does not match smaps dtype.")
        is_complex = image.is_complex()
        if not is_complex:
            if not image.shape[-1] == 2:
                raise ValueError("For real inputs, last dimension must be size 2.")
            if not kernel.shape[-1] == 2:
                raise ValueError("For real inputs, last dimension must be size 2.")
            if smaps is not None:
                if not smaps.shape[-1] == 2:
                    raise ValueError("For real inputs, last dimension must be size 2.")
                smaps = torch.view_as_complex(smaps)
            image = torch.view_as_complex(image)
            kernel = torch.view_as_complex(kernel)
        d = image.ndim - 2
        if kernel.ndim > d:
            if kernel.shape[0] == 1:
                kernel = kernel[0]
            elif not kernel.shape[0] == image.shape[0]:
                raise ValueError("If using batch dimension, kernel must have same batch size as image")
        if smaps is not None and kernel.ndim > d:
            kernel = kernel.unsqueeze(1)  # kernel b multiplies every coil of batch item b
        output = toeplitz_filter(image, smaps, kernel, norm)
        if not is_complex:
            output = torch.view_as_real(output)
        return output
