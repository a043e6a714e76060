"""NUFFT modules (``torchkbnufft/modules/kbnufft.py:11-410``), table mode.

SENSE expansion and coil combination are fused into the HIP scale/pad and
crop/deapodise passes instead of separate ``image * smaps`` and
``sum(out * conj(smaps))`` ATen ops.  With ``interp_mats`` (from
``calc_tensor_spmatrix``) the interpolation is the sparse-matrix path
(``kbnufft.py:185-197, 364-376``).
"""
from __future__ import annotations

from typing import Optional, Sequence, Tuple, Union

import torch
from torch import Tensor

from ..functional.nufft import _nufft_adjoint, _nufft_forward, _spmat_nufft_adjoint, _spmat_nufft_forward
from .._nufft.utils import compute_scaling_coefs
from ._kbmodule import KbModule


class KbNufftModule(KbModule):
    def __init__(
        self,
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
        super().__init__(
            im_size=im_size,
            grid_size=grid_size,
            numpoints=numpoints,
            n_shift=n_shift,
            table_oversamp=table_oversamp,
            kbwidth=kbwidth,
            order=order,
            dtype=dtype,
            device=device,
        )
        scaling_coef = compute_scaling_coefs(
            im_size=self.im_size.tolist(),
            grid_size=self.grid_size.tolist(),
            numpoints=self.numpoints.tolist(),
            alpha=self.alpha.tolist(),
            order=self.order.tolist(),
        )
        self.register_buffer("scaling_coef", scaling_coef.to(dtype=self.table_0.dtype, device=device))


def _complex_view(x: Tensor, smaps: Optional[Tensor]):
    if smaps is not None and not smaps.dtype == x.dtype:
        raise TypeError("image dtype does not match smaps dtype.")
    if x.is_complex():
        return x, smaps, True
    if not x.shape[-1] == 2:
        raise ValueError("For real inputs, last dimension must be size 2.")
    if smaps is not None:
        if not smaps.shape[-1] == 2:
            raise ValueError("For real inputs, last dimension must be size 2.")
        smaps = torch.view_as_complex(smaps.contiguous())
    return torch.view_as_complex(x.contiguous()), smaps, False


class KbNufft(KbNufftModule):
    r"""Non-uniform FFT layer: image (N, C, *im_size) -> k-space (N, C, klength).

    With ``smaps`` (N, C, *im_size) the image (N, 1, *im_size) is expanded to coils
    first.  ``norm`` is ``None`` or ``"ortho"``.
    """

    def forward(
        self,
        image: Tensor,
        omega: Tensor,
        interp_mats: Optional[Tuple[Tensor, Tensor]] = None,
        smaps: Optional[Tensor] = None,
        norm: Optional[str] = None,
    ) -> Tensor:
        image, smaps, is_complex = _complex_view(image, smaps)
        if interp_mats is not None:
            out = _spmat_nufft_forward(image, smaps, self.scaling_coef, self.im_size, self.grid_size, interp_mats, norm)
            return out if is_complex else torch.view_as_real(out)
        out = _nufft_forward(
            image, smaps, self.scaling_coef, self.im_size, self.grid_size, omega, self._tables(),
            self.n_shift, self.numpoints, self.table_oversamp, self.offsets.to(torch.long), norm,
        )
        return out if is_complex else torch.view_as_real(out)


class KbNufftAdjoint(KbNufftModule):
    r"""Non-uniform FFT adjoint layer: k-space (N, C, klength) -> image (N, C, *im_size).

    With ``smaps`` the coils are combined into (N, 1, *im_size).
    """

    def forward(
        self,
        data: Tensor,
        omega: Tensor,
        interp_mats: Optional[Tuple[Tensor, Tensor]] = None,
        smaps: Optional[Tensor] = None,
        norm: Optional[str] = None,
    ) -> Tensor:
        if smaps is not None and not smaps.dtype == data.dtype:
            raise TypeError("data dtype does not match smaps dtype.")
        data, smaps, is_complex = _complex_view(data, smaps)
        if interp_mats is not None:
            out = _spmat_nufft_adjoint(data, smaps, self.scaling_coef, self.im_size, self.grid_size, interp_mats, norm)
            return out if is_complex else torch.view_as_real(out)
        out = _nufft_adjoint(
            data, smaps, self.scaling_coef, self.im_size, self.grid_size, omega, self._tables(),
            self.n_shift, self.numpoints, self.table_oversamp, self.offsets.to(torch.long), norm,
        )
        return out if is_complex else torch.view_as_real(out)
