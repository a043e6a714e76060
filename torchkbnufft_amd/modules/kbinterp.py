"""Interpolation modules (``torchkbnufft/modules/kbinterp.py:10-294``).

Table interpolation by default; with ``interp_mats`` (from ``calc_tensor_spmatrix``)
the sparse-matrix path, as in the reference (``kbinterp.py:137-138, 269-272``).
"""
from __future__ import annotations

from typing import Optional, Tuple

import torch
from torch import Tensor

from .. import functional as F
from ._kbmodule import KbModule


class KbInterpModule(KbModule):
    """Parent of the interpolation modules; see ``KbModule`` for the arguments."""


class KbInterp(KbInterpModule):
    r"""Non-uniform Kaiser-Bessel interpolation layer (grid -> off-grid samples).

    Arguments as the reference's ``KbInterp``: ``im_size``, ``grid_size``,
    ``numpoints``, ``n_shift``, ``table_oversamp``, ``kbwidth``, ``order``,
    ``dtype``, ``device``.
    """

    def forward(self, image: Tensor, omega: Tensor, interp_mats: Optional[Tuple[Tensor, Tensor]] = None) -> Tensor:
        """``image`` (N, C, *grid) -> (N, C, klength); ``omega`` (d, klength) or (N, d, klength)."""
        if interp_mats is not None:
            return F.kb_spmat_interp(image=image, interp_mats=interp_mats)
        return F.kb_table_interp(
            image=image,
            omega=omega,
            tables=self._tables(),
            n_shift=self.n_shift,
            numpoints=self.numpoints,
            table_oversamp=self.table_oversamp,
            offsets=self.offsets.to(torch.long),
        )


class KbInterpAdjoint(KbInterpModule):
    r"""Non-uniform Kaiser-Bessel interpolation adjoint layer (off-grid samples -> grid)."""

    def forward(
        self,
        data: Tensor,
        omega: Tensor,
        interp_mats: Optional[Tuple[Tensor, Tensor]] = None,
        grid_size: Optional[Tensor] = None,
    ) -> Tensor:
        """``data`` (N, C, klength) -> (N, C, *grid_size)."""
        if grid_size is None:
            grid_size = self.grid_size
        if interp_mats is not None:
            return F.kb_spmat_interp_adjoint(data=data, interp_mats=interp_mats, grid_size=grid_size)
        return F.kb_table_interp_adjoint(
            data=data,
            omega=omega,
            tables=self._tables(),
            n_shift=self.n_shift,
            numpoints=self.numpoints,
            table_oversamp=self.table_oversamp,
            offsets=self.offsets,
            grid_size=grid_size,
        )
