"""Functional API (``torchkbnufft/functional/__init__.py``)."""
from .._nufft.fft import fft_and_scale, fft_filter, ifft_and_scale
from .interp import kb_spmat_interp, kb_spmat_interp_adjoint, kb_table_interp, kb_table_interp_adjoint
from .nufft import kb_spmat_nufft, kb_spmat_nufft_adjoint, kb_table_nufft, kb_table_nufft_adjoint

__all__ = [
    "fft_and_scale",
    "fft_filter",
    "ifft_and_scale",
    "kb_spmat_interp",
    "kb_spmat_interp_adjoint",
    "kb_spmat_nufft",
    "kb_spmat_nufft_adjoint",
    "kb_table_interp",
    "kb_table_interp_adjoint",
    "kb_table_nufft",
    "kb_table_nufft_adjoint",
]
