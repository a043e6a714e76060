"""torchkbnufft_amd: MI355X (gfx950) table-mode Kaiser-Bessel NUFFT.

Drop-in for torchkbnufft's table path: the same module classes
(``KbNufft``, ``KbNufftAdjoint``, ``KbInterp``, ``KbInterpAdjoint``), the same
``functional.kb_table_*`` signatures and the same ``state_dict`` buffers, executed
by hand-written HIP kernels (``libtkbn_hip.so``, loaded with ctypes) around
``torch.fft``.  Tensors must live on a HIP device; there is no CPU fallback.
"""
import torch  # noqa: F401  (load torch's HIP runtime before libtkbn_hip.so)

from . import functional, modules
from .batched import BatchedKbNufft
from ._math import absolute, complex_mult, complex_sign, conj_complex_mult, imag_exp, inner_product
from ._nufft import utils as nufft_utils
from ._nufft.dcomp import calc_density_compensation_function
from ._nufft.spmat import calc_tensor_spmatrix
from ._nufft.toep import calc_toeplitz_kernel
from ._nufft.interp import clear_plan_cache, deterministic_enabled, set_deterministic
from .modules import KbInterp, KbInterpAdjoint, KbNufft, KbNufftAdjoint, ToepNufft

__version__ = "0.1.0"

__all__ = [
    "KbInterp",
    "KbInterpAdjoint",
    "KbNufft",
    "KbNufftAdjoint",
    "ToepNufft",
    "absolute",
    "calc_density_compensation_function",
    "calc_tensor_spmatrix",
    "calc_toeplitz_kernel",
    "complex_mult",
    "complex_sign",
    "conj_complex_mult",
    "deterministic_enabled",
    "functional",
    "imag_exp",
    "inner_product",
    "modules",
    "nufft_utils",
    "set_deterministic",
    "clear_plan_cache",
    "BatchedKbNufft",
]
