from .kbinterp import KbInterp, KbInterpAdjoint
from .kbnufft import KbNufft, KbNufftAdjoint

__all__ = ["KbInterp", "KbInterpAdjoint", "KbNufft", "KbNufftAdjoint"]
