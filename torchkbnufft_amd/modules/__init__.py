from .kbinterp import KbInterp, KbInterpAdjoint
from .kbnufft import KbNufft, KbNufftAdjoint
from .toepnufft import ToepNufft

__all__ = ["KbInterp", "KbInterpAdjoint", "KbNufft", "KbNufftAdjoint", "ToepNufft"]
