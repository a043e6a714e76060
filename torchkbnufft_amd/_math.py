"""Complex helpers with the reference's names and semantics (``torchkbnufft/_math.py``).

Only ``imag_exp`` lies on the NUFFT path (the fftshift phase; inside the HIP
kernels it is computed in-register); the rest are the reference's public helpers,
kept so user code and tests written against it run unchanged.
"""
from __future__ import annotations

import torch
from torch import Tensor


def _as_complex(v: Tensor, dim: int) -> Tensor:
    if v.is_complex():
        return v
    if v.shape[dim] != 2:
        raise ValueError("Real input does not have dimension size 2 at dim.")
    return torch.view_as_complex(v.movedim(dim, -1).contiguous())


def _back(v: Tensor, dim: int, like: Tensor) -> Tensor:
    if like.is_complex():
        return v
    return torch.view_as_real(v).movedim(-1, dim)


def complex_mult(val1: Tensor, val2: Tensor, dim: int = -1) -> Tensor:
    """``val1 * val2`` (real inputs carry (re, im) along ``dim``)."""
    return _back(_as_complex(val1, dim) * _as_complex(val2, dim), dim, val1)


def conj_complex_mult(val1: Tensor, val2: Tensor, dim: int = -1) -> Tensor:
    """``val1 * conj(val2)``."""
    return _back(_as_complex(val1, dim) * _as_complex(val2, dim).conj(), dim, val1)


def absolute(val: Tensor, dim: int = -1) -> Tensor:
    """``|val|`` (real inputs: magnitude over ``dim``)."""
    if val.is_complex():
        return torch.abs(val)
    return torch.abs(_as_complex(val, dim))


def complex_sign(val: Tensor, dim: int = -1) -> Tensor:
    """``exp(i * atan2(im, re))`` (``torchkbnufft/_math.py:60-82``): 1 at 0, -1 at -0.0.

    Real inputs carry (re, im) along ``dim``; the result's pair sits in the last
    dimension, as the reference's ``imag_exp`` places it.
    """
    if torch.is_complex(val):
        ang = torch.atan2(val.imag, val.real)
        return imag_exp(ang, return_complex=True)
    if val.shape[dim] != 2:
        raise ValueError("Real input does not have dimension size 2 at dim.")
    ang = torch.atan2(val.select(dim, 1), val.select(dim, 0))
    return imag_exp(ang, dim=dim, return_complex=False)


def imag_exp(val: Tensor, dim: int = -1, return_complex: bool = True) -> Tensor:
    """``exp(i * val)``; the real form stacks (cos, sin) in the last dimension whatever
    ``dim`` is (``torchkbnufft/_math.py:120-135``)."""
    out = torch.stack((torch.cos(val), torch.sin(val)), -1)
    if return_complex:
        return torch.view_as_complex(out)
    return out


def inner_product(val1: Tensor, val2: Tensor, dim: int = -1) -> Tensor:
    """``sum(val2 * conj(val1))`` (``torchkbnufft/_math.py:138-166``: the reference
    evaluates ``conj_complex_mult(val2, val1)``); real inputs return a (re, im) pair."""
    if not val1.dtype == val2.dtype:
        raise ValueError("val1 has different dtype than val2.")
    if not torch.is_complex(val1):
        if not val1.shape[dim] == val2.shape[dim] == 2:
            raise ValueError("Real input does not have dimension size 2 at dim.")
        p = _as_complex(val2, dim) * _as_complex(val1, dim).conj()
        s = p.sum()
        return torch.stack((s.real, s.imag))
    return torch.sum(val2 * val1.conj())
