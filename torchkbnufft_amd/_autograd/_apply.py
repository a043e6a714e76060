"""Calling an autograd Function without recording it when nothing needs a gradient."""
from __future__ import annotations

import torch
from torch import Tensor


class _NoGradCtx:
    """Stand-in context for a Function's forward called directly: attribute stores and
    ``save_for_backward`` are accepted and dropped."""

    def save_for_backward(self, *tensors) -> None:
        pass


def _wants_grad(a) -> bool:
    if isinstance(a, Tensor):
        return a.requires_grad
    if isinstance(a, (list, tuple)):
        return any(isinstance(t, Tensor) and t.requires_grad for t in a)
    return False


def run(fn, *args):
    """``fn.apply(*args)``, or ``fn.forward`` on a throw-away context when autograd would
    record nothing (grad mode off, or no tensor argument requires a gradient): the same
    result, without the graph node's host work (about 10 us of a small problem's call)."""
    if torch.is_grad_enabled() and any(_wants_grad(a) for a in args):
        return fn.apply(*args)
    return fn.forward(_NoGradCtx(), *args)
