"""Table interpolation backends on the HIP library.

``table_interp`` / ``table_interp_adjoint`` keep the signatures and errors of
``torchkbnufft/_nufft/interp.py:315-403`` and ``:587-726``; the work is one HIP
launch each (``tkbn_interp_fwd`` / ``tkbn_interp_adj``) instead of J^d Python-level
iterations of ATen gathers and ``index_add_`` calls.
"""
from __future__ import annotations

import ctypes
from collections import OrderedDict
from typing import List, Optional, Sequence

import torch
from torch import Tensor

from .. import _lib
from ._host import host_values, real_dtype_of, require_gpu

# adjoint mode used when the caller does not ask for one
_DEFAULT_ADJ_MODE = _lib.TKBN_ADJ_ATOMIC
# samples x rows from which the forward uses the binned, LDS-staged gather
_BINNED_FWD_MIN_WORK = 1 << 16


# Interpolation workspaces kept between calls: the library stores the trajectory plan
# in them and reuses it while omega, the tables and the geometry stay bitwise equal
# (checked on the device).  Keyed by (kind, device, stream, trajectory tensor); the
# key only picks a buffer, correctness never depends on it.
_WORKSPACES: "OrderedDict[tuple, torch.Tensor]" = OrderedDict()
_WORKSPACES_MAX = 256  # entries
_WORKSPACES_MAX_BYTES = 2 << 30  # total bytes held


def _plan_workspace(kind: str, omega: Tensor, like: Tensor, nbytes: int) -> Tensor:
    key = (kind, like.device, _lib.stream_of(like), omega.data_ptr(), tuple(omega.shape), omega.dtype)
    ws = _WORKSPACES.get(key)
    if ws is None or ws.numel() < nbytes:
        ws = torch.empty(max(nbytes, _lib.PLAN_CACHE_HEADER_BYTES), dtype=torch.uint8, device=like.device)
        ws[: _lib.PLAN_CACHE_HEADER_BYTES].zero_()  # a zero header never matches: first call builds
        _WORKSPACES[key] = ws
        while len(_WORKSPACES) > 1 and (
            len(_WORKSPACES) > _WORKSPACES_MAX or sum(w.numel() for w in _WORKSPACES.values()) > _WORKSPACES_MAX_BYTES
        ):
            _WORKSPACES.popitem(last=False)
    else:
        _WORKSPACES.move_to_end(key)
    return ws


def clear_plan_cache() -> None:
    """Release the workspaces that hold cached trajectory plans."""
    _WORKSPACES.clear()


def set_binned_forward_threshold(work: int) -> None:
    """Samples x rows from which ``table_interp`` bins samples (0: always, huge: never)."""
    global _BINNED_FWD_MIN_WORK
    _BINNED_FWD_MIN_WORK = int(work)


def set_deterministic(flag: bool) -> None:
    """Select the fixed-order (bitwise repeatable) adjoint for subsequent calls."""
    global _DEFAULT_ADJ_MODE
    _DEFAULT_ADJ_MODE = _lib.TKBN_ADJ_DETERMINISTIC if flag else _lib.TKBN_ADJ_ATOMIC


def deterministic_enabled() -> bool:
    return _DEFAULT_ADJ_MODE == _lib.TKBN_ADJ_DETERMINISTIC


def _check_omega(omega: Tensor, batch: int, what: str) -> Tensor:
    if omega.ndim not in (2, 3):
        raise ValueError("omega must have 2 or 3 dimensions.")
    if omega.ndim == 3 and omega.shape[0] == 1:
        omega = omega[0]  # broadcast a single trajectory
    if omega.ndim == 3 and omega.shape[0] != batch:
        raise ValueError(f"If omega has batch dim, omega batch dimension must match {what}.")
    return omega


def make_desc(
    data: Tensor,
    omega: Tensor,
    tables: List[Tensor],
    n_shift,
    numpoints,
    table_oversamp,
    grid_size: Sequence[int],
    batch: int,
    coils: int,
) -> _lib.InterpDesc:
    d = len(grid_size)
    if omega.shape[-2] != d:
        raise ValueError(f"omega has {omega.shape[-2]} rows but the grid has {d} dimensions")
    if len(tables) != d:
        raise ValueError(f"{len(tables)} tables for a {d}-dimensional grid")
    jj = host_values(numpoints)
    ll = host_values(table_oversamp)
    ss = host_values(n_shift)
    desc = _lib.InterpDesc()
    desc.dtype = _lib.dtype_code(data)
    desc.ndim = d
    for i in range(d):
        desc.grid_size[i] = int(grid_size[i])
        desc.numpoints[i] = int(jj[i])
        desc.table_oversamp[i] = int(ll[i])
        desc.n_shift[i] = float(ss[i])
        t = tables[i]
        if t.dtype != data.dtype or t.device != data.device:
            raise TypeError(f"table {i} is {t.dtype} on {t.device}; data is {data.dtype} on {data.device}")
        if t.numel() != int(jj[i]) * int(ll[i]) + 1:
            raise ValueError(f"table {i} has {t.numel()} entries, expected numpoints*table_oversamp+1")
        desc.tables[i] = t.data_ptr()
    desc.batch = batch
    desc.coils = coils
    desc.klength = omega.shape[-1]
    desc.omega_batched = 1 if omega.ndim == 3 else 0
    desc.flags = 0
    return desc


class PlanHolder:
    """Caller-owned interpolation workspace (one per problem and direction).

    The library keeps the trajectory plan in it.  With ``trusted = True`` the owner
    promises that the trajectory, tables and geometry never change after the first
    call through this holder, and later calls skip all plan work (TKBN_PLAN_REUSE);
    otherwise the plan is checked on the device each call.  Use one stream per holder.
    """

    def __init__(self, trusted: bool = False, tile: Sequence[int] = (0, 0, 0), work_target: int = 0):
        self.ws: Optional[Tensor] = None
        self.built = False
        self.trusted = trusted
        # tuning of the sample-lane binned kernels (0: the library's choice); a change
        # makes the next call rebuild the plan
        self.tile = tuple(int(v) for v in tile)
        self.work_target = int(work_target)

    def tune(self, desc: _lib.InterpDesc) -> None:
        for i in range(3):
            desc.tile[i] = self.tile[i] if i < len(self.tile) else 0
        desc.work_target = self.work_target

    def set_tuning(self, tile: Sequence[int], work_target: int) -> None:
        tile = tuple(int(v) for v in tile)
        if tile != self.tile or int(work_target) != self.work_target:
            self.tile, self.work_target = tile, int(work_target)
            self.built = False

    def get(self, like: Tensor, nbytes: int) -> Tensor:
        if self.ws is None or self.ws.numel() < nbytes or self.ws.device != like.device:
            self.ws = torch.empty(max(nbytes, _lib.PLAN_CACHE_HEADER_BYTES), dtype=torch.uint8, device=like.device)
            self.ws[: _lib.PLAN_CACHE_HEADER_BYTES].zero_()
            self.built = False
        return self.ws

    def flags(self) -> int:
        return _lib.TKBN_PLAN_REUSE if (self.built and self.trusted) else 0


def _check_out(out: Optional[Tensor], shape: tuple, like: Tensor) -> Tensor:
    """``out`` if given (validated), else a new tensor of ``shape`` like ``like``."""
    if out is None:
        return torch.empty(shape, dtype=like.dtype, device=like.device)
    if tuple(out.shape) != tuple(shape) or out.dtype != like.dtype or out.device != like.device:
        raise ValueError(f"out must be a {tuple(shape)} {like.dtype} tensor on {like.device}")
    if not out.is_contiguous():
        raise ValueError("out must be contiguous")
    return out


def _prep_tables(tables: List[Tensor], like: Tensor) -> List[Tensor]:
    return [t.to(device=like.device, dtype=like.dtype).contiguous() for t in tables]


def _prep_omega(omega: Tensor, like: Tensor) -> Tensor:
    # the trajectory is evaluated in the data's real precision (tm, rint, floor)
    return omega.to(device=like.device, dtype=real_dtype_of(like)).contiguous()


def table_interp(
    image: Tensor,
    omega: Tensor,
    tables: List[Tensor],
    n_shift: Tensor,
    numpoints: Tensor,
    table_oversamp: Tensor,
    offsets: Tensor,
    min_kspace_per_fork: int = 1024,
    out: Optional[Tensor] = None,
    plan: Optional[PlanHolder] = None,
    binned: Optional[bool] = None,
) -> Tensor:
    """Interpolate gridded ``image`` (B, C, *K) to ``omega``; returns (B, C, M).

    ``out``: optional contiguous (B, C, M) tensor to write into.
    ``binned``: force (True) or forbid (False) the binned gather; None: by problem size.
    """
    omega = _check_omega(omega, image.shape[0], "image")
    require_gpu(image, omega)
    key_omega = omega
    image = image.contiguous()
    tables = _prep_tables(tables, image)
    omega = _prep_omega(omega, image)
    grid_size = tuple(image.shape[2:])
    B, C = image.shape[0], image.shape[1]
    desc = make_desc(image, omega, tables, n_shift, numpoints, table_oversamp, grid_size, B, C)
    kdata = _check_out(out, (B, C, omega.shape[-1]), image)
    lib = _lib.load()
    ws, nws = None, 0
    # binned LDS-staged gather once there is enough work to amortise the plan launches
    if plan is not None:
        plan.tune(desc)
    if binned if binned is not None else omega.shape[-1] * B * C >= _BINNED_FWD_MIN_WORK:
        nws = int(lib.tkbn_interp_fwd_workspace_bytes(ctypes.byref(desc)))
        if nws > 0:
            if plan is not None:
                ws = plan.get(image, nws)
                desc.flags = plan.flags()
            else:
                ws = _plan_workspace("fwd", key_omega, image, nws)
    with torch.cuda.device(image.device):
        st = lib.tkbn_interp_fwd(
            ctypes.byref(desc), image.data_ptr(), omega.data_ptr(), kdata.data_ptr(),
            ws.data_ptr() if ws is not None else None, ws.numel() if ws is not None else 0, _lib.stream_of(image),
        )
    _lib.check(st, "tkbn_interp_fwd")
    if plan is not None and ws is not None:
        plan.built = True
    return kdata


def table_interp_adjoint(
    data: Tensor,
    omega: Tensor,
    tables: List[Tensor],
    n_shift: Tensor,
    numpoints: Tensor,
    table_oversamp: Tensor,
    offsets: Tensor,
    grid_size,
    deterministic: Optional[bool] = None,
    out: Optional[Tensor] = None,
    plan: Optional[PlanHolder] = None,
) -> Tensor:
    """Spread off-grid ``data`` (B, C, M) onto a grid of ``grid_size``; returns (B, C, *K).

    ``out``: optional contiguous (B, C, *K) tensor to write into (it is fully overwritten).
    """
    omega = _check_omega(omega, data.shape[0], "data")
    require_gpu(data, omega)
    key_omega = omega
    data = data.contiguous()
    tables = _prep_tables(tables, data)
    omega = _prep_omega(omega, data)
    gs = tuple(int(v) for v in host_values(grid_size))
    B, C = data.shape[0], data.shape[1]
    if data.shape[-1] != omega.shape[-1]:
        raise ValueError(f"data has {data.shape[-1]} samples, omega has {omega.shape[-1]}")
    desc = make_desc(data, omega, tables, n_shift, numpoints, table_oversamp, gs, B, C)
    mode = _DEFAULT_ADJ_MODE if deterministic is None else (
        _lib.TKBN_ADJ_DETERMINISTIC if deterministic else _lib.TKBN_ADJ_ATOMIC
    )
    lib = _lib.load()
    grid = _check_out(out, (B, C) + gs, data)
    ws_bytes = lib.tkbn_interp_adj_workspace_bytes(ctypes.byref(desc), mode)
    if ws_bytes < 0:
        _lib.check(_lib.TKBN_ERR_INVALID, "tkbn_interp_adj_workspace_bytes")
    if plan is not None:
        plan.tune(desc)
        ws_bytes = lib.tkbn_interp_adj_workspace_bytes(ctypes.byref(desc), mode)
        if ws_bytes < 0:
            _lib.check(_lib.TKBN_ERR_INVALID, "tkbn_interp_adj_workspace_bytes")
        ws = plan.get(data, int(ws_bytes))
        desc.flags = plan.flags()
    else:
        ws = _plan_workspace("adj", key_omega, data, int(ws_bytes))
    with torch.cuda.device(data.device):
        st = lib.tkbn_interp_adj(
            ctypes.byref(desc),
            data.data_ptr(),
            omega.data_ptr(),
            grid.data_ptr(),
            mode,
            ws.data_ptr(),
            ws.numel(),
            _lib.stream_of(data),
        )
    _lib.check(st, "tkbn_interp_adj")
    if plan is not None:
        plan.built = True
    return grid


def table_interp_domega(
    grid: Tensor,
    grad_kdata: Tensor,
    omega: Tensor,
    tables: List[Tensor],
    n_shift: Tensor,
    numpoints: Tensor,
    table_oversamp: Tensor,
) -> Tensor:
    """Gradient of ``Re <grad_kdata, table_interp(grid, omega)>`` with respect to ``omega``.

    No reference counterpart (the reference returns ``None`` for the trajectory).
    Returns a tensor shaped and typed like ``omega``.
    """
    orig = omega
    om = _check_omega(omega, grid.shape[0], "image")
    require_gpu(grid, grad_kdata, om)
    grid = grid.contiguous()
    grad_kdata = grad_kdata.to(grid.dtype).contiguous()
    tables = _prep_tables(tables, grid)
    om = _prep_omega(om, grid)
    B, C = grid.shape[0], grid.shape[1]
    desc = make_desc(grid, om, tables, n_shift, numpoints, table_oversamp, tuple(grid.shape[2:]), B, C)
    out = torch.empty(om.shape, dtype=om.dtype, device=om.device)
    with torch.cuda.device(grid.device):
        st = _lib.load().tkbn_interp_domega(
            ctypes.byref(desc), grid.data_ptr(), grad_kdata.data_ptr(), om.data_ptr(), out.data_ptr(),
            _lib.stream_of(grid),
        )
    _lib.check(st, "tkbn_interp_domega")
    return out.reshape(orig.shape).to(orig.dtype)
