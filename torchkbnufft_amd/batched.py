"""Many independent table NUFFT problems of uneven sizes run as one unit.

No reference counterpart: with the reference such a workload is a Python loop of
``KbNufft`` / ``KbNufftAdjoint`` calls (``modules/kbnufft.py:125-410``), one per
problem.  Here

* problems with the same image size share one operator (tables, apodisation), and
  their oversampled FFTs run as one fused launch over a stacked buffer;
* every problem keeps its own trajectory plan in workspaces it owns, built once at
  the first call (the trajectories are fixed at construction);
* when every problem takes the sample-lane kernels (complex64, fewer than 32 coil
  rows), the interpolations of ALL problems run as one gather launch and one spread
  launch (``tkbn_batch_interp_fwd`` / ``_adj``): a device table holds each problem's
  geometry, plans and buffers, and each block finds its (problem, work item) from
  offsets computed on the device.  Otherwise each problem is its own launch, dealt
  over a few side streams;
* the tile edge and work-item target of each image size are picked by the
  autotuner (``_nufft/autotune.py``) on one problem of that size;
* ``capture()`` records the whole forward + adjoint pair of all problems in a HIP
  graph; ``replay()`` reruns it with one launch from the host.

Each result is the same computation as the per-problem module call (same weights,
same sums; the float summation order of the spread may differ), so it matches the
reference within the same tolerance.
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Optional, Sequence, Tuple

import torch
from torch import Tensor

from . import _lib
from ._nufft import autotune as _autotune
from ._nufft import fft as _fft
from ._nufft import interp as _interp
from .modules import KbNufft


class _BatchHandle:
    """Owner of a ``tkbn_batch`` (the library object) and of its device table."""

    def __init__(self, ptr: int, table: Tensor, workspace: Tensor, keep: list):
        self.ptr = ptr
        self.table = table
        self.workspace = workspace
        self.keep = keep  # tensors whose device pointers the table holds

    def __del__(self):
        if self.ptr:
            try:
                _lib.load().tkbn_batch_destroy(self.ptr)
            except Exception:  # interpreter shutdown
                pass
            self.ptr = 0


class BatchedKbNufft(torch.nn.Module):
    """Forward / adjoint NUFFTs of ``len(im_sizes)`` problems with fixed trajectories.

    Args:
        im_sizes: image size of each problem.
        ktrajs: trajectory of each problem, (ndim, M_p) real tensors in [-pi, pi).
        coils: coil images per problem (each problem's image is (1, coils, *im_size)).
        numpoints, table_oversamp, kbwidth, dtype: as for ``KbNufft``.
        device: device of all buffers.
        streams: side streams the problems of a group are dealt to when they run as
            separate launches (1: in order on the caller's stream).
        fused: one launch per stage for all problems (True: required, False: never,
            None: when every problem qualifies, i.e. fewer than 32 coil rows).
        autotune: pick the tile edge and work-item target per image size by timing
            the candidates on one problem of that size (choices cached per geometry).
        deterministic: adjoint mode of every problem (None: the global default when the
            plans are built).

    Buffers (write inputs into them in place; outputs are overwritten):
        ``images[p]``  : (1, coils, *im_size_p) forward input
        ``kdata[p]``   : (1, coils, M_p) forward output / adjoint input
        ``adjoint_images[p]`` : (1, coils, *im_size_p) adjoint output
    """

    def __init__(
        self,
        im_sizes: Sequence[Sequence[int]],
        ktrajs: Sequence[Tensor],
        coils: int,
        numpoints: int = 6,
        table_oversamp: int = 2**10,
        kbwidth: float = 2.34,
        dtype: torch.dtype = torch.complex64,
        device: Optional[torch.device] = None,
        streams: int = 4,
        fused: Optional[bool] = None,
        autotune: bool = True,
        deterministic: Optional[bool] = None,
    ):
        super().__init__()
        if len(im_sizes) != len(ktrajs):
            raise ValueError("one trajectory per problem is needed")
        if dtype != torch.complex64:
            raise TypeError("BatchedKbNufft is complex64 only")
        device = torch.device(device) if device is not None else torch.device("cuda")
        self.sizes = [tuple(int(v) for v in s) for s in im_sizes]
        self.coils = int(coils)
        uniq = sorted(set(self.sizes))
        self.ops = torch.nn.ModuleList(
            [
                KbNufft(im_size=s, numpoints=numpoints, table_oversamp=table_oversamp, kbwidth=kbwidth, dtype=dtype).to(device)
                for s in uniq
            ]
        )
        # group g: problems with image size uniq[g], in problem order
        self.groups: List[List[int]] = [[p for p, s in enumerate(self.sizes) if s == u] for u in uniq]
        self.where = [0] * len(self.sizes)  # problem -> (group, slot)
        for g, members in enumerate(self.groups):
            for slot, p in enumerate(members):
                self.where[p] = (g, slot)
        # private copies: the plans are built once and trusted, so a caller's later
        # in-place edit of its trajectory must not reach them
        self.ktrajs = [
            k.detach().to(device=device, dtype=torch.float32).clone(memory_format=torch.contiguous_format) for k in ktrajs
        ]
        for k, s in zip(self.ktrajs, self.sizes):
            if k.ndim != 2 or k.shape[0] != len(s):
                raise ValueError(f"trajectory of shape {tuple(k.shape)} for a {len(s)}-d image")
        C = self.coils
        self._img = []
        self._grid = []
        self._adj = []
        for g, members in enumerate(self.groups):
            im = uniq[g]
            gs = tuple(int(v) for v in self.ops[g].grid_size.tolist())
            P = len(members)
            self._img.append(torch.zeros((P, C) + im, dtype=dtype, device=device))
            self._grid.append(torch.zeros((P, C) + gs, dtype=dtype, device=device))
            self._adj.append(torch.zeros((P, C) + im, dtype=dtype, device=device))
        self.images = [self._img[g][slot : slot + 1] for (g, slot) in self.where]
        self.adjoint_images = [self._adj[g][slot : slot + 1] for (g, slot) in self.where]
        self.kdata = [torch.zeros((1, C, k.shape[1]), dtype=dtype, device=device) for k in self.ktrajs]
        # host copies of the group geometry (no device reads while a graph is captured)
        self._im = [tuple(int(v) for v in op.im_size.tolist()) for op in self.ops]
        self._gs = [tuple(int(v) for v in op.grid_size.tolist()) for op in self.ops]
        # per-problem plans; the trajectories are fixed, so once built they are reused
        # without checks
        self._fwd_plans = [_interp.PlanHolder(trusted=True) for _ in self.sizes]
        self._adj_plans = [_interp.PlanHolder(trusted=True) for _ in self.sizes]
        self._graph: Optional[torch.cuda.CUDAGraph] = None
        self._stream: Optional[torch.cuda.Stream] = None
        # side streams (default priority), created on first use on this device
        self._nside = max(1, int(streams))
        self._side: List[torch.cuda.Stream] = []
        self._device = device
        self._want_fused = fused
        self._autotune = bool(autotune)
        self._det = deterministic
        self._built = False
        self._batch: Optional[_BatchHandle] = None
        # per image size: {"fwd": (tile, target), "adj": (tile, target)} (per-problem tuner)
        self.tuning: Dict[Tuple[int, ...], Dict[str, tuple]] = {}
        # the setting the fused batch runs with, and the batch times that chose it
        self.batch_tuning: Dict[str, object] = {}

    # ------------------------------------------------------------------ set-up
    def _interp_args(self, g: int):
        op = self.ops[g]
        return op._tables(), op.n_shift, op.numpoints, op.table_oversamp, op.offsets

    def _tune(self, det: bool) -> None:
        """Autotune each size group on its median-length problem (sample-lane kernels
        only: from 32 coil rows the multi-coil kernels run, which take no tuning)."""
        for g, members in enumerate(self.groups):
            if self.coils >= 32:
                break
            tabs, n_shift, numpoints, L, offsets = self._interp_args(g)
            p = sorted(members, key=lambda q: self.ktrajs[q].shape[1])[len(members) // 2]
            grid = self._grid[g][:1]
            fwd = _autotune.tune("fwd", grid, self.ktrajs[p], tabs, n_shift, numpoints, L, offsets, self._gs[g])
            adj = _autotune.tune("adj", self.kdata[p], self.ktrajs[p], tabs, n_shift, numpoints, L, offsets,
                                 self._gs[g], deterministic=det)
            self.tuning[self.sizes[p]] = {"fwd": fwd, "adj": adj}
            for q in members:
                self._fwd_plans[q].set_tuning(*fwd)
                self._adj_plans[q].set_tuning(*adj)

    def _build_plans(self, det: bool, fused: bool) -> None:
        """Every problem's plans: one ordinary call per problem and direction, into scratch
        outputs, through the problem's plan holders (and their tuning)."""
        for g, members in enumerate(self.groups):
            tabs, n_shift, numpoints, L, offsets = self._interp_args(g)
            gs = self._gs[g]
            scratch_grid = torch.zeros((1, self.coils) + gs, dtype=self._grid[g].dtype, device=self._device)
            for p in members:
                _interp.table_interp(
                    scratch_grid, self.ktrajs[p], tabs, n_shift, numpoints, L, offsets,
                    out=torch.empty_like(self.kdata[p]), plan=self._fwd_plans[p], binned=True if fused else None,
                )
                _interp.table_interp_adjoint(
                    self.kdata[p], self.ktrajs[p], tabs, n_shift, numpoints, L, offsets, gs, deterministic=det,
                    out=scratch_grid, plan=self._adj_plans[p],
                )

    def _build(self) -> None:
        """Build every problem's plans, then the batch table when the problems qualify."""
        det = _interp.deterministic_enabled() if self._det is None else bool(self._det)
        self._det = det
        if self._autotune:
            self._tune(det)
        fused = self._want_fused is not False and self.coils < 32
        if self._want_fused and self.coils >= 32:
            raise NotImplementedError("fused batches take fewer than 32 coil rows per problem")
        self._build_plans(det, fused)
        if fused:
            self._batch = self._make_batch(det, required=bool(self._want_fused))
            if self._batch is not None and self._autotune:
                self._tune_batch(det)
        self._built = True

    def _run_batch(self, direction: str) -> None:
        b = self._batch
        fn = _lib.load().tkbn_batch_interp_fwd if direction == "fwd" else _lib.load().tkbn_batch_interp_adj
        with torch.cuda.device(self._device):
            st = fn(b.ptr, b.table.data_ptr(), b.workspace.data_ptr(), b.workspace.numel(), _lib.stream_of(b.table))
        _lib.check(st, f"tkbn_batch_interp_{direction}")

    def _time_batch(self, reps: int = 5) -> float:
        """Median ms of one batched gather + spread (HIP events, current stream)."""
        ts = []
        for i in range(reps + 2):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            self._run_batch("fwd")
            self._run_batch("adj")
            b.record()
            b.synchronize()
            if i >= 2:
                ts.append(a.elapsed_time(b))
        ts.sort()
        return ts[len(ts) // 2]

    def _tune_batch(self, det: bool) -> None:
        """Per-problem choices can lose inside one launch (small work items multiply its
        blocks), so the batch itself is timed under the per-problem choices and under a
        few settings shared by all problems (library tiles, work targets 0 = the
        library's), and the fastest is kept (another must be >= 3 % faster)."""
        n = len(self.sizes)
        own = [(self._fwd_plans[p].tile, self._fwd_plans[p].work_target, self._adj_plans[p].tile,
                self._adj_plans[p].work_target) for p in range(n)]
        required = bool(self._want_fused)
        configs = [("per-problem", None)] + [
            (f"shared: gather target {wf or 'default'}, spread target {wa or 'default'}", (wf, wa))
            for wf in (0, 2048) for wa in (0, 1024, 2048)
        ]
        results = []
        best = None
        saved = [k.clone() for k in self.kdata]  # the timed gathers overwrite kdata (an adjoint's input)
        for name, cfg in configs:
            if cfg is not None:
                for p in range(n):
                    self._fwd_plans[p].set_tuning((0, 0, 0), cfg[0])
                    self._adj_plans[p].set_tuning((0, 0, 0), cfg[1])
                self._build_plans(det, True)
                self._batch = self._make_batch(det, required)
                if self._batch is None:
                    continue
            ms = self._time_batch()
            results.append((name, round(ms, 4)))
            if best is None or ms < 0.97 * best[1]:
                best = (name, ms, cfg)
        name, ms, cfg = best
        for p in range(n):
            if cfg is None:
                self._fwd_plans[p].set_tuning(own[p][0], own[p][1])
                self._adj_plans[p].set_tuning(own[p][2], own[p][3])
            else:
                self._fwd_plans[p].set_tuning((0, 0, 0), cfg[0])
                self._adj_plans[p].set_tuning((0, 0, 0), cfg[1])
        self._build_plans(det, True)
        self._batch = self._make_batch(det, required)
        for k, v in zip(self.kdata, saved):
            k.copy_(v)
        self.batch_tuning = {"choice": name, "ms": round(ms, 4), "tried": results}

    def _make_batch(self, det: bool, required: bool) -> Optional[_BatchHandle]:
        lib = _lib.load()
        n = len(self.sizes)
        descs = (_lib.InterpDesc * n)()
        ptrs = {k: (ctypes.c_void_p * n)() for k in ("gi", "ko", "ki", "go", "om", "wf", "wa")}
        keep = []
        for p in range(n):
            g, slot = self.where[p]
            tabs, n_shift, numpoints, L, _ = self._interp_args(g)
            grid = self._grid[g][slot : slot + 1]
            tabs = _interp._prep_tables(tabs, grid)
            keep.extend(tabs)
            d = _interp.make_desc(grid, self.ktrajs[p], tabs, n_shift, numpoints, L, self._gs[g], 1, self.coils)
            self._fwd_plans[p].tune(d)
            descs[p] = d
            fw, aw = self._fwd_plans[p].ws, self._adj_plans[p].ws
            if fw is None or aw is None:  # the forward took the unbinned gather
                if required:
                    raise NotImplementedError(f"problem {p} has no binned forward plan")
                return None
            ptrs["gi"][p] = grid.data_ptr()
            ptrs["ko"][p] = self.kdata[p].data_ptr()
            ptrs["ki"][p] = self.kdata[p].data_ptr()
            ptrs["go"][p] = grid.data_ptr()
            ptrs["om"][p] = self.ktrajs[p].data_ptr()
            ptrs["wf"][p] = fw.data_ptr()
            ptrs["wa"][p] = aw.data_ptr()
        # the adjoint descriptors differ from the forward ones in their tuning only
        adescs = (_lib.InterpDesc * n)()
        for p in range(n):
            adescs[p] = descs[p]
            self._adj_plans[p].tune(adescs[p])
        handle = ctypes.c_void_p()
        mode = _lib.TKBN_ADJ_DETERMINISTIC if det else _lib.TKBN_ADJ_ATOMIC
        st = lib.tkbn_batch_create(descs, adescs, n, mode, ptrs["gi"], ptrs["ko"], ptrs["ki"], ptrs["go"], ptrs["om"],
                                   ptrs["wf"], ptrs["wa"], ctypes.byref(handle))
        if st == _lib.TKBN_ERR_UNSUPPORTED and not required:
            return None
        _lib.check(st, "tkbn_batch_create")
        try:
            nbytes = int(lib.tkbn_batch_table_bytes(handle))
            host = torch.empty(nbytes, dtype=torch.uint8, pin_memory=True)
            _lib.check(lib.tkbn_batch_write_table(handle, host.data_ptr()), "tkbn_batch_write_table")
            table = host.to(self._device)
            wsb = int(lib.tkbn_batch_workspace_bytes(handle))
            workspace = torch.empty(max(wsb, 16), dtype=torch.uint8, device=self._device)
        except Exception:
            lib.tkbn_batch_destroy(handle)
            raise
        return _BatchHandle(handle.value, table, workspace, keep)

    @property
    def fused(self) -> bool:
        """True when the interpolations run as one launch per stage (after the first call)."""
        return self._batch is not None

    # ------------------------------------------------------------------ per-problem launches
    def _streams(self) -> List[torch.cuda.Stream]:
        if self._nside > 1 and not self._side:
            self._side = [torch.cuda.Stream(device=self._device) for _ in range(self._nside)]
        return self._side

    def _fanout(self, members: List[int], work) -> None:
        """work(slot, p) for every problem of a group: problem p on side stream
        p mod S (always the same one, as its plan holders require), after the caller's
        stream; the caller's stream then waits for all side streams."""
        side = self._streams()
        if not side:
            for slot, p in enumerate(members):
                work(slot, p)
            return
        main = torch.cuda.current_stream()
        ready = torch.cuda.Event()
        ready.record(main)
        for st in side:
            st.wait_event(ready)
        for slot, p in enumerate(members):
            with torch.cuda.stream(side[p % len(side)]):
                work(slot, p)
        for st in side:
            done = torch.cuda.Event()
            done.record(st)
            main.wait_event(done)

    # ------------------------------------------------------------------ operators
    def forward(self) -> List[Tensor]:
        """kdata[p] = A_p images[p] for every problem."""
        if not self._built:
            self._build()
        grids = []
        for g in range(len(self.groups)):
            op = self.ops[g]
            grids.append(_fft.fft_grid_fwd(self._img[g], None, op.scaling_coef, self._im[g], self._gs[g], 1.0,
                                           out=self._grid[g]))
        if self._batch is not None:
            self._run_batch("fwd")
            return self.kdata
        for g, members in enumerate(self.groups):
            grid = grids[g]
            tabs, n_shift, numpoints, L, offsets = self._interp_args(g)

            def fwd(slot, p, grid=grid, tabs=tabs, n_shift=n_shift, numpoints=numpoints, L=L, offsets=offsets):
                _interp.table_interp(
                    grid[slot : slot + 1], self.ktrajs[p], tabs, n_shift, numpoints, L, offsets,
                    out=self.kdata[p], plan=self._fwd_plans[p],
                )

            self._fanout(members, fwd)
        return self.kdata

    def adjoint(self) -> List[Tensor]:
        """adjoint_images[p] = A_p^H kdata[p] for every problem."""
        if not self._built:
            self._build()
        if self._batch is not None:
            self._run_batch("adj")
        for g, members in enumerate(self.groups):
            op = self.ops[g]
            im, gs = self._im[g], self._gs[g]
            grid = self._grid[g]
            if self._batch is None:
                tabs, n_shift, numpoints, L, offsets = self._interp_args(g)

                def adj(slot, p, grid=grid, tabs=tabs, n_shift=n_shift, numpoints=numpoints, L=L, offsets=offsets,
                        gs=gs):
                    _interp.table_interp_adjoint(
                        self.kdata[p], self.ktrajs[p], tabs, n_shift, numpoints, L, offsets, gs,
                        deterministic=self._det, out=grid[slot : slot + 1], plan=self._adj_plans[p],
                    )

                self._fanout(members, adj)
            _fft.fft_grid_adj(grid, None, op.scaling_coef, im, 1.0, out=self._adj[g])
        return self.adjoint_images

    def pair(self) -> List[Tensor]:
        """adjoint_images = A^H A images (forward then adjoint of every problem)."""
        self.forward()
        return self.adjoint()

    def capture(self, warmup: int = 2) -> None:
        """Record ``pair()`` in a HIP graph on a side stream (plans are built by the
        warm-up runs, or earlier calls, and reused by the graph)."""
        self._stream = torch.cuda.Stream()
        self._stream.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(self._stream):
            for _ in range(warmup):
                self.pair()
        torch.cuda.current_stream().wait_stream(self._stream)
        self._graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self._graph, stream=self._stream):
            self.pair()

    def replay(self) -> List[Tensor]:
        """Rerun the captured pair on the current inputs."""
        if self._graph is None:
            raise RuntimeError("call capture() first")
        self._graph.replay()
        return self.adjoint_images
