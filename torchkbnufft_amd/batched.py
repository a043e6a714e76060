"""Many independent table NUFFT problems of uneven sizes run as one unit.

No reference counterpart: with the reference such a workload is a Python loop of
``KbNufft`` / ``KbNufftAdjoint`` calls (``modules/kbnufft.py:125-410``), one per
problem.  Here

* problems with the same image size share one operator (tables, apodisation), and
  their oversampled FFTs run as one fused launch over a stacked buffer;
* every problem keeps its own trajectory plan in a workspace it owns; the plans are
  built by the first call and then reused without any plan kernels (the trajectories
  are fixed at construction), so a call only runs the gather / spread kernels of each
  problem, writing into its slice of the stacked grid (no copies);
* the problems of a size group are spread over a few side streams (fork after the
  group's FFT, join before the next stage), so the many small interpolation launches
  of uneven problems run concurrently instead of one after another;
* ``capture()`` records the whole forward + adjoint pair of all problems in a HIP
  graph (the fork/join becomes graph dependencies); ``replay()`` reruns it with one
  launch from the host.

Each result is the same computation as the per-problem module call (same kernels,
same plans), so it matches the reference within the same tolerance.
"""
from __future__ import annotations

from typing import List, Optional, Sequence

import torch
from torch import Tensor

from ._nufft import fft as _fft
from ._nufft import interp as _interp
from .modules import KbNufft


class BatchedKbNufft(torch.nn.Module):
    """Forward / adjoint NUFFTs of ``len(im_sizes)`` problems with fixed trajectories.

    Args:
        im_sizes: image size of each problem.
        ktrajs: trajectory of each problem, (ndim, M_p) real tensors in [-pi, pi).
        coils: coil images per problem (each problem's image is (1, coils, *im_size)).
        numpoints, table_oversamp, kbwidth, dtype: as for ``KbNufft``.
        device: device of all buffers.
        streams: side streams the problems of a group are dealt to (1: in order on the
            caller's stream).  A problem always uses the same stream.

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

    def _interp_args(self, g: int):
        op = self.ops[g]
        return op._tables(), op.n_shift, op.numpoints, op.table_oversamp, op.offsets

    def forward(self) -> List[Tensor]:
        """kdata[p] = A_p images[p] for every problem."""
        for g, members in enumerate(self.groups):
            op = self.ops[g]
            im, gs = self._im[g], self._gs[g]
            grid = _fft.fft_grid_fwd(self._img[g], None, op.scaling_coef, im, gs, 1.0, out=self._grid[g])
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
        for g, members in enumerate(self.groups):
            op = self.ops[g]
            im, gs = self._im[g], self._gs[g]
            tabs, n_shift, numpoints, L, offsets = self._interp_args(g)
            grid = self._grid[g]

            def adj(slot, p, grid=grid, tabs=tabs, n_shift=n_shift, numpoints=numpoints, L=L, offsets=offsets, gs=gs):
                _interp.table_interp_adjoint(
                    self.kdata[p], self.ktrajs[p], tabs, n_shift, numpoints, L, offsets, gs,
                    out=grid[slot : slot + 1], plan=self._adj_plans[p],
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
        warm-up runs on that stream and reused by the graph)."""
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
