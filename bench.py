#!/usr/bin/env python
"""Throughput of the SENSE table NUFFT forward + adjoint pair on one MI355X.

Workload (one "step" = one forward + one adjoint NUFFT of the whole batch):
  batch 4, 16 coils, 320x320 complex64 images, 640x640 grid (2x), numpoints 6,
  table oversampling 1024; golden-angle radial trajectory of 400 spokes x 640
  samples (256000 points); sensitivity maps from complex N(0,1) low-passed by a
  periodic Gaussian (sigma 20 px) and normalised to unit root-sum-of-squares.
  The adjoint is applied to the forward's output (A^H A x).

Prints one JSON line (rank 0, one GPU).  `value` is pairs/s with all inputs
resident in HBM; `roofline` is for the stage that takes the most time, from HIP
events recorded on the stream the kernels run on; `cpu_baseline` times the CPU
oracle (a restatement of the reference table path) on this host's cores.
"""
from __future__ import annotations

import argparse
import itertools
import json
import math
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
B, C, N, K, J, L = 4, 16, 320, 640, 6, 1024
NSPOKES, NREAD = 400, 640
METRIC = "fwd+adj NUFFT/s, 16-coil 320² radial; % HBM peak; batched & grad vs plain"


def radial_golden(nspokes: int, nread: int) -> torch.Tensor:
    r = -np.pi + 2 * np.pi * np.arange(nread) / nread
    th = np.arange(nspokes) * np.pi / ((1 + np.sqrt(5)) / 2)
    ky = (np.sin(th)[:, None] * r[None, :]).reshape(-1)
    kx = (np.cos(th)[:, None] * r[None, :]).reshape(-1)
    return torch.tensor(np.stack([ky, kx]), dtype=torch.float32)


def make_smaps(gen: torch.Generator, dev) -> torch.Tensor:
    s = torch.complex(torch.randn(B, C, N, N, generator=gen), torch.randn(B, C, N, N, generator=gen)).to(dev)
    f = torch.fft.fftfreq(N, device=dev)
    # periodic Gaussian low-pass, sigma = 20 px in image space -> exp(-2 pi^2 sigma^2 f^2)
    g1 = torch.exp(-2 * (math.pi**2) * (20.0**2) * f**2)
    s = torch.fft.ifft2(torch.fft.fft2(s) * (g1[:, None] * g1[None, :]))
    rss = torch.sqrt(torch.sum(torch.abs(s) ** 2, dim=1, keepdim=True))
    return (s / rss).to(torch.complex64).contiguous()


def footprint_cells(ktraj: np.ndarray, grid: tuple, numpoints: int) -> int:
    """Grid cells inside some sample's J-neighbourhood (the kernel's base-offset rule)."""
    d = len(grid)
    kk = np.asarray(grid)
    gam = (np.float32(1) / kk.astype(np.float32)) * np.float32(2 * np.pi)
    tm = ktraj.astype(np.float32) / gam[:, None].astype(np.float32)
    base = 1 + np.floor(tm - np.float32(numpoints / 2)).astype(np.int64)
    mark = np.zeros(grid, dtype=bool)
    for off in itertools.product(range(numpoints), repeat=d):
        mark[tuple(np.mod(base[i] + off[i], kk[i]) for i in range(d))] = True
    return int(mark.sum())


def stage_bytes(F: int, M: int) -> dict:
    """Algorithmic bytes per stage (the FFT stages include the model's FFT read+write)."""
    P, G, BC = N * N, K * K, B * C
    return {
        "fft_fwd": B * P * 8 + BC * P * 8 + BC * G * 8 + 2 * BC * G * 8,
        "interp": BC * F * 8 + 2 * M * 4 + BC * M * 8,
        "spread": BC * M * 8 + 2 * M * 4 + BC * G * 8,
        "fft_adj": 2 * BC * G * 8 + BC * P * 8 + BC * P * 8 + B * P * 8,
    }


def time_stages(ops, x, ktraj, smaps, iters: int):
    """Average duration of each pipeline stage (the same calls the modules make) and of
    the two binned interpolation kernels, from HIP events on the current stream, which
    is the stream the kernels run on (the library records the kernel events itself,
    tkbn_profile_next_kernel)."""
    from torchkbnufft_amd import _lib
    from torchkbnufft_amd._nufft import fft as _fft
    from torchkbnufft_amd._nufft import interp as _interp

    fw, ad = ops
    tabs = fw._tables()
    lib = _lib.load()
    fused = _fft.fused_supported(x.dtype, (N, N), (K, K))
    names = ["fft_fwd", "interp", "spread", "fft_adj"]
    kern = ["gather_sep_kernel", "spread_sep_kernel"]
    acc = {n: 0.0 for n in names + kern}

    def run(ev, kev):
        ev[0].record()
        if fused:
            g = _fft.fft_grid_fwd(x, smaps, fw.scaling_coef, (N, N), (K, K), 1.0)
        else:
            g = torch.fft.fftn(_fft.apod_pad(x, smaps, fw.scaling_coef, (N, N), (K, K), 1.0), dim=(-2, -1))
        ev[1].record()
        lib.tkbn_profile_next_kernel(kev[0].cuda_event, kev[1].cuda_event)
        kd = _interp.table_interp(g, ktraj, tabs, fw.n_shift, fw.numpoints, fw.table_oversamp, fw.offsets)
        ev[2].record()
        lib.tkbn_profile_next_kernel(kev[2].cuda_event, kev[3].cuda_event)
        g2 = _interp.table_interp_adjoint(kd, ktraj, tabs, fw.n_shift, fw.numpoints, fw.table_oversamp, fw.offsets, (K, K))
        ev[3].record()
        if fused:
            _fft.fft_grid_adj(g2, smaps, fw.scaling_coef, (N, N), 1.0)
        else:
            _fft.crop_deapod(torch.fft.ifftn(g2, dim=(-2, -1), norm="forward"), smaps, fw.scaling_coef, (N, N), 1.0)
        ev[4].record()

    def events(n):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(n)]
        for e in ev:
            e.record()  # torch creates the HIP event on first record; the library re-records it
        return ev

    run(events(5), events(4))  # warm (plans, caches)
    torch.cuda.synchronize()
    for _ in range(iters):
        ev, kev = events(5), events(4)
        run(ev, kev)
        torch.cuda.synchronize()
        for i, n in enumerate(names):
            acc[n] += ev[i].elapsed_time(ev[i + 1]) * 1e3
        acc[kern[0]] += kev[0].elapsed_time(kev[1]) * 1e3
        acc[kern[1]] += kev[2].elapsed_time(kev[3]) * 1e3
    stages = {n: acc[n] / iters for n in names}
    kernels = {n: acc[n] / iters for n in kern}
    return stages, kernels


def plan_build_us(fw, x, ktraj, smaps, reps: int = 5) -> dict:
    """Trajectory-plan cost: a forward and an adjoint interpolation on a trajectory the
    library has not seen (the bins, permutation and work lists are built) against the
    same call again (plan reused after the on-device check), HIP events on the current
    stream, median over `reps` new trajectories (the radial one scaled by 1 - r 1e-6)."""
    from torchkbnufft_amd._nufft import fft as _fft
    from torchkbnufft_amd._nufft import interp as _interp

    tabs = fw._tables()
    g = _fft.fft_grid_fwd(x, smaps, fw.scaling_coef, (N, N), (K, K), 1.0)
    args = (tabs, fw.n_shift, fw.numpoints, fw.table_oversamp, fw.offsets)
    kd = _interp.table_interp(g, ktraj, *args)
    res = {"fwd_first": [], "fwd_reused": [], "adj_first": [], "adj_reused": []}
    for r in range(reps):
        # values differ from every earlier call's, so the on-device check sees a new
        # trajectory even where the allocator hands back an earlier copy's address
        om = ktraj * (1.0 - 1e-6 * (r + 1))
        for name, fn in (("fwd", lambda: _interp.table_interp(g, om, *args, out=kd)),
                         ("adj", lambda: _interp.table_interp_adjoint(kd, om, *args, (K, K), out=g))):
            for when in ("first", "reused"):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                fn()
                e1.record()
                torch.cuda.synchronize()
                res[f"{name}_{when}"].append(e0.elapsed_time(e1) * 1e3)
        del om
    med = {k: float(np.median(v)) for k, v in res.items()}
    return {
        "fwd_call_first_us": med["fwd_first"],
        "fwd_call_reused_us": med["fwd_reused"],
        "adj_call_first_us": med["adj_first"],
        "adj_call_reused_us": med["adj_reused"],
        "plan_build_us": med["fwd_first"] - med["fwd_reused"] + med["adj_first"] - med["adj_reused"],
    }


def cpu_baseline(x, ktraj, smaps, seconds: float) -> dict:
    from oracle import table_path as tp

    n = len(os.sched_getaffinity(0))
    torch.set_num_threads(n)
    plan = tp.Plan((N, N), dtype=torch.complex64)
    xc, kc, sc = x.cpu(), ktraj.cpu(), smaps.cpu()
    # warm-up on one batch entry
    tp.kb_table_nufft_adjoint(plan, tp.kb_table_nufft(plan, xc[:1], kc, sc[:1]), kc, sc[:1])
    runs, t0 = 0, time.perf_counter()
    while True:
        kd = tp.kb_table_nufft(plan, xc, kc, sc)
        tp.kb_table_nufft_adjoint(plan, kd, kc, sc)
        runs += 1
        el = time.perf_counter() - t0
        if el >= seconds or runs >= 5:
            break
    try:
        cpu_model = next(l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name"))
    except Exception:
        cpu_model = "unknown"
    return {
        "value": runs / el,
        "unit": "pairs/s",
        "cores": torch.get_num_threads(),
        "kind": "port",
        "sample": f"{runs} full pair(s) (B=4, C=16, all 256000 samples) of the oracle's complex64 "
        f"restatement of the reference table path, torch CPU ops, {torch.get_num_threads()} intra-op / "
        f"{torch.get_num_interop_threads()} inter-op threads, {cpu_model}",
    }


def vd_spiral(n_inter: int, n_read: int, turns: float = 16.0) -> torch.Tensor:
    """Variable-density spiral, r(t) = pi * t^1.5 (t = i / n_read), theta(t) = 2 pi turns t
    + 2 pi l / n_inter; rows (ky, kx); radius kept just below pi."""
    t = np.arange(n_read) / n_read
    r = np.pi * t**1.5 * 0.9999
    arms = []
    for l in range(n_inter):
        th = 2 * np.pi * turns * t + 2 * np.pi * l / n_inter
        arms.append(np.stack([r * np.sin(th), r * np.cos(th)]))
    return torch.tensor(np.concatenate(arms, axis=1), dtype=torch.float32)


def nudft_rel_err(y_sel, image, om_sel, n_shift) -> float:
    """Relative L2 error of GPU samples y_sel against the FP64 direct NUDFT of `image`
    (one coil, already multiplied by its map) at the trajectory columns om_sel."""
    from oracle import nudft

    ref = nudft.nudft(image.detach().cpu().numpy().astype(np.complex128), om_sel.double().cpu().numpy(), n_shift)
    y = y_sel.detach().cpu().numpy()
    return float(np.linalg.norm(y - ref) / np.linalg.norm(ref))


def timed(fn, iters: int) -> float:
    """Mean ms per call of fn (synchronised wall clock around `iters` calls)."""
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(iters):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / iters * 1e3


def bench_batched(dev, iters: int) -> dict:
    """64 uneven problems: image N in {128, 192, 256}^2, 8 coils, VD spiral of 8-48
    interleaves x 1024 samples; BatchedKbNufft (graph replay) vs the plain per-problem
    loop of KbNufft / KbNufftAdjoint calls."""
    import torchkbnufft_amd as tkbn

    gen = torch.Generator().manual_seed(3)
    nprob, coils = 64, 8
    sizes = [(int(s), int(s)) for s in torch.tensor([128, 192, 256])[torch.randint(0, 3, (nprob,), generator=gen)]]
    inter = torch.randint(8, 49, (nprob,), generator=gen).tolist()
    ktrajs = [vd_spiral(n, 1024) for n in inter]
    images = [torch.complex(torch.randn((1, coils) + s, generator=gen), torch.randn((1, coils) + s, generator=gen)).to(dev)
              for s in sizes]
    kt = [k.to(dev) for k in ktrajs]
    ops = {s: (tkbn.KbNufft(im_size=s, dtype=torch.complex64).to(dev), tkbn.KbNufftAdjoint(im_size=s, dtype=torch.complex64).to(dev))
           for s in sorted(set(sizes))}

    def loop():
        return [ops[s][1](ops[s][0](x, k), k) for s, x, k in zip(sizes, images, kt)]

    t_loop = timed(loop, iters)
    ref = loop()
    # separate launches (one per problem and stage, on 4 side streams) in a graph
    sep = tkbn.BatchedKbNufft(sizes, ktrajs, coils=coils, device=dev, fused=False, autotune=False)
    for p, x in enumerate(images):
        sep.images[p].copy_(x)
    sep.capture()
    t_sep_graph = timed(sep.replay, iters)
    del sep
    # one launch per stage for all problems, autotuned tiles
    bt = tkbn.BatchedKbNufft(sizes, ktrajs, coils=coils, device=dev)
    for p, x in enumerate(images):
        bt.images[p].copy_(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bt.pair()  # autotune + plans + batch table
    torch.cuda.synchronize()
    t_setup = (time.perf_counter() - t0) * 1e3
    t_eager = timed(bt.pair, iters)
    bt.capture()
    t_graph = timed(bt.replay, iters)
    out = bt.replay()
    torch.cuda.synchronize()
    diff = max(float(torch.linalg.norm(a - b) / torch.linalg.norm(b)) for a, b in zip(out, ref))
    # FP64 NUDFT error of the batched forward: the first problem of each size, coil 0,
    # every 37th sample
    fp64 = {}
    for s in sorted(set(sizes)):
        p = sizes.index(s)
        sel = torch.arange(0, kt[p].shape[1], 37, device=dev)
        fp64[f"{s[0]}x{s[1]}"] = nudft_rel_err(bt.kdata[p][0, 0, sel], images[p][0, 0], kt[p][:, sel],
                                                (s[0] // 2, s[1] // 2))
    return {
        "workload": "64 problems, N in {128,192,256}^2 (uniform), 8 coils, VD spiral 8-48 interleaves x 1024 "
        "(r = pi t^1.5, 16 turns), fwd + adj of every problem per step",
        "samples": int(sum(k.shape[1] for k in ktrajs)),
        "plain_loop_ms": t_loop,
        "separate_launches_graph_ms": t_sep_graph,
        "fused": bt.fused,
        "fused_setup_ms": t_setup,
        "batched_eager_ms": t_eager,
        "batched_graph_ms": t_graph,
        "speedup_graph_vs_loop": t_loop / t_graph,
        "autotune": {f"{k[0]}x{k[1]}": {d: {"tile": list(v[0][: len(k)]), "work_target": v[1]} for d, v in c.items()}
                     for k, c in bt.tuning.items()},
        "batch_tuning": bt.batch_tuning,
        "max_rel_diff_vs_loop": diff,
        "fp64_oracle_rel_err": fp64,
    }


def domega_torch(grid, gout, omega, tables, n_shift, numpoints: int, table_oversamp: int):
    """The trajectory gradient of Re<gout, interp(grid, omega)> in plain PyTorch ops on
    the same coil batch (2-D): the taps' grid values are gathered into one
    (rows, M, J*J) tensor, then reduced with the weights and the table derivative
    (quadratic through the three entries around round(u L), as the fused kernel
    evaluates it).  The unfused baseline of the fused trajectory-gradient kernel."""
    R = grid.shape[0] * grid.shape[1]
    K0, K1 = grid.shape[-2:]
    J, L = numpoints, table_oversamp
    center, tmax = (J * L) // 2, J * L
    g2 = grid.reshape(R, K0 * K1)
    taps = []
    for d, Kd in enumerate((K0, K1)):
        gam = (torch.tensor(1.0, dtype=torch.float32) / Kd) * torch.tensor(6.283185307179586, dtype=torch.float32)
        tm = omega[d] / gam.to(omega.device)
        base = torch.floor(tm - J / 2).to(torch.int64) + 1
        j = torch.arange(J, device=omega.device)
        gi = base[:, None] + j[None, :]
        ul = (tm[:, None] - gi.to(tm.dtype)) * L
        r = torch.round(ul)
        fr = ul - r
        ti = (r.to(torch.int64) + center).clamp(0, tmax)
        tab = tables[d]
        t0 = tab[ti]
        zero = torch.zeros((), dtype=tab.dtype, device=tab.device)
        tm1 = torch.where(ti > 0, tab[(ti - 1).clamp(min=0)], zero)
        tp1 = torch.where(ti < tmax, tab[(ti + 1).clamp(max=tmax)], zero)
        dt = ((tp1 - tm1) * 0.5 + fr * (tp1 - 2 * t0 + tm1)) * L
        taps.append((t0, dt, torch.remainder(gi, Kd)))
    (w0, dw0, i0), (w1, dw1, i1) = taps
    idx = (i0[:, :, None] * K1 + i1[:, None, :]).reshape(-1)          # (M * J * J)
    X = g2[:, idx].reshape(R, -1, J * J)                                # gathered taps
    W = (w0[:, :, None] * w1[:, None, :]).reshape(-1, J * J)
    D0 = (dw0[:, :, None] * w1[:, None, :]).reshape(-1, J * J)
    D1 = (w0[:, :, None] * dw1[:, None, :]).reshape(-1, J * J)
    S0 = (X * W).sum(-1)
    S = [(X * D0).sum(-1), (X * D1).sum(-1)]
    theta = omega[0] * float(n_shift[0]) + omega[1] * float(n_shift[1])
    q = torch.polar(torch.ones_like(theta), theta) * gout.reshape(R, -1).conj()
    out = []
    for d, Kd in enumerate((K0, K1)):
        gam = float((np.float32(1) / np.float32(Kd)) * np.float32(6.283185307179586))
        out.append((q * S[d]).real.sum(0) / gam - float(n_shift[d]) * (q * S0).imag.sum(0))
    return torch.stack(out)


def event_ms(fn, iters: int) -> float:
    """Median ms of `iters` calls of fn from HIP events on the current stream."""
    fn()
    times = []
    for _ in range(iters):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        times.append(a.elapsed_time(b))
    times.sort()
    return times[len(times) // 2]


def bench_grad(dev, iters: int) -> dict:
    """Gradients of ||A(x, ktraj) - y||^2 w.r.t. the image and the trajectory: 256^2, 12
    coils as SENSE maps, 64-interleave VD spiral x 1536 (98304 samples); one SENSE
    operator vs the plain loop of 12 single-coil operators (same gradients)."""
    import torchkbnufft_amd as tkbn

    gen = torch.Generator().manual_seed(4)
    n, coils = 256, 12
    ktraj = vd_spiral(64, 1536).to(dev)
    s = torch.complex(torch.randn(1, coils, n, n, generator=gen), torch.randn(1, coils, n, n, generator=gen)).to(dev)
    f = torch.fft.fftfreq(n, device=dev)
    g1 = torch.exp(-2 * (math.pi**2) * (20.0**2) * f**2)
    s = torch.fft.ifft2(torch.fft.fft2(s) * (g1[:, None] * g1[None, :]))
    smaps = (s / torch.sqrt(torch.sum(torch.abs(s) ** 2, dim=1, keepdim=True))).to(torch.complex64).contiguous()
    fw = tkbn.KbNufft(im_size=(n, n), dtype=torch.complex64).to(dev)
    x_true = torch.complex(torch.randn(1, 1, n, n, generator=gen), torch.randn(1, 1, n, n, generator=gen)).to(dev)
    with torch.no_grad():
        y = fw(x_true, ktraj, smaps=smaps)
        y = y + 0.01 * torch.complex(torch.randn(y.shape, generator=gen), torch.randn(y.shape, generator=gen)).to(dev)
    x0 = torch.complex(torch.randn(1, 1, n, n, generator=gen), torch.randn(1, 1, n, n, generator=gen)).to(dev)

    def grads(per_coil: bool):
        x = x0.clone().requires_grad_(True)
        k = ktraj.clone().requires_grad_(True)
        if per_coil:
            loss = sum(torch.sum(torch.abs(fw(x, k, smaps=smaps[:, c : c + 1]) - y[:, c : c + 1]) ** 2)
                       for c in range(coils))
        else:
            loss = torch.sum(torch.abs(fw(x, k, smaps=smaps) - y) ** 2)
        loss.backward()
        return x.grad, k.grad

    t_fused = timed(lambda: grads(False), iters)
    t_loop = timed(lambda: grads(True), iters)
    gx, gk = grads(False)
    lx, lk = grads(True)
    torch.cuda.synchronize()
    sel = torch.arange(0, ktraj.shape[1], 97, device=dev)
    with torch.no_grad():
        y0 = fw(x_true, ktraj, smaps=smaps[:, :1])[0, 0, sel]
    fp64 = nudft_rel_err(y0, (x_true * smaps[:, :1])[0, 0], ktraj[:, sel], (n // 2, n // 2))

    # The trajectory-gradient kernel alone (one launch) on the operator's oversampled
    # grid and an upstream gradient of the same coil batch, vs the same quantity from
    # plain PyTorch ops (taps gathered to memory, then reduced).
    from torchkbnufft_amd._nufft import interp as _interp

    K2 = 2 * n
    gridv = torch.complex(torch.randn(1, coils, K2, K2, generator=gen), torch.randn(1, coils, K2, K2, generator=gen))
    gridv = gridv.to(dev)
    gout = torch.complex(torch.randn(1, coils, ktraj.shape[1], generator=gen),
                         torch.randn(1, coils, ktraj.shape[1], generator=gen)).to(dev)
    tabs = fw._tables()
    dom = lambda: _interp.table_interp_domega(gridv, gout, ktraj, tabs, fw.n_shift, fw.numpoints,  # noqa: E731
                                              fw.table_oversamp)
    J, L = int(fw.numpoints[0]), int(fw.table_oversamp[0])
    base = lambda: domega_torch(gridv, gout, ktraj, tabs, fw.n_shift.cpu(), J, L)  # noqa: E731
    k_ms = event_ms(dom, iters)
    b_ms = event_ms(base, max(3, iters // 2))
    a, r = dom(), base()
    M = ktraj.shape[1]
    F = footprint_cells(ktraj.cpu().numpy(), (K2, K2), J)
    kbytes = coils * F * 8 + coils * M * 8 + 2 * M * 4 + 2 * M * 4  # grid cells reached, gout, omega, out
    achieved = kbytes / (k_ms * 1e-3) / 1e9
    return {
        "workload": "256^2, 12 coils (SENSE maps), 64-interleave VD spiral x 1536 (98304 samples); "
        "loss ||A(x, ktraj) - y||^2, y = A(x_true) + N(0, 0.01); d/dx and d/dktraj",
        "fused_ms": t_fused,
        "per_coil_loop_ms": t_loop,
        "speedup_vs_loop": t_loop / t_fused,
        "rel_diff_dx": float(torch.linalg.norm(gx - lx) / torch.linalg.norm(lx)),
        "rel_diff_dktraj": float(torch.linalg.norm(gk - lk) / torch.linalg.norm(lk)),
        "fp64_oracle_rel_err": fp64,
        "domega_kernel": {
            "kernel": "interp_domega_kernel (12 rows x 98304 samples, 512^2 grid, J 6)",
            "kernel_ms": k_ms,
            "algorithmic_bytes": kbytes,
            "achieved_GBps": achieved,
            "hbm_frac": achieved / 8000.0,
            "unfused_torch_ms": b_ms,
            "speedup_vs_unfused": b_ms / k_ms,
            "rel_diff_vs_unfused": float(torch.linalg.norm(a - r) / torch.linalg.norm(r)),
        },
    }


def bench_dcomp(dev, ktraj, iters: int, with_cpu: bool) -> dict:
    """Density compensation (10 spread + gather iterations, one row) on the radial
    trajectory; the CPU oracle's time for the same call when the baseline is on."""
    import torchkbnufft_amd as tkbn

    om = ktraj.to(dev)
    fn = lambda: tkbn.calc_density_compensation_function(ktraj=om, im_size=(N, N))  # noqa: E731
    out = {
        "workload": f"calc_density_compensation_function, {N}x{N} -> {K}x{K}, {om.shape[-1]} radial samples, "
        "10 iterations (interp_adjoint then interp, one row)",
        "gpu_ms": timed(fn, iters),
    }
    if with_cpu:
        from oracle import table_path as tp

        plan = tp.Plan((N, N), dtype=torch.complex64)
        t0 = time.perf_counter()
        tp.density_compensation(plan, ktraj.cpu())
        out["cpu_oracle_ms"] = (time.perf_counter() - t0) * 1e3
        out["cpu_threads"] = torch.get_num_threads()
    return out


def bench_toeplitz(dev, x, ktraj, smaps, iters: int, pair_ms: float) -> dict:
    """Toeplitz normal operator on the SENSE workload: one ToepNufft call replaces the
    forward + adjoint pair.  Algorithmic bytes are the stage boundaries of a two-pass
    2-D FFT filter: read image + smaps, write the row-transformed intermediate
    (B*C x N x 2N), read and write it in the column filter pass, read it and smaps
    again in the last pass, read the kernel once, write the image."""
    import torchkbnufft_amd as tkbn

    t0 = time.perf_counter()
    kernel = tkbn.calc_toeplitz_kernel(ktraj, (N, N))
    torch.cuda.synchronize()
    kernel_ms = (time.perf_counter() - t0) * 1e3
    toep = tkbn.ToepNufft()
    ms = timed(lambda: toep(x, kernel, smaps=smaps), iters)
    G2, P = (2 * N) ** 2, N * N
    inter = B * C * N * (2 * N) * 8
    alg = B * P * 8 + B * C * P * 8 + 4 * inter + B * C * P * 8 + G2 * 8 + B * P * 8
    return {
        "workload": f"ToepNufft on the same batch ({B} x {C} coils, {N}x{N}), kernel {2 * N}x{2 * N} from the "
        "radial trajectory; A^H A x without interpolation",
        "kernel_build_ms": kernel_ms,
        "toep_ms": ms,
        "nufft_pair_ms": pair_ms,
        "speedup_vs_pair": pair_ms / ms,
        "algorithmic_bytes": alg,
        "hbm_frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
    }


def bench_spmat(dev, x, ktraj, smaps, iters: int, pair_ms: float, with_cpu: bool = False) -> dict:
    """The SENSE pair with sparse-matrix interpolation (calc_tensor_spmatrix matrices,
    exact Kaiser-Bessel weights) instead of the table.  Algorithmic bytes of one
    interpolation: the matrix once (4-byte column + 8-byte complex value per nonzero),
    the input vectors once and the output vectors once.  Also the interpolations alone
    for one coil, sparse matrix and table."""
    import torchkbnufft_amd as tkbn

    od = ktraj.to(dev)
    M = od.shape[-1]
    t0 = time.perf_counter()
    mats = tkbn.calc_tensor_spmatrix(od, (N, N))
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    fw = tkbn.KbNufft(im_size=(N, N)).to(dev)
    ad = tkbn.KbNufftAdjoint(im_size=(N, N)).to(dev)
    t0 = time.perf_counter()
    ad(fw(x, od, mats, smaps=smaps), od, mats, smaps=smaps)  # first call: CSR forms of both directions
    torch.cuda.synchronize()
    first_ms = (time.perf_counter() - t0) * 1e3
    ms = timed(lambda: ad(fw(x, od, mats, smaps=smaps), od, mats, smaps=smaps), iters)
    gen = torch.Generator().manual_seed(9)
    grid = torch.complex(torch.randn(B, C, K, K, generator=gen), torch.randn(B, C, K, K, generator=gen)).to(dev)
    data = torch.complex(torch.randn(B, C, M, generator=gen), torch.randn(B, C, M, generator=gen)).to(dev)
    ip = tkbn.KbInterp(im_size=(N, N)).to(dev)
    ia = tkbn.KbInterpAdjoint(im_size=(N, N)).to(dev)
    fwd_ms = timed(lambda: ip(grid, od, mats), iters)
    adj_ms = timed(lambda: ia(data, od, mats), iters)
    g1, d1 = grid[:1, :1].contiguous(), data[:1, :1].contiguous()
    one = {
        "spmat_fwd_ms": timed(lambda: ip(g1, od, mats), iters),
        "spmat_adj_ms": timed(lambda: ia(d1, od, mats), iters),
        "table_fwd_ms": timed(lambda: ip(g1, od), iters),
        "table_adj_ms": timed(lambda: ia(d1, od), iters),
    }
    nnz = mats[0]._nnz()
    alg = nnz * 12 + B * C * (K * K + M) * 8
    cpu = {}
    if with_cpu:
        # the oracle's restatement of the reference's four torch.mm products, one
        # forward and one adjoint product on the same 64 rows, host threads
        from oracle import table_path as tp

        cm = tuple(m.cpu() for m in mats)
        gc, dc = grid.cpu(), data.cpu()
        t0 = time.perf_counter()
        tp.spmat_interp(gc, cm)
        cpu["cpu_oracle_fwd_ms"] = (time.perf_counter() - t0) * 1e3
        t0 = time.perf_counter()
        tp.spmat_interp_adjoint(dc, cm, (K, K))
        cpu["cpu_oracle_adj_ms"] = (time.perf_counter() - t0) * 1e3
        cpu["cpu_threads"] = torch.get_num_threads()
    return {
        "workload": f"KbNufft + KbNufftAdjoint with interp_mats on the same batch ({B} x {C} coils, {N}x{N} -> "
        f"{K}x{K}, {M} radial samples, {nnz} nonzeros)",
        "matrix_build_ms": build_ms,
        "first_pair_ms": first_ms,
        "pair_ms": ms,
        "table_pair_ms": pair_ms,
        "interp_fwd_ms": fwd_ms,
        "interp_adj_ms": adj_ms,
        "interp_algorithmic_bytes": alg,
        "interp_fwd_hbm_frac": alg / (fwd_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "interp_adj_hbm_frac": alg / (adj_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "one_coil_interp": one,
        **cpu,
    }


def bench_small(dev, iters: int, with_cpu: bool) -> dict:
    """The reference's own CPU-scale case: one coil, 256^2 -> 512^2, 402 spokes x 512
    readout points (205 824 samples, angles s*pi/402), forward + adjoint; the CPU
    oracle's time for the same pair when the baseline is on."""
    import torchkbnufft_amd as tkbn

    gen = torch.Generator().manual_seed(1)
    r = -np.pi + 2 * np.pi * np.arange(512) / 512
    th = np.arange(402) * np.pi / 402
    om = torch.tensor(np.stack([(np.sin(th)[:, None] * r).reshape(-1), (np.cos(th)[:, None] * r).reshape(-1)]),
                      dtype=torch.float32)
    x = torch.complex(torch.randn(1, 1, 256, 256, generator=gen), torch.randn(1, 1, 256, 256, generator=gen))
    fw = tkbn.KbNufft(im_size=(256, 256)).to(dev)
    ad = tkbn.KbNufftAdjoint(im_size=(256, 256)).to(dev)
    xd, od = x.to(dev), om.to(dev)
    ms = timed(lambda: ad(fw(xd, od), od), iters)
    sel = torch.arange(0, od.shape[1], 101, device=dev)
    out = {
        "workload": "1 coil, 256x256 -> 512x512, 402 x 512 radial (205824 samples), fwd + adj",
        "gpu_ms": ms,
        "gpu_pairs_per_s": 1e3 / ms,
        "fp64_oracle_rel_err": nudft_rel_err(fw(xd, od)[0, 0, sel], xd[0, 0], od[:, sel], (128, 128)),
    }
    # The same pair captured once as a HIP graph and replayed: the eager pair is bound by
    # host launch work (17 launches and the modules' Python), the replay by the kernels.
    try:
        gs = torch.cuda.Stream()
        gs.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(gs):
            for _ in range(3):
                ad(fw(xd, od), od)  # plans and workspaces of the capture stream
        torch.cuda.current_stream().wait_stream(gs)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=gs):
            yg = ad(fw(xd, od), od)
        diffs = []
        for _ in range(3):  # every replay against the eager pair, eager work in between
            graph.replay()
            torch.cuda.synchronize()
            ye = ad(fw(xd, od), od)
            diffs.append(float(torch.linalg.norm(yg - ye) / torch.linalg.norm(ye)))
        gms = timed(graph.replay, iters)
        out["graph_ms"] = gms
        out["graph_pairs_per_s"] = 1e3 / gms
        out["graph_rel_diff_max"] = max(diffs)
    except RuntimeError as e:  # reported, not fatal: the eager figures stand
        out["graph_error"] = str(e)[:200]
    if with_cpu:
        from oracle import table_path as tp

        plan = tp.Plan((256, 256), dtype=torch.complex64)
        tp.kb_table_nufft_adjoint(plan, tp.kb_table_nufft(plan, x, om), om)
        t0 = time.perf_counter()
        tp.kb_table_nufft_adjoint(plan, tp.kb_table_nufft(plan, x, om), om)
        out["cpu_oracle_ms"] = (time.perf_counter() - t0) * 1e3
        out["gpu_over_cpu"] = out["cpu_oracle_ms"] / ms
    return out


def bench_kooshball(dev, steps: int, warmup: int) -> dict:
    """3-D radial kooshball: 160^3 -> 320^3, 8 coils with Gaussian-smoothed maps, 3200
    spokes uniform on the sphere x 512 samples (1 638 400), deterministic spread;
    forward + adjoint per step, plus the image and trajectory gradient of ||A x||^2."""
    import torchkbnufft_amd as tkbn
    from oracle import nudft

    gen = torch.Generator().manual_seed(5)
    n, coils, nsp, nread = 160, 8, 3200, 512
    d = torch.randn(3, nsp, generator=gen, dtype=torch.float64)
    d = d / torch.linalg.norm(d, dim=0, keepdim=True)
    r = -np.pi + 2 * np.pi * torch.arange(nread, dtype=torch.float64) / nread
    ktraj = (d[:, :, None] * r[None, None, :]).reshape(3, -1).to(torch.float32).to(dev)
    x = torch.complex(torch.randn(1, 1, n, n, n, generator=gen), torch.randn(1, 1, n, n, n, generator=gen)).to(dev)
    s = torch.complex(torch.randn(1, coils, n, n, n, generator=gen), torch.randn(1, coils, n, n, n, generator=gen)).to(dev)
    f = torch.fft.fftfreq(n, device=dev)
    g1 = torch.exp(-2 * (math.pi**2) * (20.0**2) * f**2)
    s = torch.fft.ifftn(torch.fft.fftn(s, dim=(-3, -2, -1)) * (g1[:, None, None] * g1[None, :, None] * g1[None, None, :]),
                        dim=(-3, -2, -1))
    smaps = (s / torch.sqrt(torch.sum(torch.abs(s) ** 2, dim=1, keepdim=True))).to(torch.complex64).contiguous()
    del s
    fw = tkbn.KbNufft(im_size=(n, n, n), dtype=torch.complex64).to(dev)
    ad = tkbn.KbNufftAdjoint(im_size=(n, n, n), dtype=torch.complex64).to(dev)
    was = tkbn.deterministic_enabled()
    tkbn.set_deterministic(True)
    try:
        def pair():
            return ad(fw(x, ktraj, smaps=smaps), ktraj, smaps=smaps)

        t_pair = timed(pair, steps) if warmup >= 0 else None
        a, b = pair(), pair()
        repeat = bool(torch.equal(a, b))

        def grad():
            xx = x.clone().requires_grad_(True)
            kk = ktraj.clone().requires_grad_(True)
            torch.sum(torch.abs(fw(xx, kk, smaps=smaps)) ** 2).backward()
            return xx.grad, kk.grad

        t_grad = timed(grad, max(1, steps // 2))
        sel = torch.arange(0, ktraj.shape[1], ktraj.shape[1] // 16, device=dev)[:16]  # 16 x 4.1M-term NUDFT
        y = fw(x, ktraj, smaps=smaps[:, :1])[..., sel].cpu().numpy()
        xs = (x * smaps[:, :1]).cpu().numpy().astype(np.complex128)
        ref = nudft.nudft(xs, ktraj[:, sel].double().cpu().numpy(), (n // 2, n // 2, n // 2))
        err = float(np.linalg.norm(y - ref) / np.linalg.norm(ref))
    finally:
        tkbn.set_deterministic(was)
    return {
        "metric": "3-D kooshball fwd+adj NUFFT/s (deterministic)",
        "value": 1e3 / t_pair,
        "unit": "pairs/s",
        "ms_per_pair": t_pair,
        "grad_ms": t_grad,
        "adjoint_bitwise_repeatable": repeat,
        "fp64_oracle_rel_err": err,
        "config": {"workload": "kooshball-3d: 160^3 -> 320^3, 8 coils (smoothed SENSE maps), 3200 spokes x 512 "
                               "(1638400 samples), numpoints 6, deterministic adjoint", "coils": coils,
                   "klength": int(ktraj.shape[1])},
    }


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 runs of the TCC/EA
    read and write request counters summed over every TCC instance
    (profiles/traffic_sense2d_radial_640.*; the request sizes are checked on a plain copy
    kernel); None when that profile is absent or does not cover the kernel."""
    path = os.path.join(ROOT, "profiles", "traffic_sense2d_radial_640.json")
    try:
        with open(path) as f:
            return json.load(f)["kernels"][kernel]["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def fp64_oracle_error(fw, x, ktraj, smaps) -> float:
    """Forward NUFFT vs FP64 NUDFT on a 1-coil slice and a sample subset."""
    from oracle import nudft

    sel = torch.arange(0, ktraj.shape[1], 128, device=ktraj.device)
    y = fw(x[:1], ktraj, smaps=smaps[:1, :1])[..., sel].cpu().numpy()
    xs = (x[:1] * smaps[:1, :1]).cpu().numpy().astype(np.complex128)
    ref = nudft.nudft(xs, ktraj[:, sel].double().cpu().numpy(), (N // 2, N // 2))
    return float(np.linalg.norm(y - ref) / np.linalg.norm(ref))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--stage-iters", type=int, default=20)
    ap.add_argument("--form-iters", type=int, default=10, help="iterations of the batched / gradient forms")
    ap.add_argument("--no-forms", action="store_true", help="skip the batched and gradient forms")
    ap.add_argument("--workload", default="sense2d",
                    choices=["sense2d", "kooshball3d", "copy512", "spmat", "batched", "grad"],
                    help="sense2d: the headline SENSE pair (default); kooshball3d: the 3-D deterministic case; "
                    "batched: the 64 uneven spiral problems only (fused batch, separate launches, plain loop); "
                    "grad: the image + trajectory gradient form only; "
                    "copy512: a plain 512 MiB read + 512 MiB write kernel (checks the request sizes of the traffic counters); "
                    "spmat: the SENSE batch with sparse-matrix interpolation only")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            f"bench.py measures one GPU only (asked for --gpus {args.gpus}); sharding the path "
            "over several GPUs (coil rows per GPU, one coil-sum reduction) is not run here",
            file=sys.stderr,
        )
        return 2

    import torchkbnufft_amd as tkbn

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.workload == "kooshball3d":
        print(json.dumps(bench_kooshball(dev, max(1, args.steps // 10), args.warmup)), flush=True)
        return 0
    if args.workload == "batched":
        print(json.dumps(bench_batched(dev, args.form_iters)), flush=True)
        return 0
    if args.workload == "grad":
        print(json.dumps(bench_grad(dev, args.form_iters)), flush=True)
        return 0
    if args.workload == "copy512":
        src = torch.ones(128 * 1024 * 1024, dtype=torch.float32, device=dev)  # 512 MiB
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 2.0, out=dst)
        torch.cuda.synchronize()
        print(json.dumps({"workload": "copy512", "read_bytes": src.numel() * 4, "write_bytes": dst.numel() * 4}))
        return 0
    gen = torch.Generator().manual_seed(0)
    ktraj_cpu = radial_golden(NSPOKES, NREAD)
    M = ktraj_cpu.shape[1]
    ktraj = ktraj_cpu.to(dev)
    x = torch.complex(torch.randn(B, 1, N, N, generator=gen), torch.randn(B, 1, N, N, generator=gen)).to(dev)
    smaps = make_smaps(gen, dev)
    fw = tkbn.KbNufft(im_size=(N, N), numpoints=J, table_oversamp=L, dtype=torch.complex64).to(dev)
    ad = tkbn.KbNufftAdjoint(im_size=(N, N), numpoints=J, table_oversamp=L, dtype=torch.complex64).to(dev)
    if args.workload == "spmat":
        print(json.dumps(bench_spmat(dev, x, ktraj, smaps, args.form_iters, None, not args.no_cpu_baseline)),
              flush=True)
        return 0

    def step():
        kd = fw(x, ktraj, smaps=smaps)
        return ad(kd, ktraj, smaps=smaps)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # one HIP event between consecutive steps on the stream the kernels run on: the
    # per-step durations give the median / p10 / p90 (recording an event enqueues no work)
    marks = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
    t0 = time.perf_counter()
    marks[0].record()
    for i in range(args.steps):
        step()
        marks[i + 1].record()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = elapsed / args.steps * 1e3
    per_step = np.array([marks[i].elapsed_time(marks[i + 1]) for i in range(args.steps)])

    F = footprint_cells(ktraj_cpu.numpy(), (K, K), J)
    sb = stage_bytes(F, M)
    stages, kernels = time_stages((fw, ad), x, ktraj, smaps, args.stage_iters)
    # algorithmic bytes of one launch of each binned kernel = its stage's bytes
    kbytes = {"gather_sep_kernel": sb["interp"], "spread_sep_kernel": sb["spread"]}
    dom = max(kernels, key=kernels.get)
    achieved = kbytes[dom] / (kernels[dom] * 1e-6) / 1e9
    pair_bytes = sum(sb.values())

    out = {
        "metric": METRIC,
        "value": args.steps / elapsed,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "step_ms_events": {
            "median": float(np.median(per_step)),
            "p10": float(np.percentile(per_step, 10)),
            "p90": float(np.percentile(per_step, 90)),
            "n": int(per_step.size),
            "pairs_per_s_at_median": float(1e3 / np.median(per_step)),
        },
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {
            "workload": "sense-2d-radial: batch 4 x 16 coils, 320x320 -> 640x640, 400 x 640 golden-angle "
            "radial (256000 samples), numpoints 6, table_oversamp 1024, fwd + adj per step",
            "batch": B,
            "coils": C,
            "im_size": [N, N],
            "grid_size": [K, K],
            "klength": M,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "algorithmic_bytes": kbytes[dom],
            "avg_us": round(kernels[dom], 2),
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(dom),
        },
        "pair": {
            "algorithmic_bytes": pair_bytes,
            "hbm_frac": pair_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "footprint_frac": F / (K * K),
        },
        "stages_us": {k: round(v, 2) for k, v in stages.items()},
        "kernels_us": {k: round(v, 2) for k, v in kernels.items()},
        "stage_bytes": sb,
        "fp64_oracle_rel_err": fp64_oracle_error(fw, x, ktraj, smaps),
        "plan": plan_build_us(fw, x, ktraj, smaps),
    }
    if not args.no_forms:
        out["batched"] = bench_batched(dev, args.form_iters)
        out["grad"] = bench_grad(dev, args.form_iters)
        out["dcomp"] = bench_dcomp(dev, ktraj, args.form_iters, not args.no_cpu_baseline)
        out["toeplitz"] = bench_toeplitz(dev, x, ktraj, smaps, max(args.form_iters, 10), ms)
        out["small"] = bench_small(dev, max(args.form_iters, 20), not args.no_cpu_baseline)
        out["spmat"] = bench_spmat(dev, x, ktraj, smaps, args.form_iters, ms, not args.no_cpu_baseline)
        out["kooshball"] = bench_kooshball(dev, 3, 1)
        out["fp64_oracle_rel_err_by_config"] = {
            "c1_radial_256": out["small"]["fp64_oracle_rel_err"],
            "c2_sense_radial_320": out["fp64_oracle_rel_err"],
            "c3_batched_spirals": out["batched"]["fp64_oracle_rel_err"],
            "c4_grad_spiral_256": out["grad"]["fp64_oracle_rel_err"],
            "c5_kooshball_160": out["kooshball"]["fp64_oracle_rel_err"],
        }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(x, ktraj, smaps, args.cpu_seconds)
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
