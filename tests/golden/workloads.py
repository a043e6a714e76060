"""Seeded synthetic inputs of the five benchmark workloads (shared by bench.py and tests).

Every builder returns CPU tensors (the caller moves them to the GPU), so the oracle
and the GPU see bitwise the same inputs.

* C1  plain 2-D: 256^2 -> 512^2, 1 coil, 402 uniformly spaced radial spokes x 512.
* C2  SENSE 2-D: batch 4 x 16 coils, 320^2 -> 640^2, 400 x 640 golden-angle radial,
      maps = complex N(0,1) low-passed by a periodic Gaussian (sigma 20 px) and
      normalised to unit root-sum-of-squares.
* C3  batched uneven: 64 problems, N in {128,192,256}^2, 8 coils each, variable-density
      spirals of 8-48 interleaves x 1024 samples (r = pi t^1.5, 16 turns).
* C4  differentiation: 256^2, 12 coils as SENSE maps, 64-interleave VD spiral x 1536.
* C5  3-D kooshball: 160^3 -> 320^3, 8 coils with Gaussian-smoothed maps, 3200 spokes
      with directions uniform on the sphere x 512 samples.
"""
from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np
import torch


def cplx(shape, gen, dtype=torch.complex64):
    rd = torch.float32 if dtype == torch.complex64 else torch.float64
    return torch.complex(torch.randn(shape, generator=gen, dtype=rd), torch.randn(shape, generator=gen, dtype=rd))


def radial_uniform(nspokes: int, nread: int) -> torch.Tensor:
    """Spoke angles s*pi/nspokes, readout -pi + 2 pi i / nread; rows (ky, kx)."""
    r = -np.pi + 2 * np.pi * np.arange(nread) / nread
    th = np.arange(nspokes) * np.pi / nspokes
    return torch.tensor(np.stack([(np.sin(th)[:, None] * r).reshape(-1), (np.cos(th)[:, None] * r).reshape(-1)]),
                        dtype=torch.float32)


def radial_golden(nspokes: int, nread: int) -> torch.Tensor:
    """Golden-angle spokes (theta_s = s pi / phi); rows (ky, kx)."""
    r = -np.pi + 2 * np.pi * np.arange(nread) / nread
    th = np.arange(nspokes) * np.pi / ((1 + np.sqrt(5)) / 2)
    ky = (np.sin(th)[:, None] * r[None, :]).reshape(-1)
    kx = (np.cos(th)[:, None] * r[None, :]).reshape(-1)
    return torch.tensor(np.stack([ky, kx]), dtype=torch.float32)


def vd_spiral(n_inter: int, n_read: int, turns: float = 16.0) -> torch.Tensor:
    """Variable-density spiral, r(t) = pi t^1.5 (t = i / n_read), theta(t) = 2 pi turns t
    + 2 pi l / n_inter; rows (ky, kx); radius kept just below pi."""
    t = np.arange(n_read) / n_read
    r = np.pi * t**1.5 * 0.9999
    arms = []
    for l in range(n_inter):
        th = 2 * np.pi * turns * t + 2 * np.pi * l / n_inter
        arms.append(np.stack([r * np.sin(th), r * np.cos(th)]))
    return torch.tensor(np.concatenate(arms, axis=1), dtype=torch.float32)


def kooshball(nspokes: int, nread: int, gen: torch.Generator) -> torch.Tensor:
    """Spokes along normalised Gaussian directions (uniform on S^2), readout -pi + 2 pi i / nread."""
    d = torch.randn(3, nspokes, generator=gen, dtype=torch.float64)
    d = d / torch.linalg.norm(d, dim=0, keepdim=True)
    r = -np.pi + 2 * np.pi * torch.arange(nread, dtype=torch.float64) / nread
    return (d[:, :, None] * r[None, None, :]).reshape(3, -1).to(torch.float32)


def smooth_maps(gen: torch.Generator, batch: int, coils: int, im_size: Tuple[int, ...], sigma: float = 20.0):
    """Complex N(0,1) maps low-passed by a periodic Gaussian (sigma px, FFT-domain multiply)
    and normalised to unit root-sum-of-squares per pixel; complex64."""
    s = cplx((batch, coils) + tuple(im_size), gen)
    dims = tuple(range(-len(im_size), 0))
    filt = None
    for i, n in enumerate(im_size):
        f = torch.fft.fftfreq(n)
        g = torch.exp(-2 * (math.pi**2) * (sigma**2) * f**2)
        shape = [1] * len(im_size)
        shape[i] = n
        g = g.reshape(shape)
        filt = g if filt is None else filt * g
    s = torch.fft.ifftn(torch.fft.fftn(s, dim=dims) * filt, dim=dims)
    rss = torch.sqrt(torch.sum(torch.abs(s) ** 2, dim=1, keepdim=True))
    return (s / rss).to(torch.complex64).contiguous()


def c1(seed: int = 1):
    gen = torch.Generator().manual_seed(seed)
    ktraj = radial_uniform(402, 512)
    x = cplx((1, 1, 256, 256), gen)
    return x, ktraj


def c2(seed: int = 0, batch: int = 4, coils: int = 16):
    """x (B,1,320,320), smaps (B,C,320,320), ktraj (2, 256000)."""
    gen = torch.Generator().manual_seed(seed)
    ktraj = radial_golden(400, 640)
    x = cplx((batch, 1, 320, 320), gen)
    smaps = smooth_maps(gen, batch, coils, (320, 320))
    return x, smaps, ktraj


def c3(seed: int = 3, nprob: int = 64, coils: int = 8):
    """sizes [(N, N)], interleaves [int], ktrajs [(2, M_p)], images [(1, coils, N, N)]."""
    gen = torch.Generator().manual_seed(seed)
    sizes = [(int(s), int(s)) for s in torch.tensor([128, 192, 256])[torch.randint(0, 3, (nprob,), generator=gen)]]
    inter = torch.randint(8, 49, (nprob,), generator=gen).tolist()
    ktrajs = [vd_spiral(n, 1024) for n in inter]
    images = [cplx((1, coils) + s, gen) for s in sizes]
    return sizes, inter, ktrajs, images


def c4(seed: int = 4, n: int = 256, coils: int = 12, n_inter: int = 64, n_read: int = 1536):
    """ktraj (2, 98304), smaps (1,12,256,256), x_true, x0 (1,1,256,256), noise (1,12,M)
    (y = A(x_true) + 0.01 * noise is formed by the caller with its own operator)."""
    gen = torch.Generator().manual_seed(seed)
    ktraj = vd_spiral(n_inter, n_read)
    smaps = smooth_maps(gen, 1, coils, (n, n))
    x_true = cplx((1, 1, n, n), gen)
    noise = cplx((1, coils, ktraj.shape[1]), gen)
    x0 = cplx((1, 1, n, n), gen)
    return ktraj, smaps, x_true, noise, x0


def c5(seed: int = 5, n: int = 160, coils: int = 8, nspokes: int = 3200, nread: int = 512):
    """ktraj (3, nspokes*nread), x (1,1,n,n,n), smaps (1,coils,n,n,n)."""
    gen = torch.Generator().manual_seed(seed)
    ktraj = kooshball(nspokes, nread, gen)
    x = cplx((1, 1, n, n, n), gen)
    smaps = smooth_maps(gen, 1, coils, (n, n, n))
    return ktraj, x, smaps


def footprint_cells(ktraj: np.ndarray, grid: tuple, numpoints: int) -> int:
    """Grid cells inside some sample's J-neighbourhood (the base-offset rule of the
    reference's table interpolation, float32 arithmetic)."""
    import itertools

    d = len(grid)
    kk = np.asarray(grid)
    gam = (np.float32(1) / kk.astype(np.float32)) * np.float32(2 * np.pi)
    tm = ktraj.astype(np.float32) / gam[:, None].astype(np.float32)
    base = 1 + np.floor(tm - np.float32(numpoints / 2)).astype(np.int64)
    mark = np.zeros(grid, dtype=bool)
    for off in itertools.product(range(numpoints), repeat=d):
        mark[tuple(np.mod(base[i] + off[i], kk[i]) for i in range(d))] = True
    return int(mark.sum())


def problem_subset(sizes: List[Tuple[int, int]], inter: List[int], count: int = 8) -> List[int]:
    """Problem indices covering every image size and both interleave extremes."""
    pick = []
    for s in sorted(set(sizes)):
        idx = [p for p in range(len(sizes)) if sizes[p] == s]
        lo = min(idx, key=lambda p: inter[p])
        hi = max(idx, key=lambda p: inter[p])
        for p in (lo, hi):
            if p not in pick:
                pick.append(p)
    rest = [p for p in range(len(sizes)) if p not in pick]
    for p in rest[: max(0, count - len(pick))]:
        pick.append(p)
    return pick
