"""Seeded inputs for the golden cases (shared by the generator and the tests).

Inputs are rebuilt from ``torch.Generator`` seeds instead of being stored, so the
committed ``.npz`` files hold only the reference's outputs.  The call order of the
generator inside each builder is part of the fixture format: do not reorder.
"""
from __future__ import annotations

import zlib

import numpy as np
import torch

# name, im_size, grid_size, numpoints, table_oversamp, B, C, trajectory kind, smaps, norm, batched
C64_CASES = [
    ("i2d", (24, 32), (48, 64), 6, 1024, 1, 2, "rand3000", False, None, False),
    ("n2d", (48, 48), None, 6, 1024, 1, 3, "radial32x128", True, None, False),
    ("n2do", (48, 40), None, 6, 1024, 2, 2, "radial24x96", True, "ortho", False),
    ("n2dj", (33, 24), None, (4, 6), 1024, 2, 2, "rand2000", False, None, False),
    ("n2dl", (30, 36), None, 5, 300, 1, 2, "rand2500", False, None, False),
    ("n1d", (37,), None, 6, 1024, 2, 1, "rand200", False, None, False),
    ("n3d", (10, 12, 14), None, 6, 1024, 1, 2, "rand1500", True, None, False),
    ("n3dj", (11, 9, 8), None, (2, 4, 6), 512, 1, 2, "rand800", False, None, False),
    ("b2d", (32, 28), None, 6, 1024, 3, 2, "rand900", False, None, True),
    ("c1r", (96, 96), None, 6, 1024, 1, 1, "radial60x192", False, None, False),
]

# shape (with trailing real dim 2), klength -- tests/data/create_old_data.py:17-21,66-72
INTERP_FIXTURES = [([3, 2, 38, 2], 52), ([5, 3, 78, 96, 2], 410), ([1, 5, 17, 19, 12, 2], 112)]
NUFFT_FIXTURES = [([3, 1, 38, 2], 52), ([5, 1, 78, 96, 2], 410), ([1, 3, 17, 19, 12, 2], 112)]

# im_size, grid_size, numpoints, table_oversamp, kbwidth
PLAN_CASES = [
    ((256, 256), None, 6, 1024, 2.34),
    ((320, 320), None, 6, 1024, 2.34),
    ((13, 18, 12), (13, 18, 12), (2, 4, 6), 1024, 2.34),
    ((19,), (19,), 6, 1024, 2.34),
    ((38,), (77,), 5, 333, 2.5),
    ((160, 160, 160), None, 6, 1024, 2.34),
    ((128, 192), None, 6, 1024, 2.34),
    # odd J where om / gam rounds below J/2 - 1 on one table row (shifted sparse columns)
    ((50, 45), (100, 90), (5, 3), 1024, 2.34),
]

# sparse interpolation matrices: im_size, grid_size, numpoints, order, klength, omega dtype
SPMAT_CASES = [
    ((12, 10), None, 6, 0.0, 40, torch.float32),
    ((9, 14), (13, 20), (4, 5), 0.0, 33, torch.float32),
    ((6, 5, 4), None, 4, 0.0, 20, torch.float64),
    ((16,), (24,), 6, 1.0, 25, torch.float32),
    ((40, 36), None, 6, 0.0, 300, torch.float32),
]


def noise_input(shape, gen, dtype=torch.float64):
    """``arange + N(0,1)`` (the input recipe of the reference's tests/conftest.py:28-37)."""
    x = torch.arange(int(np.prod(shape)), dtype=dtype).reshape(shape)
    return x + torch.randn(shape, generator=gen, dtype=dtype)


def uniform_ktraj(d, m, gen, dtype=torch.float64):
    return torch.rand((d, m), generator=gen, dtype=dtype) * 2 * np.pi - np.pi


def radial(nspokes, nread, golden=True, dtype=torch.float32):
    """Radial spokes, rows (ky, kx) as profile_torchkbnufft.py:205-216 stacks them."""
    r = -np.pi + 2 * np.pi * np.arange(nread) / nread
    step = np.pi / ((1 + np.sqrt(5)) / 2) if golden else np.pi / nspokes
    th = np.arange(nspokes) * step
    ky = (np.sin(th)[:, None] * r[None, :]).reshape(-1)
    kx = (np.cos(th)[:, None] * r[None, :]).reshape(-1)
    return torch.tensor(np.stack([ky, kx]), dtype=dtype)


def cplx(shape, gen, dtype=torch.complex64):
    rd = torch.float32 if dtype == torch.complex64 else torch.float64
    return torch.complex(
        torch.randn(shape, generator=gen, dtype=rd), torch.randn(shape, generator=gen, dtype=rd)
    )


def grid_of(im, grid):
    return tuple(2 * v for v in im) if grid is None else tuple(grid)


def c64_inputs(case):
    """Returns dict(ktraj, x, smaps|None, grid_in, y_in) for a ``C64_CASES`` row."""
    name, im, grid, J, L, B, C, mk, use_smaps, norm, batched = case
    gen = torch.Generator().manual_seed(zlib.crc32(name.encode()))
    d = len(im)
    if mk.startswith("radial"):
        ns, nr = (int(v) for v in mk[6:].split("x"))
        ktraj = radial(ns, nr, golden=True)
    else:
        ktraj = uniform_ktraj(d, int(mk[4:]), gen, torch.float32)
    if batched:
        ktraj = torch.stack([uniform_ktraj(d, ktraj.shape[1], gen, torch.float32) for _ in range(B)])
    gs = grid_of(im, grid)
    cx = 1 if use_smaps else C
    x = cplx((B, cx) + tuple(im), gen)
    smaps = cplx((B, C) + tuple(im), gen) if use_smaps else None
    grid_in = cplx((B, C) + gs, gen)
    y_in = cplx((B, C, ktraj.shape[-1]), gen)
    return dict(ktraj=ktraj, x=x, smaps=smaps, grid_in=grid_in, y_in=y_in)


def interp_fixture_inputs(i):
    shape, klen = INTERP_FIXTURES[i]
    gen = torch.Generator().manual_seed(123)
    im_size = shape[2:-1]
    image = noise_input(shape, gen)
    dims = tuple(range(2, 2 + len(im_size)))
    image = torch.view_as_real(torch.fft.fftn(torch.view_as_complex(image), dim=dims)).contiguous()
    ktraj = uniform_ktraj(len(im_size), klen, gen)
    return im_size, image, ktraj


def nufft_fixture_inputs(i):
    shape, klen = NUFFT_FIXTURES[i]
    gen = torch.Generator().manual_seed(123)
    im_size = shape[2:-1]
    image = noise_input(shape, gen)
    ktraj = uniform_ktraj(len(im_size), klen, gen)
    return im_size, image, ktraj


def grad_inputs():
    gen = torch.Generator().manual_seed(7)
    im = (48, 48)
    ktraj = uniform_ktraj(2, 1500, gen, torch.float32)
    x = cplx((1, 1) + im, gen)
    smaps = cplx((1, 4) + im, gen)
    y = cplx((1, 4, ktraj.shape[1]), gen)
    return im, ktraj, x, smaps, y


def spmat_inputs(i):
    """omega (d, klength) of sparse-matrix case ``i``."""
    im, grid, J, order, m, dtype = SPMAT_CASES[i]
    gen = torch.Generator().manual_seed(7100 + i)
    return uniform_ktraj(len(im), m, gen, dtype)
