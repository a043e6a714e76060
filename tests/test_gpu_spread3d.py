"""GPU parity of the 3-D register-tile spread (complex64, J <= 6, >= 4 coil rows).

Against the oracle's table adjoint (``_nufft/interp.py:587-726`` restated) at 1e-5
relative L2, in both adjoint modes, on grids that are not tile multiples, with mixed
numpoints and a dense centre, so that hot tiles are cut into pieces (added in piece
order by the combine kernel) and, in deterministic mode, long tiles go through the
chunk sort and several merge passes.  Deterministic mode must be bitwise repeatable,
also when the plan is reused without checks.
"""
import numpy as np
import pytest
import torch

import inputs as gi
import torchkbnufft_amd as tkbn
from conftest import rel_l2
from oracle import table_path as tp
from torchkbnufft_amd import _lib
from torchkbnufft_amd._nufft import interp as _interp

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")
C64_TOL = 1e-5


def _last_kernel() -> str:
    return _lib.load().tkbn_last_interp_kernel().decode()


CASES = [
    # grid, numpoints, coil rows, samples, fraction of samples packed near the centre
    ((24, 26, 40), 6, 4, 30000, 0.5),
    ((40, 40, 40), (6, 5, 4), 8, 60000, 0.6),
    ((21, 17, 33), 6, 13, 20000, 0.3),
]


@pytest.mark.parametrize("grid,J,rows,M,frac", CASES)
@pytest.mark.parametrize("deterministic", [False, True])
def test_register_tile_spread_matches_oracle(grid, J, rows, M, frac, deterministic):
    print(f"[case] reg3 spread grid={grid} J={J} rows={rows} M={M} det={deterministic}", flush=True)
    gen = torch.Generator().manual_seed(71)
    om = (torch.rand(3, M, generator=gen) * 2 - 1) * np.pi
    om[:, : int(M * frac)] *= 0.04  # a dense centre: split tiles, long sorts
    y = gi.cplx((1, rows, M), gen)
    im = tuple(g // 2 for g in grid)
    ad = tkbn.KbInterpAdjoint(im_size=im, grid_size=grid, numpoints=J, dtype=torch.complex64).to(DEV)
    args = (om.to(DEV), ad._tables(), ad.n_shift, ad.numpoints, ad.table_oversamp, ad.offsets, grid)
    yd = y.to(DEV)
    out = _interp.table_interp_adjoint(yd, *args, deterministic=deterministic)
    assert _last_kernel().startswith("spread_reg3_kernel")
    if deterministic:
        holder = _interp.PlanHolder(trusted=True)
        outs = [_interp.table_interp_adjoint(yd, *args, deterministic=True, plan=holder) for _ in range(3)]
        torch.cuda.synchronize()
        assert all(torch.equal(out, o) for o in outs)  # checked, built, reused unchecked
    plan = tp.Plan(im, grid, J, dtype=torch.complex64)
    sel = [0, rows - 1]
    ref = tp.kb_table_interp_adjoint(plan, y[:, sel], om)
    err = rel_l2(out[:, sel].cpu(), ref)
    print(f"[reg3] rel err {err:.2e}")
    assert err < C64_TOL


def test_register_tile_spread_batched_rows_and_empty_trajectory():
    """Batch x coils rows share one trajectory (rows = batch * coils, chunks of 8 with a
    partial last chunk); an empty trajectory stores a zero grid."""
    print("[case] reg3 spread batch 3 x coils 3, and M = 0", flush=True)
    gen = torch.Generator().manual_seed(72)
    grid, im, M = (32, 30, 36), (16, 15, 18), 12000
    om = (torch.rand(3, M, generator=gen) * 2 - 1) * np.pi
    y = gi.cplx((3, 3, M), gen)
    ad = tkbn.KbInterpAdjoint(im_size=im, grid_size=grid, dtype=torch.complex64).to(DEV)
    args = (ad._tables(), ad.n_shift, ad.numpoints, ad.table_oversamp, ad.offsets, grid)
    out = _interp.table_interp_adjoint(y.to(DEV), om.to(DEV), *args)
    assert _last_kernel().startswith("spread_reg3_kernel")
    plan = tp.Plan(im, grid, dtype=torch.complex64)
    ref = tp.kb_table_interp_adjoint(plan, y[1:2, 1:3], om)
    assert rel_l2(out[1:2, 1:3].cpu(), ref) < C64_TOL
    empty = torch.zeros((3, 3, 0), dtype=torch.complex64, device=DEV)
    z = _interp.table_interp_adjoint(empty, torch.zeros((3, 0), device=DEV), *args,
                                     out=torch.full((3, 3) + grid, 7.0, dtype=torch.complex64, device=DEV))
    assert int(torch.count_nonzero(z)) == 0
