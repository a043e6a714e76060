"""Sparse interpolation matrices and their CSR segment plan (CPU).

* ``calc_tensor_spmatrix`` against the reference's own output (golden vectors from
  ``torchkbnufft/_nufft/spmat.py:10-105``, coalesced): indices and values bit-exact,
  including the float32 grid-coordinate arithmetic of a float32 trajectory;
* the matrix against the oracle's closed-form exact Kaiser-Bessel interpolation
  (``oracle.nudft.exact_kb_interp``, an independent per-tap evaluation) for float64
  trajectories;
* the row-segment plan that ``tkbn_spmv`` consumes: segments cover every row in
  order, and the segment / fix-up sums it prescribes reproduce the dense product
  (emulated here in numpy, so no GPU is needed).
"""
import numpy as np
import pytest
import torch

import inputs as gi
import torchkbnufft_amd as tkbn
from conftest import golden
from oracle import nudft
from torchkbnufft_amd._nufft import spmat as sp


def _dense(mats):
    re, im = mats
    return (re.to_dense() + 1j * im.to_dense()).to(torch.complex128).numpy()


@pytest.mark.parametrize("i", range(len(gi.SPMAT_CASES)))
def test_spmatrix_matches_reference_golden(i):
    g = golden("spmat.npz")
    im, grid, J, order, m, dtype = gi.SPMAT_CASES[i]
    re, imag = tkbn.calc_tensor_spmatrix(gi.spmat_inputs(i), im, grid_size=grid, numpoints=J, order=order)
    assert re.dtype == dtype and tuple(re.shape) == tuple(g[f"s{i}_shape"])
    assert re._nnz() == int(g[f"s{i}_nnz_raw"])
    mat = torch.sparse_coo_tensor(re._indices(), torch.complex(re._values(), imag._values()), re.shape).coalesce()
    np.testing.assert_array_equal(mat.indices().numpy(), g[f"s{i}_indices"])
    np.testing.assert_array_equal(mat.values().numpy(), g[f"s{i}_values"])


@pytest.mark.parametrize(
    "im_size,grid_size,numpoints,n_shift",
    [((10,), (23,), 6, None), ((12, 9), (20, 16), (4, 6), (3, 5)), ((6, 5, 7), None, 3, None), ((16, 16), None, 6, None)],
)
def test_spmatrix_is_exact_kb_operator(im_size, grid_size, numpoints, n_shift):
    gen = torch.Generator().manual_seed(5)
    d = len(im_size)
    om = gi.uniform_ktraj(d, 57, gen)
    mats = tkbn.calc_tensor_spmatrix(om, im_size, grid_size=grid_size, numpoints=numpoints, n_shift=n_shift)
    gs = grid_size or tuple(2 * n for n in im_size)
    J = (numpoints,) * d if isinstance(numpoints, int) else numpoints
    ns = n_shift or tuple(n // 2 for n in im_size)
    grid = np.random.default_rng(1).standard_normal((2,) + gs) + 1j * np.random.default_rng(2).standard_normal((2,) + gs)
    got = grid.reshape(2, -1) @ _dense(mats).T
    ref = nudft.exact_kb_interp(grid, om.numpy(), im_size, J, [2.34 * j for j in J], ns)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-13


def test_spmatrix_errors():
    with pytest.raises(ValueError, match="batched omega"):
        tkbn.calc_tensor_spmatrix(torch.zeros(2, 2, 5), (8, 8))
    ob = tkbn.KbInterp(im_size=(8, 8))
    with pytest.raises(TypeError, match="2-tuple"):
        ob(torch.zeros(1, 1, 16, 16, dtype=torch.complex64), torch.zeros(2, 5), interp_mats=[torch.zeros(1)] * 2)


def _emulate(plan, x):
    """The sums tkbn_spmv performs, from the plan's arrays (x: (batch, ncols))."""
    R = sp._BLOCK_ROWS
    seg, fix = plan.seg.numpy(), plan.fix.numpy()
    ucol, uval = plan.ucol.numpy(), plan.uval.numpy()
    nblk = -(-plan.nrows // R)
    yt = np.full((nblk * R, x.shape[0]), np.nan, dtype=np.complex128)
    part = np.full((plan.nslots * R, x.shape[0]), np.nan, dtype=np.complex128)
    for blk, u0, u1, slot in seg:
        s = uval[u0:u1].T @ x[:, ucol[u0:u1]].T  # (R, batch)
        if slot < 0:
            yt[blk * R : blk * R + R] = s
        else:
            part[slot * R : slot * R + R] = s
    for blk, first, cnt in fix:
        yt[blk * R : blk * R + R] = part[first * R : (first + cnt) * R].reshape(cnt, R, -1).sum(0)
    return yt[: plan.nrows].T


def test_segment_plan_covers_rows_and_reproduces_product():
    rng = np.random.default_rng(3)
    nrows, ncols = 43, 5000
    counts = rng.integers(0, 12, nrows)
    counts[[3, 17]] = [700, 300]  # blocks split into segments; one empty row below
    counts[25] = 0
    rows = np.repeat(np.arange(nrows), counts)
    cols = rng.integers(0, ncols, rows.size)  # duplicates included
    vals = rng.standard_normal(rows.size) + 1j * rng.standard_normal(rows.size)
    perm = rng.permutation(rows.size)
    rows, cols, vals = rows[perm], cols[perm], vals[perm]
    plan = sp.SpmvPlan(torch.tensor(rows), torch.tensor(cols), torch.tensor(vals), nrows, ncols)
    R = sp._BLOCK_ROWS
    nblk = -(-nrows // R)
    seg = plan.seg.numpy()
    # segments: block-ordered, contiguous per block, at most _SEG_COLS union columns
    assert np.all(np.diff(seg[:, 0]) >= 0) and set(seg[:, 0]) == set(range(nblk))
    assert np.all(seg[:, 2] - seg[:, 1] <= sp._SEG_COLS) and np.all(seg[:, 2] >= seg[:, 1])
    for q in range(nblk):
        s = seg[seg[:, 0] == q]
        assert np.all(s[1:, 1] == s[:-1, 2])
        assert (s[:, 3] < 0).all() == (len(s) == 1)
        # the block's union columns: sorted, distinct, exactly its rows' columns
        sel = (rows // R) == q
        np.testing.assert_array_equal(plan.ucol.numpy()[s[0, 1] : s[-1, 2]], np.unique(cols[sel]))
    union = np.array([np.unique(cols[(rows // R) == q]).size for q in range(nblk)])
    nseg = np.maximum(1, -(-union // sp._SEG_COLS))
    assert plan.nseg == nseg.sum() and plan.nfix == (nseg > 1).sum() >= 1
    assert plan.nslots == nseg[nseg > 1].sum()
    dense = np.zeros((nrows, ncols), dtype=np.complex128)
    np.add.at(dense, (rows, cols), vals)
    x = rng.standard_normal((3, ncols)) + 1j * rng.standard_normal((3, ncols))
    np.testing.assert_allclose(_emulate(plan, x), x @ dense.T, rtol=1e-12, atol=1e-12)


def test_plan_index_check():
    with pytest.raises(ValueError, match="outside"):
        sp.SpmvPlan(torch.tensor([0, 5]), torch.tensor([1, 2]), torch.ones(2, dtype=torch.complex64), 4, 4)


@pytest.mark.parametrize("adjoint", [False, True])
@pytest.mark.parametrize("shared", [True, False])
def test_plan_merges_real_and_imaginary_parts(adjoint, shared):
    """A = A_re + i A_im (conjugate transpose for the adjoint), also when the two parts
    have different sparsity patterns or duplicate entries."""
    rng = np.random.default_rng(8)
    nrows, ncols, nnz = 30, 50, 400
    ri = np.stack([rng.integers(0, nrows, nnz), rng.integers(0, ncols, nnz)])
    ii = ri.copy() if shared else np.stack([rng.integers(0, nrows, 300), rng.integers(0, ncols, 300)])
    rv, iv = rng.standard_normal(ri.shape[1]), rng.standard_normal(ii.shape[1])
    re = torch.sparse_coo_tensor(torch.tensor(ri), torch.tensor(rv), (nrows, ncols))
    im = torch.sparse_coo_tensor(torch.tensor(ii), torch.tensor(iv), (nrows, ncols))
    dense = re.to_dense().numpy() + 1j * im.to_dense().numpy()
    if adjoint:
        dense = dense.conj().T
    plan = sp._plan((re, im), adjoint, torch.empty(1, dtype=torch.complex128))
    assert (plan.nrows, plan.ncols) == dense.shape
    x = rng.standard_normal((2, dense.shape[1])) + 1j * rng.standard_normal((2, dense.shape[1]))
    np.testing.assert_allclose(_emulate(plan, x), x @ dense.T, rtol=1e-12, atol=1e-12)


def test_spmatrix_without_samples():
    re, im = tkbn.calc_tensor_spmatrix(torch.zeros(2, 0), (8, 6))
    assert re.shape == (0, 16 * 12) and re._nnz() == 0 and im._nnz() == 0
    re, im = tkbn.calc_tensor_spmatrix(torch.zeros(3, 0, dtype=torch.float64), (4, 4, 4))
    assert re.shape == (0, 512) and re.dtype == torch.float64


def test_oracle_spmat_products_match_dense():
    """The oracle's restatement of the reference's four-product spmat_interp pair
    against the dense matrix (pins it for the CPU baseline)."""
    from oracle import table_path as tp

    gen = torch.Generator().manual_seed(12)
    om = gi.uniform_ktraj(2, 40, gen)
    mats = tkbn.calc_tensor_spmatrix(om, (6, 5))
    A = _dense(mats)
    grid = torch.complex(torch.randn(2, 3, 12, 10, dtype=torch.float64), torch.randn(2, 3, 12, 10, dtype=torch.float64))
    data = torch.complex(torch.randn(2, 3, 40, dtype=torch.float64), torch.randn(2, 3, 40, dtype=torch.float64))
    y = tp.spmat_interp(grid, mats)
    x = tp.spmat_interp_adjoint(data, mats, (12, 10))
    np.testing.assert_allclose(y.reshape(6, -1).numpy(), grid.reshape(6, -1).numpy() @ A.T, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(x.reshape(6, -1).numpy(), data.reshape(6, -1).numpy() @ A.conj(), rtol=1e-12, atol=1e-12)
