"""GPU parity of sparse-matrix interpolation (``torchkbnufft/_nufft/interp.py:13-85``,
``_nufft/spmat.py``, ``functional/nufft.py:15-122``, ``modules/kbnufft.py:185-197, 364-376``).

* ``tkbn_spmv`` forward and adjoint against the dense product with the same matrix
  (numpy, FP64), over batch widths that take every lane split of the kernel and over
  a radial trajectory whose centre rows are cut into several segments;
* sparse-matrix NUFFTs against the oracle's FFT stages around the exact-KB
  interpolation (the matrix itself is pinned against the reference's output and the
  oracle's closed form by the CPU tests);
* the reference's own tests with ``spmat``: dot tests and autograd identities
  (``tests/conftest.py:40-87`` with ``test_interp.py``, ``test_nufft.py``,
  ``test_sense_nufft.py`` shapes), complex vs real-view inputs;
* at the 320^2 -> 640^2 / 256 000-sample radial geometry: the product against the
  table path and the dot test, and bitwise-repeatable adjoints.
"""
import numpy as np
import pytest
import torch

import inputs as gi
import torchkbnufft_amd as tkbn
from conftest import create_input_plus_noise, create_ktraj, rel_l2
from oracle import table_path as tp
from torchkbnufft_amd._nufft import spmat as sp

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")


def _dense(mats):
    re, im = mats
    return (re.to_dense() + 1j * im.to_dense()).to(torch.complex128).cpu().numpy()


def _cplx(shape, gen, dtype):
    rd = torch.float64 if dtype == torch.complex128 else torch.float32
    return torch.complex(torch.randn(shape, generator=gen, dtype=rd), torch.randn(shape, generator=gen, dtype=rd))


@pytest.mark.parametrize(
    "im_size,B,C",
    [((20,), 1, 1), ((12, 10), 1, 2), ((12, 10), 3, 1), ((9, 14), 2, 3), ((16, 12), 3, 4), ((6, 5, 4), 4, 5),
     ((10, 8), 5, 8), ((8, 8), 7, 10)],
)
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_spmv_matches_dense_product(im_size, B, C, dtype):
    print(f"[case] spmv im={im_size} B={B} C={C} {dtype}", flush=True)
    gen = torch.Generator().manual_seed(21)
    rd = torch.float64 if dtype == torch.complex128 else torch.float32
    om = create_ktraj(len(im_size), 70, gen, dtype=rd)
    mats = tkbn.calc_tensor_spmatrix(om, im_size)
    mats = tuple(m.to(DEV) for m in mats)
    A = _dense(mats)
    gs = tuple(2 * n for n in im_size)
    grid = _cplx((B, C) + gs, gen, dtype)
    data = _cplx((B, C, 70), gen, dtype)
    y = tkbn.KbInterp(im_size=im_size, dtype=dtype, device=DEV)(grid.to(DEV), om.to(DEV), interp_mats=mats)
    x = tkbn.KbInterpAdjoint(im_size=im_size, dtype=dtype, device=DEV)(data.to(DEV), om.to(DEV), interp_mats=mats)
    assert y.dtype == dtype and y.shape == (B, C, 70) and x.shape == (B, C) + gs
    y_ref = grid.reshape(B * C, -1).to(torch.complex128).numpy() @ A.T
    x_ref = data.reshape(B * C, -1).to(torch.complex128).numpy() @ A.conj()
    tol = 1e-13 if dtype == torch.complex128 else 2e-6
    assert rel_l2(y.reshape(B * C, -1), torch.from_numpy(y_ref)) < tol
    assert rel_l2(x.reshape(B * C, -1), torch.from_numpy(x_ref)) < tol


@pytest.mark.parametrize("B,C", [(1, 1), (2, 3), (1, 40)])
def test_spmv_split_rows_radial(B, C):
    """Radial centre grid points gather more nonzeros than one segment holds."""
    print(f"[case] spmv radial split B={B} C={C}", flush=True)
    im = (24, 24)
    om = gi.radial(64, 48, golden=True, dtype=torch.float64)
    mats = tuple(m.to(DEV) for m in tkbn.calc_tensor_spmatrix(om, im))
    plan = sp._plan(mats, True, torch.empty(1, dtype=torch.complex128, device=DEV))
    assert plan.nfix > 0
    A = _dense(mats)
    gen = torch.Generator().manual_seed(4)
    data = _cplx((B, C, om.shape[1]), gen, torch.complex128)
    x = tkbn.KbInterpAdjoint(im_size=im, dtype=torch.complex128, device=DEV)(data.to(DEV), om.to(DEV), mats)
    x_ref = data.reshape(B * C, -1).numpy() @ A.conj()
    assert rel_l2(x.reshape(B * C, -1), torch.from_numpy(x_ref)) < 1e-13
    x2 = tkbn.KbInterpAdjoint(im_size=im, dtype=torch.complex128, device=DEV)(data.to(DEV), om.to(DEV), mats)
    assert torch.equal(x, x2)


@pytest.mark.parametrize(
    "im_size,B,C,smaps,norm",
    [((18,), 2, 1, False, None), ((14, 12), 1, 3, True, None), ((16, 10), 2, 2, False, "ortho"),
     ((8, 6, 5), 1, 2, True, "ortho")],
)
def test_spmat_nufft_matches_oracle(im_size, B, C, smaps, norm):
    print(f"[case] spmat nufft im={im_size} B={B} C={C} smaps={smaps} norm={norm}", flush=True)
    gen = torch.Generator().manual_seed(8)
    om = create_ktraj(len(im_size), 90, gen)
    mats = tuple(m.to(DEV) for m in tkbn.calc_tensor_spmatrix(om, im_size))
    A = _dense(mats)
    plan = tp.Plan(im_size, dtype=torch.complex128)
    sm = _cplx((B, C) + im_size, gen, torch.complex128) if smaps else None
    image = _cplx((B, 1 if smaps else C) + im_size, gen, torch.complex128)
    data = _cplx((B, C, 90), gen, torch.complex128)
    kw = dict(im_size=im_size, dtype=torch.complex128, device=DEV)
    smd = sm.to(DEV) if smaps else None
    y = tkbn.KbNufft(**kw)(image.to(DEV), om.to(DEV), mats, smaps=smd, norm=norm)
    x = tkbn.KbNufftAdjoint(**kw)(data.to(DEV), om.to(DEV), mats, smaps=smd, norm=norm)
    xi = image * sm if smaps else image
    grid = tp.fft_and_scale(xi, plan.scaling_coef, plan.im_size, plan.grid_size, norm)
    y_ref = (grid.reshape(B * C, -1).numpy() @ A.T).reshape(B, C, -1)
    g_adj = torch.from_numpy(data.reshape(B * C, -1).numpy() @ A.conj()).reshape((B, C) + plan.grid_size)
    x_ref = tp.ifft_and_scale(g_adj, plan.scaling_coef, plan.im_size, plan.grid_size, norm)
    if smaps:
        x_ref = torch.sum(x_ref * sm.conj(), dim=1, keepdim=True)
    assert rel_l2(y, torch.from_numpy(y_ref)) < 1e-12
    assert rel_l2(x, x_ref) < 1e-12


def _dot_and_autograd(forw, adj, image, kdata, ktraj, spmat, **kw):
    """The reference's nufft_adjoint_test / nufft_autograd_test (conftest.py:40-87)."""
    image_forw = forw(image, ktraj, spmat, **kw)
    kdata_adj = adj(kdata, ktraj, spmat, **kw)
    assert torch.allclose(tkbn.inner_product(image_forw, kdata), tkbn.inner_product(image, kdata_adj))
    image = image.clone().requires_grad_(True)
    kdata = kdata.clone().requires_grad_(True)
    image_forw = forw(image, ktraj, spmat, **kw)
    kdata_adj = adj(kdata, ktraj, spmat, **kw)
    (torch.abs(image_forw) ** 2 / 2).sum().backward()
    (torch.abs(kdata_adj) ** 2 / 2).sum().backward()
    assert torch.allclose(image.grad, adj(image_forw.detach(), ktraj, spmat, **kw))
    assert torch.allclose(kdata.grad, forw(kdata_adj.detach(), ktraj, spmat, **kw))


# tests/test_interp.py:237-246 shapes (grid = im_size)
INTERP_SHAPES = [
    ([1, 1, 19], [1, 1, 25], True),
    ([3, 1, 13, 2], [3, 1, 18, 2], False),
    ([1, 1, 32, 16], [1, 1, 83], True),
    ([5, 1, 15, 12, 2], [5, 1, 83, 2], False),
    ([3, 2, 13, 18, 12], [3, 2, 112], True),
    ([1, 2, 17, 19, 12, 2], [1, 2, 112, 2], False),
]


@pytest.mark.parametrize("shape,kdata_shape,is_complex", INTERP_SHAPES)
def test_reference_interp_spmat_tests(shape, kdata_shape, is_complex):
    print(f"[case] interp spmat dot/autograd {shape}", flush=True)
    gen = torch.Generator().manual_seed(123)
    im_size = shape[2:] if is_complex else shape[2:-1]
    image = create_input_plus_noise(shape, is_complex, gen).to(DEV)
    kdata = create_input_plus_noise(kdata_shape, is_complex, gen).to(DEV)
    ktraj = create_ktraj(len(im_size), kdata_shape[2], gen).to(DEV)
    kw = dict(im_size=im_size, grid_size=im_size, dtype=torch.float64, device=DEV)
    spmat = tkbn.calc_tensor_spmatrix(ktraj, im_size, grid_size=im_size)
    _dot_and_autograd(tkbn.KbInterp(**kw), tkbn.KbInterpAdjoint(**kw), image, kdata, ktraj, spmat)


@pytest.mark.parametrize("shape,kdata_shape,is_complex", INTERP_SHAPES)
def test_reference_nufft_spmat_tests(shape, kdata_shape, is_complex):
    print(f"[case] nufft spmat dot/autograd {shape}", flush=True)
    gen = torch.Generator().manual_seed(123)
    im_size = shape[2:] if is_complex else shape[2:-1]
    image = create_input_plus_noise(shape, is_complex, gen).to(DEV)
    kdata = create_input_plus_noise(kdata_shape, is_complex, gen).to(DEV)
    ktraj = create_ktraj(len(im_size), kdata_shape[2], gen).to(DEV)
    kw = dict(im_size=im_size, dtype=torch.float64, device=DEV)
    spmat = tkbn.calc_tensor_spmatrix(ktraj, im_size)
    _dot_and_autograd(tkbn.KbNufft(**kw), tkbn.KbNufftAdjoint(**kw), image, kdata, ktraj, spmat)


@pytest.mark.parametrize(
    "shape,kdata_shape",
    [([1, 3, 19], [1, 3, 25]), ([3, 4, 13, 18], [3, 4, 112]), ([1, 2, 11, 13, 9], [1, 2, 83])],
)
def test_reference_sense_spmat_tests(shape, kdata_shape):
    """tests/test_sense_nufft.py:28-158 with spmat, complex inputs."""
    print(f"[case] sense spmat {shape}", flush=True)
    gen = torch.Generator().manual_seed(123)
    im_size = shape[2:]
    image = create_input_plus_noise([shape[0], 1] + shape[2:], True, gen).to(DEV)
    smaps = create_input_plus_noise(shape, True, gen).to(DEV)
    kdata = create_input_plus_noise(kdata_shape, True, gen).to(DEV)
    ktraj = create_ktraj(len(im_size), kdata_shape[2], gen).to(DEV)
    kw = dict(im_size=im_size, dtype=torch.float64, device=DEV)
    spmat = tkbn.calc_tensor_spmatrix(ktraj, im_size)
    _dot_and_autograd(tkbn.KbNufft(**kw), tkbn.KbNufftAdjoint(**kw), image, kdata, ktraj, spmat, smaps=smaps)


def test_spmat_complex_and_real_views_agree():
    """tests/test_nufft.py:120-145 with spmat."""
    gen = torch.Generator().manual_seed(3)
    im_size = (21, 17)
    image = create_input_plus_noise([2, 3, 21, 17], True, gen).to(DEV)
    ktraj = create_ktraj(2, 200, gen).to(DEV)
    spmat = tkbn.calc_tensor_spmatrix(ktraj, im_size)
    ob = tkbn.KbNufft(im_size=im_size, dtype=torch.float64, device=DEV)
    a = ob(image, ktraj, spmat)
    b = torch.view_as_complex(ob(torch.view_as_real(image), ktraj, spmat))
    assert torch.allclose(a, b)


def test_spmat_radial_config_against_table_path():
    """320^2 -> 640^2, 400 x 640 golden-angle radial, 16 coils, complex64."""
    print("[case] spmat radial 320 config", flush=True)
    im = (320, 320)
    ktraj = gi.radial(400, 640, golden=True).to(DEV)
    mats = tkbn.calc_tensor_spmatrix(ktraj, im)
    gen = torch.Generator().manual_seed(2)
    grid = _cplx((1, 16, 640, 640), gen, torch.complex64).to(DEV)
    data = _cplx((1, 16, ktraj.shape[1]), gen, torch.complex64).to(DEV)
    fw = tkbn.KbInterp(im_size=im, dtype=torch.complex64, device=DEV)
    ad = tkbn.KbInterpAdjoint(im_size=im, dtype=torch.complex64, device=DEV)
    y_sp, y_tab = fw(grid, ktraj, mats), fw(grid, ktraj)
    x_sp, x_tab = ad(data, ktraj, mats), ad(data, ktraj)
    # exact kernel vs the 1024-oversampled table: the table's own error
    assert rel_l2(y_sp, y_tab) < 2e-3 and rel_l2(x_sp, x_tab) < 2e-3
    lhs = torch.vdot(y_sp.reshape(-1).to(torch.complex128), data.reshape(-1).to(torch.complex128))
    rhs = torch.vdot(grid.reshape(-1).to(torch.complex128), x_sp.reshape(-1).to(torch.complex128))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5
    assert torch.equal(ad(data, ktraj, mats), x_sp)


@pytest.mark.parametrize("shared", [True, False])
def test_spmv_user_matrices(shared):
    """Hand-built (real, imaginary) pairs: duplicates, empty rows, different patterns."""
    print(f"[case] spmv user matrices shared={shared}", flush=True)
    rng = np.random.default_rng(8)
    nrows, ncols = 300, 12 * 10
    ri = np.stack([rng.integers(0, nrows - 20, 5000), rng.integers(0, ncols, 5000)])  # last rows empty
    ii = ri.copy() if shared else np.stack([rng.integers(0, nrows, 3000), rng.integers(0, ncols, 3000)])
    re = torch.sparse_coo_tensor(torch.tensor(ri), torch.tensor(rng.standard_normal(ri.shape[1])), (nrows, ncols))
    im = torch.sparse_coo_tensor(torch.tensor(ii), torch.tensor(rng.standard_normal(ii.shape[1])), (nrows, ncols))
    A = re.to_dense().numpy() + 1j * im.to_dense().numpy()
    mats = (re.to(DEV), im.to(DEV))
    grid = torch.complex(torch.randn(2, 3, 12, 10, dtype=torch.float64), torch.randn(2, 3, 12, 10, dtype=torch.float64))
    data = torch.complex(torch.randn(2, 3, nrows, dtype=torch.float64), torch.randn(2, 3, nrows, dtype=torch.float64))
    y = tkbn.functional.kb_spmat_interp(grid.to(DEV), mats)
    x = tkbn.functional.kb_spmat_interp_adjoint(data.to(DEV), mats, torch.tensor([12, 10]))
    assert rel_l2(y.reshape(6, -1), torch.from_numpy(grid.reshape(6, -1).numpy() @ A.T)) < 1e-13
    assert rel_l2(x.reshape(6, -1), torch.from_numpy(data.reshape(6, -1).numpy() @ A.conj())) < 1e-13


def test_spmat_without_samples():
    """An empty trajectory: empty k-space forward, zero grid adjoint."""
    im = (8, 6)
    om = torch.zeros(2, 0, device=DEV)
    mats = tkbn.calc_tensor_spmatrix(om, im)
    grid = torch.ones(2, 3, 16, 12, dtype=torch.complex64, device=DEV)
    y = tkbn.KbInterp(im_size=im, device=DEV)(grid, om, mats)
    assert y.shape == (2, 3, 0)
    x = tkbn.KbInterpAdjoint(im_size=im, device=DEV)(torch.zeros(2, 3, 0, dtype=torch.complex64, device=DEV), om, mats)
    assert x.shape == (2, 3, 16, 12) and not x.abs().any()
