"""GPU parity of the table interpolation kernels (tkbn_interp_fwd / _adj / _domega).

Tolerances: complex64 results must be within 1e-5 relative L2 of the reference
(golden vectors) and of the oracle; float64 results use the reference tests'
``torch.allclose`` defaults (rtol 1e-5, atol 1e-8).
"""
import numpy as np
import pytest
import torch

import inputs as gi
import torchkbnufft_amd as tkbn
from conftest import create_input_plus_noise, create_ktraj, golden, rel_l2
from oracle import nudft
from oracle import table_path as tp
from torchkbnufft_amd import _lib

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")
C64_TOL = 1e-5

SHAPES = [
    ([1, 1, 19], [1, 1, 25], True),
    ([3, 1, 13, 2], [3, 1, 18, 2], False),
    ([1, 1, 32, 16], [1, 1, 83], True),
    ([5, 1, 15, 12, 2], [5, 1, 83, 2], False),
    ([3, 2, 13, 18, 12], [3, 2, 112], True),
    ([1, 2, 17, 19, 12, 2], [1, 2, 112, 2], False),
]


def _im_size(shape, is_complex):
    return shape[2:] if is_complex else shape[2:-1]


@pytest.mark.parametrize("case", gi.C64_CASES, ids=[c[0] for c in gi.C64_CASES])
def test_interp_c64_matches_reference(case):
    name, im, grid, J, L, B, C, mk, use_smaps, norm, batched = case
    print(f"[case] interp {name}", flush=True)
    g = golden("golden_c64.npz")
    inp = gi.c64_inputs(case)
    kw = dict(im_size=im, grid_size=grid, numpoints=J, table_oversamp=L, dtype=torch.complex64)
    fw = tkbn.KbInterp(**kw).to(DEV)
    ad = tkbn.KbInterpAdjoint(**kw).to(DEV)
    k = inp["ktraj"].to(DEV)
    out = fw(inp["grid_in"].to(DEV), k).cpu()
    assert rel_l2(out, g[name + "_interp"]) < C64_TOL
    outa = ad(inp["y_in"].to(DEV), k).cpu()
    assert rel_l2(outa, g[name + "_interp_adj"]) < C64_TOL


def test_interp_f64_fixtures():
    g = golden("fixtures_f64.npz")
    for i in range(len(gi.INTERP_FIXTURES)):
        im_size, image, ktraj = gi.interp_fixture_inputs(i)
        ob = tkbn.KbInterp(im_size=im_size, grid_size=im_size, dtype=torch.complex128).to(DEV)
        out = ob(image.to(DEV), ktraj.to(DEV)).cpu()
        assert torch.allclose(out, torch.from_numpy(g[f"interp{i}_kdata"]))


@pytest.mark.parametrize("shape, kdata_shape, is_complex", SHAPES)
def test_interp_adjoint_dot(shape, kdata_shape, is_complex):
    gen = torch.Generator().manual_seed(123)
    im_size = _im_size(shape, is_complex)
    image = create_input_plus_noise(shape, is_complex, gen).to(DEV)
    kdata = create_input_plus_noise(kdata_shape, is_complex, gen).to(DEV)
    ktraj = create_ktraj(len(im_size), kdata_shape[2], gen).to(DEV)
    fw = tkbn.KbInterp(im_size=im_size, grid_size=im_size, dtype=torch.complex128).to(DEV)
    ad = tkbn.KbInterpAdjoint(im_size=im_size, grid_size=im_size, dtype=torch.complex128).to(DEV)
    lhs = tkbn.inner_product(fw(image, ktraj), kdata)
    rhs = tkbn.inner_product(image, ad(kdata, ktraj))
    assert torch.allclose(lhs, rhs)


@pytest.mark.parametrize("shape, kdata_shape, is_complex", SHAPES)
def test_interp_autograd_is_adjoint(shape, kdata_shape, is_complex):
    gen = torch.Generator().manual_seed(123)
    im_size = _im_size(shape, is_complex)
    image = create_input_plus_noise(shape, is_complex, gen).to(DEV).requires_grad_(True)
    kdata = create_input_plus_noise(kdata_shape, is_complex, gen).to(DEV).requires_grad_(True)
    ktraj = create_ktraj(len(im_size), kdata_shape[2], gen).to(DEV)
    fw = tkbn.KbInterp(im_size=im_size, grid_size=im_size, dtype=torch.complex128).to(DEV)
    ad = tkbn.KbInterpAdjoint(im_size=im_size, grid_size=im_size, dtype=torch.complex128).to(DEV)
    f = fw(image, ktraj)
    a = ad(kdata, ktraj)
    (torch.abs(f) ** 2 / 2).sum().backward()
    (torch.abs(a) ** 2 / 2).sum().backward()
    assert torch.allclose(image.grad, ad(f.detach(), ktraj))
    assert torch.allclose(kdata.grad, fw(a.detach(), ktraj))


@pytest.mark.parametrize("shape, kdata_shape", [([1, 1, 19], [1, 1, 25]), ([1, 1, 32, 16], [1, 1, 83]), ([3, 2, 13, 18, 12], [3, 2, 112])])
def test_interp_complex_real_match(shape, kdata_shape):
    gen = torch.Generator().manual_seed(123)
    im_size = shape[2:]
    image = create_input_plus_noise(shape, True, gen).to(DEV)
    ktraj = create_ktraj(len(im_size), kdata_shape[2], gen).to(DEV)
    fw = tkbn.KbInterp(im_size=im_size, grid_size=im_size, dtype=torch.complex128).to(DEV)
    a = fw(image, ktraj)
    b = torch.view_as_complex(fw(torch.view_as_real(image), ktraj))
    assert torch.allclose(a, b)


@pytest.mark.parametrize(
    "shape, kdata_shape, is_complex",
    [([3, 1, 19], [3, 1, 25], True), ([4, 1, 32, 16], [4, 1, 83], True), ([3, 2, 13, 18, 12], [3, 2, 112], True),
     ([2, 2, 17, 19, 12, 2], [2, 2, 112, 2], False)],
)
def test_interp_batched_omega_equals_loop(shape, kdata_shape, is_complex):
    gen = torch.Generator().manual_seed(123)
    im_size = _im_size(shape, is_complex)
    image = create_input_plus_noise(shape, is_complex, gen).to(DEV)
    kdata = create_input_plus_noise(kdata_shape, is_complex, gen).to(DEV)
    ktraj = (torch.rand((shape[0], len(im_size), kdata_shape[2]), generator=gen, dtype=torch.float64) * 2 * np.pi - np.pi).to(DEV)
    fw = tkbn.KbInterp(im_size=im_size, grid_size=im_size, dtype=torch.complex128).to(DEV)
    ad = tkbn.KbInterpAdjoint(im_size=im_size, grid_size=im_size, dtype=torch.complex128).to(DEV)
    loop = torch.cat([fw(i.unsqueeze(0), k) for i, k in zip(image, ktraj)])
    assert torch.allclose(loop, fw(image, ktraj))
    loop = torch.cat([ad(d.unsqueeze(0), k) for d, k in zip(kdata, ktraj)])
    assert torch.allclose(loop, ad(kdata, ktraj))


def test_adjoint_batched_radial_split_tiles_equals_loop():
    """Per-batch radial trajectories (dense centres: the centre tiles are split into
    pieces that are zeroed, then added) through the complex64 sample-lane adjoint: the
    zero pass, the max-|kdata| pass and the spread of all batch items in one call equal
    one call per item (to the float rounding of the pieces' global adds), and a NaN in
    one item's data reaches only that item's grid."""
    print("[case] batched radial adjoint, split tiles", flush=True)
    gen = torch.Generator().manual_seed(321)
    B, C, nsp, nread, im = 3, 2, 64, 256, (64, 64)
    angles = torch.rand(B, nsp, generator=gen, dtype=torch.float64) * np.pi
    r = torch.linspace(-np.pi, np.pi, nread, dtype=torch.float64)
    om = torch.stack([torch.cos(angles)[..., None] * r, torch.sin(angles)[..., None] * r], dim=1)
    om = om.reshape(B, 2, -1).to(torch.float32).to(DEV)
    kd = gi.cplx((B, C, om.shape[-1]), gen, torch.complex64).to(DEV)
    ad = tkbn.KbInterpAdjoint(im_size=im).to(DEV)
    together = ad(kd, om)
    loop = torch.cat([ad(kd[b:b + 1], om[b]) for b in range(B)])
    assert rel_l2(together, loop) < 1e-6
    kd2 = kd.clone()
    kd2[1, 0, 5] = complex(float("nan"), 0.0)
    g2 = ad(kd2, om)
    assert torch.isnan(g2[1]).any() and not torch.isnan(g2[0]).any() and not torch.isnan(g2[2]).any()


def test_c1_full_size_interp_matches_oracle():
    """256^2 image / 512^2 grid, 402 x 512 radial spokes (the CPU-runnable plain config)."""
    print("[case] C1 interp fwd/adj 512^2 x 205824", flush=True)
    gen = torch.Generator().manual_seed(0)
    ktraj = gi.radial(402, 512, golden=False)
    plan = tp.Plan((256, 256), dtype=torch.complex64)
    grid = gi.cplx((1, 1, 512, 512), gen)
    y = gi.cplx((1, 1, ktraj.shape[1]), gen)
    fw = tkbn.KbInterp(im_size=(256, 256), dtype=torch.complex64).to(DEV)
    ad = tkbn.KbInterpAdjoint(im_size=(256, 256), dtype=torch.complex64).to(DEV)
    assert rel_l2(fw(grid.to(DEV), ktraj.to(DEV)).cpu(), tp.kb_table_interp(plan, grid, ktraj)) < C64_TOL
    assert rel_l2(ad(y.to(DEV), ktraj.to(DEV)).cpu(), tp.kb_table_interp_adjoint(plan, y, ktraj)) < C64_TOL


def test_nonfinite_and_empty_inputs():
    ob = tkbn.KbInterp(im_size=(16, 16), dtype=torch.complex64).to(DEV)
    ad = tkbn.KbInterpAdjoint(im_size=(16, 16), dtype=torch.complex64).to(DEV)
    grid = torch.randn(1, 2, 32, 32, dtype=torch.complex64, device=DEV)
    om = (torch.rand(2, 10, device=DEV) * 2 - 1) * np.pi
    om[0, 3] = float("nan")
    om[1, 7] = float("inf")
    out = ob(grid, om)
    torch.cuda.synchronize()
    bad = torch.isnan(out.real) | torch.isnan(out.imag)
    assert bad[:, :, 3].all() and bad[:, :, 7].all()
    assert not bad[:, :, [0, 1, 2, 4, 5, 6, 8, 9]].any()
    empty = ob(grid, torch.zeros(2, 0, device=DEV))
    assert empty.shape == (1, 2, 0)
    z = ad(torch.zeros(1, 2, 0, dtype=torch.complex64, device=DEV), torch.zeros(2, 0, device=DEV))
    assert z.shape == (1, 2, 32, 32) and torch.count_nonzero(z) == 0


@pytest.mark.parametrize("ndim,rows", [(1, 2), (2, 2), (3, 2), (2, 5), (3, 5)])
def test_interp_domega_matches_exact_kb_fd(ndim, rows):
    """d/domega of Re<g, A(omega) x>: HIP gradient vs FP64 central differences of the
    exact-KB operator (smooth in omega; the table path itself is piecewise constant).
    Two rows run the per-lane kernel, five the kernel with taps shared through LDS."""
    im = {1: (40,), 2: (24, 20), 3: (10, 12, 8)}[ndim]
    gen = torch.Generator().manual_seed(10 + ndim)
    grid_size = tuple(2 * v for v in im)
    M = 300
    x = gi.cplx((1, rows) + grid_size, gen, torch.complex128)
    gout = gi.cplx((1, rows, M), gen, torch.complex128)
    om = create_ktraj(ndim, M, gen)
    fw = tkbn.KbInterp(im_size=im, dtype=torch.complex128).to(DEV)
    omd = om.to(DEV).requires_grad_(True)
    y = fw(x.to(DEV), omd)
    torch.sum((y * gout.to(DEV).conj()).real).backward()
    grad = omd.grad.cpu().numpy()
    n_shift = [v // 2 for v in im]
    alpha = [2.34 * 6] * ndim

    def f(o):
        yy = nudft.exact_kb_interp(x.numpy(), o, im, [6] * ndim, alpha, n_shift)
        return np.sum((yy * gout.numpy().conj()).real, axis=(0, 1))  # per-sample

    h = 1e-6
    fd = np.zeros_like(grad)
    for d in range(ndim):
        e = np.zeros_like(om.numpy())
        e[d] = h
        fd[d] = (f(om.numpy() + e) - f(om.numpy() - e)) / (2 * h)
    err = np.linalg.norm(grad - fd) / np.linalg.norm(fd)
    print(f"[domega {ndim}d] rel err vs exact-KB FD: {err:.2e}")
    assert err < 5e-3


@pytest.mark.parametrize("ndim,J,dtype", [(2, 6, torch.complex64), (2, 5, torch.complex128),
                                           (3, 6, torch.complex64), (3, 4, torch.complex128)])
def test_domega_shared_taps_equal_sum_of_rows(ndim, J, dtype):
    """The trajectory gradient of 6 rows (taps shared through LDS, lanes over rows) equals
    the sum of the 6 one-row gradients (per-lane kernel) to rounding, with every
    numpoints at the compile-time maximum (unguarded taps) and below it (guarded)."""
    from torchkbnufft_amd._nufft import interp as _interp

    im = {2: (30, 26), 3: (12, 10, 14)}[ndim]
    gen = torch.Generator().manual_seed(40 + ndim + J)
    grid_size = tuple(2 * v for v in im)
    M, rows = 700, 6
    print(f"[case] domega shared taps {ndim}-D J {J} {dtype}", flush=True)
    x = gi.cplx((1, rows) + grid_size, gen, dtype).to(DEV)
    gout = gi.cplx((1, rows, M), gen, dtype).to(DEV)
    real = torch.float32 if dtype == torch.complex64 else torch.float64
    om = create_ktraj(ndim, M, gen).to(real).to(DEV)
    fw = tkbn.KbInterp(im_size=im, numpoints=J, dtype=dtype).to(DEV)
    args = (fw._tables(), fw.n_shift, fw.numpoints, fw.table_oversamp)
    together = _interp.table_interp_domega(x, gout, om, *args)
    apart = sum(_interp.table_interp_domega(x[:, c:c + 1].contiguous(), gout[:, c:c + 1].contiguous(), om, *args)
                for c in range(rows))
    err = float(torch.linalg.norm(together - apart) / torch.linalg.norm(apart))
    print(f"[domega shared {ndim}d J{J}] rel diff vs sum of rows: {err:.2e}")
    assert err < (1e-5 if dtype == torch.complex64 else 1e-12)


@pytest.mark.parametrize("ndim", [2, 3])
def test_domega_shared_taps_batched_omega(ndim):
    """Per-batch trajectories (omega (B, D, M)) with 5 rows each run the shared-tap
    kernel once for all batches: equal to one call per batch item."""
    from torchkbnufft_amd._nufft import interp as _interp

    im = {2: (20, 18), 3: (8, 10, 8)}[ndim]
    gen = torch.Generator().manual_seed(70 + ndim)
    grid_size = tuple(2 * v for v in im)
    B, C, M = 3, 5, 400
    print(f"[case] domega shared taps, batched omega {ndim}-D", flush=True)
    x = gi.cplx((B, C) + grid_size, gen, torch.complex64).to(DEV)
    gout = gi.cplx((B, C, M), gen, torch.complex64).to(DEV)
    om = torch.stack([create_ktraj(ndim, M, gen) for _ in range(B)]).to(torch.float32).to(DEV)
    fw = tkbn.KbInterp(im_size=im).to(DEV)
    args = (fw._tables(), fw.n_shift, fw.numpoints, fw.table_oversamp)
    together = _interp.table_interp_domega(x, gout, om, *args)
    for b in range(B):
        one = _interp.table_interp_domega(x[b:b + 1].contiguous(), gout[b:b + 1].contiguous(), om[b], *args)
        err = float(torch.linalg.norm(together[b] - one) / torch.linalg.norm(one))
        assert err < 1e-6, (b, err)


@pytest.mark.parametrize("im_size,grid,J", [((33, 35), (65, 70), 6), ((4, 4), (4, 4), 6), ((40,), (77,), 5), ((9, 10, 7), (17, 21, 15), (4, 6, 5))])
@pytest.mark.parametrize("deterministic", [False, True])
def test_adjoint_awkward_grids_match_oracle(im_size, grid, J, deterministic):
    """Grid sizes that are not tile multiples, and grids smaller than the kernel width."""
    gen = torch.Generator().manual_seed(21)
    d = len(im_size)
    ad = tkbn.KbInterpAdjoint(im_size=im_size, grid_size=grid, numpoints=J, dtype=torch.complex64).to(DEV)
    plan = tp.Plan(im_size, grid, J, dtype=torch.complex64)
    M = 3000
    om = create_ktraj(d, M, gen, torch.float32)
    y = gi.cplx((2, 3, M), gen)
    tkbn.set_deterministic(deterministic)
    try:
        out = ad(y.to(DEV), om.to(DEV)).cpu()
    finally:
        tkbn.set_deterministic(False)
    assert rel_l2(out, tp.kb_table_interp_adjoint(plan, y, om)) < C64_TOL


@pytest.mark.parametrize("im_size,grid,J", [((33, 35), (65, 70), 6), ((4, 4), (4, 4), 6), ((40,), (77,), 5),
                                            ((9, 10, 7), (17, 21, 15), (4, 6, 5)), ((3, 4, 3), (5, 6, 4), 6)])
@pytest.mark.parametrize("rows", [1, 5])
def test_forward_awkward_grids_match_oracle(im_size, grid, J, rows):
    """Forward gather on grids that are not tile multiples and grids smaller than the
    kernel width: the staged tile + halo region wraps around the grid once or several
    times per dimension."""
    print(f"[case] forward awkward grid {grid} J {J} rows {rows}", flush=True)
    gen = torch.Generator().manual_seed(23)
    d = len(im_size)
    fw = tkbn.KbInterp(im_size=im_size, grid_size=grid, numpoints=J, dtype=torch.complex64).to(DEV)
    plan = tp.Plan(im_size, grid, J, dtype=torch.complex64)
    M = 3000
    om = create_ktraj(d, M, gen, torch.float32)
    x = gi.cplx((1, rows) + tuple(grid), gen)
    out = fw(x.to(DEV), om.to(DEV)).cpu()
    assert rel_l2(out, tp.kb_table_interp(plan, x, om)) < C64_TOL


def test_adjoint_deterministic_is_bitwise_repeatable():
    print("[case] deterministic adjoint, 320^2 -> 640^2 radial, 8 coils", flush=True)
    gen = torch.Generator().manual_seed(22)
    ktraj = gi.radial(400, 640, golden=True).to(DEV)
    y = gi.cplx((1, 8, ktraj.shape[1]), gen).to(DEV)
    ad = tkbn.KbInterpAdjoint(im_size=(320, 320), dtype=torch.complex64).to(DEV)
    from torchkbnufft_amd._nufft import interp as _interp

    args = (ktraj, ad._tables(), ad.n_shift, ad.numpoints, ad.table_oversamp, ad.offsets, (640, 640))
    outs = [_interp.table_interp_adjoint(y, *args, deterministic=True) for _ in range(3)]
    fast = _interp.table_interp_adjoint(y, *args, deterministic=False)
    torch.cuda.synchronize()
    assert all(torch.equal(outs[0], o) for o in outs[1:])
    assert rel_l2(fast, outs[0]) < 1e-6
    plan = tp.Plan((320, 320), dtype=torch.complex64)
    ref = tp.kb_table_interp_adjoint(plan, y[:, :2].cpu(), ktraj.cpu())
    assert rel_l2(outs[0][:, :2].cpu(), ref) < C64_TOL


@pytest.mark.parametrize("grid,rows,M", [((640, 640), 64, 40000), ((65, 70), 40, 6000), ((9, 11), 96, 3000), ((160, 160), 33, 20000)])
@pytest.mark.parametrize("deterministic", [False, True])
def test_multicoil_adjoint_matches_oracle(grid, rows, M, deterministic):
    """Many coil rows sharing one trajectory (the lanes-per-row spread)."""
    print(f"[case] multicoil adjoint grid={grid} rows={rows} M={M} det={deterministic}", flush=True)
    gen = torch.Generator().manual_seed(31)
    im = tuple(max(1, g // 2) for g in grid)
    om = create_ktraj(2, M, gen, torch.float32)
    y = gi.cplx((1, rows, M), gen)
    ad = tkbn.KbInterpAdjoint(im_size=im, grid_size=grid, dtype=torch.complex64).to(DEV)
    plan = tp.Plan(im, grid, dtype=torch.complex64)
    from torchkbnufft_amd._nufft import interp as _interp

    args = (om.to(DEV), ad._tables(), ad.n_shift, ad.numpoints, ad.table_oversamp, ad.offsets, grid)
    out = _interp.table_interp_adjoint(y.to(DEV), *args, deterministic=deterministic)
    if deterministic:
        again = _interp.table_interp_adjoint(y.to(DEV), *args, deterministic=True)
        assert torch.equal(out, again)
    sel = [0, rows // 2, rows - 1]
    ref = tp.kb_table_interp_adjoint(plan, y[:, sel], om)
    assert rel_l2(out[:, sel].cpu(), ref) < C64_TOL


@pytest.mark.parametrize(
    "grid,rows,M,J",
    [((640, 640), 64, 40000, 6), ((65, 70), 40, 6000, 6), ((9, 11), 96, 3000, 6), ((160, 160), 33, 20000, 4),
     ((100, 90), 70, 20000, (5, 3))],
)
def test_multicoil_forward_matches_oracle(grid, rows, M, J):
    """Many coil rows sharing one trajectory (the lanes-per-row gather), incl. wrapped tiles."""
    print(f"[case] multicoil forward grid={grid} rows={rows} M={M} J={J}", flush=True)
    gen = torch.Generator().manual_seed(37)
    im = tuple(max(1, g // 2) for g in grid)
    om = create_ktraj(2, M, gen, torch.float32)
    x = gi.cplx((1, rows) + tuple(grid), gen)
    fw = tkbn.KbInterp(im_size=im, grid_size=grid, numpoints=J, dtype=torch.complex64).to(DEV)
    out = fw(x.to(DEV), om.to(DEV))
    plan = tp.Plan(im, grid, numpoints=J, dtype=torch.complex64)
    sel = [0, rows // 2, rows - 1]
    ref = tp.kb_table_interp(plan, x[:, sel], om)
    assert rel_l2(out[:, sel].cpu(), ref) < C64_TOL
    # every row goes through the same path: compare all rows against the plain kernel
    from torchkbnufft_amd._nufft import interp as _interp

    old = _interp._BINNED_FWD_MIN_WORK
    try:
        _interp.set_binned_forward_threshold(1 << 62)
        plain = fw(x.to(DEV), om.to(DEV))
    finally:
        _interp.set_binned_forward_threshold(old)
    assert rel_l2(out, plain) < 1e-6


def test_plan_cache_reuse_and_invalidation():
    """The trajectory plan kept in a reused workspace must never go stale."""
    from torchkbnufft_amd._nufft import interp as _interp

    print("[case] plan cache reuse / invalidation", flush=True)
    gen = torch.Generator().manual_seed(41)
    grid, rows, M = (96, 96), 40, 30000
    om = create_ktraj(2, M, gen, torch.float32).to(DEV)
    x = gi.cplx((1, rows) + grid, gen).to(DEV)
    y = gi.cplx((1, rows, M), gen).to(DEV)
    fw = tkbn.KbInterp(im_size=(48, 48), grid_size=grid, dtype=torch.complex64).to(DEV)
    ad = tkbn.KbInterpAdjoint(im_size=(48, 48), grid_size=grid, dtype=torch.complex64).to(DEV)

    def fresh(fn, *a):
        _interp.clear_plan_cache()
        return fn(*a)

    f1, a1 = fw(x, om), ad(y, om)
    assert torch.equal(f1, fw(x, om)) and rel_l2(a1, ad(y, om)) < 1e-6  # reused plan
    # in-place edit that torch does not version (through .data): must be seen
    om.data[0, 5] += 0.25
    om.data[1, M - 3] -= 0.5
    f2, a2 = fw(x, om), ad(y, om)
    assert torch.equal(f2, fresh(fw, x, om))
    assert rel_l2(a2, fresh(ad, y, om)) < 1e-6
    assert not torch.equal(f1, f2)
    # a deterministic call through the same adjoint workspace, then the fast path again
    d = _interp.table_interp_adjoint(y, om, ad._tables(), ad.n_shift, ad.numpoints, ad.table_oversamp, ad.offsets, grid,
                                     deterministic=True)
    assert rel_l2(ad(y, om), d) < 1e-6
    # different table (numpoints) through the same trajectory tensor
    fw4 = tkbn.KbInterp(im_size=(48, 48), grid_size=grid, numpoints=4, dtype=torch.complex64).to(DEV)
    plan4 = tp.Plan((48, 48), grid, numpoints=4, dtype=torch.complex64)
    ref4 = tp.kb_table_interp(plan4, x[:, :2].cpu(), om.cpu())
    assert rel_l2(fw4(x, om)[:, :2].cpu(), ref4) < C64_TOL
    assert torch.equal(fw(x, om), f2)


@pytest.mark.parametrize("rows,deterministic", [(2, False), (2, True), (40, False)])
def test_adjoint_nonfinite_kdata_propagates(rows, deterministic):
    """A NaN sample lands as NaN exactly on its footprint cells (fixed-point paths hand
    non-finite inputs to the float kernel)."""
    from torchkbnufft_amd._nufft import interp as _interp

    print(f"[case] adjoint NaN propagation rows={rows} det={deterministic}", flush=True)
    gen = torch.Generator().manual_seed(43)
    grid, M = (64, 48), 5000
    om = create_ktraj(2, M, gen, torch.float32)
    y = gi.cplx((1, rows, M), gen)
    y[0, :, 17] = complex(float("nan"), 0.0)
    ad = tkbn.KbInterpAdjoint(im_size=(32, 24), grid_size=grid, dtype=torch.complex64).to(DEV)
    args = (om.to(DEV), ad._tables(), ad.n_shift, ad.numpoints, ad.table_oversamp, ad.offsets, grid)
    out = _interp.table_interp_adjoint(y.to(DEV), *args, deterministic=deterministic).cpu()
    nan = torch.isnan(out.real) | torch.isnan(out.imag)
    unit = torch.zeros(1, 1, M, dtype=torch.complex64)
    unit[0, 0, 17] = 1.0
    plan = tp.Plan((32, 24), grid, dtype=torch.complex64)
    foot = tp.kb_table_interp_adjoint(plan, unit, om) != 0
    assert torch.equal(nan[:, :1], foot) and bool(nan.all(dim=1).eq(nan.any(dim=1)).all())
    # every other cell matches the finite computation without that sample
    y2 = y.clone()
    y2[0, :, 17] = 0
    ref = _interp.table_interp_adjoint(y2.to(DEV), *args, deterministic=deterministic).cpu()
    keep = ~nan
    assert rel_l2(out[keep], ref[keep]) < 1e-6


def test_plan_cache_key_covers_tile_shape():
    """One trajectory through the module cache with many and with one coil row: the
    sample-lane tile shape depends on the row count, so the cached plan must not be
    shared between them."""
    from torchkbnufft_amd._nufft import interp as _interp

    print("[case] plan cache across row counts", flush=True)
    gen = torch.Generator().manual_seed(47)
    grid, M = (256, 256), 30000
    om = create_ktraj(2, M, gen, torch.float32).to(DEV)
    fw = tkbn.KbInterp(im_size=(128, 128), grid_size=grid, dtype=torch.complex64).to(DEV)
    ad = tkbn.KbInterpAdjoint(im_size=(128, 128), grid_size=grid, dtype=torch.complex64).to(DEV)
    _interp.clear_plan_cache()
    results = []
    for rows in (12, 1, 12, 1):
        x = gi.cplx((1, rows) + grid, gen).to(DEV)
        y = gi.cplx((1, rows, M), gen).to(DEV)
        results.append((rows, x, y, fw(x, om), ad(y, om)))
    torch.cuda.synchronize()
    for rows, x, y, f, a in results:
        _interp.clear_plan_cache()
        assert rel_l2(f, fw(x, om)) < 1e-6, rows
        _interp.clear_plan_cache()
        assert rel_l2(a, ad(y, om)) < 1e-6, rows


def _rounding_taps(om: np.ndarray, K: int, J: int = 6, L: int = 1024) -> np.ndarray:
    """Samples whose float32 tap indices (the reference's round((tm - g) L)) are not all
    i_0 - j L along this dimension (the multi-coil kernels' complex-weight entries)."""
    om = om.astype(np.float32)
    gam = np.float32(np.float32(1) / np.float32(K)) * np.float32(6.283185307179586)
    tm = (om / gam).astype(np.float32)
    base = np.floor(tm - np.float32(J / 2)).astype(np.float32) + np.float32(1)
    i0 = np.rint((tm - base).astype(np.float32) * np.float32(L))
    bad = np.zeros(om.shape, bool)
    for j in range(1, J):
        bad |= np.rint((tm - (base + np.float32(j))).astype(np.float32) * np.float32(L)) != i0 - j * L
    return bad


@pytest.mark.parametrize("table", ["reference", "perturbed"])
def test_multicoil_exception_entries_match_oracle(table):
    """The multi-coil kernels apply real tap weights on a phase-rotated grid where the
    table is the reference's KB(u) exp(-i pi (N-1) u / K) and every tap index is i_0 - j L;
    other entries take complex weights.  Samples near k = 0 hit the float32 rounding
    case, and a table with an extra non-linear phase makes every entry complex: both
    against the oracle, forward and adjoint, 40 coil rows."""
    print(f"[case] multicoil exception entries table={table}", flush=True)
    gen = torch.Generator().manual_seed(43)
    im, grid, rows = (64, 64), (128, 128), 40
    # rounding cases are rare (about 1 in 10^4 near k = 0): take them from a scan of
    # 4e5 near-centre values, paired with random second coordinates
    near = ((torch.rand(400000, generator=gen) - 0.5) * 0.4).numpy().astype(np.float32)
    hits = torch.from_numpy(near[_rounding_taps(near, 128)][:48])
    assert hits.numel() >= 16
    odd = torch.stack([hits, (torch.rand(hits.numel(), generator=gen) - 0.5) * 6.0])
    om = torch.cat([create_ktraj(2, 3000, gen, torch.float32), odd, odd.flip(0)], dim=1)
    bad = _rounding_taps(om[0].numpy(), 128) | _rounding_taps(om[1].numpy(), 128)
    assert bad.sum() >= 2 * hits.numel()
    ad = tkbn.KbInterpAdjoint(im_size=im, grid_size=grid, dtype=torch.complex64).to(DEV)
    plan = tp.Plan(im, grid, dtype=torch.complex64)
    tabs = [t.cpu() for t in plan.tables]
    if table == "perturbed":
        m = torch.arange(tabs[0].numel(), dtype=torch.float64)
        tabs = [(t.to(torch.complex128) * torch.exp(1j * 0.3 * torch.sin(m / 97.0))).to(torch.complex64) for t in tabs]
    from torchkbnufft_amd._nufft import interp as _interp

    x = gi.cplx((1, rows) + grid, gen)
    y = gi.cplx((1, rows, om.shape[1]), gen)
    dt = [t.to(DEV) for t in tabs]
    args = (ad.n_shift, ad.numpoints, ad.table_oversamp, ad.offsets)
    kd = _interp.table_interp(x.to(DEV), om.to(DEV), dt, *args)
    assert _lib.load().tkbn_last_interp_kernel().decode() == "gather_sep_kernel"
    gr = _interp.table_interp_adjoint(y.to(DEV), om.to(DEV), dt, *args, grid, deterministic=False)
    assert _lib.load().tkbn_last_interp_kernel().decode() == "spread_sep_kernel"
    ns, J, L, off = plan.t_n_shift, plan.t_numpoints, plan.t_table_oversamp, plan.offsets
    sel = [0, 17, rows - 1]
    ref_f = tp.table_interp(x[:, sel], om, tabs, ns, J, L, off)
    ref_a = tp.table_interp_adjoint(y[:, sel], om, tabs, ns, J, L, off, plan.t_grid_size)
    e_f, e_a = rel_l2(kd[:, sel].cpu(), ref_f), rel_l2(gr[:, sel].cpu(), ref_a)
    # the rounding-case samples alone
    idx = torch.nonzero(torch.as_tensor(bad)).flatten()
    e_fb = rel_l2(kd[:, sel][..., idx].cpu(), ref_f[..., idx])
    print(f"[exceptions {table}] {int(bad.sum())} rounding samples; fwd {e_f:.2e} (those {e_fb:.2e}) adj {e_a:.2e}")
    assert e_f < C64_TOL and e_a < C64_TOL and e_fb < C64_TOL


@pytest.mark.parametrize(
    "im_size,grid,rows,M",
    [((64, 60), (128, 120), 3, 20000), ((20, 18, 22), (40, 36, 44), 2, 20000), ((300,), (600,), 4, 5000)],
)
def test_tuned_tiles_and_targets_match_oracle(im_size, grid, rows, M):
    """Every autotuner candidate (tile edge x work-item target, both directions, both
    adjoint modes) gives the oracle's result; the plan is rebuilt when the tuning of a
    holder changes and reused while it stays."""
    from torchkbnufft_amd._nufft import autotune, interp as _interp

    print(f"[case] tuned tiles grid={grid} rows={rows} M={M}", flush=True)
    gen = torch.Generator().manual_seed(61)
    d = len(grid)
    om = create_ktraj(d, M, gen, torch.float32)
    om[:, : M // 2] *= 0.05  # a dense centre: hot tiles cut into several work items
    x = gi.cplx((1, rows) + grid, gen)
    y = gi.cplx((1, rows, M), gen)
    ob = tkbn.KbInterp(im_size=im_size, grid_size=grid, dtype=torch.complex64).to(DEV)
    plan = tp.Plan(im_size, grid, dtype=torch.complex64)
    ref_f = tp.kb_table_interp(plan, x, om)
    ref_a = tp.kb_table_interp_adjoint(plan, y, om)
    args = (om.to(DEV), ob._tables(), ob.n_shift, ob.numpoints, ob.table_oversamp, ob.offsets)
    xd, yd = x.to(DEV), y.to(DEV)
    jj = [6] * d
    fwd_holder = _interp.PlanHolder(trusted=True)
    for tile, target in autotune.candidates("fwd", grid, jj, False):
        fwd_holder.set_tuning(tile, target)
        for _ in range(2):  # build, then reuse
            out = _interp.table_interp(xd, *args, plan=fwd_holder, binned=True).cpu()
            assert rel_l2(out, ref_f) < C64_TOL, (tile, target)
    for det in (False, True):
        holder = _interp.PlanHolder(trusted=True)
        for tile, target in autotune.candidates("adj", grid, jj, det):
            holder.set_tuning(tile, target)
            for _ in range(2):
                out = _interp.table_interp_adjoint(yd, *args, grid, deterministic=det, plan=holder).cpu()
                assert rel_l2(out, ref_a) < C64_TOL, (tile, target, det)
