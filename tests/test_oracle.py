"""CPU: pin the oracle before it is trusted as the GPU parity checker.

* against the reference's own outputs (golden vectors written from the reference,
  ``tests/golden/make_golden.py``), complex64 and float64;
* against mathematics that is independent of the restatement: the FP64 NUDFT,
  the closed-form (exact) Kaiser-Bessel interpolation, a dense brute-force
  adjoint, dot tests, closed-form kernel/derivative identities.
"""
import numpy as np
import pytest
import torch

import inputs as gi
from conftest import create_input_plus_noise, create_ktraj, golden, rel_l2
from oracle import kb, nudft
from oracle import table_path as tp


@pytest.mark.parametrize("case", gi.C64_CASES, ids=[c[0] for c in gi.C64_CASES])
def test_oracle_matches_reference_c64(case):
    name, im, grid, J, L, B, C, mk, use_smaps, norm, batched = case
    g = golden("golden_c64.npz")
    inp = gi.c64_inputs(case)
    plan = tp.Plan(im, grid, J, None, L, dtype=torch.complex64)
    k, sm = inp["ktraj"], inp["smaps"]
    fwd_i = tp.kb_table_interp(plan, inp["grid_in"], k)
    # interpolation reproduces the reference's float32 index arithmetic: nearly all
    # outputs are bit-identical
    assert (fwd_i.numpy() == g[name + "_interp"]).mean() > 0.99
    assert rel_l2(fwd_i, g[name + "_interp"]) < 1e-7
    assert rel_l2(tp.kb_table_interp_adjoint(plan, inp["y_in"], k), g[name + "_interp_adj"]) < 1e-7
    assert rel_l2(tp.kb_table_nufft(plan, inp["x"], k, sm, norm), g[name + "_nufft"]) < 1e-6
    assert rel_l2(tp.kb_table_nufft_adjoint(plan, inp["y_in"], k, sm, norm), g[name + "_nufft_adj"]) < 1e-6


def test_oracle_matches_reference_f64_fixtures():
    g = golden("fixtures_f64.npz")
    for i in range(len(gi.INTERP_FIXTURES)):
        im_size, image, ktraj = gi.interp_fixture_inputs(i)
        plan = tp.Plan(im_size, im_size, dtype=torch.complex128)
        out = torch.view_as_real(tp.kb_table_interp(plan, torch.view_as_complex(image), ktraj))
        assert torch.allclose(out, torch.from_numpy(g[f"interp{i}_kdata"]))
    for i in range(len(gi.NUFFT_FIXTURES)):
        im_size, image, ktraj = gi.nufft_fixture_inputs(i)
        plan = tp.Plan(im_size, dtype=torch.complex128)
        kd = tp.kb_table_nufft(plan, torch.view_as_complex(image), ktraj)
        assert torch.allclose(torch.view_as_real(kd), torch.from_numpy(g[f"nufft{i}_kdata"]))
        if f"nufft{i}_adj" in g.files:
            adj = tp.kb_table_nufft_adjoint(plan, kd, ktraj)
            assert torch.allclose(torch.view_as_real(adj), torch.from_numpy(g[f"nufft{i}_adj"]))


@pytest.mark.parametrize("idx", range(len(gi.PLAN_CASES)))
def test_closed_form_kb_matches_reference_plan(idx):
    im, grid, J, L, w = gi.PLAN_CASES[idx]
    p = golden("plans_c128.npz")
    d = len(im)
    gs = gi.grid_of(im, grid)
    Js = (J,) * d if isinstance(J, int) else J
    for k in range(d):
        key = f"p{idx}_table{k}"
        t = kb.kb_table(im[k], gs[k], Js[k], L, w * Js[k])
        if key in p.files:
            assert np.abs(t - p[key]).max() < 1e-13
        # spot values: centre is KB(0)*exp(0) = 1, both ends are outside |u| < J/2
        assert abs(t[Js[k] * L // 2] - 1) < 1e-15 if (Js[k] * L) % 2 == 0 else True
        assert t[0] == 0 and t[-1] == 0
        mid = Js[k] * L // 2
        if (Js[k] * L) % 2 == 0:
            r = np.arange(1, mid)
            assert np.allclose(np.abs(t[mid + r]), np.abs(t[mid - r]), rtol=0, atol=1e-15)
    sc = kb.apodization(im, gs, Js, [w * j for j in Js], [0.0] * d)
    for k in range(d):
        line = p[f"p{idx}_scaling_line{k}"]
        sl = [0] * d
        sl[k] = slice(None)
        assert np.abs(sc[tuple(sl)] - line).max() / np.abs(line).max() < 1e-12
    if f"p{idx}_scaling_coef" in p.files:
        full = p[f"p{idx}_scaling_coef"]
        assert np.abs(sc - full).max() / np.abs(full).max() < 1e-12


def test_kb_ft_matches_bessel_j_form():
    # J sin z / (z I0(alpha)) == sqrt(2 pi)(J/2) J_{1/2}(z)/(z^{1/2} I0(alpha)) via scipy
    from scipy import special

    w = np.linspace(-0.3, 0.3, 101)
    J, a = 6, 2.34 * 6
    z = np.sqrt((np.pi * J * w) ** 2 - a**2 + 0j)
    ref = np.real(np.sqrt(2 * np.pi) * (J / 2) / special.iv(0, a) * special.jv(0.5, z) / z**0.5)
    assert np.allclose(kb.kb_ft(w, J, a), ref, rtol=1e-12, atol=0)


@pytest.mark.parametrize("im_size", [(20,), (20, 24), (8, 10, 6)])
def test_table_nufft_approximates_nudft(im_size):
    torch.manual_seed(0)
    plan = tp.Plan(im_size, dtype=torch.complex128)
    x = torch.randn((2, 1) + im_size, dtype=torch.complex128)
    om = create_ktraj(len(im_size), 300)
    y = tp.kb_table_nufft(plan, x, om)
    ref = nudft.nudft(x.numpy(), om.numpy(), plan.n_shift)
    err = rel_l2(y, ref)
    # J=6, 2x oversampling, L=1024: table quantisation dominates (~7e-4)
    assert 1e-5 < err < 2e-3
    d = torch.randn(2, 1, 300, dtype=torch.complex128)
    xa = tp.kb_table_nufft_adjoint(plan, d, om)
    ref_a = nudft.nudft_adjoint(d.numpy(), om.numpy(), im_size, plan.n_shift)
    assert rel_l2(xa, ref_a) < 2e-3


@pytest.mark.parametrize("im_size,grid", [((20,), (40,)), ((12, 16), (24, 32)), ((6, 8, 5), (12, 16, 10))])
def test_table_interp_matches_exact_kb(im_size, grid):
    # closed-form kernel at the exact distance vs table lookup at round(u*L)/L
    torch.manual_seed(1)
    plan = tp.Plan(im_size, grid, dtype=torch.complex128)
    g = torch.randn((1, 2) + grid, dtype=torch.complex128)
    om = create_ktraj(len(im_size), 400)
    y = tp.kb_table_interp(plan, g, om)
    ref = nudft.exact_kb_interp(g.numpy(), om.numpy(), im_size, plan.numpoints, plan.alpha, plan.n_shift)
    assert rel_l2(y, ref) < 2e-3


@pytest.mark.parametrize("im_size", [(7,), (5, 6)])
def test_adjoint_is_dense_conjugate_transpose(im_size):
    torch.manual_seed(2)
    plan = tp.Plan(im_size, im_size, numpoints=4, dtype=torch.complex128)
    m = 9
    om = create_ktraj(len(im_size), m)
    G = int(np.prod(im_size))
    A = torch.zeros(m, G, dtype=torch.complex128)
    for col in range(G):
        e = torch.zeros((1, 1, G), dtype=torch.complex128)
        e[0, 0, col] = 1
        A[:, col] = tp.kb_table_interp(plan, e.reshape((1, 1) + im_size), om)[0, 0]
    AH = torch.zeros(G, m, dtype=torch.complex128)
    for row in range(m):
        e = torch.zeros(1, 1, m, dtype=torch.complex128)
        e[0, 0, row] = 1
        AH[:, row] = tp.kb_table_interp_adjoint(plan, e, om).reshape(-1)
    assert torch.allclose(AH, A.conj().T, rtol=0, atol=1e-14)


@pytest.mark.parametrize("im_size,smaps", [((19,), False), ((13, 12), True), ((7, 9, 6), True)])
def test_oracle_dot_product(im_size, smaps):
    torch.manual_seed(3)
    plan = tp.Plan(im_size, dtype=torch.complex128)
    C = 3
    x = create_input_plus_noise((2, 1 if smaps else C) + im_size, True)
    s = create_input_plus_noise((2, C) + im_size, True) if smaps else None
    y = create_input_plus_noise((2, C, 50), True)
    om = create_ktraj(len(im_size), 50)
    lhs = tp.inner_product(tp.kb_table_nufft(plan, x, om, s), y)
    rhs = tp.inner_product(x, tp.kb_table_nufft_adjoint(plan, y, om, s))
    assert torch.allclose(lhs, rhs)


@pytest.mark.parametrize("im_size", [(16,), (12, 10), (6, 8, 5)])
def test_nudft_on_grid_is_shifted_fft(im_size):
    """Samples on the uniform grid, omega = 2 pi k / N: the NUDFT is numpy's FFT times
    the shift phase exp(i 2 pi k . s / N), and its adjoint is the unnormalised inverse
    FFT times the conjugate phase (closed form, 1e-12)."""
    rng = np.random.default_rng(5)
    d = len(im_size)
    n_shift = [n // 2 for n in im_size]
    x = rng.standard_normal((2,) + im_size) + 1j * rng.standard_normal((2,) + im_size)
    ks = np.meshgrid(*[np.arange(n) for n in im_size], indexing="ij")
    k = np.stack([kk.reshape(-1) for kk in ks])  # (d, P) frequency indices, row-major
    omega = 2 * np.pi * k / np.array(im_size, dtype=np.float64)[:, None]
    phase = np.exp(1j * 2 * np.pi * np.sum(k * (np.array(n_shift)[:, None] / np.array(im_size)[:, None]), axis=0))
    axes = tuple(range(1, d + 1))
    ref = np.fft.fftn(x, axes=axes).reshape(2, -1) * phase
    y = nudft.nudft(x, omega, n_shift)
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-12
    yy = rng.standard_normal((2, k.shape[1])) + 1j * rng.standard_normal((2, k.shape[1]))
    adj = nudft.nudft_adjoint(yy, omega, im_size, n_shift)
    ref_adj = np.fft.ifftn((yy * np.conj(phase)).reshape((2,) + im_size), axes=axes) * np.prod(im_size)
    assert np.linalg.norm(adj - ref_adj) / np.linalg.norm(ref_adj) < 1e-12


def test_nudft_domega_matches_fd():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((1, 6, 7)) + 1j * rng.standard_normal((1, 6, 7))
    om = rng.uniform(-np.pi, np.pi, (2, 5))
    an = nudft.nudft_domega(x, om, (3, 3))
    h = 1e-6
    for d in range(2):
        e = np.zeros_like(om)
        e[d] = h
        fd = (nudft.nudft(x, om + e, (3, 3)) - nudft.nudft(x, om - e, (3, 3))) / (2 * h)
        assert np.allclose(an[d], fd, rtol=1e-6, atol=1e-7)


def test_kb_derivative_matches_fd():
    u = np.linspace(-2.9, 2.9, 57)
    h = 1e-6
    fd = (kb.kb_value(u + h, 6, 14.04) - kb.kb_value(u - h, 6, 14.04)) / (2 * h)
    assert np.allclose(kb.kb_deriv(u, 6, 14.04), fd, rtol=1e-6, atol=1e-8)
    t = kb.kb_table_deriv(16, 32, 6, 1024, 14.04)
    v = kb.kb_table(16, 32, 6, 1024, 14.04)
    # central difference of the table itself, away from the support edges (jumps)
    cd = (v[3:-2] - v[1:-4]) * 1024 / 2
    assert np.abs(cd - t[2:-3]).max() / np.abs(t).max() < 1e-4


def _off_edge(om, grid, numpoints, margin):
    """Samples whose footprint edges are at least `margin` grid units from a grid line
    in every dimension (the exact-KB operator is smooth there)."""
    tm = om * (np.asarray(grid, dtype=np.float64)[:, None] / (2 * np.pi)) - numpoints / 2.0
    return np.all(np.abs(tm - np.rint(tm)) > margin, axis=0)


@pytest.mark.parametrize("im_size", [(9,), (8, 7), (5, 6, 4)])
def test_exact_kb_domega_matches_fd(im_size):
    rng = np.random.default_rng(11)
    d = len(im_size)
    grid = tuple(2 * n for n in im_size)
    g = rng.standard_normal((2,) + grid) + 1j * rng.standard_normal((2,) + grid)
    om = rng.uniform(-np.pi, np.pi, (d, 40))
    om = om[:, _off_edge(om, grid, 6, 1e-3)]
    alpha, J, shift = [2.34 * 6] * d, [6] * d, [n // 2 for n in im_size]
    an = nudft.exact_kb_interp_domega(g, om, im_size, J, alpha, shift)
    h = 1e-7
    for k in range(d):
        e = np.zeros_like(om)
        e[k] = h
        fd = (nudft.exact_kb_interp(g, om + e, im_size, J, alpha, shift)
              - nudft.exact_kb_interp(g, om - e, im_size, J, alpha, shift)) / (2 * h)
        assert np.linalg.norm(an[k] - fd) / np.linalg.norm(fd) < 1e-6


def test_directional_fd_protocol_matches_analytic_gradient():
    """The trajectory-gradient protocol of the differentiation workload, at a reduced
    size: loss ||A(x0, omega) - y||^2 with the exact-KB operator, FP64 central
    differences (step 1e-6) along random unit directions supported on 256 samples away
    from footprint edges, against the analytic gradient 2 Re sum_c conj(r) dA/domega.
    A spiral interleave starts at omega = 0, exactly on a footprint edge, so the edge
    filter is part of the protocol."""
    rng = np.random.default_rng(12)
    n, coils = 32, 3
    t = np.arange(96) / 96
    om = np.concatenate([np.stack([np.pi * t**1.5 * np.sin(2 * np.pi * 3 * t + l), np.pi * t**1.5 *
                                   np.cos(2 * np.pi * 3 * t + l)]) for l in range(8)], axis=1)
    grid = (2 * n, 2 * n)
    G = rng.standard_normal((coils,) + grid) + 1j * rng.standard_normal((coils,) + grid)
    y = rng.standard_normal((coils, om.shape[1])) + 1j * rng.standard_normal((coils, om.shape[1]))
    args = ((n, n), (6, 6), [2.34 * 6] * 2, (n // 2, n // 2))
    r = nudft.exact_kb_interp(G, om, *args) - y
    grad = 2 * np.sum(np.real(np.conj(r)[None] * nudft.exact_kb_interp_domega(G, om, *args)), axis=1)
    ok = _off_edge(om, grid, 6, 1e-3)
    assert not ok[0]  # omega = 0 sits on an edge
    pool = np.nonzero(ok)[0]
    h = 1e-6
    for _ in range(8):
        sub = rng.choice(pool, 256, replace=False)
        dirv = rng.standard_normal((2, sub.size))
        dirv /= np.linalg.norm(dirv)
        o = om.copy()
        o[:, sub] += h * dirv
        lp = np.sum(np.abs(nudft.exact_kb_interp(G, o, *args) - y) ** 2)
        o[:, sub] -= 2 * h * dirv
        lm = np.sum(np.abs(nudft.exact_kb_interp(G, o, *args) - y) ** 2)
        fd = (lp - lm) / (2 * h)
        assert abs(np.sum(grad[:, sub] * dirv) - fd) <= 1e-5 * abs(fd)


def _dense_interp_matrix(plan, om, grid):
    """Columns of the forward interpolation applied to unit grids: (M, prod(grid))."""
    g = int(np.prod(grid))
    cols = []
    for col in range(g):
        e = torch.zeros((1, 1, g), dtype=plan.tables[0].dtype)
        e[0, 0, col] = 1
        cols.append(tp.kb_table_interp(plan, e.reshape((1, 1) + tuple(grid)), om)[0, 0])
    return torch.stack(cols, dim=1)


def test_density_compensation_matches_dense_iteration():
    """Pipe's iteration on the dense interpolation matrix A (w <- w / |A A^H w|)."""
    torch.manual_seed(5)
    im, grid = (5, 6), (10, 12)
    plan = tp.Plan(im, grid, numpoints=4, dtype=torch.complex128)
    om = create_ktraj(2, 30)
    A = _dense_interp_matrix(plan, om, grid).numpy()
    w = np.ones(30, dtype=np.complex128)
    for _ in range(7):
        w = w / np.abs(A @ (A.conj().T @ w))
    out = tp.density_compensation(plan, om, num_iterations=7)
    assert out.shape == (1, 1, 30)
    np.testing.assert_allclose(out[0, 0].numpy(), w, rtol=1e-12, atol=0)


def test_density_compensation_uniform_grid_is_flat_fixed_point():
    """Samples on every oversampled grid point: the kernel matrix is circulant, so
    the weights are equal and |A A^H w| = 1 (the iteration's fixed point).  The
    trajectory's float rounding puts a few samples a table step off the grid point,
    hence the loose flatness bound."""
    im, grid = (6, 5), (12, 10)
    plan = tp.Plan(im, grid, numpoints=4, dtype=torch.complex128)
    k0, k1 = torch.meshgrid(torch.arange(grid[0]), torch.arange(grid[1]), indexing="ij")
    om = torch.stack([2 * np.pi * k0.reshape(-1) / grid[0], 2 * np.pi * k1.reshape(-1) / grid[1]]).double()
    om = torch.remainder(om + np.pi, 2 * np.pi) - np.pi
    w = tp.density_compensation(plan, om)
    assert float(w.real.max() / w.real.min()) - 1 < 1e-5
    back = tp.kb_table_interp(plan, tp.kb_table_interp_adjoint(plan, w, om), om)
    assert torch.allclose(back.abs(), torch.ones_like(back.real), rtol=0, atol=1e-7)


def test_density_compensation_batched_is_loop():
    torch.manual_seed(6)
    plan = tp.Plan((7, 8), numpoints=6, dtype=torch.complex128)
    om = torch.rand((3, 2, 50), dtype=torch.float64) * 2 * np.pi - np.pi
    loop = torch.cat([tp.density_compensation(plan, o) for o in om])
    assert torch.allclose(tp.density_compensation(plan, om), loop, rtol=1e-13, atol=0)


@pytest.mark.parametrize("im_size", [(9,), (6, 5), (4, 5, 3)])
@pytest.mark.parametrize("norm", [None, "ortho"])
def test_toeplitz_assembly_is_exact_with_nudft_psf(im_size, norm):
    """With the exact NUDFT adjoint as the point-spread function, the assembled kernel
    reproduces A^H W A of the NUDFT to rounding: the flips, the Hermitian half, the
    zero middle row and the scaling are all checked."""
    torch.manual_seed(7)
    d = len(im_size)
    om = create_ktraj(d, 40)
    w = torch.rand(1, 40, dtype=torch.float64)
    sc = 1.0 if norm is None else 1.0 / np.sqrt(np.prod([2 * n for n in im_size]))

    def exact_adjoint(wt, o):
        return torch.from_numpy(nudft.nudft_adjoint(wt.numpy(), o.numpy(), im_size, [0] * d)) * sc

    kernel = tp.toeplitz_kernel(om, im_size, weights=w, norm=norm, dtype=torch.complex128, adjoint=exact_adjoint)
    x = torch.complex(torch.randn((1, 1) + im_size, dtype=torch.float64), torch.randn((1, 1) + im_size, dtype=torch.float64))
    y = nudft.nudft(x.numpy(), om.numpy(), [0] * d) * w.numpy()
    ref = nudft.nudft_adjoint(y, om.numpy(), im_size, [0] * d)
    if norm == "ortho":
        ref = ref / np.prod([2 * n for n in im_size])
    out = tp.toep_nufft(x, kernel, norm=norm)
    assert rel_l2(out, torch.from_numpy(ref)) < 1e-10


@pytest.mark.parametrize(
    "shape, grid_size, M, norm, tol",
    [
        ([1, 3, 19], [57], 25, "ortho", 1e-2),
        ([1, 4, 32, 16], [64, 24], 83, None, 1e-2),
        ([2, 4, 32, 16], None, 83, "ortho", 1e-4),
        ([1, 3, 7, 6, 5], [14, 12, 10], 60, None, 1e-4),
    ],
)
def test_toeplitz_oracle_approximates_normal_operator(shape, grid_size, M, norm, tol):
    """The reference's own accuracy criterion (tests/test_toep.py:9-55) on the oracle."""
    torch.manual_seed(123)
    im_size = tuple(shape[2:])
    B, C = shape[0], shape[1]
    x = create_input_plus_noise([B, 1] + list(im_size), True)
    smaps = create_input_plus_noise(shape, True)
    om = create_ktraj(len(im_size), M)
    plan = tp.Plan(im_size, grid_size, dtype=torch.complex128)
    kernel = tp.toeplitz_kernel(om, im_size, norm=norm, grid_size=grid_size, dtype=torch.complex128)
    fbn = tp.kb_table_nufft_adjoint(plan, tp.kb_table_nufft(plan, x, om, smaps, norm), om, smaps, norm)
    fbt = tp.toep_nufft(x, kernel, smaps, norm)
    assert rel_l2(fbt, fbn) < tol
