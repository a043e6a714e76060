"""GPU parity at the benchmark workloads' own scale (complex64 unless stated).

* C2, in full: batch 4 x 16 coils, 320^2 -> 640^2, 400 x 640 golden-angle radial with
  smoothed maps.  Every one of the 64 rows of the forward, the spread grid of every
  row and the coil-combined adjoint image are compared with the oracle, in both
  adjoint modes, and the kernels that ran are asserted (the 64-row multi-coil
  kernels, not the sample-lane ones).
* Split hot tiles of the multi-coil spread on a centre-dense radial trajectory.
* C3: the batched form over the 64 uneven problems against the oracle on a subset
  covering every size and both interleave extremes; graph replay against eager.
* C4: trajectory gradient at 256^2, 12 coils, 64 x 1536 spiral, against FP64 central
  differences of the exact Kaiser-Bessel operator along 32 random directions.
* C5: a reduced kooshball against the oracle in deterministic mode (forward, adjoint,
  image and trajectory gradients), and the full 160^3 case through linearity,
  adjointness and bitwise repeatability.
"""
import numpy as np
import pytest
import torch

import torchkbnufft_amd as tkbn
import workloads as W
from conftest import rel_l2
from oracle import nudft
from oracle import table_path as tp
from torchkbnufft_amd import _lib
from torchkbnufft_amd._nufft import interp as _interp

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")
C64_TOL = 1e-5


def _last_kernel() -> str:
    return _lib.load().tkbn_last_interp_kernel().decode()


def _row_errors(a, b):
    a = a.detach().cpu().reshape(-1, int(np.prod(a.shape[2:])))
    b = torch.as_tensor(b).reshape(a.shape)
    return [rel_l2(a[i], b[i]) for i in range(a.shape[0])]


@pytest.fixture(scope="module")
def c2():
    x, smaps, ktraj = W.c2()
    plan = tp.Plan((320, 320), dtype=torch.complex64)
    kd_ref = tp.kb_table_nufft(plan, x, ktraj, smaps)
    return x, smaps, ktraj, plan, kd_ref


def test_c2_forward_all_rows_match_oracle(c2):
    x, smaps, ktraj, plan, kd_ref = c2
    print("[case] C2 forward, 4 x 16 coils, all rows", flush=True)
    fw = tkbn.KbNufft(im_size=(320, 320), dtype=torch.complex64).to(DEV)
    kd = fw(x.to(DEV), ktraj.to(DEV), smaps=smaps.to(DEV))
    assert _last_kernel() == "gather_sep_kernel"
    errs = _row_errors(kd, kd_ref)
    print(f"[C2 fwd] rel L2 all {rel_l2(kd, kd_ref):.2e}, worst row {max(errs):.2e}")
    assert len(errs) == 64 and max(errs) < C64_TOL


@pytest.mark.parametrize("deterministic", [False, True])
def test_c2_adjoint_all_rows_match_oracle(c2, deterministic):
    x, smaps, ktraj, plan, kd_ref = c2
    print(f"[case] C2 adjoint deterministic={deterministic}", flush=True)
    kt = ktraj.to(DEV)
    y = kd_ref.to(DEV)
    args = (kt,) + _module_args((320, 320)) + ((640, 640),)
    grid = _interp.table_interp_adjoint(y, *args, deterministic=deterministic)
    want = "spread_tiles_kernel<fixed,deterministic>" if deterministic else "spread_sep_kernel"
    assert _last_kernel() == want
    grid_ref = tp.kb_table_interp_adjoint(plan, kd_ref, ktraj)
    errs = _row_errors(grid, grid_ref)
    print(f"[C2 spread det={deterministic}] worst row {max(errs):.2e}")
    assert len(errs) == 64 and max(errs) < C64_TOL
    del grid, grid_ref
    ad = tkbn.KbNufftAdjoint(im_size=(320, 320), dtype=torch.complex64).to(DEV)
    tkbn.set_deterministic(deterministic)
    try:
        img = ad(y, kt, smaps=smaps.to(DEV))
    finally:
        tkbn.set_deterministic(False)
    img_ref = tp.kb_table_nufft_adjoint(plan, kd_ref, ktraj, smaps)
    err = rel_l2(img, img_ref)
    print(f"[C2 adjoint image det={deterministic}] rel L2 {err:.2e}")
    assert err < C64_TOL


def _module_args(im, grid=None, numpoints=6):
    ad = tkbn.KbInterpAdjoint(im_size=im, grid_size=grid, numpoints=numpoints, dtype=torch.complex64).to(DEV)
    return ad._tables(), ad.n_shift, ad.numpoints, ad.table_oversamp, ad.offsets


@pytest.mark.parametrize("rows", [64, 40])
def test_multicoil_split_tiles_match_oracle(rows):
    """Centre-dense radial (100 x 320 on a 160^2 grid): centre tiles hold ~4000 entries,
    so the multi-coil spread cuts them into pieces added into pre-zeroed tiles."""
    print(f"[case] split hot tiles rows={rows}", flush=True)
    gen = torch.Generator().manual_seed(61)
    om = W.radial_golden(100, 320)
    grid = (160, 160)
    y = W.cplx((1, rows, om.shape[1]), gen)
    args = (om.to(DEV),) + _module_args((80, 80)) + (grid,)
    out = _interp.table_interp_adjoint(y.to(DEV), *args, deterministic=False)
    assert _last_kernel() == "spread_sep_kernel"
    plan = tp.Plan((80, 80), grid, dtype=torch.complex64)
    sel = list(range(0, rows, 7)) + [rows - 1]
    ref = tp.kb_table_interp_adjoint(plan, y[:, sel], om)
    errs = _row_errors(out[:, sel], ref)
    assert max(errs) < C64_TOL, errs
    x = W.cplx((1, rows) + grid, gen)
    fw = tkbn.KbInterp(im_size=(80, 80), grid_size=grid, dtype=torch.complex64).to(DEV)
    kd = fw(x.to(DEV), om.to(DEV))
    assert _last_kernel() == "gather_sep_kernel"
    errs = _row_errors(kd[:, sel], tp.kb_table_interp(plan, x[:, sel], om))
    assert max(errs) < C64_TOL, errs


def test_deterministic_complex128_matches_oracle():
    """complex128 deterministic adjoint (64-bit fixed point, quantum 2 max|y| V 2^-60 per
    tile): within 1e-11 relative L2 of the float64 oracle on a centre-dense trajectory."""
    print("[case] deterministic complex128", flush=True)
    gen = torch.Generator().manual_seed(62)
    om = W.radial_golden(64, 256).double()
    y = W.cplx((1, 3, om.shape[1]), gen, torch.complex128)
    ad = tkbn.KbInterpAdjoint(im_size=(64, 64), dtype=torch.complex128).to(DEV)
    args = (om.to(DEV), ad._tables(), ad.n_shift, ad.numpoints, ad.table_oversamp, ad.offsets, (128, 128))
    out = _interp.table_interp_adjoint(y.to(DEV), *args, deterministic=True)
    again = _interp.table_interp_adjoint(y.to(DEV), *args, deterministic=True)
    assert torch.equal(out, again)
    plan = tp.Plan((64, 64), dtype=torch.complex128)
    err = rel_l2(out, tp.kb_table_interp_adjoint(plan, y, om))
    print(f"[c128 det] rel L2 {err:.2e}")
    assert err < 1e-11


# --------------------------------------------------------------------------- C3
def test_c3_batched_matches_oracle():
    sizes, inter, ktrajs, images = W.c3()
    print(f"[case] C3 batched, {len(sizes)} problems", flush=True)
    bt = tkbn.BatchedKbNufft(sizes, ktrajs, coils=8, device=DEV)
    for p, x in enumerate(images):
        bt.images[p].copy_(x.to(DEV))
    bt.pair()
    torch.cuda.synchronize()
    assert bt.fused  # one gather and one spread launch for all 64 problems
    print(f"[C3] tuning per size: {bt.tuning}")
    kd = [k.clone() for k in bt.kdata]
    eager = [a.clone() for a in bt.adjoint_images]
    bt.capture()
    for _ in range(3):
        out = bt.replay()
        torch.cuda.synchronize()
        for a, b in zip(out, eager):
            assert rel_l2(a, b) < 1e-6
    pick = W.problem_subset(sizes, inter, 8)
    assert len({sizes[p] for p in pick}) == 3
    plans = {}
    fp64 = []
    for p in pick:
        s = sizes[p]
        plan = plans.setdefault(s, tp.Plan(s, dtype=torch.complex64))
        ref = tp.kb_table_nufft(plan, images[p], ktrajs[p])
        e_f = rel_l2(kd[p], ref)
        e_a = rel_l2(eager[p], tp.kb_table_nufft_adjoint(plan, kd[p].cpu(), ktrajs[p]))
        print(f"[C3 p={p} N={s[0]} interleaves={inter[p]}] fwd {e_f:.2e} adj {e_a:.2e}")
        assert e_f < C64_TOL and e_a < C64_TOL
        if len(fp64) < 3:
            sel = torch.arange(0, ktrajs[p].shape[1], 37)
            y = kd[p][0, 0, sel].cpu().numpy()
            r = nudft.nudft(images[p][0, 0].numpy().astype(np.complex128), ktrajs[p][:, sel].double().numpy(),
                            (s[0] // 2, s[1] // 2))
            fp64.append(float(np.linalg.norm(y - r) / np.linalg.norm(r)))
    print(f"[C3] FP64 NUDFT rel err (coil 0, every 37th sample): {fp64}")
    assert max(fp64) < 2e-3


# --------------------------------------------------------------------------- C4
def test_c4_ktraj_gradient_vs_exact_kb_finite_differences():
    """Loss ||A(x0, ktraj) - y||^2 at the C4 size; the GPU's trajectory gradient along 32
    random unit directions (each supported on 1024 random samples, both coordinates)
    against FP64 central differences (step 1e-6) of the exact Kaiser-Bessel operator,
    and the image gradient along 4 random directions (the loss is quadratic in x, so the
    central difference is exact) against the same operator."""
    print("[case] C4 gradient FD protocol", flush=True)
    ktraj, smaps, x_true, noise, x0 = W.c4()
    n, coils = 256, smaps.shape[1]
    fw = tkbn.KbNufft(im_size=(n, n), dtype=torch.complex64).to(DEV)
    kt, sm = ktraj.to(DEV), smaps.to(DEV)
    with torch.no_grad():
        y = fw(x_true.to(DEV), kt, smaps=sm) + 0.01 * noise.to(DEV)
    x = x0.to(DEV).requires_grad_(True)
    k = kt.clone().requires_grad_(True)
    torch.sum(torch.abs(fw(x, k, smaps=sm) - y) ** 2).backward()
    gk = k.grad.double().cpu().numpy()
    gx = x.grad.cpu().to(torch.complex128).numpy()
    # FP64 exact-KB operator: grid = FFT(pad(x * smaps * scaling)), then closed-form taps
    plan = tp.Plan((n, n), dtype=torch.complex128)
    alpha = [2.34 * 6] * 2
    scale = plan.scaling_coef.numpy()
    yy = y.cpu().to(torch.complex128).numpy()[0]
    sm64 = smaps.to(torch.complex128).numpy()[0]

    def grid_of(img):
        pad = np.zeros((coils, 2 * n, 2 * n), dtype=np.complex128)
        pad[:, :n, :n] = img[None] * sm64 * scale[None]
        return np.fft.fft2(pad)

    G0 = grid_of(x0.to(torch.complex128).numpy()[0, 0])
    om0 = ktraj.double().numpy()
    rng = np.random.default_rng(7)

    def loss_on(om_sub, sub):
        a = nudft.exact_kb_interp(G0, om_sub, (n, n), (6, 6), alpha, (n // 2, n // 2))
        return float(np.sum(np.abs(a - yy[:, sub]) ** 2))

    # The exact-KB operator jumps (by KB(J/2) = 1/I0(alpha) of a tap) where a footprint
    # edge crosses a grid line, i.e. where omega * K / (2 pi) - J / 2 is an integer; every
    # spiral interleave starts exactly there (omega = 0), and a central difference across
    # such a jump measures the jump, not the derivative.  Directions are supported on
    # samples at least 1e-3 grid units away from it in both coordinates.
    tmk = om0 * (2 * n) / (2 * np.pi) - 3.0
    smooth = np.all(np.abs(tmk - np.rint(tmk)) > 1e-3, axis=0)
    pool = np.nonzero(smooth)[0]
    assert pool.size > 0.95 * om0.shape[1]
    h = 1e-6
    errs = []
    for _ in range(32):
        sub = rng.choice(pool, 1024, replace=False)
        dirv = rng.standard_normal((2, sub.size))
        dirv /= np.linalg.norm(dirv)
        fd = (loss_on(om0[:, sub] + h * dirv, sub) - loss_on(om0[:, sub] - h * dirv, sub)) / (2 * h)
        an = float(np.sum(gk[:, sub] * dirv))
        # |an - fd| against the RMS size of a directional derivative on this support,
        # |grad on support| / sqrt(coordinates): a ratio to fd itself is unbounded when
        # a random direction is nearly orthogonal to the gradient
        errs.append(abs(an - fd) * np.sqrt(dirv.size) / np.linalg.norm(gk[:, sub]))
    print(f"[C4 ktraj] directional err (RMS-normalised): median {np.median(errs):.2e} max {max(errs):.2e}")
    # the table path's gradient differs from the exact-KB one by the table quantisation
    # (a few 1e-3, a little more than its values' 7e-4)
    assert np.median(errs) < 6e-3 and max(errs) < 2e-2

    r0 = nudft.exact_kb_interp(G0, om0, (n, n), (6, 6), alpha, (n // 2, n // 2)) - yy
    # the whole trajectory gradient against the analytic exact-KB one (off the edges)
    dA = nudft.exact_kb_interp_domega(G0, om0, (n, n), (6, 6), alpha, (n // 2, n // 2))
    g_an = 2 * np.sum(np.real(np.conj(r0)[None] * dA), axis=1)
    e_all = float(np.linalg.norm(gk[:, pool] - g_an[:, pool]) / np.linalg.norm(g_an[:, pool]))
    print(f"[C4 ktraj] rel L2 vs analytic exact-KB gradient over {pool.size} samples: {e_all:.2e}")
    assert e_all < 6e-3
    del dA
    xerrs = []
    for _ in range(4):
        dx = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        dx /= np.linalg.norm(dx)
        adx = nudft.exact_kb_interp(grid_of(dx), om0, (n, n), (6, 6), alpha, (n // 2, n // 2))
        fd = 2 * np.real(np.vdot(adx, r0))  # (L(x0 + h dx) - L(x0 - h dx)) / 2h, exact for a quadratic
        an = float(np.real(np.vdot(gx[0, 0], dx)))  # torch's grad is 2 dL/d conj(x): Re<g, dx>
        xerrs.append(abs(an - fd) / abs(fd))
    print(f"[C4 image] directional rel err: {max(xerrs):.2e}")
    assert max(xerrs) < 5e-3


# --------------------------------------------------------------------------- C5
@pytest.fixture(scope="module")
def c5_small():
    ktraj, x, smaps = W.c5(seed=55, n=48, coils=8, nspokes=300, nread=96)
    return ktraj, x, smaps


def test_c5_reduced_kooshball_matches_oracle(c5_small):
    ktraj, x, smaps = c5_small
    print("[case] C5 reduced kooshball 48^3, 8 coils, 300 x 96, deterministic", flush=True)
    im = (48, 48, 48)
    plan = tp.Plan(im, dtype=torch.complex64)
    fw = tkbn.KbNufft(im_size=im, dtype=torch.complex64).to(DEV)
    ad = tkbn.KbNufftAdjoint(im_size=im, dtype=torch.complex64).to(DEV)
    kt, sm = ktraj.to(DEV), smaps.to(DEV)
    tkbn.set_deterministic(True)
    try:
        kd = fw(x.to(DEV), kt, smaps=sm)
        ref = tp.kb_table_nufft(plan, x, ktraj, smaps)
        assert rel_l2(kd, ref) < C64_TOL
        img = ad(ref.to(DEV), kt, smaps=sm)
        assert _last_kernel() == "spread_reg3_kernel<deterministic>"  # the 3-D register-tile spread
        img2 = ad(ref.to(DEV), kt, smaps=sm)
        assert torch.equal(img, img2)
        err = rel_l2(img, tp.kb_table_nufft_adjoint(plan, ref, ktraj, smaps))
        print(f"[C5 small] adjoint rel L2 {err:.2e}")
        assert err < C64_TOL
        # image gradient of ||A x - y||^2 = 2 A^H (A x - y), against the oracle
        y = 0.5 * ref
        xx = x.to(DEV).requires_grad_(True)
        torch.sum(torch.abs(fw(xx, kt, smaps=sm) - y.to(DEV)) ** 2).backward()
        gref = 2 * tp.kb_table_nufft_adjoint(plan, ref - y, ktraj, smaps)
        assert rel_l2(xx.grad, gref) < C64_TOL
        # trajectory gradient against the FP64 NUDFT derivative on a sample subset
        kk = kt.clone().requires_grad_(True)
        torch.sum(torch.abs(fw(x.to(DEV), kk, smaps=sm) - y.to(DEV)) ** 2).backward()
        sel = torch.arange(0, ktraj.shape[1], 97)
        xs = (x * smaps).numpy().astype(np.complex128)
        om = ktraj[:, sel].double().numpy()
        res = nudft.nudft(xs, om, (24, 24, 24)) - y[..., sel].numpy()
        dy = nudft.nudft_domega(xs, om, (24, 24, 24))
        gref = 2 * np.sum(np.real(np.conj(res)[None] * dy), axis=(1, 2))
        g = kk.grad[:, sel].double().cpu().numpy()
        e = np.linalg.norm(g - gref) / np.linalg.norm(gref)
        print(f"[C5 small] ktraj grad vs FP64 NUDFT derivative: {e:.2e}")
        assert e < 1e-2
    finally:
        tkbn.set_deterministic(False)
    sel = torch.arange(0, ktraj.shape[1], 53)
    r = nudft.nudft((x * smaps[:, :1]).numpy().astype(np.complex128), ktraj[:, sel].double().numpy(), (24, 24, 24))
    e64 = float(np.linalg.norm(kd[:, :1, sel].cpu().numpy() - r) / np.linalg.norm(r))
    print(f"[C5 small] forward vs FP64 NUDFT {e64:.2e}")
    assert e64 < 2e-3


def test_c5_full_size_properties():
    """160^3 -> 320^3, 8 coils, 3200 x 512 kooshball, deterministic: linearity of the
    forward, adjointness <A x, y> = <x, A^H y>, bitwise repeatable adjoint."""
    print("[case] C5 full 160^3 properties", flush=True)
    ktraj, x, smaps = W.c5()
    im = (160, 160, 160)
    fw = tkbn.KbNufft(im_size=im, dtype=torch.complex64).to(DEV)
    ad = tkbn.KbNufftAdjoint(im_size=im, dtype=torch.complex64).to(DEV)
    kt, sm, xd = ktraj.to(DEV), smaps.to(DEV), x.to(DEV)
    gen = torch.Generator().manual_seed(56)
    x2 = W.cplx(x.shape, gen).to(DEV)
    y = W.cplx((1, 8, kt.shape[1]), gen).to(DEV)
    tkbn.set_deterministic(True)
    try:
        a, b = 0.7 - 0.2j, -1.3 + 0.5j
        lin = fw(a * xd + b * x2, kt, smaps=sm)
        assert rel_l2(lin, a * fw(xd, kt, smaps=sm) + b * fw(x2, kt, smaps=sm)) < 1e-5
        del lin
        u = ad(y, kt, smaps=sm)
        v = ad(y, kt, smaps=sm)
        assert torch.equal(u, v)
        lhs = tkbn.inner_product(fw(xd, kt, smaps=sm).to(torch.complex128), y.to(torch.complex128))
        rhs = tkbn.inner_product(xd.to(torch.complex128), u.to(torch.complex128))
        e = float(abs(lhs - rhs) / abs(lhs))
        print(f"[C5 full] adjointness {e:.2e}")
        assert e < 1e-4
    finally:
        tkbn.set_deterministic(False)
