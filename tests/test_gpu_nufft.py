"""GPU parity of the full table NUFFT (fused scale/pad + rocFFT + interpolation) and SENSE."""
import numpy as np
import pytest
import torch

import inputs as gi
import torchkbnufft_amd as tkbn
from conftest import create_input_plus_noise, create_ktraj, golden, rel_l2
from oracle import nudft
from oracle import table_path as tp

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
SENSE_SHAPES = [
    ([1, 3, 19], [1, 3, 25], True),
    ([3, 5, 13, 2], [3, 5, 18, 2], False),
    ([1, 4, 32, 16], [1, 4, 83], True),
    ([5, 8, 15, 12, 2], [5, 8, 83, 2], False),
    ([3, 10, 13, 18, 12], [3, 10, 112], True),
    ([1, 12, 17, 19, 12, 2], [1, 12, 112, 2], False),
]


def _im_size(shape, is_complex):
    return shape[2:] if is_complex else shape[2:-1]


@pytest.mark.parametrize("case", gi.C64_CASES, ids=[c[0] for c in gi.C64_CASES])
def test_nufft_c64_matches_reference(case):
    name, im, grid, J, L, B, C, mk, use_smaps, norm, batched = case
    print(f"[case] nufft {name}", flush=True)
    g = golden("golden_c64.npz")
    inp = gi.c64_inputs(case)
    kw = dict(im_size=im, grid_size=grid, numpoints=J, table_oversamp=L, dtype=torch.complex64)
    fw = tkbn.KbNufft(**kw).to(DEV)
    ad = tkbn.KbNufftAdjoint(**kw).to(DEV)
    k = inp["ktraj"].to(DEV)
    sm = inp["smaps"].to(DEV) if inp["smaps"] is not None else None
    assert rel_l2(fw(inp["x"].to(DEV), k, smaps=sm, norm=norm).cpu(), g[name + "_nufft"]) < C64_TOL
    assert rel_l2(ad(inp["y_in"].to(DEV), k, smaps=sm, norm=norm).cpu(), g[name + "_nufft_adj"]) < C64_TOL


def test_nufft_f64_fixtures():
    g = golden("fixtures_f64.npz")
    for i in range(len(gi.NUFFT_FIXTURES)):
        im_size, image, ktraj = gi.nufft_fixture_inputs(i)
        fw = tkbn.KbNufft(im_size=im_size, dtype=torch.complex128).to(DEV)
        kd = fw(image.to(DEV), ktraj.to(DEV))
        assert torch.allclose(kd.cpu(), torch.from_numpy(g[f"nufft{i}_kdata"]))
        if f"nufft{i}_adj" in g.files:
            ad = tkbn.KbNufftAdjoint(im_size=im_size, dtype=torch.complex128).to(DEV)
            assert torch.allclose(ad(kd, ktraj.to(DEV)).cpu(), torch.from_numpy(g[f"nufft{i}_adj"]))


@pytest.mark.parametrize("shape, kdata_shape, is_complex", SHAPES)
def test_nufft_adjoint_dot_and_autograd(shape, kdata_shape, is_complex):
    gen = torch.Generator().manual_seed(123)
    im_size = _im_size(shape, is_complex)
    image = create_input_plus_noise(shape, is_complex, gen).to(DEV)
    kdata = create_input_plus_noise(kdata_shape, is_complex, gen).to(DEV)
    ktraj = create_ktraj(len(im_size), kdata_shape[2], gen).to(DEV)
    fw = tkbn.KbNufft(im_size=im_size, dtype=torch.complex128).to(DEV)
    ad = tkbn.KbNufftAdjoint(im_size=im_size, dtype=torch.complex128).to(DEV)
    assert torch.allclose(tkbn.inner_product(fw(image, ktraj), kdata), tkbn.inner_product(image, ad(kdata, ktraj)))
    image.requires_grad_(True)
    kdata.requires_grad_(True)
    f = fw(image, ktraj)
    a = ad(kdata, ktraj)
    (torch.abs(f) ** 2 / 2).sum().backward()
    (torch.abs(a) ** 2 / 2).sum().backward()
    assert torch.allclose(image.grad, ad(f.detach(), ktraj))
    assert torch.allclose(kdata.grad, fw(a.detach(), ktraj))


@pytest.mark.parametrize("shape, kdata_shape, is_complex", SENSE_SHAPES)
def test_sense_nufft_adjoint_dot_and_autograd(shape, kdata_shape, is_complex):
    gen = torch.Generator().manual_seed(123)
    im_size = _im_size(shape, is_complex)
    im_shape = list(shape)
    im_shape[1] = 1
    image = create_input_plus_noise(im_shape, is_complex, gen).to(DEV)
    kdata = create_input_plus_noise(kdata_shape, is_complex, gen).to(DEV)
    smaps = create_input_plus_noise(shape, is_complex, gen).to(DEV)
    ktraj = create_ktraj(len(im_size), kdata_shape[2], gen).to(DEV)
    fw = tkbn.KbNufft(im_size=im_size, dtype=torch.complex128).to(DEV)
    ad = tkbn.KbNufftAdjoint(im_size=im_size, dtype=torch.complex128).to(DEV)
    lhs = tkbn.inner_product(fw(image, ktraj, smaps=smaps), kdata)
    rhs = tkbn.inner_product(image, ad(kdata, ktraj, smaps=smaps))
    assert torch.allclose(lhs, rhs)
    image.requires_grad_(True)
    kdata.requires_grad_(True)
    f = fw(image, ktraj, smaps=smaps)
    a = ad(kdata, ktraj, smaps=smaps)
    (torch.abs(f) ** 2 / 2).sum().backward()
    (torch.abs(a) ** 2 / 2).sum().backward()
    assert torch.allclose(image.grad, ad(f.detach(), ktraj, smaps=smaps))
    assert torch.allclose(kdata.grad, fw(a.detach(), ktraj, smaps=smaps))


def test_smaps_gradient_matches_torch_composition():
    gen = torch.Generator().manual_seed(5)
    im = (12, 10)
    x = gi.cplx((2, 1) + im, gen, torch.complex128).to(DEV)
    s = gi.cplx((2, 3) + im, gen, torch.complex128).to(DEV).requires_grad_(True)
    k = create_ktraj(2, 40, gen).to(DEV)
    y = gi.cplx((2, 3, 40), gen, torch.complex128).to(DEV)
    fw = tkbn.KbNufft(im_size=im, dtype=torch.complex128).to(DEV)
    torch.sum(torch.abs(fw(x, k, smaps=s) - y) ** 2).backward()
    g1 = s.grad.clone()
    s.grad = None
    # same loss with the SENSE product done by torch (coil images as input)
    torch.sum(torch.abs(fw(x * s, k) - y) ** 2).backward()
    assert torch.allclose(g1, s.grad)


def test_image_gradient_matches_reference_autograd():
    im, ktraj, x, smaps, y = gi.grad_inputs()
    g = golden("grad_c64.npz")
    fw = tkbn.KbNufft(im_size=im, dtype=torch.complex64).to(DEV)
    xd = x.to(DEV).requires_grad_(True)
    loss = torch.sum(torch.abs(fw(xd, ktraj.to(DEV), smaps=smaps.to(DEV)) - y.to(DEV)) ** 2)
    loss.backward()
    assert abs(loss.item() - float(g["loss"])) / float(g["loss"]) < 1e-5
    assert rel_l2(xd.grad.cpu(), g["grad_x"]) < C64_TOL


def test_c1_full_size_nufft_matches_oracle_and_states_nudft_error():
    """Plain 2-D config: 256^2 image, 512^2 grid, 402 x 512 radial, complex64."""
    print("[case] C1 nufft fwd/adj", flush=True)
    gen = torch.Generator().manual_seed(0)
    ktraj = gi.radial(402, 512, golden=False)
    x = gi.cplx((1, 1, 256, 256), gen)
    y = gi.cplx((1, 1, ktraj.shape[1]), gen)
    plan = tp.Plan((256, 256), dtype=torch.complex64)
    fw = tkbn.KbNufft(im_size=(256, 256), dtype=torch.complex64).to(DEV)
    ad = tkbn.KbNufftAdjoint(im_size=(256, 256), dtype=torch.complex64).to(DEV)
    kd = fw(x.to(DEV), ktraj.to(DEV)).cpu()
    xa = ad(y.to(DEV), ktraj.to(DEV)).cpu()
    assert rel_l2(kd, tp.kb_table_nufft(plan, x, ktraj)) < C64_TOL
    assert rel_l2(xa, tp.kb_table_nufft_adjoint(plan, y, ktraj)) < C64_TOL
    # FP64 NUDFT on a sample subset (stated error of the table path, ~7e-4)
    sel = torch.arange(0, ktraj.shape[1], 97)
    ref = nudft.nudft(x.numpy().astype(np.complex128), ktraj[:, sel].double().numpy(), (128, 128))
    err = rel_l2(kd[..., sel], ref)
    print(f"[C1] forward rel err vs FP64 NUDFT: {err:.3e}")
    assert err < 2e-3


def test_c2_geometry_matches_oracle_and_properties():
    """SENSE multi-coil config geometry (320^2 -> 640^2, 400 x 640 golden-angle radial).

    Parity against the oracle on a coil subset (the oracle must finish in seconds), then
    size-independent properties at the full batch 4 x 16 coils: linearity and adjointness.
    """
    print("[case] C2 geometry", flush=True)
    gen = torch.Generator().manual_seed(1)
    ktraj = gi.radial(400, 640, golden=True)
    im = (320, 320)
    fw = tkbn.KbNufft(im_size=im, dtype=torch.complex64).to(DEV)
    ad = tkbn.KbNufftAdjoint(im_size=im, dtype=torch.complex64).to(DEV)
    plan = tp.Plan(im, dtype=torch.complex64)
    x = gi.cplx((1, 1) + im, gen)
    s = gi.cplx((1, 2) + im, gen)
    y = gi.cplx((1, 2, ktraj.shape[1]), gen)
    kd = fw(x.to(DEV), ktraj.to(DEV), smaps=s.to(DEV)).cpu()
    assert rel_l2(kd, tp.kb_table_nufft(plan, x, ktraj, s)) < C64_TOL
    xa = ad(y.to(DEV), ktraj.to(DEV), smaps=s.to(DEV)).cpu()
    assert rel_l2(xa, tp.kb_table_nufft_adjoint(plan, y, ktraj, s)) < C64_TOL
    # full size: 4 x 16 coils
    kt = ktraj.to(DEV)
    X = torch.randn(4, 1, *im, dtype=torch.complex64, device=DEV)
    X2 = torch.randn(4, 1, *im, dtype=torch.complex64, device=DEV)
    S = torch.randn(4, 16, *im, dtype=torch.complex64, device=DEV)
    Y = torch.randn(4, 16, kt.shape[1], dtype=torch.complex64, device=DEV)
    a, b = 0.7 - 0.2j, -1.3 + 0.5j
    lin = fw(a * X + b * X2, kt, smaps=S)
    ref = a * fw(X, kt, smaps=S) + b * fw(X2, kt, smaps=S)
    assert rel_l2(lin, ref) < 1e-5
    lhs = tkbn.inner_product(fw(X, kt, smaps=S).to(torch.complex128), Y.to(torch.complex128))
    rhs = tkbn.inner_product(X.to(torch.complex128), ad(Y, kt, smaps=S).to(torch.complex128))
    assert abs(lhs - rhs) / abs(lhs) < 1e-4


def test_ktraj_gradient_matches_nudft_derivative():
    """Trajectory gradient of ||A(x, omega) - y||^2 through the whole NUFFT vs the FP64
    analytic NUDFT derivative (differences are the NUFFT approximation error)."""
    gen = torch.Generator().manual_seed(11)
    im = (32, 32)
    M = 600
    x = gi.cplx((1, 1) + im, gen, torch.complex128)
    s = gi.cplx((1, 3) + im, gen, torch.complex128)
    om = create_ktraj(2, M, gen) * 0.9
    y = gi.cplx((1, 3, M), gen, torch.complex128)
    fw = tkbn.KbNufft(im_size=im, dtype=torch.complex128).to(DEV)
    omd = om.to(DEV).requires_grad_(True)
    loss = torch.sum(torch.abs(fw(x.to(DEV), omd, smaps=s.to(DEV)) - y.to(DEV)) ** 2)
    loss.backward()
    grad = omd.grad.cpu().numpy()
    xs = (x * s).numpy()
    n_shift = (16, 16)
    res = nudft.nudft(xs, om.numpy(), n_shift) - y.numpy()
    dy = nudft.nudft_domega(xs, om.numpy(), n_shift)  # (2, 1, 3, M)
    ref = 2 * np.sum(np.real(np.conj(res)[None] * dy), axis=(1, 2))
    err = np.linalg.norm(grad - ref) / np.linalg.norm(ref)
    print(f"[ktraj grad] rel err vs FP64 NUDFT derivative: {err:.2e}")
    assert err < 1e-2


@pytest.mark.parametrize("coils", [1, 40])
def test_pair_graph_replay_matches_eager(coils):
    """A forward + adjoint pair captured as a HIP graph (after warm-up calls on the
    capture stream build the plans) replays to the eager result, replay after replay,
    with eager work in between: the library's per-call work (plan-cache check, early-exit
    plan kernels, fill kernels) is capturable.  40 coil rows run the multi-coil kernels."""
    print(f"[case] graph replay, {coils} coil rows", flush=True)
    gen = torch.Generator().manual_seed(61)
    im = (48, 40)
    x = gi.cplx((1, coils) + im, gen, torch.complex64).to(DEV)
    om = create_ktraj(2, 3000, gen).to(DEV)
    fw = tkbn.KbNufft(im_size=im).to(DEV)
    ad = tkbn.KbNufftAdjoint(im_size=im).to(DEV)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(3):
            ad(fw(x, om), om)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        y = ad(fw(x, om), om)
    for _ in range(3):
        g.replay()
        torch.cuda.synchronize()
        ref = ad(fw(x, om), om)
        assert rel_l2(y, ref) < 1e-6
