"""GPU parity of density compensation (``torchkbnufft/_nufft/dcomp.py:10-119``).

The reference's tests (``tests/test_dcomp.py``) run the function on six shapes and
check that a batched trajectory gives the loop over its batch; here the GPU result
is also compared with the oracle's restatement of the iteration.
"""
import numpy as np
import pytest
import torch

import torchkbnufft_amd as tkbn
from conftest import create_input_plus_noise, create_ktraj, rel_l2
from oracle import table_path as tp

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")

SHAPES = [
    ([1, 1, 19], [1, 1, 25], True),
    ([3, 1, 13, 2], [3, 1, 18, 2], False),
    ([1, 1, 32, 16], [1, 1, 83], True),
    ([5, 1, 15, 12, 2], [5, 1, 83, 2], False),
    ([3, 2, 13, 18, 12], [3, 2, 112], True),
    ([1, 2, 17, 19, 12, 2], [1, 2, 112, 2], False),
]


@pytest.mark.parametrize("shape, kdata_shape, is_complex", SHAPES)
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_dcomp_run_matches_oracle(shape, kdata_shape, is_complex, dtype):
    print(f"[case] dcomp shape={shape} dtype={dtype}", flush=True)
    gen = torch.Generator().manual_seed(123)
    im_size = shape[2:] if is_complex else shape[2:-1]
    cdtype = torch.complex128 if dtype == torch.float64 else torch.complex64
    kdata = create_input_plus_noise(kdata_shape, is_complex, gen, dtype=dtype)
    ktraj = create_ktraj(len(im_size), kdata_shape[2], gen, dtype=dtype)
    dcomp = tkbn.calc_density_compensation_function(ktraj=ktraj.to(DEV), im_size=im_size)
    assert dcomp.shape == (1, 1, kdata_shape[2]) and dcomp.dtype == cdtype
    ref = tp.density_compensation(tp.Plan(im_size, dtype=cdtype), ktraj)
    if dtype == torch.float64:
        assert torch.allclose(dcomp.cpu(), ref)
    else:
        assert rel_l2(dcomp.cpu(), ref) < 1e-5
    # as the reference test: weight k-space and run the adjoint NUFFT
    adj_ob = tkbn.KbNufftAdjoint(im_size=im_size, dtype=cdtype).to(DEV)
    w = dcomp if is_complex else torch.view_as_real(dcomp)
    image = adj_ob(kdata.to(DEV) * w, ktraj.to(DEV))
    assert torch.isfinite(torch.view_as_real(image) if image.is_complex() else image).all()


@pytest.mark.parametrize(
    "shape, kdata_shape",
    [
        ([2, 1, 19], [2, 1, 25]),
        ([3, 1, 13], [3, 1, 18]),
        ([6, 1, 32, 16], [6, 1, 83]),
        ([5, 1, 15, 12], [5, 1, 83]),
        ([3, 2, 13, 18, 12], [3, 2, 112]),
        ([2, 2, 17, 19, 12], [2, 2, 112]),
    ],
)
def test_batched_dcomp_is_loop(shape, kdata_shape):
    print(f"[case] batched dcomp shape={shape}", flush=True)
    gen = torch.Generator().manual_seed(123)
    im_size = shape[2:]
    ktraj = (torch.rand((shape[0], len(im_size), kdata_shape[2]), generator=gen, dtype=torch.float64) * 2 - 1) * np.pi
    ktraj = ktraj.to(DEV)
    loop = torch.cat([tkbn.calc_density_compensation_function(ktraj=k, im_size=im_size) for k in ktraj])
    batched = tkbn.calc_density_compensation_function(ktraj=ktraj, im_size=im_size)
    assert torch.allclose(loop, batched)


def test_dcomp_radial_full_size_matches_oracle():
    """The 256 000-sample golden-angle radial trajectory on 320^2 -> 640^2 (complex64)."""
    print("[case] dcomp radial 320^2, 256000 samples", flush=True)
    spokes, n = 400, 640
    ga = np.pi / ((1 + np.sqrt(5)) / 2)
    th = torch.arange(spokes, dtype=torch.float64) * ga
    r = -np.pi + 2 * np.pi * torch.arange(n, dtype=torch.float64) / n
    om = torch.stack([(r[None, :] * torch.sin(th[:, None])).reshape(-1), (r[None, :] * torch.cos(th[:, None])).reshape(-1)])
    om = om.float()
    dcomp = tkbn.calc_density_compensation_function(ktraj=om.to(DEV), im_size=(320, 320))
    ref = tp.density_compensation(tp.Plan((320, 320), dtype=torch.complex64), om)
    assert rel_l2(dcomp.cpu(), ref) < 1e-5
    # radial density grows with |k|: the weights of the outer samples exceed the centre's
    mag = om.norm(dim=0)
    w = dcomp.cpu()[0, 0].real
    assert float(w[mag > 2.5].mean()) > 10 * float(w[mag < 0.2].mean())
