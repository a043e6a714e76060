"""GPU parity of the Toeplitz path (``torchkbnufft/_nufft/toep.py``, ``fft.py:121-173``,
``modules/kbnufft.py:413-547``).

* the kernel against the oracle's restatement of the reference's assembly;
* ``ToepNufft`` against the oracle on the same kernel (fused HIP FFT path for
  complex64 and 2/3/5-smooth sizes, rocFFT otherwise);
* the reference's own accuracy tests (``tests/test_toep.py``): Toeplitz vs the
  NUFFT normal operator, unbatched, batched and with batched weights.
"""
import numpy as np
import pytest
import torch

import torchkbnufft_amd as tkbn
from conftest import create_input_plus_noise, create_ktraj, rel_l2
from oracle import table_path as tp

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")


@pytest.mark.parametrize("im_size,grid_size", [((19,), (57,)), ((32, 16), (64, 24)), ((20, 24), None), ((7, 6, 5), (14, 12, 10))])
@pytest.mark.parametrize("norm", [None, "ortho"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_kernel_matches_oracle(im_size, grid_size, norm, dtype):
    print(f"[case] toeplitz kernel im={im_size} grid={grid_size} norm={norm} {dtype}", flush=True)
    gen = torch.Generator().manual_seed(11)
    om = create_ktraj(len(im_size), 90, gen, dtype=dtype)
    w = torch.rand((1, 90), generator=gen, dtype=dtype)
    cdtype = torch.complex128 if dtype == torch.float64 else torch.complex64
    k = tkbn.calc_toeplitz_kernel(om.to(DEV), im_size, weights=w.to(DEV), norm=norm, grid_size=grid_size)
    ref = tp.toeplitz_kernel(om, im_size, weights=w, norm=norm, grid_size=grid_size, dtype=cdtype)
    assert k.shape == ref.shape and k.dtype == cdtype
    assert rel_l2(k.cpu(), ref) < (1e-12 if dtype == torch.float64 else 1e-5)


@pytest.mark.parametrize(
    "B,C,im_size,batched_kernel",
    [(1, 4, (32, 16), False), (3, 5, (20, 24), True), (2, 3, (40,), False), (2, 2, (10, 12, 8), True)],
)
@pytest.mark.parametrize("norm", [None, "ortho"])
def test_toep_nufft_matches_oracle(B, C, im_size, batched_kernel, norm):
    """complex64: the fused FFT pipeline with the kernel product in between."""
    print(f"[case] toep nufft B={B} C={C} im={im_size} batched={batched_kernel} norm={norm}", flush=True)
    gen = torch.Generator().manual_seed(12)
    d = len(im_size)
    om = create_ktraj(d, 120, gen, dtype=torch.float32)
    if batched_kernel:
        om = torch.stack([create_ktraj(d, 120, gen, dtype=torch.float32) for _ in range(B)])
    kernel = tp.toeplitz_kernel(om, im_size, norm=norm, dtype=torch.complex64)
    x = create_input_plus_noise([B, 1] + list(im_size), True, gen, dtype=torch.float32)
    smaps = create_input_plus_noise([B, C] + list(im_size), True, gen, dtype=torch.float32)
    ref = tp.toep_nufft(x, kernel, smaps, norm)
    out = tkbn.ToepNufft()(x.to(DEV), kernel.to(DEV), smaps=smaps.to(DEV), norm=norm)
    assert out.shape == ref.shape
    assert rel_l2(out.cpu(), ref) < 1e-5
    # without smaps: the per-coil filter (fft_filter)
    xc = create_input_plus_noise([B, C] + list(im_size), True, gen, dtype=torch.float32)
    k1 = kernel[0] if batched_kernel else kernel
    out = tkbn.functional.fft_filter(xc.to(DEV), k1.to(DEV), norm=norm)
    assert rel_l2(out.cpu(), tp.fft_filter(xc, k1, norm)) < 1e-5


TOEP_CASES = [
    ([1, 3, 19], [57], [1, 3, 25], True, "ortho"),
    ([3, 5, 13, 2], [19], [3, 5, 18, 2], False, None),
    ([1, 4, 32, 16], [64, 24], [1, 4, 83], True, None),
    ([5, 8, 15, 12, 2], [30, 24], [5, 8, 83, 2], False, "ortho"),
    ([3, 10, 13, 18, 12], [20, 26, 37], [3, 10, 112], True, None),
    ([1, 12, 17, 19, 12, 2], [25, 28, 24], [1, 12, 112, 2], False, "ortho"),
]


@pytest.mark.parametrize("shape, grid_size, kdata_shape, is_complex, norm", TOEP_CASES)
def test_toeplitz_nufft_accuracy(shape, grid_size, kdata_shape, is_complex, norm):
    """Reference tests/test_toep.py:9-55 (complex128, general grid sizes, 1e-2)."""
    print(f"[case] toeplitz accuracy {shape} grid={grid_size} norm={norm}", flush=True)
    gen = torch.Generator().manual_seed(123)
    im_size = shape[2:] if is_complex else shape[2:-1]
    im_shape = list(shape)
    im_shape[1] = 1
    image = create_input_plus_noise(im_shape, is_complex, gen).to(DEV)
    smaps = create_input_plus_noise(shape, is_complex, gen).to(DEV)
    ktraj = create_ktraj(len(im_size), kdata_shape[2], gen).to(DEV)
    forw = tkbn.KbNufft(im_size=im_size, grid_size=grid_size, dtype=torch.complex128).to(DEV)
    adj = tkbn.KbNufftAdjoint(im_size=im_size, grid_size=grid_size, dtype=torch.complex128).to(DEV)
    kernel = tkbn.calc_toeplitz_kernel(ktraj, im_size, grid_size=grid_size, norm=norm)
    if not is_complex:
        kernel = torch.view_as_real(kernel)
    fbn = adj(forw(image, ktraj, smaps=smaps, norm=norm), ktraj, smaps=smaps, norm=norm)
    fbt = tkbn.ToepNufft()(image, kernel, smaps=smaps, norm=norm)
    if is_complex:
        fbn, fbt = torch.view_as_real(fbn), torch.view_as_real(fbt)
    assert float(torch.norm(fbn - fbt) / torch.norm(fbn)) < 1e-2


BATCH_CASES = [
    ([4, 3, 19], [4, 3, 25], True),
    ([3, 5, 13, 2], [3, 5, 18, 2], False),
    ([2, 4, 32, 16], [2, 4, 83], True),
    ([5, 8, 15, 12, 2], [5, 8, 83, 2], False),
    ([3, 10, 13, 18, 12], [3, 10, 112], True),
    ([2, 12, 17, 19, 12, 2], [2, 12, 112, 2], False),
]


@pytest.mark.parametrize("shape, kdata_shape, is_complex", BATCH_CASES)
@pytest.mark.parametrize("with_weights", [False, True])
def test_batched_toeplitz_nufft_accuracy(shape, kdata_shape, is_complex, with_weights):
    """Reference tests/test_toep.py:58-175 (batched trajectories, optional weights, 1e-4)."""
    print(f"[case] batched toeplitz {shape} weights={with_weights}", flush=True)
    gen = torch.Generator().manual_seed(123)
    im_size = shape[2:] if is_complex else shape[2:-1]
    im_shape = list(shape)
    im_shape[1] = 1
    image = create_input_plus_noise(im_shape, is_complex, gen).to(DEV)
    smaps = create_input_plus_noise(shape, is_complex, gen).to(DEV)
    ktraj = (torch.rand((shape[0], len(im_size), kdata_shape[2]), generator=gen, dtype=torch.float64) * 2 - 1) * np.pi
    ktraj = ktraj.to(DEV)
    weights = torch.rand((kdata_shape[0], 1, kdata_shape[2]), generator=gen, dtype=torch.float64).to(DEV)
    forw = tkbn.KbNufft(im_size=im_size, dtype=torch.complex128).to(DEV)
    adj = tkbn.KbNufftAdjoint(im_size=im_size, dtype=torch.complex128).to(DEV)
    kernel = tkbn.calc_toeplitz_kernel(ktraj, im_size, weights=weights if with_weights else None, norm="ortho")
    if not is_complex:
        kernel = torch.view_as_real(kernel)
    w = weights if with_weights else torch.ones_like(weights)
    if len(kdata_shape) == 4:
        w = w.unsqueeze(-1)
    fbn = adj(w * forw(image, ktraj, smaps=smaps, norm="ortho"), ktraj, smaps=smaps, norm="ortho")
    fbt = tkbn.ToepNufft()(image, kernel, smaps=smaps, norm="ortho")
    if is_complex:
        fbn, fbt = torch.view_as_real(fbn), torch.view_as_real(fbt)
    assert float(torch.norm(fbn - fbt) / torch.norm(fbn)) < 1e-4


def test_toeplitz_sense_radial_full_size():
    """16-coil 320^2 SENSE with the 256 000-sample radial trajectory (complex64): the
    Toeplitz operator against the NUFFT normal operator and against the oracle."""
    print("[case] toeplitz SENSE 320^2 radial, B=4 C=16", flush=True)
    gen = torch.Generator().manual_seed(21)
    spokes, n = 400, 640
    ga = np.pi / ((1 + np.sqrt(5)) / 2)
    th = torch.arange(spokes, dtype=torch.float64) * ga
    r = -np.pi + 2 * np.pi * torch.arange(n, dtype=torch.float64) / n
    om = torch.stack([(r[None, :] * torch.sin(th[:, None])).reshape(-1), (r[None, :] * torch.cos(th[:, None])).reshape(-1)])
    om = om.float()
    x = torch.complex(torch.randn((4, 1, 320, 320), generator=gen), torch.randn((4, 1, 320, 320), generator=gen))
    smaps = torch.complex(torch.randn((4, 16, 320, 320), generator=gen), torch.randn((4, 16, 320, 320), generator=gen))
    smaps = smaps / smaps.abs().pow(2).sum(1, keepdim=True).sqrt()
    xd, sd, od = x.to(DEV), smaps.to(DEV), om.to(DEV)
    kernel = tkbn.calc_toeplitz_kernel(od, (320, 320))
    assert kernel.shape == (640, 640)
    fbt = tkbn.ToepNufft()(xd, kernel, smaps=sd)
    fw = tkbn.KbNufft(im_size=(320, 320)).to(DEV)
    aj = tkbn.KbNufftAdjoint(im_size=(320, 320)).to(DEV)
    fbn = aj(fw(xd, od, smaps=sd), od, smaps=sd)
    assert rel_l2(fbt.cpu(), fbn.cpu()) < 1e-3
    ref = tp.toep_nufft(x, kernel.cpu(), smaps)
    assert rel_l2(fbt.cpu(), ref) < 1e-5


@pytest.mark.parametrize("im_size,batched", [((32, 16), False), ((20, 24), True), ((10, 12, 8), False)])
def test_toep_one_pass_gradient_matches_staged_path(im_size, batched):
    """The one-pass filter's image gradient (the same filter with conj(kernel)) against
    autograd through the staged pad/FFT, product, iFFT/crop path."""
    print(f"[case] toep gradient im={im_size} batched={batched}", flush=True)
    gen = torch.Generator().manual_seed(14)
    d = len(im_size)
    B, C = 2, 3
    om = torch.stack([create_ktraj(d, 100, gen, dtype=torch.float32) for _ in range(B)]) if batched else create_ktraj(d, 100, gen, dtype=torch.float32)
    kernel = tp.toeplitz_kernel(om, im_size, norm="ortho", dtype=torch.complex64).to(DEV)
    x = create_input_plus_noise([B, 1] + list(im_size), True, gen, dtype=torch.float32).to(DEV)
    smaps = create_input_plus_noise([B, C] + list(im_size), True, gen, dtype=torch.float32).to(DEV)
    w = create_input_plus_noise([B, 1] + list(im_size), True, gen, dtype=torch.float32).to(DEV)
    toep = tkbn.ToepNufft()
    x1 = x.clone().requires_grad_(True)
    y1 = toep(x1, kernel, smaps=smaps, norm="ortho")
    (y1 * w.conj()).real.sum().backward()
    x2 = x.clone().requires_grad_(True)
    k2 = kernel.clone().requires_grad_(True)  # forces the staged path
    y2 = toep(x2, k2, smaps=smaps, norm="ortho")
    (y2 * w.conj()).real.sum().backward()
    assert rel_l2(y1.detach().cpu(), y2.detach().cpu()) < 1e-5
    assert rel_l2(x1.grad.cpu(), x2.grad.cpu()) < 1e-5
