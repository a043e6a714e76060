"""GPU parity of the fused oversampled FFT (tkbn_fft_grid_fwd / tkbn_fft_grid_adj)."""
import pytest
import torch

import inputs as gi
import torchkbnufft_amd as tkbn
from conftest import rel_l2
from oracle import table_path as tp
from torchkbnufft_amd._nufft import fft as F

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")
TOL = 1e-5

GEOMS = [
    ((320, 320), (640, 640), 2, 3),
    ((256, 256), (512, 512), 1, 1),
    ((37,), (80,), 2, 3),
    ((48, 40), (96, 80), 2, 2),
    ((20, 18, 10), (40, 36, 20), 1, 3),
    ((64, 96), (128, 192), 1, 2),
    ((7, 5), (10, 6), 3, 2),
    # lengths with compiled passes (256, 320, 384, 512, 640, 768) in every pass kind
    ((192, 128), (384, 256), 1, 2),
    ((160, 300), (320, 768), 1, 2),
    ((200,), (512,), 2, 2),
    ((60, 10, 100), (320, 20, 640), 1, 2),
]


@pytest.mark.parametrize("im,gs,B,C", GEOMS)
@pytest.mark.parametrize("smaps", [False, True])
@pytest.mark.parametrize("norm", [None, "ortho"])
def test_fused_fft_matches_oracle(im, gs, B, C, smaps, norm):
    print(f"[case] fused fft {im}->{gs} B={B} C={C} smaps={smaps} norm={norm}", flush=True)
    assert F.fused_supported(torch.complex64, im, gs)
    gen = torch.Generator().manual_seed(hash((im, gs, B, C)) % 1000)
    x = gi.cplx((B, 1 if smaps else C) + im, gen)
    s = gi.cplx((B, C) + im, gen) if smaps else None
    g = gi.cplx((B, C) + gs, gen)
    plan = tp.Plan(im, gs, dtype=torch.complex64)
    scale = F._norm_scale(norm, gs)
    got = F.fft_grid_fwd(x.to(DEV), s.to(DEV) if smaps else None, plan.scaling_coef.to(DEV), im, gs, scale).cpu()
    ref = tp.fft_and_scale(x * s if smaps else x, plan.scaling_coef, im, gs, norm)
    assert rel_l2(got, ref) < TOL
    got_a = F.fft_grid_adj(g.to(DEV), s.to(DEV) if smaps else None, plan.scaling_coef.to(DEV), im, scale).cpu()
    ref_a = tp.ifft_and_scale(g, plan.scaling_coef, im, gs, norm)
    if smaps:
        ref_a = torch.sum(ref_a * s.conj(), dim=1, keepdim=True)
    assert rel_l2(got_a, ref_a) < TOL


def test_unsupported_sizes_fall_back():
    assert not F.fused_supported(torch.complex64, (33, 24), (66, 48))  # 11 is not a radix
    assert not F.fused_supported(torch.complex128, (32, 32), (64, 64))
    assert F.fused_supported(torch.complex64, (32, 32), (64, 64))


def test_fused_and_rocfft_paths_agree_through_modules():
    gen = torch.Generator().manual_seed(3)
    im = (96, 80)
    k = gi.radial(60, 192, golden=True).to(DEV)
    x = gi.cplx((2, 1) + im, gen).to(DEV)
    s = gi.cplx((2, 4) + im, gen).to(DEV)
    fw = tkbn.KbNufft(im_size=im, dtype=torch.complex64).to(DEV)
    ad = tkbn.KbNufftAdjoint(im_size=im, dtype=torch.complex64).to(DEV)
    y1 = fw(x, k, smaps=s)
    a1 = ad(y1, k, smaps=s)
    F.set_fused_fft(False)
    try:
        y2 = fw(x, k, smaps=s)
        a2 = ad(y1, k, smaps=s)
    finally:
        F.set_fused_fft(True)
    assert rel_l2(y1, y2) < TOL
    assert rel_l2(a1, a2) < TOL
