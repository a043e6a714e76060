"""GPU: many uneven problems through BatchedKbNufft (stacked FFTs, owned plans, one
launch per stage over all problems, autotuned tiles, HIP graph).

Each problem's result must equal the plain per-problem module call (same weights;
float summation order of the spread may differ, so 1e-6 relative), and every graph
replay must equal the eager result.  Oracle parity of the batched form at the C3
workload's scale is in test_gpu_configs.py.
"""
import numpy as np
import pytest
import torch

import torchkbnufft_amd as tkbn
from conftest import rel_l2

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")


def _last_kernel() -> str:
    from torchkbnufft_amd import _lib

    return _lib.load().tkbn_last_interp_kernel().decode()


def spiral(n_inter: int, n_read: int, turns: float = 8.0) -> torch.Tensor:
    t = np.arange(n_read) / n_read
    r = np.pi * t**1.5 * 0.999
    out = []
    for l in range(n_inter):
        th = 2 * np.pi * turns * t + 2 * np.pi * l / n_inter
        out.append(np.stack([r * np.sin(th), r * np.cos(th)]))
    return torch.tensor(np.concatenate(out, axis=1), dtype=torch.float32)


def _problems(seed=5, n=7, coils=8):
    gen = torch.Generator().manual_seed(seed)
    sizes = [(32, 32), (48, 48), (32, 32), (40, 40), (48, 48), (32, 32), (40, 40)][:n]
    inter = torch.randint(3, 9, (n,), generator=gen).tolist()
    ktrajs = [spiral(i, 512) for i in inter]
    images = [torch.complex(torch.randn((1, coils) + s, generator=gen), torch.randn((1, coils) + s, generator=gen))
              for s in sizes]
    return sizes, ktrajs, images


@pytest.mark.parametrize("coils", [8, 40])
def test_batched_matches_per_problem_modules(coils):
    print(f"[case] batched vs loop coils={coils}", flush=True)
    sizes, ktrajs, images = _problems(coils=coils)
    bt = tkbn.BatchedKbNufft(sizes, ktrajs, coils=coils)
    for p, x in enumerate(images):
        bt.images[p].copy_(x.to(DEV))
    kd = [k.clone() for k in bt.forward()]
    assert bt.fused == (coils < 32)
    if bt.fused:
        assert _last_kernel() == "batch_fwd_kernel"
    adj = [a.clone() for a in bt.adjoint()]
    if bt.fused:
        assert _last_kernel().startswith("batch_spread_kernel")
    for p, (s, k, x) in enumerate(zip(sizes, ktrajs, images)):
        fw = tkbn.KbNufft(im_size=s, dtype=torch.complex64).to(DEV)
        ad = tkbn.KbNufftAdjoint(im_size=s, dtype=torch.complex64).to(DEV)
        ref_k = fw(x.to(DEV), k.to(DEV))
        assert rel_l2(kd[p], ref_k) < 1e-6, p
        ref_a = ad(kd[p], k.to(DEV))
        assert rel_l2(adj[p], ref_a) < 1e-6, p
    # second call reuses the plans without any check: same results
    kd2 = bt.forward()
    assert all(torch.equal(a, b) for a, b in zip(kd, kd2))


def test_batched_graph_replays_equal_eager():
    print("[case] batched graph replay", flush=True)
    sizes, ktrajs, images = _problems(seed=9)
    bt = tkbn.BatchedKbNufft(sizes, ktrajs, coils=8)
    for p, x in enumerate(images):
        bt.images[p].copy_(x.to(DEV))
    eager = [a.clone() for a in bt.pair()]
    bt.capture()
    gen = torch.Generator().manual_seed(11)
    for rep in range(3):
        out = [a.clone() for a in bt.replay()]
        torch.cuda.synchronize()
        assert all(rel_l2(a, b) < 1e-6 for a, b in zip(out, eager)), rep
        # eager work between replays
        _ = torch.randn(1000, device=DEV).sum().item()
    # new inputs written in place are picked up by the replay
    for p, s in enumerate(sizes):
        bt.images[p].copy_(torch.complex(torch.randn((1, 8) + s, generator=gen), torch.randn((1, 8) + s, generator=gen)).to(DEV))
    out = [a.clone() for a in bt.replay()]
    torch.cuda.synchronize()
    again = [a.clone() for a in bt.pair()]
    assert all(rel_l2(a, b) < 1e-6 for a, b in zip(out, again))
    assert not all(rel_l2(a, b) < 1e-3 for a, b in zip(out, eager))


@pytest.mark.parametrize("deterministic", [False, True])
def test_fused_equals_separate_launches(deterministic):
    """One launch per stage (fused) against one launch per problem on side streams:
    same results, and the deterministic fused adjoint is bitwise repeatable."""
    print(f"[case] fused vs separate det={deterministic}", flush=True)
    sizes, ktrajs, images = _problems(seed=13)
    outs = {}
    for fused in (True, False):
        bt = tkbn.BatchedKbNufft(sizes, ktrajs, coils=8, fused=fused, deterministic=deterministic)
        for p, x in enumerate(images):
            bt.images[p].copy_(x.to(DEV))
        a1 = [a.clone() for a in bt.pair()]
        a2 = [a.clone() for a in bt.pair()]
        assert bt.fused == fused
        if deterministic:
            assert all(torch.equal(u, v) for u, v in zip(a1, a2))
        outs[fused] = ([k.clone() for k in bt.kdata], a1)
    for u, v in zip(outs[True][0], outs[False][0]):
        assert rel_l2(u, v) < 1e-6
    for u, v in zip(outs[True][1], outs[False][1]):
        assert rel_l2(u, v) < 1e-6


def test_fused_3d_and_untuned_match_modules():
    """3-D problems of three sizes in one fused batch (tuning off, library defaults)
    against the per-problem modules (3 coil rows: the sample-lane kernels; from 4 rows
    3-D problems take the register-tile spread, one launch each)."""
    print("[case] fused 3-D batch", flush=True)
    gen = torch.Generator().manual_seed(17)
    sizes = [(16, 16, 16), (20, 20, 20), (16, 16, 16), (24, 20, 16)]
    ktrajs = [(torch.rand(3, m, generator=gen) * 2 - 1) * np.pi for m in (3000, 5000, 800, 4000)]
    images = [torch.complex(torch.randn((1, 3) + s, generator=gen), torch.randn((1, 3) + s, generator=gen))
              for s in sizes]
    bt = tkbn.BatchedKbNufft(sizes, ktrajs, coils=3, autotune=False)
    for p, x in enumerate(images):
        bt.images[p].copy_(x.to(DEV))
    kd = [k.clone() for k in bt.forward()]
    adj = [a.clone() for a in bt.adjoint()]
    assert bt.fused
    for p, (s, k, x) in enumerate(zip(sizes, ktrajs, images)):
        fw = tkbn.KbNufft(im_size=s, dtype=torch.complex64).to(DEV)
        ad = tkbn.KbNufftAdjoint(im_size=s, dtype=torch.complex64).to(DEV)
        assert rel_l2(kd[p], fw(x.to(DEV), k.to(DEV))) < 1e-6, p
        assert rel_l2(adj[p], ad(kd[p], k.to(DEV))) < 1e-6, p
