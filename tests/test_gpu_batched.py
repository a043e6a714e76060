"""GPU: many uneven problems through BatchedKbNufft (stacked FFTs, owned plans, HIP graph).

Each problem's result must equal the plain per-problem module call (same kernels and
plans; float summation order of the spread may differ, so 1e-6 relative), and every
graph replay must equal the eager result.
"""
import numpy as np
import pytest
import torch

import torchkbnufft_amd as tkbn
from conftest import rel_l2

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda")


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
    adj = [a.clone() for a in bt.adjoint()]
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
