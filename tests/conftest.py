import os
import sys

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, GOLDEN):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and the built libtkbn_hip.so")


def pytest_runtest_setup(item):
    if "gpu" in item.keywords:
        # name the case before any GPU work starts (a fault can end the process)
        print(f"\n[gpu-test] {item.nodeid}", flush=True)
        if not torch.cuda.is_available():
            pytest.fail("GPU test selected but torch.cuda.is_available() is False", pytrace=False)


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def rel_l2(a, b) -> float:
    a = torch.as_tensor(a).detach().cpu().to(torch.complex128)
    b = torch.as_tensor(b).detach().cpu().to(torch.complex128)
    den = torch.linalg.vector_norm(b)
    return float(torch.linalg.vector_norm(a - b) / (den if den > 0 else 1.0))


def create_input_plus_noise(shape, is_complex, gen=None, dtype=torch.float64):
    """``arange + N(0,1)`` inputs as the reference's tests/conftest.py:28-37 builds them."""
    x = torch.arange(int(np.prod(shape)), dtype=dtype).reshape(shape)
    if is_complex:
        cd = torch.complex128 if dtype == torch.float64 else torch.complex64
        noise = torch.complex(torch.randn(shape, generator=gen, dtype=dtype), torch.randn(shape, generator=gen, dtype=dtype))
        return x.to(cd) + noise
    return x + torch.randn(shape, generator=gen, dtype=dtype)


def create_ktraj(ndims, klength, gen=None, dtype=torch.float64):
    return torch.rand((ndims, klength), generator=gen, dtype=dtype) * 2 * np.pi - np.pi
