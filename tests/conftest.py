import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct MI355X (gfx950)")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def torch_dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def h2d(torch, a):
    """numpy -> device through a freshly allocated pinned host buffer (the
    HIP runtime is never handed pageable memory)"""
    import numpy as np
    t = torch.from_numpy(np.ascontiguousarray(a)).pin_memory()
    return t.to("cuda")


def d2h(torch, t):
    """device -> numpy through a pinned host buffer"""
    out = torch.empty(tuple(t.shape), dtype=t.dtype, pin_memory=True)
    out.copy_(t)
    torch.cuda.synchronize()
    return out.numpy()
