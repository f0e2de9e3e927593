"""Test configuration: import paths, the `gpu` marker, shared helpers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "pbf-cuda_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (an MI355X; select with -m gpu)")
    config.addinivalue_line("markers", "slow: CPU test that takes tens of seconds")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle

    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def pbf_mod():
    import pbf

    return pbf


@pytest.fixture(scope="session")
def device():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test selected but no HIP device is visible")
    return torch.device("cuda:0")


def to_dev(a: np.ndarray, device):
    import torch

    if a.dtype == np.uint32:
        a = a.astype(np.int32)
    return torch.from_numpy(np.ascontiguousarray(a).ravel().copy()).to(device)


def from_dev(t, shape=None, dtype=None):
    a = t.detach().cpu().numpy()
    if dtype is not None:
        a = a.astype(dtype)
    return a.reshape(shape) if shape is not None else a
