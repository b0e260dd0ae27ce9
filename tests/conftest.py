"""Shared pytest configuration: the `gpu` marker and device helpers."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X); run with -m gpu")


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    return torch.device("cuda:0")


def to_dev(a: np.ndarray, dev):
    """numpy uint32/uint64 array -> device tensor with the same bytes."""
    import torch
    a = np.ascontiguousarray(a)
    view = {4: np.int32, 8: np.int64}[a.dtype.itemsize]
    return torch.from_numpy(a.view(view)).to(dev)


def to_np(t, dtype=np.uint32) -> np.ndarray:
    return t.cpu().numpy().view(dtype)
