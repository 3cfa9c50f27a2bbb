"""Shared test set-up: import paths, the ``gpu`` marker and small helpers."""
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("ark-cpp-mpi_amd", "oracle"):
    p = os.path.join(REPO, sub)
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long CPU test (oracle at 64^3 and above)")


def bits_equal(a, b) -> bool:
    """Bitwise equality of two float64 arrays (distinguishes -0.0 and NaN payloads)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def describe_mismatch(name, a, b) -> str:
    d = np.abs(np.asarray(a) - np.asarray(b))
    idx = np.argwhere(~(a.view(np.int64) == b.view(np.int64)))
    return f"{name}: {len(idx)} elements differ, max |diff| {np.nanmax(d):.3e}, first at {idx[:3].tolist()}"


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def golden():
    return GOLDEN
