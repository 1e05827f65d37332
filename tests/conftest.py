"""Shared test set-up: the `gpu` marker, paths, and the relative-error measure of the reference's own
check (validate_gpu.py:83: max|a - b| / max|b|)."""
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs through libcavity_solver.so on the device)')


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


FIELDS = ('u', 'v', 'w', 'mx', 'my', 'mz', 'phi')


def gpu_available():
    try:
        import torch  # noqa: F401  (device count only; does not initialise HIP)
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope='session')
def hip_lib():
    """The built HIP library; on a GPU run a missing library is an error, never a skip."""
    from les_impulse3d_64_amd.python_api import load_library
    return load_library()
