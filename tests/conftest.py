import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    # torch (device-memory plumbing in some GPU tests) must initialise its own
    # HIP runtime before libbart_amd.so loads; importing it here makes that so
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def relerr(x, ref):
    """||x - ref|| / ||ref|| in FP64 (BART's nrmse without scaling)."""
    x = np.asarray(x, np.complex128).ravel()
    ref = np.asarray(ref, np.complex128).ravel()
    return float(np.linalg.norm(x - ref) / np.linalg.norm(ref))


def crandn(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
