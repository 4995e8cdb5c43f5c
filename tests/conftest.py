"""Shared test setup: import paths, the ``gpu`` marker and a few helpers.

``-m "not gpu"`` tests run on the CPU: the oracle against its independent checks and
the golden fixtures, the autograd wiring (with the oracle standing in for the HIP
library), and the C ABI's loading and host-side validation.  ``-m gpu`` tests call the
HIP library and compare it with the oracle.
"""
from __future__ import annotations

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cuda-knn-aniso-gaussian-feature-aggregation_amd")
for path in (ROOT, PKG):
    if path not in sys.path:
        sys.path.insert(0, path)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")
