"""Shared pytest setup: marker registration and repo-root import path.

``gpu`` tests run on an MI355X and call the HIP library through the C ABI;
everything else runs on the CPU (oracle checks, host logic, library loading).
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built HIP library")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    return oracle.lib()


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch
    import ymrt_amd
    assert torch.cuda.is_available(), "GPU test needs a GPU"
    ymrt_amd.lib()  # raises when the HIP library is missing: no fallback
    return ymrt_amd.Context(0)
