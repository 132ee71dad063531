"""Test configuration: repo root on sys.path, the `gpu` marker, shared fixtures."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and libsedona_gpu.so")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import oracle_py
    return oracle_py.load()


@pytest.fixture(scope="session")
def gpu():
    """GpuBackend on cuda:0 issuing work on torch's current stream."""
    import torch
    from pubsed_amd.runtime import GpuBackend
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    b = GpuBackend(0)
    b.use_torch_stream()
    yield b
    b.close()
