"""pytest configuration: the `gpu` marker and shared fixtures.

Tests marked `gpu` need an MI355X and run the HIP library through its C ABI;
everything else runs on the CPU (oracle, argument checking, ABI exports).
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("", "ops-blas_amd", "oracle", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (HIP library through the C ABI)")


@pytest.fixture(scope="session")
def handle():
    """One aclblas handle bound to torch's current stream on cuda:0."""
    import torch
    import aclblas
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.init()
    h = aclblas.Handle()
    assert h.set_stream(torch.cuda.current_stream().cuda_stream) == 0
    yield h
    torch.cuda.synchronize()
    h.destroy()
