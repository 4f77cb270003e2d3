import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "wgls-rs-fft_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C-ABI)")


@pytest.fixture(scope="session")
def gpu():
    """One GpuFft handle for the whole GPU session (fails loudly without a device)."""
    import wgfft

    h = wgfft.GpuFft.new()
    yield h
    h.close()


@pytest.fixture(scope="session")
def torch_cuda():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


_PARITY = {}


@pytest.fixture(scope="session")
def parity():
    """Collects measured errors per configuration; written as JSON to the
    path in WGFFT_PARITY_REPORT (if set) when the session ends."""
    return _PARITY


def pytest_sessionfinish(session, exitstatus):
    path = os.environ.get("WGFFT_PARITY_REPORT")
    if path and _PARITY:
        import json

        with open(path, "w") as f:
            json.dump(_PARITY, f, indent=1, sort_keys=True, default=float)
