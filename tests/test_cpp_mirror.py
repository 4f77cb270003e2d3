"""The C++ mirror of GpuFft / FftExecutor (wgls-rs-fft_amd/cpp/gpu_fft.hpp)
running the reference's own tests restated in C++ (gpu_fft_tests.cpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "wgls-rs-fft_amd", "wgfft", "_lib", "gpu_fft_tests")


def _binary():
    if not os.path.exists(BIN):
        import sys

        sys.path.insert(0, os.path.join(ROOT, "wgls-rs-fft_amd"))
        import build

        build.build()
    return BIN


def _run(mode):
    r = subprocess.run([_binary(), mode], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    return r.stdout


def test_cpp_mirror_error_contract_without_device():
    import wgfft

    if wgfft.is_gpu_available():
        pytest.skip("a device is present")
    out = _run("cpu")
    assert out.count("PASS") >= 20


@pytest.mark.gpu
def test_cpp_mirror_reference_tests_on_gpu():
    out = _run("gpu")
    assert out.count("PASS") >= 35
