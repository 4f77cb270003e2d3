"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
oracle (every function, driven by oracle/sanitize_check.c) and the
library's host runtime (every csrc/*.hip with the sanitizers on host code
only, wgls-rs-fft_amd/sanitize.mk), exercised by the C++ mirror tests: the
host logic on the CPU here, and on the GPU the whole runtime -- the pinned
staging pipeline, plans and global passes, the autotuner, mixed-length plans
freed with their handle."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "wgls-rs-fft_amd")
SAN_BIN = os.path.join(PKG, "wgfft", "_lib", "san", "gpu_fft_tests")
# the HIP runtime reserves address ranges that overlap ASan's protected
# shadow gap; leaks of runtime-owned allocations at exit are not ours
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:protect_shadow_gap=0:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


def _run(args, timeout):
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout, env=ENV)
    print(r.stdout[-3000:], r.stderr[-3000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "FAIL" not in r.stdout and "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
    return r.stdout


def test_oracle_under_sanitizers():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    out = _run([os.path.join(ROOT, "oracle", "_san", "oracle_check")], 300)
    assert "ALL PASSED" in out and out.count("PASS") >= 20


def _san_binary():
    if not os.path.exists(SAN_BIN):
        subprocess.run(["make", "-s", "-f", os.path.join(PKG, "sanitize.mk"), f"-j{min(8, os.cpu_count() or 1)}"],
                       check=True)
    return SAN_BIN


def test_host_runtime_under_sanitizers_without_device():
    import wgfft

    if wgfft.is_gpu_available():
        pytest.skip("a device is present (the GPU test runs the whole runtime)")
    out = _run([_san_binary(), "cpu"], 300)
    assert out.count("PASS") >= 20


@pytest.mark.gpu
def test_host_runtime_under_sanitizers_on_gpu():
    assert os.path.exists(SAN_BIN), "build the sanitizer variant first (__graft_entry__.build)"
    out = _run([SAN_BIN, "gpu"], 600)
    assert out.count("PASS") >= 35
