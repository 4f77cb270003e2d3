"""C-ABI library checks that need no GPU: it loads, exports every symbol
include/wgfft.h declares, and its host-side logic (validation, error codes
and messages, input synthesis) behaves like the reference's GpuFft
(src/lib.rs:506-533)."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
import wgfft
from wgfft import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "wgfft.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(wgfft_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    declared = header_functions()
    assert len(declared) >= 18
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(_lib.EXPORTS) == declared


def test_no_device_here_is_a_clean_error():
    if wgfft.is_gpu_available():
        pytest.skip("a device is present")
    with pytest.raises(wgfft.WgfftError) as e:
        wgfft.GpuFft.new()
    assert e.value.code == wgfft.WGFFT_ERR_NO_DEVICE
    L = _lib.load()
    h = ctypes.c_void_p(123)
    assert L.wgfft_create(ctypes.byref(h), 0) == wgfft.WGFFT_ERR_NO_DEVICE
    assert h.value is None


def test_null_handle_paths():
    L = _lib.load()
    assert L.wgfft_exec_host(None, None, None, 8, 1, -1) == wgfft.WGFFT_ERR_INVALID_ARGUMENT
    assert L.wgfft_exec_host_rows(None, None, None, 8, 1, -1) == wgfft.WGFFT_ERR_INVALID_ARGUMENT
    assert L.wgfft_exec_device(None, None, None, 8, 1, -1, None) == wgfft.WGFFT_ERR_INVALID_ARGUMENT
    assert L.wgfft_filter_device(None, None, None, None, 8, 1, 0, 0.0, None) == wgfft.WGFFT_ERR_INVALID_ARGUMENT
    assert L.wgfft_exec_strided_device(None, None, None, 8, 1, 1, 8, 1, 8, -1, None) == wgfft.WGFFT_ERR_INVALID_ARGUMENT
    assert L.wgfft_add_noise_device(None, None, 8, 1, 0, 0.1, None) == wgfft.WGFFT_ERR_INVALID_ARGUMENT
    assert "handle" in _lib.last_error()
    L.wgfft_destroy(None)  # no-op


def test_validation_contract():
    assert wgfft.validate_lengths([]) == 0
    assert wgfft.validate_lengths([1024] * 3) == 1024
    assert wgfft.validate_lengths([1]) == 1
    with pytest.raises(wgfft.WgfftError) as e:
        wgfft.validate_lengths([1024, 512])
    assert e.value.code == wgfft.WGFFT_ERR_RAGGED_BATCH
    assert str(e.value) == "All input vectors in a batch must have the same length"
    for bad in ([0], [3], [1000, 1000], [6]):
        with pytest.raises(wgfft.WgfftError) as e:
            wgfft.validate_lengths(bad)
        assert e.value.code == wgfft.WGFFT_ERR_INVALID_LENGTH
        assert str(e.value) == "Transform length must be a non-zero power of two"
    # per-signal order of src/lib.rs:528-533: the first signal decides
    with pytest.raises(wgfft.WgfftError) as e:
        wgfft.validate_lengths([3, 4])
    assert e.value.code == wgfft.WGFFT_ERR_INVALID_LENGTH
    with pytest.raises(wgfft.WgfftError) as e:
        wgfft.validate_lengths([4, 3])
    assert e.value.code == wgfft.WGFFT_ERR_RAGGED_BATCH


def test_variant_registry_counts():
    # host-side registry queries need no device
    for log2n in range(5, 15):
        assert wgfft.GpuFft.variant_count(1 << log2n) >= 1
    assert wgfft.GpuFft.variant_count(16) == 0  # register kernel
    assert wgfft.GpuFft.variant_count(1 << 20) == 0  # global passes
    assert wgfft.GpuFft.variant_count(12) == 0
    assert wgfft.GpuFft.pass_count(4096) == 0
    assert wgfft.GpuFft.pass_count(1 << 22) == 2
    assert wgfft.GpuFft.pass_count(1 << 27) == 3
    assert wgfft.GpuFft.pass_count(1 << 24) == 3  # the measured default: three radix-256 passes
    assert wgfft.GpuFft.pass_count(1 << 26) == 3
    for p in range(2):
        assert wgfft.GpuFft.pass_variant_count(1 << 22, p) >= 2
    assert wgfft.GpuFft.pass_variant_count(1 << 22, 2) == 0
    for log2n in range(9, 15):
        assert wgfft.GpuFft.adjoint_variant_count(1 << log2n) >= 1
    assert wgfft.GpuFft.adjoint_variant_count(1 << 20) == 0
    assert wgfft.GpuFft.adjoint_variant_count(12) == 0
    L = _lib.load()
    assert L.wgfft_select_variant(None, 4096, 0, None, 0) == wgfft.WGFFT_ERR_INVALID_ARGUMENT
    assert L.wgfft_select_pass_variant(None, 1 << 20, 0, 0, None, 0) == wgfft.WGFFT_ERR_INVALID_ARGUMENT
    assert L.wgfft_select_adjoint_variant(None, 4096, 0, None, 0) == wgfft.WGFFT_ERR_INVALID_ARGUMENT


def test_synth_uniform_matches_oracle_stream():
    for seed in (0, 1, 0x5EED0001):
        assert np.array_equal(wgfft.synth_uniform((5, 33), seed), oracle.synth_uniform((5, 33), seed))


def test_name_and_alloc_cap_knob():
    L = _lib.load()
    assert "Stockham" in L.wgfft_name().decode()
    wgfft.set_alloc_cap(1 << 20)
    wgfft.set_alloc_cap(0)


def test_python_mirror_validates_before_device_work():
    # the mirror validates through the C library, so these errors need no GPU
    from wgfft.gpu_fft import _flatten

    flat, n = _flatten([])
    assert flat.shape == (0, 0) and n == 0
    with pytest.raises(wgfft.WgfftError):
        _flatten([np.zeros(4, np.complex64), np.zeros(8, np.complex64)])
    with pytest.raises(wgfft.WgfftError):
        _flatten([np.zeros(0, np.complex64)])


def _unbound_executor():
    """A GpuFft whose handle was never created: its device wrappers must
    reject bad tensors before they reach the C-ABI (which would get NULL)."""
    g = wgfft.GpuFft.__new__(wgfft.GpuFft)
    g._h = None
    g._L = _lib.load()
    g._plans = []
    return g


def test_device_wrappers_reject_host_tensors():
    import torch

    g = _unbound_executor()
    x = torch.zeros((2, 64), dtype=torch.complex64)  # host memory: never a kernel pointer
    w = torch.ones(64, dtype=torch.float32)
    cases = [
        lambda: g.exec_device(x, x),
        lambda: g.exec_windowed_device(x, w, x),
        lambda: g.vjp_windowed_device(x, w, x, x, w),
        lambda: g.fwd_adjoint_masked_device(x, w, w, None, x, w, torch.zeros(1)),
        lambda: g.filter_device(x, w, x),
        lambda: g.add_noise_device(x),
        lambda: g.exec_strided_device(x, x),
        lambda: g.bench_device(x, x),
    ]
    for i, f in enumerate(cases):
        with pytest.raises(wgfft.WgfftError) as e:
            f()
        assert e.value.code == wgfft.WGFFT_ERR_INVALID_ARGUMENT, i
        assert "GPU" in str(e.value), (i, str(e.value))


def test_real_operand_checks():
    import torch
    from wgfft.gpu_fft import _check_real, _check_tensor

    with pytest.raises(wgfft.WgfftError) as e:
        _check_real(torch.ones(63, dtype=torch.float32), 64, "window")
    assert "window" in str(e.value) and "64" in str(e.value)
    for bad in (torch.ones(64, dtype=torch.float16), torch.ones(128)[::2], [1.0] * 64, None):
        with pytest.raises(wgfft.WgfftError):
            _check_real(bad, 64, "window")
    with pytest.raises(wgfft.WgfftError):
        _check_tensor(np.zeros((2, 64), np.complex64))


def test_select_split_validation():
    L = _lib.load()
    arr = (ctypes.c_int * 3)(7, 7, 6)
    assert L.wgfft_select_split(None, 1 << 20, arr, 3, None, 0) == wgfft.WGFFT_ERR_INVALID_ARGUMENT
