"""ctypes binding of libwgfft.so (declared in include/wgfft.h).

The library is loaded from ``wgfft/_lib/libwgfft.so``.  When it is missing
the import raises: there is no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libwgfft.so")

WGFFT_OK = 0
WGFFT_ERR_INVALID_LENGTH = -1
WGFFT_ERR_RAGGED_BATCH = -2
WGFFT_ERR_NO_DEVICE = -3
WGFFT_ERR_OUT_OF_MEMORY = -4
WGFFT_ERR_HIP = -5
WGFFT_ERR_INVALID_ARGUMENT = -6
WGFFT_FORWARD = -1
WGFFT_INVERSE = 1

# Every symbol include/wgfft.h declares.
EXPORTS = (
    "wgfft_device_available", "wgfft_create", "wgfft_destroy", "wgfft_last_error", "wgfft_name",
    "wgfft_validate_lengths", "wgfft_exec_host", "wgfft_exec_host_rows", "wgfft_exec_device", "wgfft_exec_device_noalloc",
    "wgfft_exec_windowed_host",
    "wgfft_exec_windowed_device", "wgfft_vjp_windowed_device", "wgfft_fwd_adjoint_masked_device",
    "wgfft_filter_device", "wgfft_exec_strided_device", "wgfft_add_noise_device",
    "wgfft_ragged_create", "wgfft_ragged_destroy", "wgfft_ragged_exec_device", "wgfft_ragged_samples", "wgfft_ragged_describe",
    "wgfft_bench_device",
    "wgfft_autotune", "wgfft_plan_describe", "wgfft_variant_count", "wgfft_select_variant",
    "wgfft_pass_count", "wgfft_pass_variant_count", "wgfft_select_pass_variant", "wgfft_select_split",
    "wgfft_adjoint_variant_count", "wgfft_select_adjoint_variant",
    "wgfft_synth_uniform", "wgfft_set_alloc_cap",
)


class WgfftError(RuntimeError):
    """Error returned through the C-ABI (code + wgfft_last_error message)."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code
        self.message = message


_u64 = ctypes.c_uint64
_vp = ctypes.c_void_p
_i = ctypes.c_int
_lib = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    # One HIP runtime per process: torch ships its own libamdhip64 (soname
    # libamdhip64.so.7, but linked by the unversioned name), so when torch is
    # present it is imported first and libwgfft.so binds to that runtime;
    # loading libwgfft.so first would make torch load a second runtime that
    # sees no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python wgls-rs-fft_amd/build.py` "
            "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    L.wgfft_device_available.restype = _i
    L.wgfft_create.argtypes = [ctypes.POINTER(_vp), _i]
    L.wgfft_destroy.argtypes = [_vp]
    L.wgfft_destroy.restype = None
    L.wgfft_last_error.restype = ctypes.c_char_p
    L.wgfft_name.restype = ctypes.c_char_p
    L.wgfft_validate_lengths.argtypes = [ctypes.POINTER(_u64), _u64, ctypes.POINTER(_u64)]
    L.wgfft_exec_host.argtypes = [_vp, _vp, _vp, _u64, _u64, _i]
    L.wgfft_exec_host_rows.argtypes = [_vp, _vp, _vp, _u64, _u64, _i]
    L.wgfft_exec_device.argtypes = [_vp, _vp, _vp, _u64, _u64, _i, _vp]
    L.wgfft_exec_device_noalloc.argtypes = [_vp, _vp, _vp, _u64, _u64, _i, _vp]
    L.wgfft_exec_windowed_host.argtypes = [_vp, _vp, _vp, _vp, _u64, _u64]
    L.wgfft_exec_windowed_device.argtypes = [_vp, _vp, _vp, _vp, _u64, _u64, _vp]
    L.wgfft_vjp_windowed_device.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _u64, _u64, _vp]
    L.wgfft_fwd_adjoint_masked_device.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u64, _u64, _vp]
    L.wgfft_filter_device.argtypes = [_vp, _vp, _vp, _vp, _u64, _u64, ctypes.c_uint32, ctypes.c_float, _vp]
    L.wgfft_exec_strided_device.argtypes = [_vp, _vp, _vp, _u64, _u64, _u64, _u64, _u64, _u64, _i, _vp]
    L.wgfft_add_noise_device.argtypes = [_vp, _vp, _u64, _u64, ctypes.c_uint32, ctypes.c_float, _vp]
    L.wgfft_ragged_create.argtypes = [_vp, _vp, _vp, _u64, ctypes.POINTER(_vp)]
    L.wgfft_ragged_destroy.argtypes = [_vp]
    L.wgfft_ragged_destroy.restype = None
    L.wgfft_ragged_exec_device.argtypes = [_vp, _vp, _vp, _i, _i, _vp]
    L.wgfft_ragged_samples.argtypes = [_vp]
    L.wgfft_ragged_samples.restype = _u64
    L.wgfft_ragged_describe.argtypes = [_vp, ctypes.c_char_p, _u64]
    L.wgfft_bench_device.argtypes = [_vp, _vp, _vp, _u64, _u64, _i, _i, _i, ctypes.POINTER(ctypes.c_double)]
    L.wgfft_autotune.argtypes = [_vp, _u64, _u64, ctypes.c_char_p, _u64]
    L.wgfft_plan_describe.argtypes = [_vp, _u64, ctypes.c_char_p, _u64]
    L.wgfft_variant_count.argtypes = [_u64]
    L.wgfft_select_variant.argtypes = [_vp, _u64, _i, ctypes.c_char_p, _u64]
    L.wgfft_pass_count.argtypes = [_u64]
    L.wgfft_pass_variant_count.argtypes = [_u64, _i]
    L.wgfft_select_pass_variant.argtypes = [_vp, _u64, _i, _i, ctypes.c_char_p, _u64]
    L.wgfft_select_split.argtypes = [_vp, _u64, ctypes.POINTER(_i), _i, ctypes.c_char_p, _u64]
    L.wgfft_adjoint_variant_count.argtypes = [_u64]
    L.wgfft_select_adjoint_variant.argtypes = [_vp, _u64, _i, ctypes.c_char_p, _u64]
    L.wgfft_synth_uniform.argtypes = [_vp, _u64, _u64]
    L.wgfft_synth_uniform.restype = None
    L.wgfft_set_alloc_cap.argtypes = [_u64]
    L.wgfft_set_alloc_cap.restype = None
    _lib = L
    return L


_rows = None


def rows_bridge():
    """The _rows extension (wgfft/_lib/_rows*.so, built with the library:
    calls wgfft_exec_host_rows on a list of numpy rows, reading their
    addresses through the buffer protocol) and that entry point's address."""
    global _rows
    if _rows is None:
        import importlib.machinery
        import importlib.util
        import sysconfig

        path = os.path.join(_HERE, "_lib", "_rows" + sysconfig.get_config_var("EXT_SUFFIX"))
        if not os.path.exists(path):
            raise ImportError(f"{path} is missing: build it with `python wgls-rs-fft_amd/build.py`")
        loader = importlib.machinery.ExtensionFileLoader("wgfft._rows", path)
        mod = importlib.util.module_from_spec(importlib.util.spec_from_file_location("wgfft._rows", path, loader=loader))
        loader.exec_module(mod)
        _rows = (mod, ctypes.cast(load().wgfft_exec_host_rows, ctypes.c_void_p).value)
    return _rows


def last_error() -> str:
    return load().wgfft_last_error().decode()


def check(rc: int) -> None:
    if rc != WGFFT_OK:
        raise WgfftError(rc, last_error())
