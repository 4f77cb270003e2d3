"""MI355X-native batched complex64 Stockham FFT (drop-in for wgls-rs-fft's
``GpuFft`` path).  Python side: a ctypes mirror of the reference interface
over the C-ABI in ``include/wgfft.h``; all arithmetic runs in the HIP
kernels of ``wgfft/_lib/libwgfft.so``."""
from ._lib import (WGFFT_ERR_HIP, WGFFT_ERR_INVALID_ARGUMENT, WGFFT_ERR_INVALID_LENGTH, WGFFT_ERR_NO_DEVICE,
                   WGFFT_ERR_OUT_OF_MEMORY, WGFFT_ERR_RAGGED_BATCH, WGFFT_FORWARD, WGFFT_INVERSE, WgfftError)
from .gpu_fft import (FftExecutor, GpuFft, RaggedPlan, is_gpu_available, ragged_layout, set_alloc_cap, synth_uniform,
                      validate_lengths)

__all__ = [
    "GpuFft", "FftExecutor", "RaggedPlan", "ragged_layout", "WgfftError", "is_gpu_available", "validate_lengths", "synth_uniform",
    "set_alloc_cap", "WGFFT_FORWARD", "WGFFT_INVERSE", "WGFFT_ERR_INVALID_LENGTH", "WGFFT_ERR_RAGGED_BATCH",
    "WGFFT_ERR_NO_DEVICE", "WGFFT_ERR_OUT_OF_MEMORY", "WGFFT_ERR_HIP", "WGFFT_ERR_INVALID_ARGUMENT",
]
