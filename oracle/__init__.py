"""CPU checker for the batched complex64 Stockham FFT path.

TEST INFRASTRUCTURE ONLY.  The product library never imports this package;
only ``tests/``, the smoke check and the bench's ``cpu_baseline`` leg do.

Thin numpy wrappers over ``oracle/liboracle.so`` (built from ``oracle.c`` by
``oracle/Makefile``):

* ``ref_fft_f32``    -- restatement of the reference f32 Stockham R4/R2 path
  (wgls-rs-fft ``src/shaders.rs:8-106``, ``src/lib.rs:517-669,722-903``).
* ``fft64`` / ``dft64`` -- FP64 FFT (radix-2) and direct DFT.
* ``windowed_fwd64``, ``vjp64``, ``masked_loss64`` -- FP64 forward, analytic
  adjoint and loss for ``Y = FFT(w * x)``.
* ``synth_*`` -- seeded input generators (SplitMix64 -> PCG32).

Every function here is pinned in ``tests/test_oracle.py`` against golden
vectors, closed forms and invariants.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64 = ctypes.c_uint64


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_splitmix64_next.argtypes = [ctypes.POINTER(_u64)]
        L.oracle_splitmix64_next.restype = _u64
        L.oracle_pcg32_raw.argtypes = [_u64, _u64, _u32p, _u64]
        L.oracle_synth_uniform.argtypes = [_f32p, _u64, _u64]
        L.oracle_synth_gauss.argtypes = [_f64p, _u64, _u64]
        L.oracle_synth_sinusoids.argtypes = [_f32p, _u64, _u64, _u64, ctypes.c_int]
        L.oracle_dft_bins64.argtypes = [_f32p, _u64, ctypes.POINTER(_u64), _u64, _f64p]
        L.oracle_synth_ragged_log2.argtypes = [_u32p, _u64, ctypes.c_uint32, ctypes.c_uint32, _u64]
        L.oracle_synth_window_mask.argtypes = [_f32p, _f32p, _u64, _u64]
        L.oracle_ref_fft_f32.argtypes = [_f32p, _f32p, _u64, _u64, ctypes.c_int, ctypes.c_int]
        L.oracle_ref_fft_f32.restype = ctypes.c_int
        L.oracle_ref_twiddles.argtypes = [_f32p, _u64]
        L.oracle_dft64.argtypes = [_f64p, _f64p, _u64, _u64, ctypes.c_int]
        L.oracle_fft64_inplace.argtypes = [_f64p, _u64, _u64, ctypes.c_int, ctypes.c_int]
        L.oracle_fft64_inplace.restype = ctypes.c_int
        L.oracle_windowed_fwd64.argtypes = [_f32p, _f32p, _f64p, _u64, _u64, ctypes.c_int]
        L.oracle_vjp64.argtypes = [_f32p, _f32p, _f32p, _f64p, _f64p, _u64, _u64, ctypes.c_int]
        L.oracle_masked_loss64.argtypes = [_f32p, _f32p, _f32p, _f64p, _u64, _u64, ctypes.c_int]
        L.oracle_masked_loss64.restype = ctypes.c_double
        _lib = L
    return _lib


def _threads() -> int:
    """Threads for the checker: the CPUs this process may run on, capped by
    OMP_NUM_THREADS when set (a GPU box declares its per-GPU CPU share there)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def _as_c64(x) -> np.ndarray:
    return np.ascontiguousarray(x, dtype=np.complex64)


def _as_c128(x) -> np.ndarray:
    return np.ascontiguousarray(x, dtype=np.complex128)


# ---------------------------------------------------------------- inputs
def splitmix64(seed: int, count: int) -> np.ndarray:
    s = _u64(seed)
    return np.array([lib().oracle_splitmix64_next(ctypes.byref(s)) for _ in range(count)],
                    dtype=np.uint64)


def pcg32_raw(initstate: int, initseq: int, count: int) -> np.ndarray:
    out = np.empty(count, dtype=np.uint32)
    lib().oracle_pcg32_raw(initstate, initseq, _p(out, _u32p), count)
    return out


def synth_uniform(shape, seed: int) -> np.ndarray:
    """complex64 array of the given shape, re/im i.i.d. U[-1,1) (re first)."""
    n = int(np.prod(shape))
    out = np.empty(2 * n, dtype=np.float32)
    lib().oracle_synth_uniform(_p(out, _f32p), 2 * n, seed)
    return out.view(np.complex64).reshape(shape)


def synth_gauss(count: int, seed: int) -> np.ndarray:
    out = np.empty(count, dtype=np.float64)
    lib().oracle_synth_gauss(_p(out, _f64p), count, seed)
    return out


def synth_sinusoids(n: int, batch: int, seed: int) -> np.ndarray:
    out = np.empty(2 * n * batch, dtype=np.float32)
    lib().oracle_synth_sinusoids(_p(out, _f32p), n, batch, seed, _threads())
    return out.view(np.complex64).reshape(batch, n)


def sinusoid_tones(n: int, batch: int, seed: int):
    """(f, phi) of the 32 tones per transform, as drawn by synth_sinusoids."""
    f = np.empty((batch, 32))
    ph = np.empty((batch, 32))
    for b in range(batch):
        u = _uniform24(seed + b, 64)
        f[b] = u[0::2] * n
        ph[b] = 2 * np.pi * u[1::2]
    return f, ph


def _uniform24(seed: int, count: int) -> np.ndarray:
    """count draws (u32 >> 8) * 2^-24 of the seeded stream (FP64)."""
    s = _u64(seed)
    st = lib().oracle_splitmix64_next(ctypes.byref(s))
    sq = lib().oracle_splitmix64_next(ctypes.byref(s))
    return (pcg32_raw(st, sq, count) >> 8).astype(np.float64) / 16777216.0


def dft_bins64(x, bins) -> np.ndarray:
    """Selected FP64 DFT bins of one complex64 signal by direct summation."""
    x = _as_c64(x).ravel()
    bins = np.ascontiguousarray(bins, dtype=np.uint64)
    out = np.empty(len(bins), dtype=np.complex128)
    lib().oracle_dft_bins64(_p(x, _f32p), len(x), bins.ctypes.data_as(ctypes.POINTER(_u64)), len(bins),
                            _p(out, _f64p))
    return out


def synth_ragged_log2(count: int, kmin: int, kmax: int, seed: int) -> np.ndarray:
    out = np.empty(count, dtype=np.uint32)
    lib().oracle_synth_ragged_log2(_p(out, _u32p), count, kmin, kmax, seed)
    return out


def synth_window_mask(n: int, seed: int):
    w = np.empty(n, dtype=np.float32)
    m = np.empty(n, dtype=np.float32)
    lib().oracle_synth_window_mask(_p(w, _f32p), _p(m, _f32p), n, seed)
    return w, m


# ------------------------------------------------- reference restatement
def ref_twiddles(n: int) -> np.ndarray:
    out = np.empty(2 * n, dtype=np.float32)
    lib().oracle_ref_twiddles(_p(out, _f32p), n)
    return out.view(np.complex64)


def ref_fft_f32(x, inverse: bool = False, threads: int | None = None) -> np.ndarray:
    """Reference GpuFft::fft / ifft arithmetic on a (batch, n) complex64 array."""
    x = _as_c64(x)
    if x.ndim == 1:
        x = x[None, :]
    batch, n = x.shape
    out = np.empty_like(x)
    rc = lib().oracle_ref_fft_f32(_p(x, _f32p), _p(out, _f32p), n, batch, int(inverse),
                                  threads or _threads())
    if rc != 0:
        raise ValueError("Transform length must be a non-zero power of two")
    return out


# ------------------------------------------------------------ FP64 oracle
def dft64(x, sign: int = -1) -> np.ndarray:
    x = _as_c128(x)
    if x.ndim == 1:
        x = x[None, :]
    batch, n = x.shape
    out = np.empty_like(x)
    lib().oracle_dft64(_p(x, _f64p), _p(out, _f64p), n, batch, sign)
    return out


def fft64(x, sign: int = -1, threads: int | None = None) -> np.ndarray:
    """FP64 FFT along the last axis; sign=+1 is the unnormalised inverse."""
    y = np.array(x, dtype=np.complex128, order="C", copy=True)
    if y.ndim == 1:
        y = y[None, :]
    batch, n = y.shape
    rc = lib().oracle_fft64_inplace(_p(y, _f64p), n, batch, sign, threads or _threads())
    if rc != 0:
        raise ValueError("Transform length must be a non-zero power of two")
    return y


def ifft64(x) -> np.ndarray:
    """FP64 inverse with the 1/N scale of GpuFft::ifft."""
    y = fft64(x, sign=+1)
    return y / y.shape[-1]


def windowed_fwd64(x, w) -> np.ndarray:
    x = _as_c64(x)
    w = np.ascontiguousarray(w, dtype=np.float32)
    batch, n = x.shape
    y = np.empty((batch, n), dtype=np.complex128)
    lib().oracle_windowed_fwd64(_p(x, _f32p), _p(w, _f32p), _p(y, _f64p), n, batch, _threads())
    return y


def vjp64(x, w, gy):
    x = _as_c64(x)
    gy = _as_c64(gy)
    w = np.ascontiguousarray(w, dtype=np.float32)
    batch, n = x.shape
    gx = np.empty((batch, n), dtype=np.complex128)
    gw = np.empty(n, dtype=np.float64)
    lib().oracle_vjp64(_p(x, _f32p), _p(w, _f32p), _p(gy, _f32p), _p(gx, _f64p), _p(gw, _f64p),
                       n, batch, _threads())
    return gx, gw


def masked_loss64(x, w, mask, want_grad_y: bool = True):
    x = _as_c64(x)
    w = np.ascontiguousarray(w, dtype=np.float32)
    mask = np.ascontiguousarray(mask, dtype=np.float32)
    batch, n = x.shape
    gy = np.empty((batch, n), dtype=np.complex128) if want_grad_y else None
    loss = lib().oracle_masked_loss64(_p(x, _f32p), _p(w, _f32p), _p(mask, _f32p),
                                      _p(gy, _f64p) if gy is not None else None, n, batch,
                                      _threads())
    return loss, gy


def rel_l2(a, b) -> float:
    """||a - b||_2 / ||b||_2 in FP64."""
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))
