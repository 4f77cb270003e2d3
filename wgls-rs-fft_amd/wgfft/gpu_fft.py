"""Host-side mirror of wgls-rs-fft's ``GpuFft`` / ``FftExecutor`` interface.

Same method names, argument meaning and error behaviour as the reference
(``src/lib.rs:32-45`` trait, ``src/lib.rs:285-568`` runtime):

* ``GpuFft.new()`` creates the executor, raising when no GPU is usable
  (``src/lib.rs:285-330``; ``Default`` panics the same way, ``:946-950``).
* ``fft(inputs)`` / ``ifft(inputs)`` take a batch (list) of equal-length,
  power-of-two complex signals and return a new list of complex64 arrays;
  the inverse is 1/N scaled (``src/lib.rs:433-488``).
* An empty batch returns ``[]``; a ragged batch raises "All input vectors in
  a batch must have the same length"; a zero or non-power-of-two length
  raises "Transform length must be a non-zero power of two"
  (``src/lib.rs:506-533``).  Validation runs in the C library before any
  device work, so it is exercised without a GPU.

Every transform runs in the HIP kernels of libwgfft.so through the C-ABI.
"""
from __future__ import annotations

import ctypes
import weakref
from typing import Iterable, List, Sequence

import numpy as np

from . import _lib
from ._lib import WGFFT_FORWARD, WGFFT_INVERSE, WgfftError, check


def validate_lengths(lengths: Sequence[int]) -> int:
    """Return the common length n (0 for an empty batch) or raise WgfftError."""
    L = _lib.load()
    arr = np.zeros(1, dtype=np.uint64) if not len(lengths) else np.asarray(lengths, dtype=np.uint64)
    n = ctypes.c_uint64(0)
    check(L.wgfft_validate_lengths(arr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), len(lengths),
                                   ctypes.byref(n)))
    return int(n.value)


def is_gpu_available() -> bool:
    """GpuFft::is_gpu_available (src/lib.rs:381-389)."""
    return bool(_lib.load().wgfft_device_available())


def _flatten(inputs: Sequence, n: int = None) -> tuple[np.ndarray, int]:
    if n is None:
        n = validate_lengths([len(x) for x in inputs])
    if not inputs:
        return np.empty((0, 0), dtype=np.complex64), 0
    flat = np.empty((len(inputs), n), dtype=np.complex64)
    for i, x in enumerate(inputs):
        flat[i] = np.asarray(x, dtype=np.complex64)
    return flat, n


class FftExecutor:
    """Plugin interface of the reference (src/lib.rs:32-45)."""

    def name(self) -> str:  # pragma: no cover - interface
        raise NotImplementedError

    def fft(self, inputs):  # pragma: no cover - interface
        raise NotImplementedError

    def ifft(self, inputs):  # pragma: no cover - interface
        raise NotImplementedError

    def as_any(self):
        return self


class GpuFft(FftExecutor):
    """Batched complex64 FFT on one MI355X through libwgfft.so."""

    def __init__(self, device: int = 0):
        L = _lib.load()
        h = ctypes.c_void_p()
        check(L.wgfft_create(ctypes.byref(h), int(device)))
        self._h = h
        self._L = L
        self._plans = []  # RaggedPlans made on this handle (weak references)

    @classmethod
    def new(cls, device: int = 0) -> "GpuFft":
        return cls(device)

    is_gpu_available = staticmethod(is_gpu_available)

    def close(self) -> None:
        """Destroy the handle; wgfft_destroy also frees every mixed-length plan
        still open on it, so those RaggedPlans are marked closed first."""
        for ref in getattr(self, "_plans", ()):
            plan = ref()
            if plan is not None:
                plan._p = None
        self._plans = []
        if getattr(self, "_h", None):
            self._L.wgfft_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- FftExecutor -------------------------------------------------
    def name(self) -> str:
        return self._L.wgfft_name().decode()

    def fft(self, inputs: Sequence) -> List[np.ndarray]:
        return self._transform(inputs, WGFFT_FORWARD)

    def ifft(self, inputs: Sequence) -> List[np.ndarray]:
        return self._transform(inputs, WGFFT_INVERSE)

    def fft_windowed(self, inputs: Sequence, window) -> List[np.ndarray]:
        """FFT(w * x) per signal, w real of length n (multiplied at load)."""
        flat, n = _flatten(inputs)
        if flat.shape[0] == 0:
            return []
        w = np.ascontiguousarray(window, dtype=np.float32)
        if w.shape != (n,):
            raise WgfftError(_lib.WGFFT_ERR_INVALID_ARGUMENT, "window length must equal the signal length")
        out = np.empty_like(flat)
        check(self._L.wgfft_exec_windowed_host(self._h, flat.ctypes.data, w.ctypes.data, out.ctypes.data, n,
                                               flat.shape[0]))
        return list(out)

    def _transform(self, inputs: Sequence, direction: int) -> List[np.ndarray]:
        """One wgfft_exec_host_rows call over the signals where they lie (the
        staging copies gather them: no flattened copy of the batch), results
        as views of one output array.  Signals that are not contiguous
        complex64 arrays are converted first.  The lengths are checked in
        full (validate_lengths, the reference's errors in its order) only when
        a row does not have the first row's length or the call fails: the
        bridge rejects rows of another length and the library a length that
        is not a power of two."""
        if len(inputs) == 0:
            return []
        n = len(inputs[0])
        bridge, fn = _lib.rows_bridge()
        out = np.empty((len(inputs), n), dtype=np.complex64)
        h = self._h.value if self._h else 0
        rc, bad = bridge.exec_rows(fn, h, inputs, out.ctypes.data, n, direction)
        if bad >= 0 or rc != _lib.WGFFT_OK:
            validate_lengths(list(map(len, inputs)))  # raises for a ragged batch or a bad length
        if bad >= 0:
            rows = [np.ascontiguousarray(x, dtype=np.complex64) for x in inputs]
            rc, bad = bridge.exec_rows(fn, h, rows, out.ctypes.data, n, direction)
            assert bad < 0
        check(rc)
        return list(out)

    # ---- batched arrays (one contiguous buffer, TODO.md:7 of the reference)
    def fft_array(self, x: np.ndarray, inverse: bool = False) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.complex64)
        if x.ndim == 1:
            x = x[None, :]
        out = np.empty_like(x)
        check(self._L.wgfft_exec_host(self._h, x.ctypes.data, out.ctypes.data, x.shape[1], x.shape[0],
                                      WGFFT_INVERSE if inverse else WGFFT_FORWARD))
        return out

    # ---- device buffers (torch tensors on the GPU) ---------------------
    @staticmethod
    def _stream(stream, t=None) -> int:
        """The HIP stream to enqueue on: `stream` as given, else torch's
        current stream of tensor t's device (through torch's raw-stream
        query where it exists: 0.14 us against 2.8 us for
        torch.cuda.current_stream() per call)."""
        if stream is not None:
            return int(stream)
        import torch

        raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)
        if raw is not None:
            return raw(t.get_device() if t is not None else torch.cuda.current_device())
        return torch.cuda.current_stream().cuda_stream

    def exec_device(self, x, out, inverse: bool = False, stream=None) -> None:
        """Transform a (batch, n) complex64 CUDA tensor into `out`."""
        _check_tensor(x)
        _check_tensor(out, x.shape)
        batch, n = x.shape
        check(self._L.wgfft_exec_device(self._h, x.data_ptr(), out.data_ptr(), n, batch,
                                        WGFFT_INVERSE if inverse else WGFFT_FORWARD, self._stream(stream, x)))

    def exec_device_noalloc(self, x, out, inverse: bool = False, stream=None) -> None:
        """exec_device without any allocation (wgfft_exec_device_noalloc): the
        plan for n must exist (describe(n) makes it) and n <= 2^23."""
        _check_tensor(x)
        _check_tensor(out, x.shape)
        batch, n = x.shape
        check(self._L.wgfft_exec_device_noalloc(self._h, x.data_ptr(), out.data_ptr(), n, batch,
                                                WGFFT_INVERSE if inverse else WGFFT_FORWARD, self._stream(stream, x)))

    def exec_windowed_device(self, x, w, out, stream=None) -> None:
        """out = FFT(w * x) per row: x, out (batch, n) complex64, w n floats."""
        _check_tensor(x)
        _check_tensor(out, x.shape)
        batch, n = x.shape
        _check_real(w, n, "window")
        check(self._L.wgfft_exec_windowed_device(self._h, x.data_ptr(), w.data_ptr(), out.data_ptr(), n, batch,
                                                 self._stream(stream, x)))

    def vjp_windowed_device(self, x, w, gy, gx, gw, stream=None) -> None:
        """gx = w * F^H gy and gw = sum over rows of Re(conj(F^H gy) x)."""
        _check_tensor(x)
        _check_tensor(gy, x.shape)
        _check_tensor(gx, x.shape)
        batch, n = x.shape
        _check_real(w, n, "window")
        _check_real(gw, n, "gw")
        check(self._L.wgfft_vjp_windowed_device(self._h, x.data_ptr(), w.data_ptr(), gy.data_ptr(),
                                                gx.data_ptr(), gw.data_ptr(), n, batch, self._stream(stream, x)))

    def fwd_adjoint_masked_device(self, x, w, mask, y, gx, gw, loss, stream=None) -> None:
        """Fused forward + adjoint of L = sum m |FFT(w x)|^2: writes gx, gw,
        the loss (one float) and, when y is a tensor, Y = FFT(w x)."""
        _check_tensor(x)
        _check_tensor(gx, x.shape)
        if y is not None:
            _check_tensor(y, x.shape)
        batch, n = x.shape
        _check_real(w, n, "window")
        _check_real(mask, n, "mask")
        _check_real(gw, n, "gw")
        _check_real(loss, 1, "loss")
        check(self._L.wgfft_fwd_adjoint_masked_device(
            self._h, x.data_ptr(), w.data_ptr(), mask.data_ptr(), 0 if y is None else y.data_ptr(),
            gx.data_ptr(), gw.data_ptr(), loss.data_ptr(), n, batch, self._stream(stream, x)))

    def exec_strided_device(self, x, out, inverse: bool = False, stream=None) -> None:
        """Transform each row of a (batch, n) complex64 CUDA tensor view with
        any element strides (a slice, a transpose) into `out`, another such
        view; the tensors' own strides give the layouts."""
        if x.dtype != out.dtype or str(x.dtype) != "torch.complex64" or not x.is_cuda or not out.is_cuda:
            raise WgfftError(_lib.WGFFT_ERR_INVALID_ARGUMENT, "expected complex64 tensors on the GPU")
        if x.dim() != 2 or tuple(out.shape) != tuple(x.shape):
            raise WgfftError(_lib.WGFFT_ERR_INVALID_ARGUMENT, "expected two (batch, n) views of the same shape")
        batch, n = x.shape
        idist, istride = x.stride()
        odist, ostride = out.stride()
        check(self._L.wgfft_exec_strided_device(self._h, x.data_ptr(), out.data_ptr(), n, batch, istride, idist,
                                                ostride, odist, WGFFT_INVERSE if inverse else WGFFT_FORWARD,
                                                self._stream(stream, x)))

    def add_noise_device(self, x, seed: int = 0, sigma: float = 0.1, stream=None) -> None:
        """The reference pipeline's Gaussian noise stage, in place on a
        (batch, n) complex64 CUDA tensor."""
        _check_tensor(x)
        batch, n = x.shape
        check(self._L.wgfft_add_noise_device(self._h, x.data_ptr(), n, batch, seed & 0xFFFFFFFF, sigma,
                                             self._stream(stream, x)))

    def filter_device(self, x, s, out, noise_sigma: float = 0.0, seed: int = 0, stream=None) -> None:
        """out = ifft(fft(x + noise) * s), 1/N scaled: a (batch, n) complex64
        CUDA tensor filtered by the real spectral window s (n floats), with the
        pipeline's Gaussian noise of add_noise_device when noise_sigma != 0;
        out may be x."""
        _check_tensor(x)
        _check_tensor(out, x.shape)
        batch, n = x.shape
        _check_real(s, n, "spectral window")
        check(self._L.wgfft_filter_device(self._h, x.data_ptr(), s.data_ptr(), out.data_ptr(), n, batch,
                                          seed & 0xFFFFFFFF, noise_sigma, self._stream(stream, x)))

    def bench_device(self, x, out, inverse: bool = False, warmup: int = 3, iters: int = 50) -> float:
        """Seconds per exec_device call on resident buffers, timed with HIP
        events on the handle's stream (wgfft_bench_device)."""
        _check_tensor(x)
        _check_tensor(out, x.shape)
        batch, n = x.shape
        sec = ctypes.c_double(0.0)
        check(self._L.wgfft_bench_device(self._h, x.data_ptr(), out.data_ptr(), n, batch,
                                         WGFFT_INVERSE if inverse else WGFFT_FORWARD, warmup, iters,
                                         ctypes.byref(sec)))
        return sec.value

    def autotune(self, n: int, batch: int) -> str:
        buf = ctypes.create_string_buffer(512)
        check(self._L.wgfft_autotune(self._h, n, batch, buf, 512))
        return buf.value.decode()

    def describe(self, n: int) -> str:
        buf = ctypes.create_string_buffer(256)
        check(self._L.wgfft_plan_describe(self._h, n, buf, 256))
        return buf.value.decode()

    @staticmethod
    def variant_count(n: int) -> int:
        """Compiled single-kernel variants for length n (wgfft_variant_count)."""
        return int(_lib.load().wgfft_variant_count(n))

    @staticmethod
    def pass_count(n: int) -> int:
        """Global passes of the plan for length n (wgfft_pass_count)."""
        return int(_lib.load().wgfft_pass_count(n))

    @staticmethod
    def pass_variant_count(n: int, p: int) -> int:
        """Compiled candidates for global pass p of length n."""
        return int(_lib.load().wgfft_pass_variant_count(n, p))

    def select_pass_variant(self, n: int, p: int, index: int) -> str:
        """Pin candidate `index` for global pass p of length n."""
        buf = ctypes.create_string_buffer(256)
        check(self._L.wgfft_select_pass_variant(self._h, n, p, index, buf, 256))
        return buf.value.decode()

    def select_split(self, n: int, log2r) -> str:
        """Pin the radix split 2^log2r[0] x 2^log2r[1] x ... of length n
        (wgfft_select_split); returns the plan description."""
        r = [int(v) for v in log2r]
        arr = (ctypes.c_int * max(1, len(r)))(*r)
        buf = ctypes.create_string_buffer(256)
        check(self._L.wgfft_select_split(self._h, n, arr, len(r), buf, 256))
        return buf.value.decode()

    def select_variant(self, n: int, index: int) -> str:
        """Pin compiled variant `index` for length n (wgfft_select_variant)."""
        buf = ctypes.create_string_buffer(256)
        check(self._L.wgfft_select_variant(self._h, n, index, buf, 256))
        return buf.value.decode()

    @staticmethod
    def adjoint_variant_count(n: int) -> int:
        """Compiled fused forward+adjoint / VJP variants for length n."""
        return int(_lib.load().wgfft_adjoint_variant_count(n))

    def select_adjoint_variant(self, n: int, index: int) -> str:
        """Pin fused variant `index` for length n (wgfft_select_adjoint_variant)."""
        buf = ctypes.create_string_buffer(256)
        check(self._L.wgfft_select_adjoint_variant(self._h, n, index, buf, 256))
        return buf.value.decode()


class RaggedPlan:
    """Mixed-length batch: transform i has length 2^log2n[i] at complex
    offset offsets[i] of one contiguous buffer (wgfft_ragged_create)."""

    MODE_PERSISTENT, MODE_GRAPH, MODE_PER_SIZE = 0, 1, 2

    def __init__(self, gpu: GpuFft, offsets, log2n):
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        k = np.ascontiguousarray(log2n, dtype=np.uint32)
        if off.shape != k.shape:
            raise WgfftError(_lib.WGFFT_ERR_INVALID_ARGUMENT, "offsets and log2n differ in length")
        if not getattr(gpu, "_h", None):
            raise WgfftError(_lib.WGFFT_ERR_INVALID_ARGUMENT, "the GpuFft handle is closed")
        p = ctypes.c_void_p()
        check(gpu._L.wgfft_ragged_create(gpu._h, off.ctypes.data, k.ctypes.data, len(off), ctypes.byref(p)))
        self._p = p
        self._gpu = gpu
        self.samples = int(gpu._L.wgfft_ragged_samples(p))
        # complex elements a buffer must hold: the end of the last signal
        self.extent = int((off + (np.uint64(1) << k.astype(np.uint64))).max()) if len(off) else 0
        gpu._plans.append(weakref.ref(self))

    def exec_device(self, x, out, inverse: bool = False, mode: int = 0, stream=None) -> None:
        """x, out: contiguous complex64 CUDA tensors holding the packed signals."""
        import torch

        if not getattr(self, "_p", None):
            raise WgfftError(_lib.WGFFT_ERR_INVALID_ARGUMENT, "the mixed-length plan is closed")
        for t in (x, out):
            if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.complex64 and t.is_contiguous()
                    and t.numel() >= self.extent):
                raise WgfftError(_lib.WGFFT_ERR_INVALID_ARGUMENT,
                                 f"expected contiguous complex64 tensors of at least {self.extent} elements on the GPU")
        check(self._gpu._L.wgfft_ragged_exec_device(self._p, x.data_ptr(), out.data_ptr(),
                                                    WGFFT_INVERSE if inverse else WGFFT_FORWARD, int(mode),
                                                    GpuFft._stream(stream, x)))

    def describe(self) -> str:
        """Kernel choice of the plan (wgfft_ragged_describe)."""
        if not getattr(self, "_p", None):
            raise WgfftError(_lib.WGFFT_ERR_INVALID_ARGUMENT, "the mixed-length plan is closed")
        buf = ctypes.create_string_buffer(256)
        check(self._gpu._L.wgfft_ragged_describe(self._p, buf, 256))
        return buf.value.decode()

    def close(self) -> None:
        # a plan whose handle was closed was freed with it (_p already None)
        if getattr(self, "_p", None) and getattr(self._gpu, "_h", None):
            self._gpu._L.wgfft_ragged_destroy(self._p)
        self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ragged_layout(log2n) -> tuple[np.ndarray, int]:
    """Contiguous packing of lengths 2^log2n: (offsets, total samples)."""
    lens = (np.uint64(1) << np.asarray(log2n, dtype=np.uint64))
    offsets = np.zeros(len(lens), dtype=np.uint64)
    if len(lens) > 1:
        offsets[1:] = np.cumsum(lens)[:-1]
    return offsets, int(lens.sum())


def _check_tensor(t, shape=None) -> None:
    """A contiguous (batch, n) complex64 CUDA tensor (of `shape` if given):
    anything else never reaches a kernel as a pointer."""
    import torch

    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.complex64 and t.is_contiguous()
            and t.dim() == 2):
        raise WgfftError(_lib.WGFFT_ERR_INVALID_ARGUMENT,
                         "expected a contiguous (batch, n) complex64 tensor on the GPU")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise WgfftError(_lib.WGFFT_ERR_INVALID_ARGUMENT,
                         f"expected a tensor of shape {tuple(shape)}, got {tuple(t.shape)}")


def _check_real(t, n: int, what: str) -> None:
    """A contiguous float32 CUDA tensor of exactly n elements."""
    import torch

    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()
            and t.numel() == n):
        raise WgfftError(_lib.WGFFT_ERR_INVALID_ARGUMENT,
                         f"{what} must be a contiguous float32 tensor of {n} elements on the GPU")


def synth_uniform(shape: Iterable[int], seed: int) -> np.ndarray:
    """complex64 array, re/im i.i.d. U[-1,1) from the library's PCG32 stream."""
    shape = tuple(int(s) for s in shape)
    count = int(np.prod(shape))
    out = np.empty(2 * count, dtype=np.float32)
    _lib.load().wgfft_synth_uniform(out.ctypes.data, 2 * count, seed)
    return out.view(np.complex64).reshape(shape)


def set_alloc_cap(nbytes: int) -> None:
    _lib.load().wgfft_set_alloc_cap(int(nbytes))
