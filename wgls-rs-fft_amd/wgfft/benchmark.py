"""Timing and validation of any ``FftExecutor`` (the reference's
``src/benchmark.rs:37-297`` and its leaderboard,
``examples/rivalry_leaderboard.rs:58-177``), for this library and for
comparators run outside it.

* ``deterministic_inputs(n, b)``: every signal is (t * 0.001,
  sin(2 pi t) * 0.001), t = i / N in f32 (``src/benchmark.rs:58-69``).
* ``validate``: max complex-norm error of two batches, a pass at <= 1e-3;
  a shape mismatch fails with an infinite error (``:272-297``).
* ``validate_rival`` / ``benchmark_rival`` / ``sweep_rival``: one forward
  FFT compared with a reference executor; 1 warm-up and 10 timed calls;
  MSamples/s and GFLOPS = 5 B N log2 N / t (``:155-270``).
* ``benchmark_gpu_pipeline``: wall clock of whole ``fft`` calls (host to
  device, transform, device to host) (``:106-149``); ``benchmark_gpu_only``:
  resident buffers, 3 warm-up and 50 timed calls between HIP events
  (``:52-100``, through ``wgfft_bench_device``).
* ``choose_batch_size``: the leaderboard's batch per length.
* ``RocfftExecutor``: torch.fft (hipFFT -> rocFFT) as an ``FftExecutor``, a
  comparator kept outside the library (libwgfft.so never calls it).
* ``leaderboard()``: the reference leaderboard's table for the executors
  given, validated against this library's GpuFft (``python -m
  wgfft.benchmark``).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from .gpu_fft import FftExecutor

WARMUP_ITERS = 1
BENCH_ITERS = 10
GPU_WARMUP_ITERS = 3
GPU_BENCH_ITERS = 50
VALIDATION_TOLERANCE = 1e-3
MAX_TOTAL_SAMPLES = 16 * 1024 * 1024
LEADERBOARD_SIZES = (256, 1024, 16384, 65536, 1048576)


@dataclass
class ValidationOutcome:
    """Pass, or Fail with the max complex-norm error (inf on a failed call or
    a shape mismatch)."""
    passed: bool
    max_error: float

    def status(self) -> str:
        return "PASS" if self.passed else f"FAIL({self.max_error:.2e})"


@dataclass
class BenchmarkResult:
    rival_name: str
    n: int
    batch_size: int
    msamples_per_sec: float
    gflops: float
    validation: ValidationOutcome


@dataclass
class GpuOnlyResult:
    rival_name: str
    n: int
    batch_size: int
    gpu_duration_sec: float
    gpu_msamples_per_sec: float
    gpu_gflops: float


def deterministic_inputs(n: int, batch_size: int) -> List[np.ndarray]:
    """The benchmark signal, computed in f32 as the reference does."""
    t = np.arange(n, dtype=np.float32) / np.float32(n)
    re = t * np.float32(0.001)
    im = np.sin(t * np.float32(2.0 * math.pi)).astype(np.float32) * np.float32(0.001)
    one = (re + 1j * im).astype(np.complex64)
    return [one.copy() for _ in range(batch_size)]


def rates(n: int, batch_size: int, seconds: float) -> tuple:
    """(MSamples/s, GFLOPS) with the 5 N log2 N flop model."""
    total = float(n * batch_size)
    return total / seconds / 1e6, 5.0 * total * math.log2(n) / seconds / 1e9


def validate(result: Sequence, reference: Sequence) -> ValidationOutcome:
    if len(result) != len(reference):
        return ValidationOutcome(False, math.inf)
    max_err = 0.0
    for r, w in zip(result, reference):
        r = np.asarray(r, dtype=np.complex64)
        w = np.asarray(w, dtype=np.complex64)
        if r.shape != w.shape:
            return ValidationOutcome(False, math.inf)
        if r.size:
            max_err = max(max_err, float(np.abs(r - w).max()))
    return ValidationOutcome(max_err <= VALIDATION_TOLERANCE, max_err)


def validate_rival(rival: FftExecutor, reference: FftExecutor, n: int, batch_size: int) -> ValidationOutcome:
    inputs = deterministic_inputs(n, batch_size)
    try:
        got = rival.fft(inputs)
        want = reference.fft(inputs)
    except Exception:
        return ValidationOutcome(False, math.inf)
    return validate(got, want)


def benchmark_rival(rival: FftExecutor, reference: FftExecutor, n: int, batch_size: int) -> BenchmarkResult:
    inputs = deterministic_inputs(n, batch_size)
    for _ in range(WARMUP_ITERS):
        rival.fft(inputs)
    t0 = time.perf_counter()
    for _ in range(BENCH_ITERS):
        rival.fft(inputs)
    sec = (time.perf_counter() - t0) / BENCH_ITERS
    ms, gf = rates(n, batch_size, sec)
    return BenchmarkResult(rival.name(), n, batch_size, ms, gf, validate(rival.fft(inputs), reference.fft(inputs)))


def sweep_rival(rival: FftExecutor, reference: FftExecutor, n: int, batch_sizes: Sequence[int]) -> BenchmarkResult:
    best: Optional[BenchmarkResult] = None
    for b in batch_sizes:
        if n * b > MAX_TOTAL_SAMPLES:
            continue
        r = benchmark_rival(rival, reference, n, b)
        if best is None or r.msamples_per_sec > best.msamples_per_sec:
            best = r
    return best if best is not None else benchmark_rival(rival, reference, n, 1)


def benchmark_gpu_pipeline(rival: FftExecutor, n: int, batch_size: int) -> GpuOnlyResult:
    inputs = deterministic_inputs(n, batch_size)
    for _ in range(WARMUP_ITERS):
        rival.fft(inputs)
    t0 = time.perf_counter()
    for _ in range(BENCH_ITERS):
        rival.fft(inputs)
    sec = (time.perf_counter() - t0) / BENCH_ITERS
    ms, gf = rates(n, batch_size, sec)
    return GpuOnlyResult(rival.name(), n, batch_size, sec, ms, gf)


def benchmark_gpu_only(gpu, n: int, batch_size: int) -> GpuOnlyResult:
    """This library's executor on resident device buffers (upload not timed)."""
    import torch

    x = torch.from_numpy(np.stack(deterministic_inputs(n, batch_size))).cuda()
    y = torch.empty_like(x)
    sec = gpu.bench_device(x, y, warmup=GPU_WARMUP_ITERS, iters=GPU_BENCH_ITERS)
    ms, gf = rates(n, batch_size, sec)
    return GpuOnlyResult(gpu.name(), n, batch_size, sec, ms, gf)


def choose_batch_size(n: int) -> int:
    if n <= 1024:
        batch = 1024
    elif n <= 16384:
        batch = 256
    elif n <= 65536:
        batch = 64
    elif n <= 262144:
        batch = 16
    else:
        batch = 1
    while n * batch > MAX_TOTAL_SAMPLES and batch > 1:
        batch //= 2
    return max(batch, 1)


class RocfftExecutor(FftExecutor):
    """torch.fft.fft / ifft (hipFFT -> rocFFT on ROCm) over the same list
    interface: host arrays in, host arrays out, 1/N-scaled inverse.  A
    comparator only; the library never calls it."""

    def __init__(self):
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError("rocFFT comparator needs a GPU")
        self._torch = torch

    def name(self) -> str:
        return "rocFFT (torch.fft, hipFFT)"

    def _run(self, inputs, inverse: bool):
        if len(inputs) == 0:
            return []
        torch = self._torch
        x = torch.from_numpy(np.stack([np.asarray(v, dtype=np.complex64) for v in inputs])).cuda()
        y = torch.fft.ifft(x, dim=1) if inverse else torch.fft.fft(x, dim=1)
        return list(y.cpu().numpy())

    def fft(self, inputs):
        return self._run(inputs, False)

    def ifft(self, inputs):
        return self._run(inputs, True)


def leaderboard(rivals: Sequence[FftExecutor], reference: FftExecutor, sizes: Sequence[int] = LEADERBOARD_SIZES,
                out=print) -> List[dict]:
    """The reference leaderboard (``examples/rivalry_leaderboard.rs:107-150``):
    per size, the whole-pipeline rate of each rival at the tier's batch and
    its validation against `reference`."""
    rows = []
    for n in sizes:
        b = choose_batch_size(n)
        out(f"\n--- N = {n} ---")
        out(f"{'Implementation':>32} | {'Batch':>8} | {'MSamples/s':>14} | {'GFLOPS':>10} | {'Status':>8}")
        out("-" * 84)
        for r in rivals:
            try:
                g = benchmark_gpu_pipeline(r, n, b)
                v = validate_rival(r, reference, n, b)
                out(f"{g.rival_name:>32} | {b:>8} | {g.gpu_msamples_per_sec:>14.2f} | {g.gpu_gflops:>10.2f} | "
                    f"{v.status():>8}")
                rows.append({"rival": g.rival_name, "n": n, "batch": b, "msamples_s": g.gpu_msamples_per_sec,
                             "gflops": g.gpu_gflops, "status": v.status(), "max_error": v.max_error})
            except Exception as e:  # a rival that fails keeps the table going, as the reference's does
                out(f"{r.name():>32} | {b:>8} | {'ERROR':>14} | {'-':>10} | {e}")
                rows.append({"rival": r.name(), "n": n, "batch": b, "error": str(e)})
    return rows


def main() -> int:
    from .gpu_fft import GpuFft

    gpu = GpuFft.new()
    rivals: List[FftExecutor] = [gpu]
    try:
        rivals.append(RocfftExecutor())
    except Exception as e:
        print(f"rocFFT comparator unavailable: {e}")
    print("\n=== wgls-rs-fft (MI355X) leaderboard: host to device + transform + device to host ===")
    leaderboard(rivals, gpu)
    gpu.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
