#!/usr/bin/env python3
"""Headline benchmark: batched complex64 forward FFT, N = 4096 x 4096.

One step = one forward transform of 4096 signals of length 4096 (16.8 M
samples, 128 MiB in + 128 MiB out) with inputs resident in HBM, launched
through the C-ABI entry wgfft_exec_device on torch's current stream and
timed with HIP events on that stream.

The same JSON line also carries, measured in the same run:
  * roofline: algorithmic bytes (16 B/sample) / kernel time vs 8 TB/s, and
    the bandwidth of a plain device copy of the same 256 MiB;
  * fwd_adjoint: the fused forward + adjoint of L = sum m |FFT(w x)|^2 on the
    same shape and its time relative to the plain forward;
  * spectral_filter: y = IFFT(FFT(x) * s) on the same shape in one fused
    pass, against the plain forward and against the unfused forward, spectral
    multiply and inverse (the reference pipeline's stages), and with the
    pipeline's noise stage fused at the load;
  * other_shapes: the same 16.8 M samples as 256 x 65536 and 16384 x 1024;
  * mixed_batch: 20,000 transforms of lengths 2^k, k ~ U{6..16}, as one
    persistent launch, its hipGraph replay, and one launch per length; its
    traffic floor (the scratch of the two-pass lengths included) as a rate
    beside a device copy over the same buffers;
  * large_n: N = 2^22 x 16 and N = 2^26 x 1 forward transforms; the passes'
    combined traffic as a rate beside a device copy of the same buffers;
  * c5_windowed: N = 8192 x 1024 with a real window: the windowed forward,
    the VJP (gx, gw from gY) and the fused masked forward + adjoint, each
    against the plain forward of the same shape;
  * c1_and_host_api: N=1024 x 8 device latency, and the host-buffer API
    (H2D + transform + D2H, PCIe-inclusive) at 1024 x 8 and 4096 x 4096;
  * rocfft: the same shapes through torch.fft (hipFFT/rocFFT), timed the same
    way, as an external comparator outside the library path;
  * headline_rotating: the headline transform over four buffer pairs (1 GiB
    working set, beyond the memory-side cache);
  * gpu_only_benchmark: the reference's benchmark_gpu_only protocol through
    wgfft_bench_device (HIP events on the library handle's stream);
  * cpu_baseline: the reference algorithm (the oracle's C restatement of the
    wgls-rs-fft f32 Stockham R4/R2 path) on the host's CPU share, bounded
    sample, median of 5; cpu_standin_pocketfft: scipy.fft (pocketfft) as a
    rustfft-class stand-in; reference_cpu_path: the probe for cargo, its
    crates and a Vulkan ICD (the reference's own CPU path).
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "wgls-rs-fft_amd"))

N, BATCH = 4096, 4096
HBM_PEAK_GBS = 8000.0
METRIC = "complex64 GSamples/s & %HBM-roofline, N=4096×4096; fwd+adjoint vs plain fwd"


def _host_cpu() -> dict:
    """The host's CPU share and model.  The threads used are the CPUs this
    process may run on, capped by OMP_NUM_THREADS when set (a GPU box declares
    its per-GPU CPU share there), so a CPU baseline never oversubscribes the
    share the box gives one GPU."""
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS")
    cores = affinity
    if omp and omp.isdigit() and int(omp) > 0:
        cores = min(cores, int(omp))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cores": max(1, cores), "nproc": os.cpu_count(), "affinity": affinity, "omp_num_threads": omp,
            "model": model}


def _median_rate(fn, samples: int, runs: int = 5) -> tuple:
    """1 warm-up call, then the median of `runs` timed calls: (samples/s, s)."""
    fn()
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    times.sort()
    med = times[len(times) // 2]
    return samples / med, med


def reference_cpu_probe() -> dict:
    """The reference's own CPU path (rustfft 6.4.1 and the wgpu fallback
    adapter) needs cargo, its crates and a Vulkan ICD: probed, never fetched."""
    import shutil
    import subprocess

    out = {}
    cargo = shutil.which("cargo")
    out["cargo"] = "unavailable offline" if not cargo else subprocess.run(
        [cargo, "--version"], capture_output=True, text=True).stdout.strip()
    reg = os.path.expanduser("~/.cargo/registry")
    out["cargo_registry"] = reg if os.path.isdir(reg) else "unavailable offline"
    vk = shutil.which("vulkaninfo")
    icd = [d for d in ("/usr/share/vulkan/icd.d", "/etc/vulkan/icd.d") if os.path.isdir(d) and os.listdir(d)]
    out["vulkan_icd"] = ",".join(icd) if icd else "unavailable offline"
    out["vulkaninfo"] = vk or "unavailable offline"
    out["ran"] = bool(cargo and os.path.isdir(reg) and icd)
    return out


def cpu_baseline() -> dict:
    """The reference algorithm (f32 Stockham R4/R2, the oracle's C restatement
    of the wgpu path) on the host's CPU share: the same seeded N=4096 inputs,
    1 warm-up and the median of 5 timed runs over a bounded sample, transform
    time only."""
    sys.path.insert(0, ROOT)
    import oracle

    host = _host_cpu()
    cores = host["cores"]
    x = oracle.synth_uniform((2048, N), 0x5EED0002)
    rate, med = _median_rate(lambda: oracle.ref_fft_f32(x, threads=cores), x.shape[0] * N)
    return {"value": round(rate / 1e9, 4), "unit": "GSamples/s", "cores": cores, "kind": "port",
            "sample": f"{x.shape[0]} transforms of N={N} per run, median of 5 runs after 1 warm-up "
                      f"({med * 1e3:.0f} ms per run; f32 Stockham R4/R2 restatement of the reference path, "
                      f"one pthread per core)",
            "host": host}


def cpu_standin_pocketfft() -> dict:
    """A rustfft-class optimised CPU FFT as a stand-in for rustfft 6.4.1, which
    is not installable offline: scipy.fft (pocketfft) in complex64 on the same
    inputs and threads, same protocol.  Labelled a stand-in, not the
    reference."""
    try:
        import scipy.fft as sfft
    except ImportError:
        return {"available": False}
    import numpy as np

    host = _host_cpu()
    x = np.asarray(__import__("wgfft").synth_uniform((4096, N), 0x5EED0002))
    rate, med = _median_rate(lambda: sfft.fft(x, axis=-1, workers=host["cores"]), x.shape[0] * N)
    return {"value": round(rate / 1e9, 4), "unit": "GSamples/s", "cores": host["cores"],
            "kind": "stand-in (scipy.fft pocketfft, complex64) for rustfft",
            "sample": f"{x.shape[0]} transforms of N={N} per run, median of 5 runs after 1 warm-up "
                      f"({med * 1e3:.0f} ms per run)"}


SETTLE_S = 0.25
# calls timed per secondary measurement whatever --steps is: their kernels
# take 20-100 us, so fewer calls would time well under a millisecond
SECONDARY_STEPS = 200


def settle(torch, fn, seconds: float = SETTLE_S) -> None:
    """Run fn() back to back for `seconds` of wall time (untimed).

    The GPU's shader clock ramps up over ~20 ms of continuous load after any
    idle gap (host-side set-up, autotuning, synchronisation); a short warmup
    leaves the timed steps on that ramp."""
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(50):
            fn()
        torch.cuda.synchronize()


def timed(torch, fn, steps: int, warmup: int) -> float:
    """Average ms per call of fn() with HIP events on the current stream."""
    stream = torch.cuda.current_stream()
    settle(torch, fn)
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def bench_rotating(torch, gpu, x, steps: int, pairs: int = 4) -> dict:
    """The forward transform of x's shape cycling through `pairs` distinct
    input/output buffer pairs (at 16.8 M samples a working set of 4 x 256 MiB,
    four times the 256 MB memory-side cache), so no step re-reads what the
    previous one left in that cache; beside it, a device copy over the same
    rotation."""
    ins = [x] + [torch.empty_like(x).copy_(x) for _ in range(pairs - 1)]
    outs = [torch.empty_like(x) for _ in range(pairs)]
    state = {"i": 0}

    def step():
        i = state["i"]
        gpu.exec_device(ins[i], outs[i])
        state["i"] = (i + 1) % pairs

    ms = timed(torch, step, max(pairs * 10, steps - steps % pairs), 2 * pairs)
    samples = x.shape[0] * x.shape[1]
    gbs = 16.0 * samples / (ms * 1e-3) / 1e9

    def copy_step():  # the same rotation through a plain device copy: the practical streaming rate
        i = state["i"]
        outs[i].copy_(ins[i])
        state["i"] = (i + 1) % pairs

    copy_ms = timed(torch, copy_step, 10 * pairs, 2 * pairs)
    copy_gbs = 16.0 * samples / (copy_ms * 1e-3) / 1e9
    del ins, outs
    torch.cuda.empty_cache()
    return {"ms": round(ms, 5), "gsamples_s": round(samples / (ms * 1e-3) / 1e9, 2), "achieved_gbs": round(gbs, 1),
            "frac": round(gbs / HBM_PEAK_GBS, 4), "working_set_mib": pairs * 2 * x.numel() * 8 >> 20,
            "copy_gbs": round(copy_gbs, 1), "frac_of_copy": round(gbs / copy_gbs, 4)}


def bench_mixed(torch, gpu, wgfft, steps: int) -> dict:
    k = _ragged_log2(20000, 6, 16, 0x5EED0003)
    offsets, total = wgfft.ragged_layout(k)
    plan = wgfft.RaggedPlan(gpu, offsets, k)
    x = (torch.rand(2 * total, device="cuda") * 2 - 1).view(torch.complex64)
    y = torch.empty_like(x)
    out = {"transforms": 20000, "samples": int(total), "bytes": 16 * int(total)}
    for mode, name in ((0, "persistent"), (1, "graph"), (2, "per_length")):
        ms = timed(torch, lambda: plan.exec_device(x, y, mode=mode), max(20, steps // 10), 3)
        gbs = 16.0 * total / (ms * 1e-3) / 1e9
        out[name] = {"ms": round(ms, 4), "gsamples_s": round(total / (ms * 1e-3) / 1e9, 2),
                     "achieved_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    out["kernel"] = plan.describe()
    plan.close()
    # the traffic floor of this design set beside a device copy over the
    # same two buffers
    floor = traffic_floor_bytes(k)
    copy_ms = timed(torch, lambda: y.copy_(x), 20, 3)
    copy_gbs = 16.0 * total / (copy_ms * 1e-3) / 1e9
    floor_gbs = floor / (out["persistent"]["ms"] * 1e-3) / 1e9
    out["traffic_floor"] = {"bytes": floor, "bytes_per_sample": round(floor / total, 2),
                            "gbs_at_persistent": round(floor_gbs, 1), "copy_gbs": round(copy_gbs, 1),
                            "frac_of_copy": round(floor_gbs / copy_gbs, 4)}
    # the copy alone, as a fraction of the same 16 B/sample roofline: the
    # bound of any kernel that reads and writes these two buffers once
    out["copy_same_buffers"] = {"ms": round(copy_ms, 4), "frac": round(copy_gbs / HBM_PEAK_GBS, 4)}
    del x, y
    return out


def traffic_floor_bytes(log2n) -> int:
    """HBM bytes one launch of the mixed-length kernel moves at least: 16 B
    per sample, plus 16 B per sample of scratch written and re-read for the
    lengths above 2^14, which run as two passes through a per-workgroup slot
    (the lengths up to 2^14 stay in the LDS)."""
    return int(sum((32 if int(k) > 14 else 16) << int(k) for k in log2n))


def pass_count(plan: str) -> int:
    """Global passes of a multipass plan description ("... passes=R2048(...),R2048(...)"): each pass
    reads and writes every sample once."""
    return len(re.findall(r"R\d+\(", plan)) or 1


def _ragged_log2(count, kmin, kmax, seed):
    """log2 lengths kmin + (u32 >> 8) % span, one PCG32 draw per transform
    (the draws of the parity tests), recovered exactly from the library's
    uniform generator, which emits (u32 >> 8) * 2^-23 - 1 per draw."""
    import numpy as np

    import wgfft

    raw = np.empty(count, dtype=np.float32)
    wgfft._lib.load().wgfft_synth_uniform(raw.ctypes.data, count, seed)
    v = np.round((raw.astype(np.float64) + 1.0) * 8388608.0).astype(np.int64)
    return (kmin + v % (kmax - kmin + 1)).astype(np.uint32)


def bench_shapes(torch, gpu, steps: int, tune: bool) -> dict:
    """The same 16.8 M samples as 256 x 65536 and 16384 x 1024, on one
    buffer pair and rotating through four."""
    out = {}
    for n, batch in ((256, 65536), (16384, 1024)):
        x = (torch.rand(2 * n * batch, device="cuda") * 2 - 1).view(torch.complex64).view(batch, n)
        y = torch.empty_like(x)
        if tune:
            gpu.autotune(n, batch)
        ms = timed(torch, lambda: gpu.exec_device(x, y), max(SECONDARY_STEPS, steps // 4), 5)
        s = n * batch
        gbs = 16.0 * s / (ms * 1e-3) / 1e9
        out[f"{n}x{batch}"] = {"ms": round(ms, 4), "gsamples_s": round(s / (ms * 1e-3) / 1e9, 2),
                               "achieved_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
                               "plan": gpu.describe(n), "rotating": bench_rotating(torch, gpu, x, steps)}
        del x, y
    return out


def bench_c1_and_host(torch, gpu, wgfft) -> dict:
    """C1 (N=1024 x 8, launch/latency-bound: device us per forward and
    inverse call) and the host-buffer API (H2D + transform + D2H,
    synchronous, PCIe-inclusive) at C1 and at 4096 x 4096: on one flat array
    (wgfft_exec_host) and on a list of separate signals, as GpuFft::fft takes
    them (wgfft_exec_host_rows), each call returning new arrays."""
    out = {}
    x = torch.from_numpy(wgfft.synth_uniform((8, 1024), 0x5EED0001)).cuda()
    y = torch.empty_like(x)
    # back-to-back calls from Python (the wrapper's checks and ctypes call are
    # ~8 us of host time per call, which bounds this figure) and from C++
    # (wgfft_bench_device: calls in a loop between HIP events, the device's
    # own rate for this launch)
    out["c1_device_us"] = {"fwd": round(1000 * timed(torch, lambda: gpu.exec_device(x, y), 2000, 100), 2),
                           "inv": round(1000 * timed(torch, lambda: gpu.exec_device(y, x, inverse=True), 2000, 100), 2),
                           "via": "python exec_device"}
    out["c1_bench_device_us"] = {"fwd": round(1e6 * gpu.bench_device(x, y, warmup=100, iters=2000), 2),
                                 "inv": round(1e6 * gpu.bench_device(y, x, inverse=True, warmup=100, iters=2000), 2),
                                 "via": "wgfft_bench_device"}
    for (n, batch, reps) in ((1024, 8, 201), (4096, 4096, 9)):
        xh = wgfft.synth_uniform((batch, n), 0x5EED0002)
        rows = [xh[b].copy() for b in range(batch)]  # separate buffers, like Vec<Vec<Complex<f32>>>
        for key, call in (("host_api", lambda: gpu.fft_array(xh)), ("host_list", lambda: gpu.fft(rows))):
            call()  # warm (plan, staging)
            call()
            ts = []
            for _ in range(reps):  # the median call (host jitter spreads the tail)
                t0 = time.perf_counter()
                call()
                ts.append(time.perf_counter() - t0)
            dt = sorted(ts)[reps // 2]
            out[f"{key}_{n}x{batch}"] = {"ms": round(dt * 1e3, 4), "gsamples_s": round(n * batch / dt / 1e9, 3)}
    return out


def bench_c5(torch, gpu, wgfft, steps: int, tune: bool) -> dict:
    """N = 8192 x 1024 with a window (Hann): windowed forward, VJP and fused
    masked forward + adjoint; algorithmic bytes 16 B/sample for the forward
    transforms, ratios against the plain forward of the same shape."""
    import numpy as np

    n, batch = 8192, 1024
    x = (torch.rand(2 * n * batch, device="cuda") * 2 - 1).view(torch.complex64).view(batch, n)
    y = torch.empty_like(x)
    gy = (torch.rand(2 * n * batch, device="cuda") * 2 - 1).view(torch.complex64).view(batch, n)
    gx = torch.empty_like(x)
    w = torch.from_numpy(np.hanning(n).astype(np.float32)).cuda()
    m = (torch.rand(n, device="cuda") < 0.3).float()
    gw = torch.empty(n, dtype=torch.float32, device="cuda")
    loss = torch.empty(1, dtype=torch.float32, device="cuda")
    if tune:
        gpu.autotune(n, batch)
    k = max(SECONDARY_STEPS, steps // 4)
    fwd = timed(torch, lambda: gpu.exec_device(x, y), k, 5)
    win = timed(torch, lambda: gpu.exec_windowed_device(x, w, y), k, 5)
    vjp = timed(torch, lambda: gpu.vjp_windowed_device(x, w, gy, gx, gw), k, 5)
    fused = timed(torch, lambda: gpu.fwd_adjoint_masked_device(x, w, m, None, gx, gw, loss), k, 5)
    gbs = 16.0 * n * batch / (win * 1e-3) / 1e9
    return {"fwd_ms": round(fwd, 4), "windowed_fwd_ms": round(win, 4), "windowed_frac": round(gbs / HBM_PEAK_GBS, 4),
            "vjp_ms": round(vjp, 4), "vjp_ratio": round(vjp / fwd, 3),
            "fused_fwd_adjoint_ms": round(fused, 4), "fused_ratio": round(fused / fwd, 3), "plan": gpu.describe(n)}


def bench_fused(torch, gpu, x, y, kern_ms: float, steps: int, warmup: int) -> dict:
    """On the headline shape: the fused masked forward + adjoint and the
    spectral filter IFFT(FFT(x) * s) (with and without the pipeline's noise at
    its load, and as separate forward, multiply and inverse launches), as
    ratios to the plain forward time kern_ms."""
    import numpy as np

    n = x.shape[1]
    w = torch.from_numpy(np.hanning(n).astype(np.float32)).cuda()
    m = (torch.rand(n, device="cuda") < 0.3).float()
    gx = torch.empty_like(x)
    gw = torch.empty(n, dtype=torch.float32, device="cuda")
    loss = torch.empty(1, dtype=torch.float32, device="cuda")
    k, kw = max(SECONDARY_STEPS, steps // 4), max(3, warmup // 4)
    fa_ms = timed(torch, lambda: gpu.fwd_adjoint_masked_device(x, w, m, None, gx, gw, loss), k, kw)
    out = {"fwd_adjoint": {"ms": round(fa_ms, 5), "plain_fwd_ms": round(kern_ms, 5), "ratio": round(fa_ms / kern_ms, 3)}}
    sp = torch.from_numpy(np.hanning(n + 1)[:n].astype(np.float32)).cuda()
    filt_ms = timed(torch, lambda: gpu.filter_device(x, sp, y), k, kw)

    def unfused():
        gpu.exec_device(x, y)
        y.mul_(sp)
        gpu.exec_device(y, y, inverse=True)

    unf_ms = timed(torch, unfused, k, kw)
    noisy_ms = timed(torch, lambda: gpu.filter_device(x, sp, y, noise_sigma=0.1, seed=1), k, kw)
    out["spectral_filter"] = {"ms": round(filt_ms, 5), "ratio_to_fwd": round(filt_ms / kern_ms, 3),
                              "unfused_ms": round(unf_ms, 5), "speedup_vs_unfused": round(unf_ms / filt_ms, 3),
                              "with_noise_ms": round(noisy_ms, 5)}
    return out


def bench_rocfft(torch, gpu, steps: int) -> dict:
    """torch.fft.fft (hipFFT -> rocFFT on ROCm) on the bench shapes, out of
    place, against this library's default plan on the same tensors."""
    out = {}
    for n, batch in ((4096, 4096), (256, 65536), (16384, 1024), (8192, 1024), (1 << 22, 16)):
        x = (torch.rand(2 * n * batch, device="cuda") * 2 - 1).view(torch.complex64).view(batch, n)
        y = torch.empty_like(x)
        k = max(50, steps // 10)
        ours = timed(torch, lambda: gpu.exec_device(x, y), k, 5)
        theirs = timed(torch, lambda: torch.fft.fft(x, dim=1), k, 5)
        out[f"{n}x{batch}"] = {"rocfft_ms": round(theirs, 4), "wgfft_ms": round(ours, 4),
                               "speedup": round(theirs / ours, 3)}
        del x, y
        torch.cuda.empty_cache()
    return out


def bench_strided(torch, gpu, steps: int) -> dict:
    """Strided layouts on the headline shape (4096 x 4096): rows padded to
    4160 complex (unit element stride, signal distance 4160: one pass) and the
    column layout of a (4096 x 4096) matrix (element stride 4096, distance
    1: gathered, transformed and scattered), against the plain forward of the
    same data; algorithmic bytes 16 B/sample for each."""
    n, batch, pad = N, BATCH, N + 64
    k = max(SECONDARY_STEPS, steps // 4)
    x = torch.from_numpy(synth_rows(batch, n)).cuda()
    y = torch.empty_like(x)
    plain = timed(torch, lambda: gpu.exec_device(x, y), k, 5)
    xp = torch.empty((batch, pad), dtype=torch.complex64, device="cuda")
    yp = torch.empty_like(xp)
    xp[:, :n].copy_(x)
    padded = timed(torch, lambda: gpu.exec_strided_device(xp[:, :n], yp[:, :n]), k, 5)
    xc = x.t().contiguous().t()  # element stride 4096, distance 1
    yc = torch.empty((n, batch), dtype=torch.complex64, device="cuda").t()
    column = timed(torch, lambda: gpu.exec_strided_device(xc, yc), k, 5)
    out = {"plain_ms": round(plain, 4)}
    for name, ms in (("padded_rows", padded), ("column_layout", column)):
        gbs = 16.0 * n * batch / (ms * 1e-3) / 1e9
        out[name] = {"ms": round(ms, 4), "frac": round(gbs / HBM_PEAK_GBS, 4), "ratio_to_plain": round(ms / plain, 3)}
    out["padded_rows"]["layout"] = f"distance {pad}, element stride 1 (one pass)"
    out["column_layout"]["layout"] = f"element stride {batch}, distance 1 (gather, transform, scatter)"
    return out


def synth_rows(batch: int, n: int):
    import wgfft

    return wgfft.synth_uniform((batch, n), 0x5EED0002)


def bench_large(torch, gpu, steps: int, tune: bool) -> dict:
    out = {}
    for log2n, batch in ((22, 16), (26, 1)):
        n = 1 << log2n
        x = (torch.rand(2 * n * batch, device="cuda") * 2 - 1).view(torch.complex64).view(batch, n)
        y = torch.empty_like(x)
        if tune:
            gpu.autotune(n, batch)
        ms = timed(torch, lambda: gpu.exec_device(x, y), max(20, steps // 20), 2)
        ms_inv = timed(torch, lambda: gpu.exec_device(y, x, inverse=True), max(20, steps // 20), 2)
        s = n * batch
        gbs = 16.0 * s / (ms * 1e-3) / 1e9
        plan = gpu.describe(n)
        # the passes together move 16 B/sample per pass, set beside a device
        # copy of x
        passes = pass_count(plan)
        copy_ms = timed(torch, lambda: y.copy_(x), 20, 2)
        copy_gbs = 16.0 * s / (copy_ms * 1e-3) / 1e9
        pass_gbs = passes * gbs
        out[f"2^{log2n}x{batch}"] = {"fwd_ms": round(ms, 4), "inv_ms": round(ms_inv, 4),
                                     "gsamples_s": round(s / (ms * 1e-3) / 1e9, 2),
                                     "achieved_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
                                     "plan": plan, "passes": passes, "pass_traffic_gbs": round(pass_gbs, 1),
                                     "copy_gbs": round(copy_gbs, 1), "frac_of_copy": round(pass_gbs / copy_gbs, 4)}
        del x, y
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--headline-only", action="store_true", help="only the N=4096x4096 forward (profiling)")
    ap.add_argument("--only", choices=("shapes", "mixed", "large", "c1", "c5", "rocfft", "fused", "strided"),
                    help="profiling: skip the headline and run one secondary section")
    ap.add_argument("--no-tune", action="store_true", help="use the default kernel variants")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import wgfft

    gpu = wgfft.GpuFft.new()
    if args.only == "fused":  # the headline shape's fused kernels, timed against its default forward
        if not args.no_tune:
            gpu.autotune(N, BATCH)
        x = torch.from_numpy(wgfft.synth_uniform((BATCH, N), 0x5EED0002)).cuda()
        y = torch.empty_like(x)
        kern_ms = timed(torch, lambda: gpu.exec_device(x, y), args.steps, args.warmup)
        print(json.dumps(bench_fused(torch, gpu, x, y, kern_ms, args.steps, args.warmup)))
        gpu.close()
        return 0
    if args.only:
        section = {"shapes": lambda: bench_shapes(torch, gpu, args.steps, not args.no_tune),
                   "mixed": lambda: bench_mixed(torch, gpu, wgfft, args.steps),
                   "large": lambda: bench_large(torch, gpu, args.steps, not args.no_tune),
                   "c1": lambda: bench_c1_and_host(torch, gpu, wgfft),
                   "c5": lambda: bench_c5(torch, gpu, wgfft, args.steps, not args.no_tune),
                   "rocfft": lambda: bench_rocfft(torch, gpu, args.steps),
                   "strided": lambda: bench_strided(torch, gpu, args.steps)}[args.only]
        print(json.dumps({args.only: section()}))
        gpu.close()
        return 0
    tuned = None if args.no_tune else gpu.autotune(N, BATCH)
    x = torch.from_numpy(wgfft.synth_uniform((BATCH, N), 0x5EED0002)).cuda()
    y = torch.empty_like(x)

    kern_ms = timed(torch, lambda: gpu.exec_device(x, y), args.steps, args.warmup)
    samples = N * BATCH
    ms_per_step = kern_ms
    value = samples / (ms_per_step * 1e-3) / 1e9
    achieved = 16.0 * samples / (kern_ms * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GSamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex64 (f32)",
        "data": "synthetic: re/im i.i.d. U[-1,1) from SplitMix64->PCG32, resident in HBM",
        "settle_s": SETTLE_S,
        "config": {"workload": "batched forward FFT, N=4096, batch 4096", "n": N, "batch": BATCH,
                   "samples": samples, "plan": gpu.describe(N), "autotuned": tuned},
    }
    # HBM bytes per launch of the headline kernel: not measurable inside this
    # process (counters need rocprofv3 --pmc); read from the committed
    # FETCH_SIZE / WRITE_SIZE runs of this kernel and workload, each scaled by
    # what the counter reads for a copy kernel of known size
    traffic = None
    traffic_src = "profiles/pmc_fft4096x4096.json (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; copy-kernel scaled)"
    prof = os.path.join(ROOT, "profiles", "pmc_fft4096x4096.json")
    if os.path.exists(prof):
        with open(prof) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    copy_ms = timed(torch, lambda: y.copy_(x), 50, 5)
    out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                       "traffic_source": traffic_src if traffic is not None else None,
                       "copy_gbs": round(2 * x.numel() * 8 / (copy_ms * 1e-3) / 1e9, 1)}
    # the measurement protocol's repeat statistics: five more timed regions of
    # the same workload (at least 200 calls each), median and spread
    reps = [timed(torch, lambda: gpu.exec_device(x, y), max(SECONDARY_STEPS, args.steps // 4), 5) for _ in range(5)]
    reps_sorted = sorted(reps)
    out["headline_repeats"] = {"ms": [round(r, 5) for r in reps], "median_ms": round(reps_sorted[2], 5),
                               "median_gsamples_s": round(samples / (reps_sorted[2] * 1e-3) / 1e9, 2),
                               "spread": round((reps_sorted[-1] - reps_sorted[0]) / reps_sorted[2], 4)}
    out["headline_rotating"] = bench_rotating(torch, gpu, x, args.steps)
    # GpuFftTrait::benchmark_gpu_only semantics (src/lib.rs:109-143): 3 warm-up
    # calls with waits, then 50 back-to-back calls between HIP events on the
    # library handle's own stream
    sec = gpu.bench_device(x, y, warmup=3, iters=50)
    out["gpu_only_benchmark"] = {"ms": round(sec * 1e3, 5), "gsamples_s": round(samples / sec / 1e9, 2),
                                 "gflops": round(5.0 * samples * 12 / sec / 1e9, 1),
                                 "via": "wgfft_bench_device"}

    if not args.headline_only:
        out.update(bench_fused(torch, gpu, x, y, kern_ms, args.steps, args.warmup))
        x_keep, y_keep = x, y
        del x, y
        out["other_shapes"] = bench_shapes(torch, gpu, args.steps, not args.no_tune)
        out["mixed_batch"] = bench_mixed(torch, gpu, wgfft, args.steps)
        out["large_n"] = bench_large(torch, gpu, args.steps, not args.no_tune)
        out["c5_windowed"] = bench_c5(torch, gpu, wgfft, args.steps, not args.no_tune)
        out["c1_and_host_api"] = bench_c1_and_host(torch, gpu, wgfft)
        out["rocfft"] = bench_rocfft(torch, gpu, args.steps)
        out["strided_layouts"] = bench_strided(torch, gpu, args.steps)
        del x_keep, y_keep
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
        out["cpu_standin_pocketfft"] = cpu_standin_pocketfft()
        out["reference_cpu_path"] = reference_cpu_probe()
    print(json.dumps(out))
    gpu.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
