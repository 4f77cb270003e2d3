#!/usr/bin/env python3
"""Headline benchmark: batched complex64 forward FFT, N = 4096 x 4096.

One step = one forward transform of 4096 signals of length 4096 (16.8 M
samples, 128 MiB in + 128 MiB out) with inputs resident in HBM, through the
C-ABI entry wgfft_exec_device on torch's current stream.  Also timed: the
fused forward + adjoint of L = sum m |FFT(w x)|^2 on the same shape (ratio
to the plain forward), and the reference algorithm on the host CPU (the
oracle's restatement of the wgls-rs-fft f32 Stockham R4/R2 path) on a
bounded sample.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "wgls-rs-fft_amd"))

N, BATCH = 4096, 4096
HBM_PEAK_GBS = 8000.0
METRIC = "complex64 GSamples/s & %HBM-roofline, N=4096×4096; fwd+adjoint vs plain fwd"


def cpu_baseline(seconds_target: float = 12.0) -> dict:
    """Reference algorithm (f32 Stockham R4/R2, restated in C) on host cores."""
    sys.path.insert(0, ROOT)
    import numpy as np

    import oracle

    cores = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()))
    x = oracle.synth_uniform((1024, N), 0x5EED0002)
    oracle.ref_fft_f32(x[:64], threads=cores)  # warm
    total = 0
    t0 = time.perf_counter()
    while True:
        oracle.ref_fft_f32(x, threads=cores)
        total += x.shape[0]
        dt = time.perf_counter() - t0
        if dt >= seconds_target:
            break
    return {"value": total * N / dt / 1e9, "unit": "GSamples/s", "cores": cores, "kind": "port",
            "sample": f"{total} transforms of N={N} (f32 Stockham R4/R2 restatement, {dt:.1f} s)"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import wgfft

    gpu = wgfft.GpuFft.new()
    x = torch.from_numpy(wgfft.synth_uniform((BATCH, N), 0x5EED0002)).cuda()
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream()

    for _ in range(args.warmup):
        gpu.exec_device(x, y)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        gpu.exec_device(x, y)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = ev0.elapsed_time(ev1) / args.steps
    ms_per_step = max(kern_ms, wall * 1e3 / args.steps)
    samples = N * BATCH
    value = samples / (ms_per_step * 1e-3) / 1e9
    achieved = 16.0 * samples / (kern_ms * 1e-3) / 1e9

    # fused forward + adjoint on the same shape
    w = torch.from_numpy(np.hanning(N).astype(np.float32)).cuda()
    m = (torch.rand(N, device="cuda") < 0.3).float()
    gx = torch.empty_like(x)
    gw = torch.empty(N, dtype=torch.float32, device="cuda")
    loss = torch.empty(1, dtype=torch.float32, device="cuda")
    fa_steps = max(10, args.steps // 4)
    for _ in range(max(3, args.warmup // 4)):
        gpu.fwd_adjoint_masked_device(x, w, m, None, gx, gw, loss)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(fa_steps):
        gpu.fwd_adjoint_masked_device(x, w, m, None, gx, gw, loss)
    ev1.record(stream)
    torch.cuda.synchronize()
    fa_ms = ev0.elapsed_time(ev1) / fa_steps

    traffic = None
    prof = os.path.join(ROOT, "profiles", "pmc_fft4096x4096.json")
    if os.path.exists(prof):
        with open(prof) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

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
        "config": {"workload": "batched forward FFT", "n": N, "batch": BATCH, "samples": samples,
                   "plan": gpu.describe(N)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
        "fwd_adjoint": {"ms": round(fa_ms, 5), "plain_fwd_ms": round(kern_ms, 5),
                        "ratio": round(fa_ms / kern_ms, 3)},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))
    gpu.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
