"""Write the FP64 golden vectors in tests/golden/fft_golden.npz.

Inputs: complex64 signals from the oracle's SplitMix64->PCG32 uniform stream
(seed 0x5EED0000 + N).  Outputs: numpy's pocketfft in FP64 (an independent
FFT implementation) -- forward (unnormalised) and, for N <= 1024, inverse (1/N scaled).
A SHA-256 of every array is stored beside it and checked by the tests.
Run from the repo root: python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import oracle  # noqa: E402

SIZES = {1: 2, 2: 2, 4: 2, 8: 2, 16: 2, 32: 2, 64: 2, 256: 2, 1024: 2, 4096: 1}


def main():
    arrays = {}
    for n, count in SIZES.items():
        x = oracle.synth_uniform((count, n), 0x5EED0000 + n)
        x128 = x.astype(np.complex128)
        arrays[f"x_{n}"] = x
        arrays[f"fwd_{n}"] = np.fft.fft(x128, axis=-1)
        if n <= 1024:
            arrays[f"inv_{n}"] = np.fft.ifft(x128, axis=-1)
    digest = {k: hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest() for k, v in arrays.items()}
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_golden.npz")
    np.savez_compressed(out, **arrays, _sha256=np.array([f"{k}:{v}" for k, v in sorted(digest.items())]))
    print("wrote", out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
