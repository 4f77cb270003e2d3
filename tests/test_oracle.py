"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Each oracle function is checked against something it does not compute
itself: published PRNG outputs, numpy pocketfft FP64 golden vectors, closed
forms (impulse, constant, complex exponential, shifted impulse, N = 2 and 4
by hand), invariants (Parseval, linearity, batch = single), the reference's
own test inputs and tolerances, and FP64 central finite differences for the
adjoint.  A few checks plant a plausible mistake in the oracle's output and
show the check rejects it.
"""
import hashlib
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "fft_golden.npz")


# ----------------------------------------------------------------- PRNG
def test_pcg32_published_outputs():
    # pcg32-demo (O'Neill): pcg32_srandom(42, 54) -> first six outputs
    got = [int(v) for v in oracle.pcg32_raw(42, 54, 6)]
    assert got == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


def test_splitmix64_published_outputs():
    # SplitMix64 (Vigna) from state 0
    got = [int(v) for v in oracle.splitmix64(0, 3)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_uniform_grid_and_moments():
    x = oracle.synth_uniform((1 << 16,), 7).view(np.float32)
    assert x.min() >= -1.0 and x.max() < 1.0
    k = (x.astype(np.float64) + 1.0) * 2.0**23
    assert np.all(k == np.round(k))  # exact multiples of 2^-23
    assert abs(x.mean()) < 0.01 and abs(x.var() - 1.0 / 3.0) < 0.01
    # determinism and seed sensitivity
    assert np.array_equal(oracle.synth_uniform((64,), 7), oracle.synth_uniform((64,), 7))
    assert not np.array_equal(oracle.synth_uniform((64,), 7), oracle.synth_uniform((64,), 8))


def test_gauss_moments():
    g = oracle.synth_gauss(1 << 16, 3)
    assert abs(g.mean()) < 0.02 and abs(g.std() - 1.0) < 0.02


def test_sinusoid_signal_integer_bins():
    # with N = 2^k the frequencies f_j = U*N land on multiples of N/2^24;
    # at N = 2^24 they are integers, so each tone is one exact DFT bin.  At a
    # small N check the signal's power instead: 32 unit tones + noise 0.1.
    x = oracle.synth_sinusoids(4096, 2, 11)
    p = np.mean(np.abs(x.astype(np.complex128)) ** 2)
    assert 15.0 < p < 60.0
    assert np.array_equal(x, oracle.synth_sinusoids(4096, 2, 11))


def test_ragged_lengths_and_window_mask():
    k = oracle.synth_ragged_log2(20000, 6, 16, 0x5EED0003)
    assert k.min() == 6 and k.max() == 16
    counts = np.bincount(k)[6:]
    assert counts.min() > 1600 and counts.max() < 2050
    w, m = oracle.synth_window_mask(8192, 0x5EED0004)
    hann = 0.5 * (1 - np.cos(2 * np.pi * np.arange(8192) / 8191))
    assert np.abs(w - hann).max() < 0.06 and 0.005 < np.std(w - hann) < 0.015
    assert set(np.unique(m)) <= {0.0, 1.0} and 0.28 < m.mean() < 0.32


# ------------------------------------------------------ golden vectors
def _golden():
    return np.load(GOLDEN, allow_pickle=False)


def test_golden_integrity():
    g = _golden()
    for entry in g["_sha256"]:
        key, digest = str(entry).split(":")
        assert hashlib.sha256(np.ascontiguousarray(g[key]).tobytes()).hexdigest() == digest, key


@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32, 64, 256, 1024, 4096])
def test_fp64_oracle_matches_pocketfft_golden(n):
    g = _golden()
    x = g[f"x_{n}"]
    assert oracle.rel_l2(oracle.fft64(x, -1), g[f"fwd_{n}"]) < 1e-13
    if n <= 1024:
        assert oracle.rel_l2(oracle.dft64(x, -1), g[f"fwd_{n}"]) < 1e-12
        assert oracle.rel_l2(oracle.ifft64(x), g[f"inv_{n}"]) < 1e-13


@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32, 64, 256, 1024, 4096])
def test_reference_restatement_matches_golden(n):
    g = _golden()
    x = g[f"x_{n}"]
    y = oracle.ref_fft_f32(x)
    assert oracle.rel_l2(y, g[f"fwd_{n}"]) < 1e-6
    assert np.abs(y - g[f"fwd_{n}"]).max() < 1e-3  # tests/gpu_fft_test.rs:6 bar
    if n <= 1024:
        assert oracle.rel_l2(oracle.ref_fft_f32(x, inverse=True), g[f"inv_{n}"]) < 1e-6


# ------------------------------------------------------ known answers
def kat_errors(fwd, n, dtype=np.complex64):
    """Max error of `fwd` (batched forward FFT) on four closed forms."""
    k = np.arange(n)
    errs = {}
    imp = np.zeros((1, n), dtype)
    imp[0, 0] = 1
    errs["impulse"] = np.abs(fwd(imp) - 1).max()
    c = np.full((1, n), 0.5 - 0.25j, dtype)
    want = np.zeros(n, complex)
    want[0] = n * (0.5 - 0.25j)
    errs["constant"] = np.abs(fwd(c) - want).max() / n
    k0 = 3 % n
    e = np.exp(2j * np.pi * k0 * k / n).astype(dtype)[None]
    want = np.zeros(n, complex)
    want[k0] = n
    errs["exponential"] = np.abs(fwd(e) - want).max() / n
    n0 = 5 % n
    sh = np.zeros((1, n), dtype)
    sh[0, n0] = 1
    errs["shift"] = np.abs(fwd(sh) - np.exp(-2j * np.pi * k * n0 / n)).max()
    return errs


@pytest.mark.parametrize("n", [8, 64, 1024, 4096, 1 << 15])
def test_known_answers(n):
    for name, f, tol, dt in (("fft64", lambda x: oracle.fft64(x, -1), 1e-12, np.complex128),
                             ("ref_f32", lambda x: oracle.ref_fft_f32(x), 2e-6, np.complex64)):
        errs = kat_errors(f, n, dt)
        assert max(errs.values()) < tol, (name, errs)


def test_known_answers_reject_planted_mistakes():
    n = 64
    good = lambda x: oracle.fft64(x, -1)  # noqa: E731
    assert max(kat_errors(good, n, np.complex128).values()) < 1e-12
    # wrong sign convention (conjugated twiddles)
    assert kat_errors(lambda x: oracle.fft64(x, +1), n)["exponential"] > 0.5
    # dropped output reordering (bit-reversed bins)
    rev = np.array([int(format(i, "06b")[::-1], 2) for i in range(n)])
    assert kat_errors(lambda x: good(x)[:, rev], n)["shift"] > 0.5
    # off-by-one index (k+1 instead of k)
    assert kat_errors(lambda x: np.roll(good(x), 1, axis=-1), n)["exponential"] > 0.5


def test_tiny_n_by_hand():
    a, b, c, d = 1 + 2j, -3 + 0.5j, 0.25 - 1j, 2 + 0j
    y2 = oracle.fft64(np.array([[a, b]]), -1)[0]
    assert np.allclose(y2, [a + b, a - b], atol=1e-15)
    y4 = oracle.fft64(np.array([[a, b, c, d]]), -1)[0]
    want = [a + b + c + d, a - 1j * b - c + 1j * d, a - b + c - d, a + 1j * b - c - 1j * d]
    assert np.allclose(y4, want, atol=1e-14)
    for f in (lambda x: oracle.ref_fft_f32(x)[0], lambda x: oracle.dft64(x)[0]):
        assert np.allclose(f(np.array([[a, b, c, d]], np.complex64)), want, atol=1e-6)


# ----------------------------------------------------------- invariants
@pytest.mark.parametrize("n", [128, 2048])
def test_parseval_linearity_batch(n):
    x = oracle.synth_uniform((4, n), 21).astype(np.complex128)
    y = oracle.fft64(x)
    assert np.allclose(np.sum(np.abs(y) ** 2, -1), n * np.sum(np.abs(x) ** 2, -1), rtol=1e-12)
    u, v = x[0], x[1]
    assert np.allclose(oracle.fft64(2.5 * u - 1j * v)[0], 2.5 * y[0] - 1j * y[1], atol=1e-10)
    xf = x.astype(np.complex64)
    batch = oracle.ref_fft_f32(xf)
    single = np.stack([oracle.ref_fft_f32(xf[i:i + 1])[0] for i in range(4)])
    assert np.array_equal(batch, single)
    assert np.array_equal(oracle.ref_fft_f32(xf, threads=1), oracle.ref_fft_f32(xf, threads=3))


def test_real_input_conjugate_symmetry():
    n = 512
    x = oracle.synth_uniform((1, n), 2).real.astype(np.complex64)
    y = oracle.fft64(x)[0]
    assert np.allclose(y[1:], np.conj(y[1:][::-1]), atol=1e-10)


# --------------------------------- the reference's own tests, on the oracle
def make_input(n):
    """tests/gpu_fft_test.rs:8-15 (f32 arithmetic)."""
    i = np.arange(n, dtype=np.float32)
    return (np.sin(i * np.float32(0.1)) + 1j * np.cos(i * np.float32(0.07))).astype(np.complex64)


def generate_test_signal(n):
    """tests/test_utils.rs:21-34 (f32 arithmetic)."""
    t = np.arange(n, dtype=np.float32) / np.float32(n)
    two_pi = np.float32(2.0 * np.pi)
    s = np.float32(0.7) * np.sin(two_pi * np.float32(10.0) * t) + np.float32(0.3) * np.sin(
        two_pi * np.float32(50.0) * t)
    return s.astype(np.float32).astype(np.complex64)


def apply_hann_window(x):
    """tests/test_utils.rs:39-50 (f32, denominator N-1)."""
    n = len(x)
    i = np.arange(n, dtype=np.float32)
    w = np.float32(0.5) * (np.float32(1.0) - np.cos(np.float32(2.0 * np.pi) * i / np.float32(n - 1)))
    return (x * w.astype(np.float32)).astype(np.complex64)


def test_reference_gpu_fft_test_on_restatement():
    # tests/gpu_fft_test.rs:17-42: every element within 1e-3 of the f32 oracle
    x = make_input(1024)
    y = oracle.ref_fft_f32(x)[0]
    assert np.abs(y - oracle.fft64(x)[0]).max() < 1e-3


def test_reference_roundtrips_on_restatement():
    # tests/ifft_test.rs:22-107 and src/lib.rs:207-227
    x = make_input(1024)
    y = oracle.ref_fft_f32(x)
    r = oracle.ref_fft_f32(y, inverse=True)
    assert np.abs(r[0] - x).max() < 1e-3
    assert np.abs(oracle.ref_fft_f32(r)[0] - y[0]).max() < 1e-3
    ramp = (np.arange(1024, dtype=np.float32) * np.float32(0.1)).astype(np.complex64)
    assert np.abs(oracle.ref_fft_f32(oracle.ref_fft_f32(ramp), inverse=True)[0] - ramp).max() < 1e-4


def test_reference_windowed_tests_on_restatement():
    # tests/windowed_fft_test.rs:10-133 at N = 2048
    n = 2048
    sig = generate_test_signal(n)
    win = apply_hann_window(sig)
    spec = oracle.ref_fft_f32(win)[0]
    assert np.abs(oracle.ref_fft_f32(spec[None], inverse=True)[0] - win).max() < 1e-3
    ratio = np.sum(np.abs(spec) ** 2) / np.sum(np.abs(oracle.ref_fft_f32(sig)[0]) ** 2)
    assert 0.1 < ratio < 10.0
    p = np.abs(spec) ** 2
    peaks = [i for i in range(1, n // 2) if p[i] > p[i - 1] and p[i] > p[i + 1] and p[i] > 10.0]
    assert any(8 <= f < 12 for f in peaks) and any(45 <= f < 55 for f in peaks)


def test_reference_twiddle_table_is_f32_angle():
    # src/lib.rs:764-768: angle rounded in f32 before cos/sin
    n = 1024
    tw = oracle.ref_twiddles(n)
    j = np.arange(n, dtype=np.float32)
    ang = (np.float32(-2.0 * np.pi) * j / np.float32(n)).astype(np.float32)
    assert np.allclose(tw.real, np.cos(ang.astype(np.float64)), atol=2e-7)
    assert np.allclose(tw.imag, np.sin(ang.astype(np.float64)), atol=2e-7)
    exact = np.exp(-2j * np.pi * np.arange(n) / n)
    assert np.abs(tw - exact).max() < 1e-6


# ------------------------------------------------------------- adjoint
def _loss_np(x, w, m):
    """FP64 loss by numpy.fft (independent of the oracle's FFT)."""
    y = np.fft.fft(w[None, :] * x, axis=-1)
    return float(np.sum(m[None, :] * np.abs(y) ** 2))


def test_vjp_matches_finite_differences():
    n, b = 64, 3
    x = oracle.synth_uniform((b, n), 31)
    w, m = oracle.synth_window_mask(n, 32)
    loss, gy = oracle.masked_loss64(x, w, m)
    assert abs(loss - _loss_np(x.astype(np.complex128), w.astype(np.float64), m.astype(np.float64))) < 1e-9 * loss
    gx, gw = oracle.vjp64(x, w, gy.astype(np.complex64))
    # gY rounded to complex64 for the oracle call: compare against the FP64
    # chain with the same rounding, then FD at a tolerance covering it
    rng = np.random.default_rng(5)
    x64, w64, m64 = x.astype(np.complex128), w.astype(np.float64), m.astype(np.float64)
    for _ in range(16):
        bi, i = rng.integers(b), rng.integers(n)
        for part in (1.0, 1j):
            h = 1e-6
            xp, xm = x64.copy(), x64.copy()
            xp[bi, i] += h * part
            xm[bi, i] -= h * part
            fd = (_loss_np(xp, w64, m64) - _loss_np(xm, w64, m64)) / (2 * h)
            an = gx[bi, i].real if part == 1.0 else gx[bi, i].imag
            assert abs(fd - an) < 1e-4 * max(1.0, abs(fd)), (bi, i, part, fd, an)
        wp, wm = w64.copy(), w64.copy()
        wp[i] += 1e-6
        wm[i] -= 1e-6
        fd = (_loss_np(x64, wp, m64) - _loss_np(x64, wm, m64)) / 2e-6
        assert abs(fd - gw[i]) < 1e-4 * max(1.0, abs(fd)), (i, fd, gw[i])


def test_adjoint_identity():
    n = 256
    u = oracle.synth_uniform((1, n), 41).astype(np.complex128)
    v = oracle.synth_uniform((1, n), 42).astype(np.complex128)
    lhs = np.vdot(oracle.fft64(u, -1), v)
    rhs = np.vdot(u, oracle.fft64(v, +1))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


# ------------------------------------------------ selected FP64 DFT bins
@pytest.mark.parametrize("n", [1, 2, 8, 64, 1024, 4096])
def test_dft_bins_match_pocketfft_golden(n):
    # every bin of the golden signals: direct sums against pocketfft FP64
    g = _golden()
    x = g[f"x_{n}"]
    bins = np.arange(n, dtype=np.uint64)
    got = oracle.dft_bins64(x[0], bins)
    want = g[f"fwd_{n}"][0]
    assert oracle.rel_l2(got, want) < 1e-12
    if n >= 8:
        # planted mistakes: the inverse sign, and the neighbouring bin
        wrong_sign = np.conj(oracle.dft_bins64(np.conj(x[0]), bins))  # sum x exp(+2 pi i k j / N)
        assert oracle.rel_l2(wrong_sign, want) > 0.5
        assert oracle.rel_l2(oracle.dft_bins64(x[0], (bins + 1) % n), want) > 0.5


def test_dft_bins_match_fft64_at_2pow15():
    # random bins of a 2^15 signal against the FP64 radix-2 FFT (itself pinned
    # against the golden vectors above), and the index wrap k mod N
    n = 1 << 15
    x = oracle.synth_uniform((1, n), 0x5EED0410)
    y = oracle.fft64(x)[0]
    rng = np.random.default_rng(12)
    bins = np.unique(np.concatenate([[0, 1, n // 2, n - 1], rng.integers(0, n, 28)])).astype(np.uint64)
    got = oracle.dft_bins64(x[0], bins)
    assert np.abs(got - y[bins.astype(np.int64)]).max() < 1e-9 * np.sqrt(n)
    assert np.array_equal(oracle.dft_bins64(x[0], bins + np.uint64(n)), got)


def test_dft_bins_closed_forms_at_large_n():
    # a complex exponential at bin k0 and a shifted impulse, at N = 2^20
    n = 1 << 20
    k0 = 123457
    j = np.arange(n, dtype=np.float64)
    e = np.exp(2j * np.pi * ((k0 * j) % n) / n).astype(np.complex64)
    got = oracle.dft_bins64(e, np.array([k0, k0 + 1, 0], dtype=np.uint64))
    assert abs(got[0] - n) < 1e-3 * n ** 0.5 and abs(got[1]) < 1e-2 * n ** 0.5 and abs(got[2]) < 1e-2 * n ** 0.5
    imp = np.zeros(n, np.complex64)
    imp[777] = 1
    ks = np.array([0, 1, 5000, n - 3], dtype=np.uint64)
    want = np.exp(-2j * np.pi * ((ks.astype(np.float64) * 777) % n) / n)
    assert np.abs(oracle.dft_bins64(imp, ks) - want).max() < 1e-12


def test_fft64_threaded_single_transform_bit_identical():
    # a transform split over threads computes every butterfly as the
    # one-thread transform does
    for n in (1 << 16, 1 << 18):
        x = oracle.synth_uniform((2, n), 0x5EED0420)
        one = oracle.fft64(x, threads=1)
        for t in (3, 8):
            assert np.array_equal(oracle.fft64(x, threads=t), one)
            assert np.array_equal(oracle.fft64(x, +1, threads=t), oracle.fft64(x, +1, threads=1))
    assert oracle.rel_l2(one, np.fft.fft(x.astype(np.complex128), axis=-1)) < 1e-14
