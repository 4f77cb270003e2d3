"""Spectral filter y = IFFT(FFT(x) * s) (1/N scaled) on the GPU vs FP64: the
reference pipeline's FFT, spectral Hann window and IFFT + 1/N stages
(examples/parallel_pipeline.rs:68-105, 586-642), fused into one pass where
the length has fused kernels and run as three launches elsewhere."""
import numpy as np
import pytest

import oracle
import wgfft

pytestmark = pytest.mark.gpu


def filter64(x, s):
    n = x.shape[1]
    return oracle.fft64(oracle.fft64(x, -1) * s, +1) / n


def hann(n):
    # the pipeline demo's spectral window
    k = np.arange(n)
    return (0.5 * (1 - np.cos(2 * np.pi * k / n))).astype(np.float32)


@pytest.mark.parametrize("log2n,batch", [(6, 33), (9, 37), (10, 37), (11, 37), (12, 37), (13, 37), (14, 9),
                                         (17, 3)])
def test_filter_vs_fp64(gpu, torch_cuda, log2n, batch):
    torch = torch_cuda
    n = 1 << log2n
    x = oracle.synth_uniform((batch, n), 0x5EED0800 + log2n)
    s = hann(n)
    s[::7] = 0.0  # a notch pattern on top of the window
    tx = torch.from_numpy(x).cuda()
    ts = torch.from_numpy(s).cuda()
    ty = torch.empty_like(tx)
    gpu.filter_device(tx, ts, ty)
    torch.cuda.synchronize()
    y = ty.cpu().numpy()
    want = filter64(x, s)
    assert oracle.rel_l2(y, want) <= 5e-6, oracle.rel_l2(y, want)
    gpu.filter_device(tx, ts, tx)  # in place
    torch.cuda.synchronize()
    assert np.array_equal(tx.cpu().numpy(), y)


def test_filter_identity_and_every_fused_variant(torch_cuda):
    """s = 1 is a round trip; every compiled fused variant's filter kernel,
    pinned in turn, matches FP64."""
    torch = torch_cuda
    for log2n in range(9, 15):
        n = 1 << log2n
        x = oracle.synth_uniform((5, n), 0x5EED0810 + log2n)
        s = hann(n)
        want = filter64(x, s)
        tx = torch.from_numpy(x).cuda()
        ts = torch.from_numpy(s).cuda()
        ones = torch.ones(n, dtype=torch.float32, device="cuda")
        for i in range(wgfft.GpuFft.adjoint_variant_count(n)):
            g = wgfft.GpuFft.new()
            try:
                desc = g.select_adjoint_variant(n, i)
                ty = torch.empty_like(tx)
                g.filter_device(tx, ts, ty)
                torch.cuda.synchronize()
                assert oracle.rel_l2(ty.cpu().numpy(), want) <= 5e-6, desc
                g.filter_device(tx, ones, ty)
                torch.cuda.synchronize()
                assert np.abs(ty.cpu().numpy() - x).max() < 1e-5, desc
            finally:
                g.close()


def test_filter_errors(gpu, torch_cuda):
    torch = torch_cuda
    x = torch.zeros((2, 1000), dtype=torch.complex64, device="cuda")
    s = torch.ones(1000, dtype=torch.float32, device="cuda")
    with pytest.raises(wgfft.WgfftError) as e:
        gpu.filter_device(x, s, x)
    assert e.value.code == wgfft.WGFFT_ERR_INVALID_LENGTH


def _noise_ref(n, rows, cols, seed, sigma):
    """The pipeline demo's noise formula in FP64 from its 32-bit hash, for the
    samples (rows x cols) of a batch: a check of the index mapping and the hash,
    not of the arithmetic (which the moment tests below pin)."""
    M = 0xFFFFFFFF

    def pcg(v):
        s = (v * 747796405 + 2891336453) & M
        w = (((s >> ((s >> 28) + 4)) ^ s) * 277803737) & M
        return (w >> 22) ^ w

    out = np.empty((len(rows), len(cols)), np.complex128)
    for i, b in enumerate(rows):
        for j, t in enumerate(cols):
            h1 = pcg((t * 1234567 + b * 7654321 + seed) & M)
            h2 = pcg((h1 + 999983) & M)
            u1 = max(h1 / 2.0 ** 32, 1e-7)
            u2 = h2 / 2.0 ** 32 * 2 * np.pi
            out[i, j] = sigma * np.sqrt(-2 * np.log(u1)) * np.exp(1j * u2)
    return out


def test_noise_stage(gpu, torch_cuda):
    """Deterministic per seed, N(0, sigma^2) per component (mean, variance,
    re/im correlation over 2^22 samples), added in place, and the demo's index
    mapping: sample t of signal b hashed from (t, b, seed)."""
    torch = torch_cuda
    n, batch, sigma = 4096, 1024, 0.1
    z = torch.zeros((batch, n), dtype=torch.complex64, device="cuda")
    gpu.add_noise_device(z, seed=0, sigma=sigma)
    a = z.cpu().numpy().astype(np.complex128)
    z2 = torch.zeros_like(z)
    gpu.add_noise_device(z2, seed=0, sigma=sigma)
    assert np.array_equal(z2.cpu().numpy(), z.cpu().numpy())
    z3 = torch.zeros_like(z)
    gpu.add_noise_device(z3, seed=12345, sigma=sigma)
    assert not np.array_equal(z3.cpu().numpy(), z.cpu().numpy())
    m = a.size
    for comp in (a.real, a.imag):
        assert abs(comp.mean()) < 4 * sigma / np.sqrt(m)
        assert abs(comp.std() / sigma - 1) < 0.01
    assert abs(np.mean(a.real * a.imag)) / sigma ** 2 < 4 / np.sqrt(m)
    rows, cols = [0, 1, 513, batch - 1], [0, 1, 7, 2048, n - 1]
    ref = _noise_ref(n, rows, cols, 0, sigma)
    assert np.abs(a[np.ix_(rows, cols)] - ref).max() < 1e-5
    # added in place on top of a signal
    x = oracle.synth_uniform((batch, n), 0x5EED0820)
    tx = torch.from_numpy(x).cuda()
    gpu.add_noise_device(tx, seed=0, sigma=sigma)
    assert np.abs(tx.cpu().numpy() - (x + a)).max() < 1e-6


@pytest.mark.parametrize("n,batch", [(4096, 64), (256, 64), (1 << 17, 2)])
def test_pipeline_noise_then_filter(gpu, torch_cuda, n, batch):
    """The demo's chain on device: noise, then FFT, spectral Hann window
    (N - 1 denominator), IFFT and 1/N, as two calls and as the one filter call
    with the noise at its load (fused at 4096, launches at 256 and 2^17),
    against FP64; x itself is left unchanged by the one-call form."""
    torch = torch_cuda
    x = oracle.synth_uniform((batch, n), 0x5EED0830 + n)
    k = np.arange(n)
    s = (0.5 * (1 - np.cos(2 * np.pi * k / (n - 1)))).astype(np.float32)
    ts = torch.from_numpy(s).cuda()
    tx = torch.from_numpy(x).cuda()
    gpu.add_noise_device(tx, seed=7, sigma=0.1)
    noisy = tx.cpu().numpy()
    want = filter64(noisy, s)
    ty = torch.empty_like(tx)
    gpu.filter_device(tx, ts, ty)
    torch.cuda.synchronize()
    two_calls = ty.cpu().numpy()
    assert oracle.rel_l2(two_calls, want) <= 5e-6
    tx2 = torch.from_numpy(x).cuda()
    ty2 = torch.empty_like(tx2)
    gpu.filter_device(tx2, ts, ty2, noise_sigma=0.1, seed=7)
    torch.cuda.synchronize()
    assert np.array_equal(tx2.cpu().numpy(), x)
    one_call = ty2.cpu().numpy()
    assert oracle.rel_l2(one_call, want) <= 5e-6
    assert oracle.rel_l2(one_call, two_calls) <= 1e-6
