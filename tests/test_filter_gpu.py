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
