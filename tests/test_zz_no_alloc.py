"""wgfft_exec_device_noalloc (the entry point that allocates nothing) against
wgfft_exec_device, called through the C-ABI: identical outputs (the same
kernels and plan; only where the multipass intermediates live differs), on
two calls in a row with the second smaller, for LDS-resident lengths, two-pass
lengths and a three-pass plan, forward and inverse, in place; and the limits
(no plan yet, a length whose intermediates exceed the static scratch)."""
import ctypes

import numpy as np
import pytest

import oracle
import wgfft
from wgfft import _lib

pytestmark = pytest.mark.gpu


def _noalloc(g, x, out, inverse=False):
    _lib.check(g._L.wgfft_exec_device_noalloc(g._h, x.data_ptr(), out.data_ptr(), x.shape[1], x.shape[0],
                                              _lib.WGFFT_INVERSE if inverse else _lib.WGFFT_FORWARD,
                                              ctypes.c_void_p(0)))


def _plain(g, x, out, inverse=False):
    _lib.check(g._L.wgfft_exec_device(g._h, x.data_ptr(), out.data_ptr(), x.shape[1], x.shape[0],
                                      _lib.WGFFT_INVERSE if inverse else _lib.WGFFT_FORWARD, ctypes.c_void_p(0)))


# (log2n, first batch, second smaller batch): the 2^22 batch of 3 runs in two
# chunks of the 64 MiB scratch (one transform's two intermediates fill half)
CASES = [(12, 64, 5), (17, 9, 2), (20, 6, 1), (22, 3, 1), (23, 2, 1)]


@pytest.mark.parametrize("log2n,b1,b2", CASES)
def test_noalloc_matches_exec_device(torch_cuda, log2n, b1, b2):
    torch = torch_cuda
    n = 1 << log2n
    g = wgfft.GpuFft.new()
    try:
        g.describe(n)  # makes the plan (allocates its twiddle tables)
        for call, batch in enumerate((b1, b2)):
            x = torch.from_numpy(oracle.synth_uniform((batch, n), 0x5EED0A00 + 16 * log2n + call)).cuda()
            for inverse in (False, True):
                want, got = torch.empty_like(x), torch.full_like(x, float("nan"))
                _plain(g, x, want, inverse)
                _noalloc(g, x, got, inverse)
                torch.cuda.synchronize()
                assert torch.equal(got, want), (log2n, batch, inverse)
            inplace = x.clone()
            _noalloc(g, inplace, inplace)
            torch.cuda.synchronize()
            _plain(g, x, want)
            torch.cuda.synchronize()
            assert torch.equal(inplace, want), (log2n, batch, "in place")
        xs = x.cpu().numpy()
        e = oracle.rel_l2(inplace.cpu().numpy(), oracle.fft64(xs))
        assert e <= (5e-6 if n <= (1 << 20) else 2e-5), e
    finally:
        g.close()


@pytest.mark.parametrize("log2n,split", [(21, (7, 7, 7)), (22, (8, 7, 7))])
def test_noalloc_three_pass_plan(torch_cuda, log2n, split):
    """Three-pass splits pinned with wgfft_select_split: both intermediates
    live in the static scratch (at 2^22 each fills half of it)."""
    torch = torch_cuda
    n = 1 << log2n
    g = wgfft.GpuFft.new()
    try:
        desc = g.select_split(n, split)
        assert desc.count("R") == 3, desc
        for call, batch in enumerate((3, 2)):
            x = torch.from_numpy(oracle.synth_uniform((batch, n), 0x5EED0A80 + call)).cuda()
            want, got = torch.empty_like(x), torch.full_like(x, float("nan"))
            _plain(g, x, want)
            _noalloc(g, x, got)
            torch.cuda.synchronize()
            assert torch.equal(got, want), (desc, batch)
        assert oracle.rel_l2(got.cpu().numpy(), oracle.fft64(x.cpu().numpy())) <= 2e-5
    finally:
        g.close()


def test_noalloc_limits(torch_cuda):
    torch = torch_cuda
    g = wgfft.GpuFft.new()
    try:
        x = torch.zeros((1, 1 << 18), dtype=torch.complex64, device="cuda")
        with pytest.raises(wgfft.WgfftError) as e:
            _noalloc(g, x, x)  # no plan for 2^18 yet
        assert e.value.code == wgfft.WGFFT_ERR_INVALID_ARGUMENT and "plan" in str(e.value)
        big = torch.zeros((1, 1 << 24), dtype=torch.complex64, device="cuda")
        g.describe(1 << 24)
        with pytest.raises(wgfft.WgfftError) as e:
            _noalloc(g, big, big)  # three passes of 128 MiB intermediates
        assert e.value.code == wgfft.WGFFT_ERR_INVALID_ARGUMENT and "static scratch" in str(e.value)
        g.describe(1 << 18)
        _noalloc(g, x, x)
        torch.cuda.synchronize()
    finally:
        g.close()
