"""torch.autograd test harness over the C-ABI differentiation entry points.

``windowed_fft(x, w, handle)`` computes Y = FFT(w * x) for a (batch, n)
complex64 CUDA tensor x and a real (n,) float32 window w with the HIP
windowed kernel; its backward is ``wgfft_vjp_windowed_device``:
gx = w * F^H gY and gw = sum_b Re(conj(F^H gY) x).  Gradients follow
torch's convention for complex tensors (dL/dRe + i dL/dIm).
"""
from __future__ import annotations

import torch


class WindowedFFT(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, w, handle):
        x = x.contiguous()
        w = w.contiguous()
        y = torch.empty_like(x)
        handle.exec_windowed_device(x, w, y)
        ctx.save_for_backward(x, w)
        ctx.handle = handle
        return y

    @staticmethod
    def backward(ctx, gy):
        x, w = ctx.saved_tensors
        gy = gy.contiguous()
        gx = torch.empty_like(x)
        gw = torch.empty_like(w)
        ctx.handle.vjp_windowed_device(x, w, gy, gx, gw)
        return gx, gw, None


def windowed_fft(x, w, handle):
    return WindowedFFT.apply(x, w, handle)


def masked_power_loss(x, w, mask, handle, want_y: bool = False):
    """Fused forward + adjoint of L = sum m |FFT(w x)|^2 in one library call.

    Returns (loss, gx, gw[, y]) as CUDA tensors."""
    batch, n = x.shape
    gx = torch.empty_like(x)
    gw = torch.empty_like(w)
    loss = torch.empty(1, dtype=torch.float32, device=x.device)
    y = torch.empty_like(x) if want_y else None
    handle.fwd_adjoint_masked_device(x, w, mask, y, gx, gw, loss)
    return (loss, gx, gw, y) if want_y else (loss, gx, gw)
