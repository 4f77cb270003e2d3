"""Differentiation of Y = FFT(w * x) on the GPU vs the FP64 analytic adjoint
and FP64 central finite differences (windowed workload: N = 8192,
batch 1024, w = Hann + N(0, 0.01^2) noise, mask m_k ~ Bernoulli(0.3),
L = sum_b sum_k m_k |Y_bk|^2)."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

N, B = 8192, 1024


@pytest.fixture(scope="module")
def c5(torch_cuda):
    torch = torch_cuda
    x = oracle.synth_uniform((B, N), 0x5EED0004)
    w, m = oracle.synth_window_mask(N, 0x5EED0014)
    dev = dict(x=torch.from_numpy(x).cuda(), w=torch.from_numpy(w).cuda(), m=torch.from_numpy(m).cuda())
    loss64, gy64 = oracle.masked_loss64(x, w, m)
    gx64, gw64 = oracle.vjp64(x, w, gy64.astype(np.complex64))
    return dict(x=x, w=w, m=m, dev=dev, loss64=loss64, gy64=gy64, gx64=gx64, gw64=gw64)


def test_vjp_entry_vs_fp64_adjoint(gpu, torch_cuda, c5, parity):
    torch = torch_cuda
    gy = c5["gy64"].astype(np.complex64)
    tgy = torch.from_numpy(gy).cuda()
    gx = torch.empty_like(c5["dev"]["x"])
    gw = torch.empty_like(c5["dev"]["w"])
    gpu.vjp_windowed_device(c5["dev"]["x"], c5["dev"]["w"], tgy, gx, gw)
    torch.cuda.synchronize()
    e_gx = oracle.rel_l2(gx.cpu().numpy(), c5["gx64"])
    e_gw = np.linalg.norm(gw.cpu().numpy() - c5["gw64"]) / np.linalg.norm(c5["gw64"])
    print(f"VJP rel-L2: gx {e_gx:.3e} gw {e_gw:.3e}")
    parity["C5 VJP entry vs FP64 adjoint"] = {"gx_rel_l2": e_gx, "gw_rel_l2": e_gw, "bars": [5e-6, 1e-5]}
    assert e_gx <= 5e-6 and e_gw <= 1e-5


def test_fused_fwd_adjoint_vs_fp64(gpu, torch_cuda, c5, parity):
    torch = torch_cuda
    from wgfft.autograd import masked_power_loss

    d = c5["dev"]
    loss, gx, gw, y = masked_power_loss(d["x"], d["w"], d["m"], gpu, want_y=True)
    torch.cuda.synchronize()
    e_l = abs(float(loss.item()) - c5["loss64"]) / c5["loss64"]
    e_y = oracle.rel_l2(y.cpu().numpy(), oracle.windowed_fwd64(c5["x"], c5["w"]))
    gx64, gw64 = oracle.vjp64(c5["x"], c5["w"], c5["gy64"].astype(np.complex64))
    e_gx = oracle.rel_l2(gx.cpu().numpy(), gx64)
    e_gw = np.linalg.norm(gw.cpu().numpy() - gw64) / np.linalg.norm(gw64)
    print(f"fused: loss {e_l:.3e} Y {e_y:.3e} gx {e_gx:.3e} gw {e_gw:.3e}")
    parity["C5 fused forward+adjoint vs FP64"] = {"loss_rel": e_l, "y_rel_l2": e_y, "gx_rel_l2": e_gx,
                                                 "gw_rel_l2": e_gw, "bars": [1e-5, 5e-6, 1e-5, 1e-5]}
    assert e_l <= 1e-5 and e_y <= 5e-6 and e_gx <= 1e-5 and e_gw <= 1e-5


def test_fused_vs_finite_differences(gpu, torch_cuda, c5, parity):
    """64 coordinates: 48 of x (re/im) and 16 of w; FP64 central differences
    of the FP64 loss (numpy.fft) against the GPU's f32 gradients.  L is
    exactly quadratic in x and in w, so the central difference has no
    truncation error and h = 0.01 keeps FP64 rounding noise (eps*L/h) far
    below the f32 error.  Tolerance: 1e-5 of the gradient's RMS plus 1e-4
    relative."""
    torch = torch_cuda
    from wgfft.autograd import masked_power_loss

    d = c5["dev"]
    _, gx, gw = masked_power_loss(d["x"], d["w"], d["m"], gpu)
    torch.cuda.synchronize()
    gx, gw = gx.cpu().numpy(), gw.cpu().numpy()
    x64 = c5["x"].astype(np.complex128)
    w64 = c5["w"].astype(np.float64)
    m64 = c5["m"].astype(np.float64)

    def row_loss(xr, wv):
        y = np.fft.fft(wv * xr)
        return float(np.sum(m64 * np.abs(y) ** 2))

    rng = np.random.default_rng(0x5EED)
    h = 1e-2
    worst_x = worst_w = 0.0  # max |fd - analytic| / gradient RMS
    for _ in range(24):
        b, n = int(rng.integers(B)), int(rng.integers(N))
        scale = float(np.sqrt(np.mean(np.abs(gx[b]) ** 2)))
        for part in (1.0, 1j):
            xp, xm = x64[b].copy(), x64[b].copy()
            xp[n] += h * part
            xm[n] -= h * part
            fd = (row_loss(xp, w64) - row_loss(xm, w64)) / (2 * h)
            an = gx[b, n].real if part == 1.0 else gx[b, n].imag
            worst_x = max(worst_x, abs(fd - an) / scale)
            assert abs(fd - an) <= 1e-5 * scale + 1e-4 * abs(fd), (b, n, part, fd, an, scale)
    scale_w = float(np.sqrt(np.mean(gw**2)))
    for _ in range(16):
        n = int(rng.integers(N))
        losses = []
        for sgn in (1, -1):
            wv = w64.copy()
            wv[n] += sgn * h
            y = np.fft.fft(wv[None] * x64, axis=-1)
            losses.append(float(np.sum(m64[None] * np.abs(y) ** 2)))
        fd = (losses[0] - losses[1]) / (2 * h)
        worst_w = max(worst_w, abs(fd - gw[n]) / scale_w)
        assert abs(fd - gw[n]) <= 1e-5 * scale_w + 1e-4 * abs(fd), (n, fd, gw[n], scale_w)
    parity["C5 fused gradients vs FP64 central differences (64 coordinates)"] = {
        "gx_max_abs_err_over_rms": worst_x, "gw_max_abs_err_over_rms": worst_w}


def test_autograd_harness(gpu, torch_cuda):
    torch = torch_cuda
    from wgfft.autograd import windowed_fft

    n, b = 4096, 16
    x = oracle.synth_uniform((b, n), 77)
    w, m = oracle.synth_window_mask(n, 78)
    tx = torch.from_numpy(x).cuda().requires_grad_(True)
    tw = torch.from_numpy(w).cuda().requires_grad_(True)
    tm = torch.from_numpy(m).cuda()
    y = windowed_fft(tx, tw, gpu)
    loss = (tm * (y.real**2 + y.imag**2)).sum()
    loss.backward()
    _, gy = oracle.masked_loss64(x, w, m)
    gx64, gw64 = oracle.vjp64(x, w, gy.astype(np.complex64))
    assert oracle.rel_l2(tx.grad.cpu().numpy(), gx64) <= 1e-5
    assert np.linalg.norm(tw.grad.cpu().numpy() - gw64) / np.linalg.norm(gw64) <= 1e-5


@pytest.mark.parametrize("log2n", list(range(9, 15)))
def test_every_fused_variant(torch_cuda, log2n):
    """Every compiled fused forward+adjoint / VJP variant, pinned in turn,
    matches the FP64 loss, Y, gx and gw (and the VJP entry its FP64 adjoint)
    on a batch that leaves a ragged last round of transforms."""
    import wgfft

    torch = torch_cuda
    n, batch = 1 << log2n, 37
    x = oracle.synth_uniform((batch, n), 0x5EED0600 + log2n)
    w, m = oracle.synth_window_mask(n, 0x5EED0610 + log2n)
    loss64, gy64 = oracle.masked_loss64(x, w, m)
    y64 = oracle.windowed_fwd64(x, w)
    gx64, gw64 = oracle.vjp64(x, w, gy64.astype(np.complex64))
    tx, tw_, tm = (torch.from_numpy(a).cuda() for a in (x, w, m))
    tgy = torch.from_numpy(gy64.astype(np.complex64)).cuda()
    for i in range(wgfft.GpuFft.adjoint_variant_count(n)):
        g = wgfft.GpuFft.new()
        try:
            desc = g.select_adjoint_variant(n, i)
            y, gx = torch.empty_like(tx), torch.empty_like(tx)
            gw, loss = torch.empty_like(tw_), torch.zeros(1, dtype=torch.float32, device="cuda")
            g.fwd_adjoint_masked_device(tx, tw_, tm, y, gx, gw, loss)
            torch.cuda.synchronize()
            e_l = abs(float(loss.item()) - loss64) / loss64
            e_y = oracle.rel_l2(y.cpu().numpy(), y64)
            e_gx = oracle.rel_l2(gx.cpu().numpy(), gx64)
            e_gw = np.linalg.norm(gw.cpu().numpy() - gw64) / np.linalg.norm(gw64)
            assert e_l <= 1e-5 and e_y <= 5e-6 and e_gx <= 1e-5 and e_gw <= 1e-5, (desc, e_l, e_y, e_gx, e_gw)
            # without Y (its own instantiation in the split-form kernels)
            gx.zero_()
            gw.zero_()
            loss.zero_()
            g.fwd_adjoint_masked_device(tx, tw_, tm, None, gx, gw, loss)
            torch.cuda.synchronize()
            e_l = abs(float(loss.item()) - loss64) / loss64
            e_gx = oracle.rel_l2(gx.cpu().numpy(), gx64)
            e_gw = np.linalg.norm(gw.cpu().numpy() - gw64) / np.linalg.norm(gw64)
            assert e_l <= 1e-5 and e_gx <= 1e-5 and e_gw <= 1e-5, (desc, "no y", e_l, e_gx, e_gw)
            gx.zero_()
            gw.zero_()
            g.vjp_windowed_device(tx, tw_, tgy, gx, gw)
            torch.cuda.synchronize()
            e_gx = oracle.rel_l2(gx.cpu().numpy(), gx64)
            e_gw = np.linalg.norm(gw.cpu().numpy() - gw64) / np.linalg.norm(gw64)
            assert e_gx <= 5e-6 and e_gw <= 1e-5, (desc, "vjp", e_gx, e_gw)
        finally:
            g.close()

