"""Mixed-length batches (20,000 transforms, N = 2^k, k ~ U{6..16}, packed
contiguously) in one persistent launch, its hipGraph replay and the
one-launch-per-length baseline, against the FP64 oracle."""
import numpy as np
import pytest

import oracle
import wgfft

pytestmark = pytest.mark.gpu


def _check(y, x, offsets, k, idx, bar=5e-6):
    for i in idx:
        n = 1 << int(k[i])
        o = int(offsets[i])
        err = oracle.rel_l2(y[o:o + n], oracle.fft64(x[o:o + n])[0])
        assert err <= bar, (i, n, err)


@pytest.mark.parametrize("count,seed", [(300, 0x5EED0103), (2000, 0x5EED0203)])
def test_ragged_small_batches_all_modes(gpu, torch_cuda, count, seed):
    torch = torch_cuda
    k = oracle.synth_ragged_log2(count, 6, 16, seed)
    offsets, total = wgfft.ragged_layout(k)
    x = oracle.synth_uniform((total,), seed + 1)
    plan = wgfft.RaggedPlan(gpu, offsets, k)
    assert plan.samples == total
    tx = torch.from_numpy(x).cuda()
    outs = []
    for mode in (0, 2):
        ty = torch.zeros_like(tx)
        plan.exec_device(tx, ty, mode=mode)
        torch.cuda.synchronize()
        outs.append(ty.cpu().numpy())
    assert np.array_equal(outs[0], outs[1])
    _check(outs[0], x, offsets, k, range(count))
    # inverse round trip through the same plan
    ty = torch.from_numpy(outs[0]).cuda()
    tr = torch.empty_like(ty)
    plan.exec_device(ty, tr, inverse=True)
    torch.cuda.synchronize()
    assert oracle.rel_l2(tr.cpu().numpy(), x) <= 5e-6
    plan.close()


@pytest.mark.parametrize("kmax", [13, 14, 15])
def test_ragged_scratch_sized_by_largest_length(gpu, torch_cuda, kmax):
    """Batches whose largest length is 2^13 or 2^14 (no two-pass unit, no
    scratch) or 2^15 (scratch slots of that size) in every launch mode."""
    torch = torch_cuda
    k = oracle.synth_ragged_log2(700, 6, kmax, 0x5EED0503 + kmax)
    offsets, total = wgfft.ragged_layout(k)
    x = oracle.synth_uniform((total,), 0x5EED0504 + kmax)
    plan = wgfft.RaggedPlan(gpu, offsets, k)
    tx = torch.from_numpy(x).cuda()
    outs = []
    for mode in (0, 1, 2):
        ty = torch.zeros_like(tx)
        plan.exec_device(tx, ty, mode=mode)
        torch.cuda.synchronize()
        outs.append(ty.cpu().numpy())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    _check(outs[0], x, offsets, k, range(0, 700, 3))
    plan.close()


def test_ragged_graph_replays(gpu, torch_cuda):
    torch = torch_cuda
    k = oracle.synth_ragged_log2(500, 6, 16, 0x5EED0303)
    offsets, total = wgfft.ragged_layout(k)
    x = oracle.synth_uniform((total,), 0x5EED0304)
    plan = wgfft.RaggedPlan(gpu, offsets, k)
    tx = torch.from_numpy(x).cuda()
    ref = torch.empty_like(tx)
    plan.exec_device(tx, ref, mode=0)
    torch.cuda.synchronize()
    eager = ref.cpu().numpy()
    ty = torch.empty_like(tx)
    for rep in range(4):
        ty.zero_()
        plan.exec_device(tx, ty, mode=1)
        torch.cuda.synchronize()
        assert np.array_equal(ty.cpu().numpy(), eager), rep
        # eager work between replays
        plan.exec_device(tx, ref, mode=0)
        torch.cuda.synchronize()
    _check(eager, x, offsets, k, range(0, 500, 7))
    plan.close()


def test_ragged_full_config(gpu, torch_cuda):
    torch = torch_cuda
    k = oracle.synth_ragged_log2(20000, 6, 16, 0x5EED0003)
    offsets, total = wgfft.ragged_layout(k)
    x = wgfft.synth_uniform((total,), 0x5EED0013)
    plan = wgfft.RaggedPlan(gpu, offsets, k)
    tx = torch.from_numpy(x).cuda()
    ty = torch.empty_like(tx)
    plan.exec_device(tx, ty, mode=0)
    torch.cuda.synchronize()
    y = ty.cpu().numpy()
    rng = np.random.default_rng(3)
    _check(y, x, offsets, k, rng.choice(20000, 300, replace=False))
    tr = torch.empty_like(ty)
    plan.exec_device(ty, tr, inverse=True)
    torch.cuda.synchronize()
    assert oracle.rel_l2(tr.cpu().numpy(), x) <= 5e-6
    plan.close()


def test_ragged_rejects_bad_tables(gpu):
    with pytest.raises(wgfft.WgfftError):
        wgfft.RaggedPlan(gpu, [0], [5])  # below 2^6
    with pytest.raises(wgfft.WgfftError):
        wgfft.RaggedPlan(gpu, [0, 32], [6, 6])  # overlap
    with pytest.raises(wgfft.WgfftError):
        wgfft.RaggedPlan(gpu, [1], [6])  # misaligned


def test_plan_lifetime_follows_handle(torch_cuda):
    """Closing the GpuFft frees its open mixed-length plans (wgfft_destroy),
    after which the Python plan is closed too: no use after free whichever is
    closed first."""
    torch = torch_cuda
    k = np.array([6, 10, 14], dtype=np.uint32)
    offsets, total = wgfft.ragged_layout(k)
    g = wgfft.GpuFft.new()
    p1 = wgfft.RaggedPlan(g, offsets, k)
    p2 = wgfft.RaggedPlan(g, offsets, k)
    tx = torch.from_numpy(oracle.synth_uniform((total,), 0x5EED0305)).cuda()
    ty = torch.empty_like(tx)
    p1.exec_device(tx, ty)
    p2.close()  # plan first, then the handle
    g.close()  # frees p1 with the handle
    p1.close()  # no-op
    with pytest.raises(wgfft.WgfftError):
        p1.exec_device(tx, ty)
    with pytest.raises(wgfft.WgfftError):
        wgfft.RaggedPlan(g, offsets, k)
    torch.cuda.synchronize()


def test_plan_rejects_short_buffers(gpu, torch_cuda):
    torch = torch_cuda
    k = np.array([8, 12], dtype=np.uint32)
    offsets, total = wgfft.ragged_layout(k)
    plan = wgfft.RaggedPlan(gpu, offsets, k)
    ok = torch.zeros(total, dtype=torch.complex64, device="cuda")
    short = torch.zeros(total - 2, dtype=torch.complex64, device="cuda")
    with pytest.raises(wgfft.WgfftError):
        plan.exec_device(short, ok)
    with pytest.raises(wgfft.WgfftError):
        plan.exec_device(ok, short)
    plan.close()
