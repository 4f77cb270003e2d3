"""Parity of the HIP path (through the C-ABI) with the oracle.

Bars (written next to each check):
  * FP64 oracle: relative L2 <= 5e-6 for N <= 2^20, <= 2e-5 above;
  * reference restatement (tests/gpu_fft_test.rs:6): max |diff| < 1e-3 where
    max|Y| * 1.2e-7 << 1e-3;
  * round trips: 1e-3 (tests/ifft_test.rs:9), 1e-4 (src/lib.rs:225).
"""
import numpy as np
import pytest

import oracle
import wgfft

pytestmark = pytest.mark.gpu

GOLDEN = "tests/golden/fft_golden.npz"


def rel_bar(n):
    return 5e-6 if n <= (1 << 20) else 2e-5


def fp64_refs(x):
    """FP64 forward and 1/N-scaled inverse of x (the oracle's radix-2 FFT)."""
    n = x.shape[-1]
    return oracle.fft64(x, -1), oracle.fft64(x, +1) / n


def check_fwd_inv(gpu, x, refs=None):
    """Whole-vector forward and inverse of the batch against FP64; returns the
    forward output and the two errors."""
    batch, n = x.shape
    want_f, want_i = refs if refs is not None else fp64_refs(x)
    y = gpu.fft_array(x)
    e_f = oracle.rel_l2(y, want_f)
    assert e_f <= rel_bar(n), (n, batch, e_f)
    yi = gpu.fft_array(x, inverse=True)
    e_i = oracle.rel_l2(yi, want_i)
    assert e_i <= rel_bar(n), (n, batch, e_i)
    return y, e_f, e_i


# ----------------------------------------------------------- golden vectors
@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32, 64, 256, 1024, 4096])
def test_golden_vectors(gpu, n):
    g = np.load(GOLDEN, allow_pickle=False)
    x = g[f"x_{n}"]
    assert oracle.rel_l2(gpu.fft_array(x), g[f"fwd_{n}"]) <= 5e-6
    if f"inv_{n}" in g:
        assert oracle.rel_l2(gpu.fft_array(x, inverse=True), g[f"inv_{n}"]) <= 5e-6


# --------------------------------------------------- every length, ragged batches
@pytest.mark.parametrize("log2n", list(range(0, 17)))
def test_all_lengths_vs_fp64_and_reference(gpu, log2n):
    n = 1 << log2n
    batch = max(1, min(37, (1 << 19) // n))  # odd batch exercises the group tail
    x = oracle.synth_uniform((batch, n), 0x5EED0100 + log2n)
    y, _, _ = check_fwd_inv(gpu, x)
    ref = oracle.ref_fft_f32(x)
    assert np.abs(y - ref).max() < 1e-3  # reference restatement, README bar
    single = gpu.fft_array(x[-1:])
    assert np.array_equal(single[0], y[-1])  # batch output == single output


@pytest.mark.parametrize("log2n", list(range(5, 15)))
def test_every_compiled_variant(log2n):
    """Each compiled single-kernel variant (radix split, E, G, store and
    twiddle-source flags), pinned in turn, matches FP64 forward and inverse
    on a batch that leaves a partial group, and the windowed forward."""
    n = 1 << log2n
    count = wgfft.GpuFft.variant_count(n)
    assert count >= 1
    x = oracle.synth_uniform((max(3, min(45, (1 << 18) // n)), n), 0x5EED0150 + log2n)
    w, _ = oracle.synth_window_mask(n, 0x5EED0151 + log2n)
    want_w = oracle.windowed_fwd64(x[:2], w)
    for i in range(count):
        g = wgfft.GpuFft.new()
        desc = g.select_variant(n, i)
        assert g.describe(n) == desc
        try:
            check_fwd_inv(g, x)
            err = oracle.rel_l2(np.stack(g.fft_windowed(list(x[:2]), w)), want_w)
            assert err <= rel_bar(n), ("windowed", err)
        except AssertionError as e:
            raise AssertionError(f"variant {i}: {desc}: {e}") from None
        finally:
            g.close()


def test_select_variant_errors(gpu):
    with pytest.raises(wgfft.WgfftError):
        gpu.select_variant(4096, wgfft.GpuFft.variant_count(4096))
    with pytest.raises(wgfft.WgfftError):
        gpu.select_variant(1 << 20, 0)  # served by global passes
    assert wgfft.GpuFft.variant_count(3) == 0
    with pytest.raises(wgfft.WgfftError):
        gpu.select_adjoint_variant(4096, wgfft.GpuFft.adjoint_variant_count(4096))
    with pytest.raises(wgfft.WgfftError):
        gpu.select_adjoint_variant(1 << 20, 0)  # no fused kernel at this length
    with pytest.raises(wgfft.WgfftError) as e:
        gpu.select_adjoint_variant(3, 0)
    assert e.value.code == wgfft.WGFFT_ERR_INVALID_LENGTH


@pytest.mark.parametrize("log2n,batch", [(17, 3), (18, 2), (20, 2), (22, 2)])
def test_multipass_lengths(gpu, log2n, batch):
    n = 1 << log2n
    x = oracle.synth_uniform((batch, n), 0x5EED0200 + log2n)
    check_fwd_inv(gpu, x)


@pytest.mark.parametrize("log2n,batch", [(17, 3), (20, 2), (22, 2), (23, 1), (24, 1), (25, 1)])
def test_every_global_pass_variant(log2n, batch):
    """Every compiled candidate of every global pass (LDS-staged and direct
    I/O, with and without the register prefetch), pinned in turn with the
    other passes on their defaults, matches FP64 forward and inverse (at
    2^24 and 2^25 the three passes of the default plans)."""
    n = 1 << log2n
    x = oracle.synth_uniform((batch, n), 0x5EED0180 + log2n)
    refs = fp64_refs(x)
    for p in range(wgfft.GpuFft.pass_count(n)):
        for i in range(wgfft.GpuFft.pass_variant_count(n, p)):
            g = wgfft.GpuFft.new()
            try:
                desc = g.select_pass_variant(n, p, i)
                check_fwd_inv(g, x, refs)
            except AssertionError as e:
                raise AssertionError(f"pass {p} variant {i}: {desc}: {e}") from None
            finally:
                g.close()


@pytest.mark.parametrize("log2n,batch", [(17, 2), (20, 2)])
def test_every_first_pass_variant_windowed(log2n, batch):
    """The first global pass applies the window: every candidate of that pass,
    through the windowed entry, matches the FP64 transform of w * x."""
    n = 1 << log2n
    x = oracle.synth_uniform((batch, n), 0x5EED0190 + log2n)
    w, _ = oracle.synth_window_mask(n, 0x5EED0191 + log2n)
    want = oracle.windowed_fwd64(x, w)
    for i in range(wgfft.GpuFft.pass_variant_count(n, 0)):
        g = wgfft.GpuFft.new()
        try:
            desc = g.select_pass_variant(n, 0, i)
            err = oracle.rel_l2(np.stack(g.fft_windowed(list(x), w)), want)
            assert err <= rel_bar(n), (desc, err)
        finally:
            g.close()


@pytest.mark.parametrize("n,batch", [(8192, 256), (4096, 512), (16384, 64)])
def test_autotuned_windowed_entry(n, batch):
    """The autotuner keeps the windowed entry's fastest variant apart when it
    is not the plain transform's (reported after "; windowed: "): the windowed
    call then runs that variant (bit-identical to pinning it), matches the
    FP64 transform of w * x, and pinning a variant drops the separate choice."""
    x = oracle.synth_uniform((batch, n), 0x5EED0270 + n)
    w, _ = oracle.synth_window_mask(n, 0x5EED0271 + n)
    want = oracle.windowed_fwd64(x[:4], w)
    g = wgfft.GpuFft.new()
    try:
        desc = g.autotune(n, batch)
        y = np.stack(g.fft_windowed(list(x), w))
        assert oracle.rel_l2(y[:4], want) <= rel_bar(n), desc
        plain = desc.split(";")[0]
        chosen = desc.split("; windowed: ")[1].split(";")[0] if "; windowed: " in desc else plain
        pinned = wgfft.GpuFft.new()
        try:
            names = [pinned.select_variant(n, i) for i in range(wgfft.GpuFft.variant_count(n))]
            assert chosen in names, (chosen, names)
            pinned.select_variant(n, names.index(chosen))
            assert np.array_equal(np.stack(pinned.fft_windowed(list(x), w)), y)
        finally:
            pinned.close()
        assert g.describe(n).startswith(plain)
        g.select_variant(n, 0)
        assert "; windowed: " not in g.describe(n)
        assert oracle.rel_l2(np.stack(g.fft_windowed(list(x[:4]), w)), want) <= rel_bar(n)
    finally:
        g.close()


# The autotuner's picks on MI355X at about 2^24 samples per call, which a
# handle uses without tuning (a typo in the runtime's table would silently
# fall back to the first registered variant).
DEFAULT_PLANS = {
    5: 'plan=8x4 E=16 G=128',
    6: 'plan=8x8 E=16 G=64',
    7: 'plan=4x4x8 E=16 G=32',
    8: 'plan=8x4x8 E=16 G=16',
    9: 'plan=8x8x8 E=16 G=8',
    10: 'plan=16x4x16 E=16 G=4 prio',
    11: 'plan=16x8x16 E=16 G=2 prio',
    12: 'plan=16x16x16 E=16 G=1 nt-store prio 3/4-prefetch',
    13: 'plan=2x16x16x16 E=16 G=1 split2',
    14: 'plan=2x16x8x4x16 E=16 G=1 split2 cached-loads',
    15: 'passes=R256(16x16,G32,direct),R128(4x8x4,G32,direct)',
    16: 'passes=R256(16x16,G32,direct),R256(16x16,G32,direct)',
    17: 'passes=R512(8x8x8,G16,direct),R256(16x16,G32,direct)',
    18: 'passes=R512(8x8x8,G16,direct),R512(8x8x8,G16,direct)',
    19: 'passes=R1024(16x4x16,G16,direct,halves),R512(8x8x8,G16,direct)',
    20: 'passes=R1024(16x4x16,G16,direct,halves),R1024(16x4x16,G16,direct,halves)',
    21: 'passes=R2048(16x8x16,G16,direct,halves),R1024(16x4x16,G16,direct,halves)',
    22: 'passes=R2048(16x8x16,G16,direct,halves),R2048(16x8x16,G16,direct,halves)',
    23: 'passes=R4096(16x16x16,G4),R2048(16x8x16,G16,direct,halves)',
    24: 'passes=R256(16x16,G32,direct),R256(16x16,G32,direct),R256(16x16,G32,direct)',
    25: 'passes=R512(8x8x8,G16,direct),R256(16x16,G32,direct),R256(16x16,G16,direct)',
    26: 'passes=R512(16x16x2,G16),R512(16x16x2,G16),R256(16x16,G32,direct)',
}


def test_default_plans():
    """Without autotuning, a handle uses the plans measured fastest on MI355X."""
    g = wgfft.GpuFft.new()
    try:
        for log2n, want in DEFAULT_PLANS.items():
            assert want in g.describe(1 << log2n), (log2n, g.describe(1 << log2n))
        x = oracle.synth_uniform((2, 1 << 16), 0x5EED0260)
        check_fwd_inv(g, x)
        # 512 .. 2048: small batches (at most 4 per CU) run one transform per
        # workgroup, until a variant is pinned or tuned
        for log2n in (9, 10, 11):
            d = g.describe(1 << log2n)
            assert "; batches <= " in d and " G=1 " in d.split("; batches <= ")[1], d
        assert "; batches <= " not in g.describe(1 << 12)
    finally:
        g.close()
    p = wgfft.GpuFft.new()
    try:
        p.select_variant(1024, 0)
        assert "; batches <= " not in p.describe(1024)
        x = oracle.synth_uniform((8, 1024), 0x5EED0261)
        check_fwd_inv(p, x)
    finally:
        p.close()


@pytest.mark.parametrize("log2n", [17, 20, 23])
def test_multipass_autotuned(log2n):
    """The global-pass combination the autotuner keeps matches FP64."""
    n = 1 << log2n
    g = wgfft.GpuFft.new()
    try:
        desc = g.autotune(n, 2)
        assert desc.startswith("multipass"), desc
        check_fwd_inv(g, oracle.synth_uniform((2, n), 0x5EED0250 + log2n))
    finally:
        g.close()


@pytest.mark.parametrize("n,batch", [(4096, 1), (4096, 2), (4096, 3), (1024, 9), (256, 33), (64, 129),
                                     (32, 257), (8192, 3), (16384, 2)])
def test_group_tails(gpu, n, batch):
    x = oracle.synth_uniform((batch, n), 0x5EED0300 + batch)
    check_fwd_inv(gpu, x)


def test_known_answers(gpu):
    for n in (16, 256, 4096, 1 << 15):
        k = np.arange(n)
        imp = np.zeros((1, n), np.complex64)
        imp[0, 0] = 1
        assert np.abs(gpu.fft_array(imp) - 1).max() < 1e-6
        sh = np.zeros((1, n), np.complex64)
        sh[0, 7 % n] = 1
        assert np.abs(gpu.fft_array(sh)[0] - np.exp(-2j * np.pi * k * (7 % n) / n)).max() < 2e-6
        e = np.exp(2j * np.pi * 5 * k / n).astype(np.complex64)[None]
        want = np.zeros(n)
        want[5 % n] = n
        assert np.abs(gpu.fft_array(e)[0] - want).max() / n < 2e-6


# ------------------------------------- the reference's tests, on the GPU path
def make_input(n):
    i = np.arange(n, dtype=np.float32)
    return (np.sin(i * np.float32(0.1)) + 1j * np.cos(i * np.float32(0.07))).astype(np.complex64)


def generate_test_signal(n):
    t = np.arange(n, dtype=np.float32) / np.float32(n)
    two_pi = np.float32(2.0 * np.pi)
    s = np.float32(0.7) * np.sin(two_pi * np.float32(10.0) * t) + np.float32(0.3) * np.sin(
        two_pi * np.float32(50.0) * t)
    return s.astype(np.float32).astype(np.complex64)


def hann(n):
    i = np.arange(n, dtype=np.float32)
    return (np.float32(0.5) * (np.float32(1.0) - np.cos(np.float32(2.0 * np.pi) * i / np.float32(n - 1)))).astype(
        np.float32)


def test_gpu_fft_matches_rustfft(gpu):
    # tests/gpu_fft_test.rs:17-42
    x = make_input(1024)
    out = gpu.fft([x])
    assert len(out) == 1 and out[0].shape == (1024,)
    assert np.abs(out[0] - oracle.fft64(x)[0]).max() < 1e-3


def test_ifft_roundtrip_and_properties(gpu):
    # tests/ifft_test.rs:22-107
    x = make_input(1024)
    spec = gpu.fft([x])
    rec = gpu.ifft(spec)
    assert np.abs(rec[0] - x).max() < 1e-3
    assert np.abs(rec[0] - oracle.ifft64(oracle.fft64(x))[0]).max() < 1e-3
    spec2 = gpu.fft(rec)
    assert np.abs(spec2[0] - spec[0]).max() < 1e-3


def test_unit_roundtrip_consistency(gpu):
    # src/lib.rs:207-227: ramp 0.1 i, N = 1024, error < 1e-4
    x = (np.arange(1024, dtype=np.float32) * np.float32(0.1)).astype(np.complex64)
    rec = gpu.ifft(gpu.fft([x]))[0]
    assert np.abs(rec - x).max() < 1e-4


def test_windowed_reference_tests(gpu):
    # tests/windowed_fft_test.rs:10-133 at N = 2048, windowing on the CPU as
    # the reference does, and through the fused windowed entry
    n = 2048
    sig = generate_test_signal(n)
    win = (sig * hann(n)).astype(np.complex64)
    spec = gpu.fft([win])[0]
    assert np.abs(gpu.ifft([spec])[0] - win).max() < 1e-3
    ratio = np.sum(np.abs(spec) ** 2) / np.sum(np.abs(gpu.fft([sig])[0]) ** 2)
    assert 0.1 < ratio < 10.0
    p = np.abs(spec) ** 2
    peaks = [i for i in range(1, n // 2) if p[i] > p[i - 1] and p[i] > p[i + 1] and p[i] > 10.0]
    assert any(8 <= f < 12 for f in peaks) and any(45 <= f < 55 for f in peaks)
    fused = gpu.fft_windowed([sig], hann(n))[0]
    assert np.abs(fused - spec).max() < 1e-3
    assert oracle.rel_l2(fused, oracle.windowed_fwd64(sig[None], hann(n))[0]) < 5e-6


def test_impulse_and_batch_semantics(gpu):
    # examples/test_parallel.rs:90-128, examples/test_batch_processing.rs:130-211
    n = 256
    imp = np.zeros(n, np.complex64)
    imp[0] = 1
    assert np.abs(np.abs(gpu.fft([imp])[0]) - 1).max() < 1e-3
    sigs = [oracle.synth_uniform((n,), 40 + i) for i in range(8)]
    batch = gpu.fft(sigs)
    assert len(batch) == 8
    for i, s in enumerate(sigs):
        assert np.abs(batch[i] - gpu.fft([s])[0]).max() < 1e-4
    const = [np.full(n, 1.0 + 0j, np.complex64)] * 2
    rt = gpu.ifft(gpu.fft(const))
    assert np.abs(rt[0] - const[0]).max() < 1e-4


def test_error_contract_on_device(gpu):
    assert gpu.fft([]) == [] and gpu.ifft([]) == []
    with pytest.raises(wgfft.WgfftError) as e:
        gpu.fft([np.zeros(8, np.complex64), np.zeros(16, np.complex64)])
    assert str(e.value) == "All input vectors in a batch must have the same length"
    with pytest.raises(wgfft.WgfftError) as e:
        gpu.fft([np.zeros(12, np.complex64)])
    assert str(e.value) == "Transform length must be a non-zero power of two"
    # the same errors for rows that are converted before the call, and for a
    # ragged batch whose first row is the longer one
    with pytest.raises(wgfft.WgfftError) as e:
        gpu.ifft([[0j] * 12, [0j] * 12])
    assert str(e.value) == "Transform length must be a non-zero power of two"
    with pytest.raises(wgfft.WgfftError) as e:
        gpu.fft([np.zeros(16, np.complex64), [0j] * 8])
    assert str(e.value) == "All input vectors in a batch must have the same length"
    with pytest.raises(wgfft.WgfftError) as e:
        gpu.fft([np.zeros(0, np.complex64)] * 2)
    assert e.value.code == wgfft.WGFFT_ERR_INVALID_LENGTH
    assert gpu.name()
    assert wgfft.is_gpu_available()


def test_list_inputs_of_every_form(gpu):
    """fft() takes any sequence of signals: contiguous complex64 rows go to
    the library where they lie, other rows (complex128, strided views,
    Python lists) are converted first; every form gives the same result."""
    x = wgfft.synth_uniform((5, 256), 0x5EED0900)
    want = gpu.fft(list(x))
    assert oracle.rel_l2(np.stack(want), oracle.fft64(x, -1)) <= rel_bar(256)
    wide = np.zeros((5, 512), np.complex64)
    wide[:, ::2] = x
    for form in (x, tuple(x), [r.astype(np.complex128) for r in x], list(wide[:, ::2]), [list(r) for r in x],
                 [x[0], x[1].astype(np.complex128), x[2], list(x[3]), x[4]]):
        got = gpu.fft(form)
        assert len(got) == 5 and all(g.dtype == np.complex64 and np.array_equal(g, w) for g, w in zip(got, want))
    back = gpu.ifft(want)
    assert oracle.rel_l2(np.stack(back), x) <= rel_bar(256)


def test_alloc_cap_out_of_memory(gpu):
    fresh = wgfft.GpuFft.new()
    try:
        wgfft.set_alloc_cap(1 << 16)
        with pytest.raises(wgfft.WgfftError) as e:
            fresh.fft_array(np.zeros((64, 1024), np.complex64))
        assert e.value.code == wgfft.WGFFT_ERR_OUT_OF_MEMORY
    finally:
        wgfft.set_alloc_cap(0)
        fresh.close()
    x = oracle.synth_uniform((2, 1024), 3)
    check_fwd_inv(gpu, x)


# ------------------------------------------------ full-size configurations
@pytest.mark.parametrize("n,batch", [(4096, 4096), (256, 65536), (16384, 1024)])
def test_headline_shapes_vs_fp64(gpu, parity, n, batch):
    x = wgfft.synth_uniform((batch, n), 0x5EED0002)
    y = gpu.fft_array(x)
    err = oracle.rel_l2(y, oracle.fft64(x))
    print(f"N={n} batch={batch} fwd rel-L2 vs FP64 = {err:.3e}")
    assert err <= 5e-6
    rt = gpu.fft_array(y, inverse=True)
    e_rt = oracle.rel_l2(rt, x)
    parity[f"C2 {n}x{batch}"] = {"fwd_rel_l2_vs_fp64": err, "round_trip_rel_l2": e_rt, "bar": 5e-6}
    assert e_rt <= 5e-6


def test_c1_forward_then_inverse(gpu, parity):
    x = wgfft.synth_uniform((8, 1024), 0x5EED0001)
    y = gpu.fft_array(x)
    e_f = oracle.rel_l2(y, oracle.fft64(x))
    r = gpu.fft_array(y, inverse=True)
    e_i = oracle.rel_l2(r, oracle.ifft64(y.astype(np.complex128)))
    ref_diff = float(np.abs(y - oracle.ref_fft_f32(x)).max())
    parity["C1 1024x8"] = {"fwd_rel_l2_vs_fp64": e_f, "inv_rel_l2_vs_fp64": e_i,
                           "round_trip_max_abs": float(np.abs(r - x).max()),
                           "fwd_max_abs_vs_reference_restatement": ref_diff, "bar": 5e-6}
    assert e_f <= 5e-6 and e_i <= 5e-6
    assert ref_diff < 1e-3
    assert np.abs(r - x).max() < 1e-4


def test_c3_large_n_2pow22(gpu, parity):
    n, batch = 1 << 22, 16
    x = oracle.synth_sinusoids(n, batch, 0x5EED0003)
    y = gpu.fft_array(x)
    err = oracle.rel_l2(y, oracle.fft64(x))
    print(f"N=2^22 batch=16 fwd rel-L2 vs FP64 = {err:.3e}")
    assert err <= 2e-5
    r = gpu.fft_array(y, inverse=True)
    e_rt = oracle.rel_l2(r, x)
    parity["C3 2^22x16"] = {"fwd_rel_l2_vs_fp64": err, "round_trip_rel_l2": e_rt, "bar": 2e-5,
                            "plan": gpu.describe(n)}
    assert e_rt <= 2e-5
    # the plan the bench times (autotuned radix split and pass variants)
    g = wgfft.GpuFft.new()
    try:
        plan = g.autotune(n, batch)
        y2 = g.fft_array(x)
        e2 = oracle.rel_l2(y2, oracle.fft64(x))
        e2_rt = oracle.rel_l2(g.fft_array(y2, inverse=True), x)
        parity["C3 2^22x16 autotuned"] = {"fwd_rel_l2_vs_fp64": e2, "round_trip_rel_l2": e2_rt, "bar": 2e-5,
                                          "plan": plan}
        assert e2 <= 2e-5 and e2_rt <= 2e-5
    finally:
        g.close()


@pytest.mark.parametrize("log2n", [24, 25])
def test_three_pass_plans_vs_fp64(parity, log2n):
    """The default three-pass plans of 2^24 and 2^25 and the plans the
    autotuner keeps there, forward and inverse, as whole vectors against the
    FP64 FFT on the large-N workload's signal (32 tones + noise)."""
    n = 1 << log2n
    x = oracle.synth_sinusoids(n, 1, 0x5EED0003)
    refs = fp64_refs(x)
    for tuned in (False, True):
        g = wgfft.GpuFft.new()
        try:
            plan = g.autotune(n, 1) if tuned else g.describe(n)
            assert plan.startswith("multipass"), plan
            if not tuned:
                assert wgfft.GpuFft.pass_count(n) == 3, plan
            _, e_f, e_i = check_fwd_inv(g, x, refs)
            parity[f"2^{log2n}x1 {'autotuned' if tuned else 'default'}"] = {
                "fwd_rel_l2_vs_fp64": e_f, "inv_rel_l2_vs_fp64": e_i, "bar": 2e-5, "plan": plan}
        finally:
            g.close()


def test_c3_large_n_2pow26(gpu, parity):
    """N = 2^26, batch 1, on the workload's signal (32 random-frequency,
    random-phase unit tones + Gaussian noise sigma 0.1): the whole forward and
    inverse against the FP64 FFT, for the default three-pass plan and the plan
    the autotuner keeps; plus bins against the FP64 direct DFT, Parseval and
    the round trip."""
    n = 1 << 26
    x = oracle.synth_sinusoids(n, 1, 0x5EED0003)
    want_f, want_i = fp64_refs(x)
    bins = np.array([0, 1, 12345, n // 2, n - 1, 33554431 - 777], dtype=np.uint64)
    want_bins = oracle.dft_bins64(x[0], bins)
    scale = np.sqrt(np.sum(np.abs(x.astype(np.complex128)) ** 2))  # ||x||, the RMS size of a bin |Y_k|
    assert np.abs(want_f[0, bins.astype(np.int64)] - want_bins).max() / scale < 1e-9  # the two FP64 oracles agree
    for tuned in (False, True):
        g = wgfft.GpuFft.new()
        try:
            plan = g.autotune(n, 1) if tuned else g.describe(n)
            y, e_f, e_i = check_fwd_inv(g, x, (want_f, want_i))
            e_bins = float(np.abs(y[0, bins.astype(np.int64)] - want_bins).max() / scale)
            assert e_bins < 2e-5
            py = np.sum(np.abs(y.astype(np.complex128)) ** 2)
            e_pars = float(abs(py / (n * scale**2) - 1))
            assert e_pars < 1e-5
            e_rt = oracle.rel_l2(g.fft_array(y, inverse=True), x)
            assert e_rt <= 2e-5
            parity["C3 2^26x1" + (" autotuned" if tuned else "")] = {
                "fwd_rel_l2_vs_fp64": e_f, "inv_rel_l2_vs_fp64": e_i, "bins_max_err_over_norm_vs_fp64_dft": e_bins,
                "parseval_rel_err": e_pars, "round_trip_rel_l2": e_rt, "bar": 2e-5, "plan": plan}
        finally:
            g.close()


# ------------------------------------------------------- device entry points
def test_device_api_inplace_and_stream(gpu, torch_cuda):
    torch = torch_cuda
    x = wgfft.synth_uniform((64, 4096), 9)
    t = torch.from_numpy(x).cuda()
    out = torch.empty_like(t)
    gpu.exec_device(t, out)
    torch.cuda.synchronize()
    y = out.cpu().numpy()
    assert oracle.rel_l2(y, oracle.fft64(x)) <= 5e-6
    gpu.exec_device(t, t)  # in place
    torch.cuda.synchronize()
    assert np.array_equal(t.cpu().numpy(), y)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        z = torch.empty_like(out)
        gpu.exec_device(out, z, inverse=True)
    s.synchronize()
    assert oracle.rel_l2(z.cpu().numpy(), x) <= 5e-6


@pytest.mark.parametrize("n,batch", [(1024, 600), (4096, 4096 + 37), (16384, 700), (64, 300000), (1 << 21, 9),
                                     (1 << 22, 3)])
def test_host_pipeline_matches_device(gpu, torch_cuda, n, batch):
    """Host-buffer calls of 4 MB and more stream through pinned staging in
    chunks of whole transforms (a quarter of the call, 1 to 32 MiB), four
    slots deep (here with partial last chunks, at 2^21 two transforms per
    chunk and at 2^22 one): bit-identical to the device entry point,
    forward, inverse, in place and windowed."""
    torch = torch_cuda
    L = wgfft._lib
    x = wgfft.synth_uniform((batch, n), 0x5EED0700 + n)
    t = torch.from_numpy(x).cuda()
    d = torch.empty_like(t)
    for inverse in (False, True):
        gpu.exec_device(t, d, inverse=inverse)
        torch.cuda.synchronize()
        want = d.cpu().numpy()
        assert np.array_equal(gpu.fft_array(x, inverse=inverse), want), inverse
        buf = x.copy()  # in == out through the C-ABI
        L.check(gpu._L.wgfft_exec_host(gpu._h, buf.ctypes.data, buf.ctypes.data, n, batch,
                                           L.WGFFT_INVERSE if inverse else L.WGFFT_FORWARD))
        assert np.array_equal(buf, want), inverse
    rows = [0, batch // 2, batch - 1]
    assert oracle.rel_l2(want[rows], oracle.fft64(x[rows], +1) / n) <= rel_bar(n)
    w, _ = oracle.synth_window_mask(n, 0x5EED0701 + n)
    gpu.exec_windowed_device(t, torch.from_numpy(w).cuda(), d)
    torch.cuda.synchronize()
    yw = np.empty_like(x)
    L.check(gpu._L.wgfft_exec_windowed_host(gpu._h, x.ctypes.data, w.ctypes.data, yw.ctypes.data, n, batch))
    assert np.array_equal(yw, d.cpu().numpy())


def test_host_calls_from_several_threads(gpu):
    """fft()/ifft() release the GIL for the library call, so Python threads
    can call one handle at once: the handle's lock serializes them and every
    result is bit-identical to the same call made alone, on a shared handle
    and on a handle per thread, with calls small (one staged chunk) and
    large (pipelined)."""
    import threading

    cases = [(wgfft.synth_uniform((b, n), 0x5EED0A00 + n), inv) for n, b, inv in
             ((1024, 64, False), (4096, 600, True), (16384, 128, False), (256, 3000, True))]
    want = [(gpu.ifft if inv else gpu.fft)(list(x)) for x, inv in cases]
    own = [wgfft.GpuFft.new() for _ in range(2)]
    errors = []

    def worker(k, h):
        try:
            for rep in range(6):
                i = (k + rep) % len(cases)
                x, inv = cases[i]
                got = (h.ifft if inv else h.fft)(list(x))
                if not all(np.array_equal(g, w) for g, w in zip(got, want[i])):
                    errors.append((k, rep, i))
        except Exception as e:  # reported below with the worker's index
            errors.append((k, repr(e)))

    try:
        threads = [threading.Thread(target=worker, args=(k, gpu if k < 4 else own[k - 4])) for k in range(6)]
        for t in threads:
            t.start()
        for t in threads:
            t.join(timeout=120)
        assert not any(t.is_alive() for t in threads)
        assert not errors, errors
    finally:
        for h in own:
            h.close()


@pytest.mark.parametrize("n,batch", [(1024, 8), (256, 700), (4096, 4096 + 37), (1 << 22, 3), (1 << 23, 2)])
def test_host_rows_match_flat(gpu, n, batch):
    """wgfft_exec_host_rows, the entry behind fft()/ifft() on lists of long
    signals: the staging copies gather each signal from its own buffer and
    scatter each result into its own (on the calling thread below 256 KB,
    split over the copy threads above, with rows straddling the threads'
    shares; below 4 MB as one chunk on the handle's stream, above through the
    pipeline; at 2^23 each transform is larger than a staging chunk and moves
    on its own).  Bit-identical to
    wgfft_exec_host on the flattened batch, forward and inverse, also in
    place (out_rows == in_rows) on buffers allocated last row first."""
    L = wgfft._lib
    x = wgfft.synth_uniform((batch, n), 0x5EED0800 + n)
    rows = [x[b].copy() for b in range(batch)]  # separate buffers
    for inverse in (False, True):
        dirn = L.WGFFT_INVERSE if inverse else L.WGFFT_FORWARD
        want = gpu.fft_array(x, inverse=inverse)
        got = gpu.ifft(rows) if inverse else gpu.fft(rows)
        assert len(got) == batch and all(np.array_equal(g, w) for g, w in zip(got, want)), inverse
        bufs = [None] * batch
        for b in reversed(range(batch)):  # allocated last row first
            bufs[b] = rows[b].copy()
        ptrs = np.array([b.ctypes.data for b in bufs], dtype=np.uint64)
        L.check(gpu._L.wgfft_exec_host_rows(gpu._h, ptrs.ctypes.data, ptrs.ctypes.data, n, batch, dirn))
        assert all(np.array_equal(b, w) for b, w in zip(bufs, want)), inverse
    assert np.array_equal(rows[batch // 2], x[batch // 2])  # inputs untouched
    null_rows = np.zeros(batch, dtype=np.uint64)
    with pytest.raises(wgfft.WgfftError):
        L.check(gpu._L.wgfft_exec_host_rows(gpu._h, null_rows.ctypes.data, null_rows.ctypes.data, n, batch, 1))


@pytest.mark.parametrize("n,batch", [(256, 33), (4096, 17), (1 << 17, 3)])
def test_strided_layouts(gpu, torch_cuda, n, batch):
    """Every other element of a wider buffer in, a transposed layout out, and
    the plain layout through the same entry: the same results as the
    contiguous device call, forward and inverse."""
    torch = torch_cuda
    x = oracle.synth_uniform((batch, n), 0x5EED0900 + n)
    t = torch.from_numpy(x).cuda()
    wide = torch.zeros((batch, 2 * n), dtype=torch.complex64, device="cuda")
    wide[:, ::2] = t
    src = wide[:, ::2]  # strides (2n, 2)
    for inverse in (False, True):
        ref = torch.empty_like(t)
        gpu.exec_device(t, ref, inverse=inverse)
        dst = torch.empty((n, batch), dtype=torch.complex64, device="cuda").t()  # strides (1, batch)
        gpu.exec_strided_device(src, dst, inverse=inverse)
        plain = torch.empty_like(t)
        gpu.exec_strided_device(t, plain, inverse=inverse)
        torch.cuda.synchronize()
        want = ref.cpu().numpy()
        assert np.array_equal(dst.cpu().numpy(), want), inverse
        assert np.array_equal(plain.cpu().numpy(), want), inverse
    assert oracle.rel_l2(want, oracle.fft64(x, +1) / n) <= rel_bar(n)
    with pytest.raises(wgfft.WgfftError):  # a zero stride
        wgfft._lib.check(gpu._L.wgfft_exec_strided_device(gpu._h, t.data_ptr(), t.data_ptr(), n, batch, 0, n, 1, n,
                                                          -1, None))


@pytest.mark.parametrize("n,batch,pitch", [(8, 50, 50), (4096, 3, 7), (256, 33, 40), (1 << 17, 2, 2)])
def test_column_layouts(gpu, torch_cuda, n, batch, pitch):
    """Signals stored as the columns of a matrix (adjacent signals: distance
    1, element stride = the row pitch >= batch) in and out, through the tiled
    gather and scatter, with edges where batch and n are not multiples of the
    32 x 32 tile: bit-identical to the contiguous device call, forward and
    inverse, and against FP64."""
    torch = torch_cuda
    x = oracle.synth_uniform((batch, n), 0x5EED0940 + n + batch)
    t = torch.from_numpy(x).cuda()
    base = torch.zeros((n, pitch), dtype=torch.complex64, device="cuda")
    base[:, :batch] = t.t()
    src = base[:, :batch].t()  # strides (1, pitch)
    assert src.stride() == (1, pitch)
    for inverse in (False, True):
        ref = torch.empty_like(t)
        gpu.exec_device(t, ref, inverse=inverse)
        out_base = torch.zeros((n, pitch + 1), dtype=torch.complex64, device="cuda")
        dst = out_base[:, :batch].t()  # strides (1, pitch + 1)
        gpu.exec_strided_device(src, dst, inverse=inverse)
        rows = torch.empty_like(t)
        gpu.exec_strided_device(src, rows, inverse=inverse)  # column layout in, rows out
        torch.cuda.synchronize()
        want = ref.cpu().numpy()
        assert np.array_equal(dst.cpu().numpy(), want), inverse
        assert np.array_equal(rows.cpu().numpy(), want), inverse
        assert not out_base[:, batch:].abs().sum().item()  # nothing written outside the columns
        if not inverse:
            assert oracle.rel_l2(want, oracle.fft64(x, -1)) <= rel_bar(n)


def _energy(torch, t, rows_per_chunk):
    """sum |t|^2 in FP64, a chunk of rows at a time (no full-size temporaries)."""
    total = 0.0
    for r0 in range(0, t.shape[0], rows_per_chunk):
        total += float(torch.view_as_real(t[r0:r0 + rows_per_chunk]).double().pow(2).sum())
    return total


@pytest.mark.parametrize("n,batch", [(4096, (1 << 20) + 37), (1 << 22, 1025)])
def test_batches_beyond_2pow32_elements(gpu, torch_cuda, n, batch):
    """More than 2^32 complex elements in one call (34-35 GB in and as much
    out; the batched kernel at 4096, the global passes at 2^22 with the batch
    in 2 GiB chunks): every transform keeps its energy (Parseval over the
    whole batch), the transforms at the start, on both sides of the 2^31- and
    2^32-element offsets and at the end match FP64, and the in-place inverse
    returns them."""
    torch = torch_cuda
    assert n * batch > (1 << 32)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0x5EED0950 + n)
    x = torch.rand((batch, 2 * n), device="cuda", generator=gen)
    x.mul_(2.0).sub_(1.0)
    x = x.view(torch.complex64)
    y = torch.empty_like(x)
    try:
        gpu.exec_device(x, y)
        torch.cuda.synchronize()
        chunk = max(1, (1 << 27) // n)
        e_x, e_y = _energy(torch, x, chunk), _energy(torch, y, chunk)
        assert abs(e_y / (n * e_x) - 1.0) <= 1e-5, (e_x, e_y)
        rows = sorted({0, (1 << 31) // n - 1, (1 << 31) // n, (1 << 32) // n - 1, (1 << 32) // n, batch - 1})
        xr = x[rows].cpu().numpy()
        assert oracle.rel_l2(y[rows].cpu().numpy(), oracle.fft64(xr, -1)) <= rel_bar(n)
        gpu.exec_device(y, y, inverse=True)
        torch.cuda.synchronize()
        assert oracle.rel_l2(y[rows].cpu().numpy(), xr) <= 2e-5
    finally:
        del x, y
        torch.cuda.empty_cache()


def test_scratch_shared_across_streams(gpu, torch_cuda):
    """Calls of one handle on two streams that both use the handle's scratch
    (global passes at 2^20, the strided gather/scatter buffer, the VJP's
    reductions) give the results of the same calls made one after another."""
    torch = torch_cuda
    n, batch = 1 << 20, 4
    xa = torch.from_numpy(oracle.synth_uniform((batch, n), 0x5EED0910)).cuda()
    xb = torch.from_numpy(oracle.synth_uniform((batch, n), 0x5EED0911)).cuda()
    m, mb = 4096, 64
    xs = torch.from_numpy(oracle.synth_uniform((mb, m), 0x5EED0912)).cuda()
    w = torch.from_numpy(np.hanning(m).astype(np.float32)).cuda()
    gy = torch.from_numpy(oracle.synth_uniform((mb, m), 0x5EED0913)).cuda()
    ra, rb, rs = torch.empty_like(xa), torch.empty_like(xb), torch.empty_like(xs)
    gx_ref, gw_ref = torch.empty_like(xs), torch.empty(m, dtype=torch.float32, device="cuda")
    gpu.exec_device(xa, ra)
    gpu.exec_device(xb, rb, inverse=True)
    gpu.exec_strided_device(xs.t().contiguous().t(), rs)
    gpu.vjp_windowed_device(xs, w, gy, gx_ref, gw_ref)
    torch.cuda.synchronize()
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    ya, yb = torch.empty_like(xa), torch.empty_like(xb)
    ys = torch.empty((m, mb), dtype=torch.complex64, device="cuda").t()
    gx, gw = torch.empty_like(xs), torch.empty_like(gw_ref)
    xst = xs.t().contiguous().t()
    torch.cuda.synchronize()
    for _ in range(3):
        gpu.exec_device(xa, ya, stream=sa.cuda_stream)
        gpu.exec_strided_device(xst, ys, stream=sb.cuda_stream)
        gpu.exec_device(xb, yb, inverse=True, stream=sb.cuda_stream)
        gpu.vjp_windowed_device(xs, w, gy, gx, gw, stream=sa.cuda_stream)
    torch.cuda.synchronize()
    assert torch.equal(ya, ra) and torch.equal(yb, rb)
    assert torch.equal(ys, rs)
    assert torch.equal(gx, gx_ref) and torch.equal(gw, gw_ref)


def test_scratch_ordering_host_calls_and_stream_switches(torch_cuda):
    """A fresh handle used from one stream records nothing; its scratch use is
    then ordered by a host-blocking call (host transform of the same
    two-pass length while a stream call is still in flight), by the first
    switch to another stream (a wait for the device), and from then on by
    events.  Every result equals the same calls made one after another on a
    separate handle."""
    torch = torch_cuda
    n, batch = 1 << 20, 8
    xa = wgfft.synth_uniform((batch, n), 0x5EED0920)
    xb = wgfft.synth_uniform((batch, n), 0x5EED0921)
    m, mb = 4096, 64
    xs = torch.from_numpy(wgfft.synth_uniform((mb, m), 0x5EED0922)).cuda()
    gy = torch.from_numpy(wgfft.synth_uniform((mb, m), 0x5EED0923)).cuda()
    w = torch.from_numpy(np.hanning(m).astype(np.float32)).cuda()
    ta, tb = torch.from_numpy(xa).cuda(), torch.from_numpy(xb).cuda()

    ref = wgfft.GpuFft.new()
    try:
        ra, rb = torch.empty_like(ta), torch.empty_like(tb)
        ref.exec_device(ta, ra)
        ref.exec_device(tb, rb, inverse=True)
        gx_ref, gw_ref = torch.empty_like(xs), torch.empty(m, dtype=torch.float32, device="cuda")
        ref.vjp_windowed_device(xs, w, gy, gx_ref, gw_ref)
        torch.cuda.synchronize()
        host_ref = ref.fft_array(xb)
    finally:
        ref.close()

    g = wgfft.GpuFft.new()
    try:
        s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
        ya, yb, ya2 = torch.empty_like(ta), torch.empty_like(tb), torch.empty_like(ta)
        gx, gw = torch.empty_like(xs), torch.empty_like(gw_ref)
        torch.cuda.synchronize()
        g.exec_device(ta, ya, stream=s1.cuda_stream)  # left in flight
        host = g.fft_array(xb)                        # host-blocking, same scratch
        g.exec_device(tb, yb, inverse=True, stream=s1.cuda_stream)
        g.vjp_windowed_device(xs, w, gy, gx, gw, stream=s2.cuda_stream)  # first switch
        g.exec_device(ta, ya2, stream=s1.cuda_stream)  # ordered by an event from here on
        torch.cuda.synchronize()
        assert np.array_equal(host, host_ref)
        assert torch.equal(ya, ra) and torch.equal(ya2, ra) and torch.equal(yb, rb)
        assert torch.equal(gx, gx_ref) and torch.equal(gw, gw_ref)
    finally:
        g.close()


def test_windowed_device_c5_shape(gpu, torch_cuda, parity):
    torch = torch_cuda
    n, batch = 8192, 1024
    x = wgfft.synth_uniform((batch, n), 0x5EED0004)
    w, _ = oracle.synth_window_mask(n, 0x5EED0014)
    tx, tw = torch.from_numpy(x).cuda(), torch.from_numpy(w).cuda()
    ty = torch.empty_like(tx)
    gpu.exec_windowed_device(tx, tw, ty)
    torch.cuda.synchronize()
    err = oracle.rel_l2(ty.cpu().numpy(), oracle.windowed_fwd64(x, w))
    parity["C5 8192x1024 windowed forward"] = {"rel_l2_vs_fp64": err, "bar": 5e-6}
    assert err <= 5e-6


def test_device_wrappers_reject_bad_operands(gpu, torch_cuda):
    """Short, wrongly typed or host-resident windows, masks and gradients, and
    outputs of another shape, are rejected before any kernel sees them."""
    torch = torch_cuda
    n, batch = 256, 2
    x = torch.zeros((batch, n), dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    w = torch.ones(n, dtype=torch.float32, device="cuda")
    bad_w = (torch.ones(n - 1, device="cuda"), torch.ones(n, dtype=torch.float16, device="cuda"), torch.ones(n),
             torch.ones(2 * n, device="cuda")[::2])
    for bw in bad_w:
        with pytest.raises(wgfft.WgfftError):
            gpu.exec_windowed_device(x, bw, y)
        with pytest.raises(wgfft.WgfftError):
            gpu.filter_device(x, bw, y)
        with pytest.raises(wgfft.WgfftError):
            gpu.vjp_windowed_device(x, bw, x, y, w)
    loss = torch.zeros(1, device="cuda")
    with pytest.raises(wgfft.WgfftError):
        gpu.fwd_adjoint_masked_device(x, w, torch.ones(n - 1, device="cuda"), None, y, w.clone(), loss)
    with pytest.raises(wgfft.WgfftError):
        gpu.vjp_windowed_device(x, w, x[:1], y, w.clone())  # gy of another shape
    with pytest.raises(wgfft.WgfftError):
        gpu.exec_device(x, torch.empty((batch, n // 2), dtype=torch.complex64, device="cuda"))
    gpu.exec_windowed_device(x, w, y)  # the well-formed call still runs
    torch.cuda.synchronize()


@pytest.mark.parametrize("log2n,split", [(21, (7, 7, 7)), (22, (8, 7, 7)), (23, (8, 8, 7))])
def test_pinned_splits_vs_fp64(log2n, split):
    """Radix splits pinned with wgfft_select_split (three passes at
    lengths where the default has two) match FP64 forward and inverse, in
    place and out of place."""
    n = 1 << log2n
    x = oracle.synth_uniform((2, n), 0x5EED0270 + log2n)
    refs = fp64_refs(x)
    g = wgfft.GpuFft.new()
    try:
        desc = g.select_split(n, split)
        assert desc.count("(") == len(split), desc
        check_fwd_inv(g, x, refs)
        buf = x.copy()
        wgfft._lib.check(g._L.wgfft_exec_host(g._h, buf.ctypes.data, buf.ctypes.data, n, 2, wgfft.WGFFT_FORWARD))
        assert oracle.rel_l2(buf, refs[0]) <= rel_bar(n), desc
    finally:
        g.close()


def test_bench_device_matches_event_timing(gpu, torch_cuda):
    """wgfft_bench_device (GpuFftTrait::benchmark_gpu_only, src/lib.rs:109-143:
    warm-up calls with waits, then back-to-back calls between HIP events on
    the handle's stream) agrees with HIP-event timing of the same device calls
    on the caller's stream, and leaves the transform in the output."""
    torch = torch_cuda
    import time

    n, batch = 4096, 2048
    xs = wgfft.synth_uniform((batch, n), 0x5EED0B00)
    x = torch.from_numpy(xs).cuda()
    y = torch.empty_like(x)
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < 0.3:  # settle the shader clock
        for _ in range(20):
            gpu.exec_device(x, y)
        torch.cuda.synchronize()
    ev = []
    lib = []
    for _ in range(3):
        lib.append(gpu.bench_device(x, y, warmup=3, iters=50) * 1e3)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(50):
            gpu.exec_device(x, y)
        e1.record()
        torch.cuda.synchronize()
        ev.append(e0.elapsed_time(e1) / 50)
    a, b = min(lib), min(ev)
    assert a > 0 and abs(a - b) / b < 0.12, (lib, ev)
    assert oracle.rel_l2(y[:4].cpu().numpy(), oracle.fft64(xs[:4])) <= 5e-6
    with pytest.raises(wgfft.WgfftError):
        gpu.bench_device(x, y, iters=0)


@pytest.mark.parametrize("n,batch", [(8, 100), (256, 33), (4096, 64), (16384, 6)])
def test_padded_rows_one_pass(torch_cuda, n, batch):
    """Padded rows (unit stride, even distances >= n) are transformed in one
    pass with no scratch: the call still works under an allocation cap that
    the gather/scatter path would exceed, and matches the contiguous call bit
    for bit; odd distances and rows overlapping in another layout take the
    gather/scatter path and match too."""
    torch = torch_cuda
    x = oracle.synth_uniform((batch, n), 0x5EED0920 + n)
    t = torch.from_numpy(x).cuda()
    g = wgfft.GpuFft.new()
    ref_fwd = None
    try:
        for inverse in (False, True):
            ref = torch.empty_like(t)
            g.exec_device(t, ref, inverse=inverse)
            if not inverse:
                ref_fwd = ref
            src = torch.zeros((batch, n + 64), dtype=torch.complex64, device="cuda")
            src[:, :n] = t
            dst = torch.zeros((batch, n + 32), dtype=torch.complex64, device="cuda")
            torch.cuda.synchronize()
            wgfft.set_alloc_cap(1 << 20 if n * batch * 8 > (1 << 20) else 0)
            try:
                g.exec_strided_device(src[:, :n], dst[:, :n], inverse=inverse)
            finally:
                wgfft.set_alloc_cap(0)
            torch.cuda.synchronize()
            assert torch.equal(dst[:, :n], ref), inverse
            assert not dst[:, n:].abs().any()  # the padding is untouched
            # in place on padded rows
            inpl = src.clone()
            g.exec_strided_device(inpl[:, :n], inpl[:, :n], inverse=inverse)
            torch.cuda.synchronize()
            assert torch.equal(inpl[:, :n], ref) and torch.equal(inpl[:, n:], src[:, n:]), inverse
            # odd input distance: gathered
            odd = torch.zeros((batch, n + 1), dtype=torch.complex64, device="cuda")
            odd[:, :n] = t
            out = torch.empty_like(t)
            g.exec_strided_device(odd[:, :n], out, inverse=inverse)
            torch.cuda.synchronize()
            assert torch.equal(out, ref), inverse
        assert oracle.rel_l2(ref.cpu().numpy(), oracle.fft64(x, +1) / n) <= 5e-6
    finally:
        g.close()
    if n * batch * 8 > (1 << 20):
        # on a fresh handle (no scratch yet) under a 1 MiB cap: padded rows
        # run, an element stride of 2 needs the scratch copy and is refused
        g = wgfft.GpuFft.new()
        try:
            g.describe(n)
            strided = torch.zeros((batch, 2 * n), dtype=torch.complex64, device="cuda")[:, ::2]
            out = torch.zeros((batch, n + 32), dtype=torch.complex64, device="cuda")
            torch.cuda.synchronize()
            wgfft.set_alloc_cap(1 << 20)
            try:
                g.exec_strided_device(src[:, :n], out[:, :n])
                with pytest.raises(wgfft.WgfftError) as e:
                    g.exec_strided_device(strided, out[:, :n])
                assert e.value.code == wgfft.WGFFT_ERR_OUT_OF_MEMORY
            finally:
                wgfft.set_alloc_cap(0)
            torch.cuda.synchronize()
            assert torch.equal(out[:, :n], ref_fwd)
        finally:
            g.close()
