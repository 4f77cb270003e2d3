"""bench.py's host-side pieces, on the CPU: the one-GPU contract of the
command line, the mixed batch's length draw (the parity tests' draw), and
the traffic figures the mixed and large-N legs derive from a plan."""
import os
import subprocess
import sys

import numpy as np

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


def test_mixed_lengths_are_the_oracles_draw():
    for seed in (0x5EED0003, 0x5EED0004):
        k = bench._ragged_log2(2000, 6, 16, seed)
        assert np.array_equal(k, oracle.synth_ragged_log2(2000, 6, 16, seed))
        assert k.min() >= 6 and k.max() <= 16


def test_traffic_floor_counts_the_two_pass_scratch():
    # 2^14 stays in the LDS (16 B/sample), 2^15 and 2^16 add their scratch
    # round trip (32 B/sample)
    assert bench.traffic_floor_bytes([6, 14]) == 16 * (64 + 16384)
    assert bench.traffic_floor_bytes([15, 16]) == 32 * (32768 + 65536)
    k = oracle.synth_ragged_log2(20000, 6, 16, 0x5EED0003)
    samples = int(sum(1 << int(v) for v in k))
    floor = bench.traffic_floor_bytes(k)
    assert 16 * samples < floor < 32 * samples
    assert abs(floor / samples - 28.0) < 0.2  # three quarters of the samples are in 2^15 / 2^16 units


def test_pass_count_reads_the_plan():
    two = "multipass N=4194304 passes=R2048(16x8x16,G16,direct,halves),R2048(16x8x16,G16,direct,halves)"
    three = "multipass N=67108864 passes=R512(16x16x2,G16),R512(16x16x2,G16),R256(16x16,G32,direct)"
    assert bench.pass_count(two) == 2
    assert bench.pass_count(three) == 3
    assert bench.pass_count("lds N=4096 plan=16x16x16 E=16 G=1") == 1
