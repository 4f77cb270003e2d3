//! Timing and validation of any `FftExecutor`, with the semantics of the
//! reference's `src/benchmark.rs:37-297` and the leaderboard's batch-size
//! tiers (`examples/rivalry_leaderboard.rs:159-177`): the same deterministic
//! input, warm-up and iteration counts, the max complex-norm error against a
//! reference executor with a pass at 1e-3, MSamples/s and
//! GFLOPS = 5 B N log2 N / t.

use std::error::Error;
use std::time::Instant;

use num_complex::Complex;

use crate::{FftExecutor, GpuFft};

/// Outcome of validating a rival's output against the reference implementation.
pub enum ValidationOutcome {
    Pass,
    /// Max absolute complex-norm error across all outputs.
    Fail { max_error: f32 },
}

/// Performance and correctness result for one (rival, N, batch_size) triple.
pub struct BenchmarkResult {
    pub rival_name: String,
    pub n: usize,
    pub batch_size: usize,
    pub msamples_per_sec: f64,
    pub gflops: f64,
    pub validation: ValidationOutcome,
}

/// GPU-only (or whole-pipeline) timing result.
pub struct GpuOnlyResult {
    pub rival_name: String,
    pub n: usize,
    pub batch_size: usize,
    pub gpu_duration_sec: f64,
    pub gpu_msamples_per_sec: f64,
    pub gpu_gflops: f64,
}

pub const WARMUP_ITERS: usize = 1;
pub const BENCH_ITERS: usize = 10;
pub const GPU_WARMUP_ITERS: usize = 3;
pub const GPU_BENCH_ITERS: usize = 50;
/// Absolute complex-norm tolerance for correctness.
pub const VALIDATION_TOLERANCE: f32 = 1e-3;
/// Cap on total samples (N x batch) of one benchmark call.
pub const MAX_TOTAL_SAMPLES: usize = 16 * 1024 * 1024;

/// The benchmark signal: sample i of every signal is
/// (t * 0.001, sin(2 pi t) * 0.001) with t = i / N in f32, so bin
/// magnitudes stay O(1) at any N.
pub fn deterministic_inputs(n: usize, batch_size: usize) -> Vec<Vec<Complex<f32>>> {
    let n_f = n as f32;
    let one: Vec<Complex<f32>> = (0..n)
        .map(|i| {
            let t = i as f32 / n_f;
            Complex::new(t * 0.001, (t * std::f32::consts::TAU).sin() * 0.001)
        })
        .collect();
    vec![one; batch_size]
}

fn rates(n: usize, batch_size: usize, seconds: f64) -> (f64, f64) {
    let total = (n * batch_size) as f64;
    (total / seconds / 1e6, 5.0 * total * (n as f64).log2() / seconds / 1e9)
}

/// GPU-only performance of this library's executor: resident data, HIP
/// events around back-to-back transforms (`GpuFft::benchmark_gpu_only`).
pub fn benchmark_gpu_only(gpu_fft: &GpuFft, n: usize, batch_size: usize) -> Result<GpuOnlyResult, Box<dyn Error>> {
    let sec = gpu_fft.benchmark_gpu_only(n, batch_size, GPU_WARMUP_ITERS, GPU_BENCH_ITERS)?;
    let (ms, gf) = rates(n, batch_size, sec);
    Ok(GpuOnlyResult {
        rival_name: FftExecutor::name(gpu_fft).to_string(),
        n,
        batch_size,
        gpu_duration_sec: sec,
        gpu_msamples_per_sec: ms,
        gpu_gflops: gf,
    })
}

/// Whole-pipeline performance (host to device, transform, device to host)
/// of any executor: wall clock of `fft` calls.
pub fn benchmark_gpu_pipeline(
    rival: &dyn FftExecutor,
    n: usize,
    batch_size: usize,
) -> Result<GpuOnlyResult, Box<dyn Error>> {
    let inputs = deterministic_inputs(n, batch_size);
    for _ in 0..WARMUP_ITERS {
        let _ = rival.fft(&inputs);
    }
    let start = Instant::now();
    for _ in 0..BENCH_ITERS {
        let _ = rival.fft(&inputs);
    }
    let sec = start.elapsed().as_secs_f64() / BENCH_ITERS as f64;
    let (ms, gf) = rates(n, batch_size, sec);
    Ok(GpuOnlyResult {
        rival_name: rival.name().to_string(),
        n,
        batch_size,
        gpu_duration_sec: sec,
        gpu_msamples_per_sec: ms,
        gpu_gflops: gf,
    })
}

/// One forward FFT of the benchmark signal by each executor, compared
/// sample by sample; a failing call fails with an infinite error.
pub fn validate_rival(
    rival: &dyn FftExecutor,
    reference: &dyn FftExecutor,
    n: usize,
    batch_size: usize,
) -> ValidationOutcome {
    let inputs = deterministic_inputs(n, batch_size);
    let got = match rival.fft(&inputs) {
        Ok(out) => out,
        Err(_) => return ValidationOutcome::Fail { max_error: f32::INFINITY },
    };
    let want = match reference.fft(&inputs) {
        Ok(out) => out,
        Err(_) => return ValidationOutcome::Fail { max_error: f32::INFINITY },
    };
    validate(&got, &want)
}

/// Wall-clock throughput of `rival` at (n, batch_size) after the warm-up,
/// validated against `reference`.  Panics when an executor fails, as the
/// reference's does (`src/benchmark.rs:198-245`).
pub fn benchmark_rival(
    rival: &dyn FftExecutor,
    reference: &dyn FftExecutor,
    n: usize,
    batch_size: usize,
) -> BenchmarkResult {
    let inputs = deterministic_inputs(n, batch_size);
    for _ in 0..WARMUP_ITERS {
        rival.fft(&inputs).expect("rival fft failed");
    }
    let start = Instant::now();
    for _ in 0..BENCH_ITERS {
        rival.fft(&inputs).expect("rival fft failed");
    }
    let sec = start.elapsed().as_secs_f64() / BENCH_ITERS as f64;
    let (msamples_per_sec, gflops) = rates(n, batch_size, sec);
    let got = rival.fft(&inputs).expect("rival fft failed");
    let want = reference.fft(&inputs).expect("reference fft failed");
    let validation = validate(&got, &want);
    BenchmarkResult { rival_name: rival.name().to_string(), n, batch_size, msamples_per_sec, gflops, validation }
}

/// The fastest of `batch_sizes` (skipping n x batch > MAX_TOTAL_SAMPLES; batch
/// 1 when every size is skipped).
pub fn sweep_rival(
    rival: &dyn FftExecutor,
    reference: &dyn FftExecutor,
    n: usize,
    batch_sizes: &[usize],
) -> BenchmarkResult {
    let mut best: Option<BenchmarkResult> = None;
    for &b in batch_sizes {
        if n * b > MAX_TOTAL_SAMPLES {
            continue;
        }
        let r = benchmark_rival(rival, reference, n, b);
        if best.as_ref().map_or(true, |x| r.msamples_per_sec > x.msamples_per_sec) {
            best = Some(r);
        }
    }
    best.unwrap_or_else(|| benchmark_rival(rival, reference, n, 1))
}

/// Max complex-norm difference of two batches; Pass at <= 1e-3, and an
/// infinite error when the shapes differ.
pub fn validate(result: &[Vec<Complex<f32>>], reference: &[Vec<Complex<f32>>]) -> ValidationOutcome {
    if result.len() != reference.len() {
        return ValidationOutcome::Fail { max_error: f32::INFINITY };
    }
    let mut max_err = 0.0f32;
    for (r, w) in result.iter().zip(reference.iter()) {
        if r.len() != w.len() {
            return ValidationOutcome::Fail { max_error: f32::INFINITY };
        }
        for (a, b) in r.iter().zip(w.iter()) {
            max_err = max_err.max((*a - *b).norm());
        }
    }
    if max_err <= VALIDATION_TOLERANCE {
        ValidationOutcome::Pass
    } else {
        ValidationOutcome::Fail { max_error: max_err }
    }
}

/// The leaderboard's batch per length: 1024 up to N = 1024, 256 up to 16384,
/// 64 up to 65536, 16 up to 262144, else 1; halved while n x batch exceeds
/// MAX_TOTAL_SAMPLES.
pub fn choose_batch_size(n: usize) -> usize {
    let mut batch = if n <= 1024 {
        1024
    } else if n <= 16384 {
        256
    } else if n <= 65536 {
        64
    } else if n <= 262_144 {
        16
    } else {
        1
    };
    while n.saturating_mul(batch) > MAX_TOTAL_SAMPLES && batch > 1 {
        batch /= 2;
    }
    batch.max(1)
}

#[cfg(test)]
mod tests {
    use super::*;

    #[test]
    fn batch_size_tiers() {
        assert_eq!(choose_batch_size(256), 1024);
        assert_eq!(choose_batch_size(1024), 1024);
        assert_eq!(choose_batch_size(16384), 256);
        assert_eq!(choose_batch_size(65536), 64);
        assert_eq!(choose_batch_size(262_144), 16);
        assert_eq!(choose_batch_size(1_048_576), 1);
        for n in [1usize, 2, 256, 1024, 16384, 65536, 262_144, 1_048_576, 1 << 26] {
            let b = choose_batch_size(n);
            assert!(b >= 1);
            assert!(n * b <= MAX_TOTAL_SAMPLES || b == 1);
        }
    }

    #[test]
    fn validate_tolerance_and_shapes() {
        let a = vec![vec![Complex::new(1.0f32, 0.0); 4]];
        let mut b = a.clone();
        assert!(matches!(validate(&a, &b), ValidationOutcome::Pass));
        b[0][2] = Complex::new(1.0, 0.0009);
        assert!(matches!(validate(&a, &b), ValidationOutcome::Pass));
        b[0][2] = Complex::new(1.0, 0.002);
        assert!(matches!(validate(&a, &b), ValidationOutcome::Fail { .. }));
        assert!(matches!(validate(&a, &[]), ValidationOutcome::Fail { max_error } if max_error.is_infinite()));
    }
}
