//! The reference crate's own GpuFft tests, run against this crate: parity
//! with rustfft at the README's 1e-3 (tests/gpu_fft_test.rs:17-42), round
//! trips (tests/ifft_test.rs:22-107, src/lib.rs:207-227), the windowed entry
//! and the error contract (src/lib.rs:506-533).  They need a HIP device.

use num_complex::Complex;
use rustfft::FftPlanner;
use wgls_rs_fft::{FftExecutor, GpuFft};

fn make_input(n: usize) -> Vec<Complex<f32>> {
    (0..n).map(|i| Complex::new((i as f32 * 0.1).sin(), (i as f32 * 0.07).cos())).collect()
}

fn max_err(a: &[Complex<f32>], b: &[Complex<f32>]) -> f32 {
    a.iter().zip(b.iter()).map(|(x, y)| (*x - *y).norm()).fold(0.0, f32::max)
}

#[test]
fn gpu_fft_matches_rustfft() {
    let fft = GpuFft::new().expect("GPU required");
    let x = make_input(1024);
    let got = fft.fft(&[x.clone()]).expect("fft failed");
    let mut want = x.clone();
    FftPlanner::<f32>::new().plan_fft_forward(1024).process(&mut want);
    assert_eq!(got.len(), 1);
    assert!(max_err(&got[0], &want) < 1e-3);
}

#[test]
fn ifft_roundtrip_and_batch() {
    let fft = GpuFft::new().expect("GPU required");
    let batch: Vec<Vec<Complex<f32>>> = (0..8).map(|k| make_input(4096).iter().map(|v| *v * (k as f32 + 1.0)).collect()).collect();
    let spec = fft.fft(&batch).expect("fft failed");
    let back = fft.ifft(&spec).expect("ifft failed");
    for (a, b) in batch.iter().zip(back.iter()) {
        assert!(max_err(a, b) < 1e-3);
    }
    let ramp: Vec<Complex<f32>> = (0..1024).map(|i| Complex::new(i as f32 * 0.1, 0.0)).collect();
    let rt = fft.ifft(&fft.fft(&[ramp.clone()]).unwrap()).unwrap();
    assert!(max_err(&rt[0], &ramp) < 1e-4);
}

#[test]
fn windowed_entry_equals_window_then_fft() {
    let fft = GpuFft::new().expect("GPU required");
    let n = 2048;
    let x = make_input(n);
    let w: Vec<f32> = (0..n)
        .map(|i| 0.5 * (1.0 - (2.0 * std::f32::consts::PI * i as f32 / (n as f32 - 1.0)).cos()))
        .collect();
    let xw: Vec<Complex<f32>> = x.iter().zip(w.iter()).map(|(v, s)| *v * *s).collect();
    let fused = fft.fft_windowed(&[x], &w).unwrap();
    let plain = fft.fft(&[xw]).unwrap();
    assert!(max_err(&fused[0], &plain[0]) < 1e-3);
}

#[test]
fn error_contract() {
    let fft = GpuFft::new().expect("GPU required");
    assert!(fft.fft(&[]).unwrap().is_empty());
    let e = fft.fft(&[vec![Complex::new(0.0, 0.0); 8], vec![Complex::new(0.0, 0.0); 16]]).unwrap_err();
    assert_eq!(e.to_string(), "All input vectors in a batch must have the same length");
    let e = fft.fft(&[vec![Complex::new(0.0, 0.0); 12]]).unwrap_err();
    assert_eq!(e.to_string(), "Transform length must be a non-zero power of two");
    let exec: &dyn FftExecutor = &fft;
    assert!(!exec.name().is_empty());
    assert!(exec.as_any().downcast_ref::<GpuFft>().is_some());
}

#[test]
fn executor_moves_to_another_thread() {
    let fft = GpuFft::new().expect("GPU required");
    let out = std::thread::spawn(move || fft.fft(&[make_input(256)]).unwrap().len()).join().unwrap();
    assert_eq!(out, 1);
}
