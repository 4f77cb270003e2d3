//! Batched complex64 FFT on an AMD MI355X, behind the `GpuFft` /
//! `FftExecutor` API of wgls-rs-fft (its `src/lib.rs:32-45` trait,
//! `:285-488` executor and `:946-950` `Default`).
//!
//! The transforms run in the hand-written gfx950 HIP kernels of
//! `libwgfft.so`; this crate is the host side: it validates a batch with the
//! reference's rules and messages and passes each `Vec<Complex<f32>>` of the
//! batch (already interleaved `re, im` f32) where it lies to the C-ABI of
//! `include/wgfft.h` (`wgfft_exec_host_rows`: the library's staging copies
//! gather the rows and scatter the results into new `Vec`s).  Forward is unnormalised; the inverse is 1/N scaled,
//! as in the reference.  There is no batch cap (the reference panics past
//! `min(128 MiB / 8N, 1024)` transforms, `src/lib.rs:736-739`).
//!
//! ```no_run
//! use num_complex::Complex;
//! use wgls_rs_fft::GpuFft;
//!
//! let fft = GpuFft::new().expect("no HIP device");
//! let spectrum = fft.fft(&[vec![Complex::new(1.0f32, 0.0); 1024]]).unwrap();
//! let signal = fft.ifft(&spectrum).unwrap();
//! ```

pub mod benchmark;
pub mod sys;

use std::any::Any;
use std::error::Error;
use std::ffi::{c_char, c_int, c_void, CStr};
use std::marker::PhantomData;

use num_complex::Complex;

/// The reference's plugin interface (`src/lib.rs:32-45`): every executor of
/// the benchmark and leaderboard code implements it.
pub trait FftExecutor {
    fn name(&self) -> &str;
    fn fft(&self, inputs: &[Vec<Complex<f32>>]) -> Result<Vec<Vec<Complex<f32>>>, Box<dyn Error>>;
    fn ifft(&self, inputs: &[Vec<Complex<f32>>]) -> Result<Vec<Vec<Complex<f32>>>, Box<dyn Error>>;
    /// Get a reference to the underlying type for downcasting.
    fn as_any(&self) -> &dyn Any;
}

/// The message of the library's last failure on this thread, as the error
/// (the validation messages are byte-for-byte the reference's,
/// `src/lib.rs:508,530`).
fn last_error() -> Box<dyn Error> {
    let msg = unsafe { CStr::from_ptr(sys::wgfft_last_error()) }.to_string_lossy().into_owned();
    msg.into()
}

fn check(rc: c_int) -> Result<(), Box<dyn Error>> {
    if rc == sys::WGFFT_OK {
        Ok(())
    } else {
        Err(last_error())
    }
}

fn desc_string(buf: &[u8]) -> String {
    let end = buf.iter().position(|&b| b == 0).unwrap_or(buf.len());
    String::from_utf8_lossy(&buf[..end]).into_owned()
}

/// GPU FFT executor on one HIP device.
///
/// The library handle owns the device stream, the per-length plans
/// (kernel choice and twiddle tables) and the device scratch.  It serialises
/// its calls with an internal lock and is tied to no thread, so a `GpuFft`
/// may be moved to another thread (`Send`), like the reference's executor; it
/// is not shared between threads (`!Sync`, as the reference's `RefCell`
/// cache makes its own).
pub struct GpuFft {
    h: *mut sys::wgfft_t,
    name: String,
}

// SAFETY: the handle is plain heap state plus a HIP stream; wgfft calls
// select the handle's device themselves (hipSetDevice) and lock the handle,
// so using it from whichever thread owns the GpuFft is sound.
unsafe impl Send for GpuFft {}

impl GpuFft {
    /// Executor on HIP device 0 (`GpuFft::new`, `src/lib.rs:285-330`); an
    /// error, not a panic, when there is no usable device.
    pub fn new() -> Result<Self, Box<dyn Error>> {
        Self::with_device(0)
    }

    /// Executor on HIP device `device`.
    pub fn with_device(device: i32) -> Result<Self, Box<dyn Error>> {
        let mut h: *mut sys::wgfft_t = std::ptr::null_mut();
        check(unsafe { sys::wgfft_create(&mut h, device as c_int) })?;
        let name = unsafe { CStr::from_ptr(sys::wgfft_name()) }.to_string_lossy().into_owned();
        Ok(Self { h, name })
    }

    /// Whether a HIP device is usable (`src/lib.rs:381-389`).
    pub fn is_gpu_available() -> bool {
        unsafe { sys::wgfft_device_available() != 0 }
    }

    /// Forward FFT of a batch of equal-length, power-of-two signals
    /// (`src/lib.rs:433-438`); an empty batch gives an empty result.
    pub fn fft(&self, inputs: &[Vec<Complex<f32>>]) -> Result<Vec<Vec<Complex<f32>>>, Box<dyn Error>> {
        self.transform(inputs, sys::WGFFT_FORWARD)
    }

    /// Inverse FFT, scaled by 1/N (`src/lib.rs:483-488`).
    pub fn ifft(&self, inputs: &[Vec<Complex<f32>>]) -> Result<Vec<Vec<Complex<f32>>>, Box<dyn Error>> {
        self.transform(inputs, sys::WGFFT_INVERSE)
    }

    /// `FFT(w * x)` per signal with a real window `w` of the signal length,
    /// multiplied as the samples are loaded (the windowed FFT the reference
    /// builds from `apply_hann_window` + `fft`, `tests/test_utils.rs:39-50`).
    pub fn fft_windowed(
        &self,
        inputs: &[Vec<Complex<f32>>],
        window: &[f32],
    ) -> Result<Vec<Vec<Complex<f32>>>, Box<dyn Error>> {
        let n = Self::validate(inputs)?;
        if inputs.is_empty() {
            return Ok(Vec::new());
        }
        if window.len() != n {
            return Err("window length must equal the signal length".into());
        }
        let flat: Vec<Complex<f32>> = inputs.concat();
        let mut out = vec![Complex::new(0.0f32, 0.0f32); flat.len()];
        check(unsafe {
            sys::wgfft_exec_windowed_host(
                self.h,
                flat.as_ptr() as *const f32,
                window.as_ptr(),
                out.as_mut_ptr() as *mut f32,
                n as u64,
                inputs.len() as u64,
            )
        })?;
        Ok(out.chunks_exact(n).map(|c| c.to_vec()).collect())
    }

    /// The reference's checks in its order (`src/lib.rs:506-533`): all
    /// lengths equal to the first, each a non-zero power of two.  Returns the
    /// common length (0 for an empty batch).
    fn validate(inputs: &[Vec<Complex<f32>>]) -> Result<usize, Box<dyn Error>> {
        let lens: Vec<u64> = inputs.iter().map(|v| v.len() as u64).collect();
        let mut n: u64 = 0;
        check(unsafe { sys::wgfft_validate_lengths(lens.as_ptr(), lens.len() as u64, &mut n) })?;
        Ok(n as usize)
    }

    fn transform(&self, inputs: &[Vec<Complex<f32>>], dir: c_int) -> Result<Vec<Vec<Complex<f32>>>, Box<dyn Error>> {
        let n = Self::validate(inputs)?;
        if inputs.is_empty() {
            return Ok(Vec::new());
        }
        // Complex<f32> is #[repr(C)] { re, im }: each Vec is the interleaved
        // f32 row the library takes; the staging copies gather the input rows
        // and scatter into the output Vecs, so nothing is flattened or re-boxed
        let mut out: Vec<Vec<Complex<f32>>> = (0..inputs.len()).map(|_| vec![Complex::new(0.0f32, 0.0f32); n]).collect();
        let in_rows: Vec<*const f32> = inputs.iter().map(|v| v.as_ptr() as *const f32).collect();
        let out_rows: Vec<*mut f32> = out.iter_mut().map(|v| v.as_mut_ptr() as *mut f32).collect();
        check(unsafe {
            sys::wgfft_exec_host_rows(self.h, in_rows.as_ptr(), out_rows.as_ptr(), n as u64, inputs.len() as u64, dir)
        })?;
        Ok(out)
    }

    /// Transform `batch` signals of length `n` already in device memory,
    /// enqueued on the HIP stream `stream` (null: the default stream).
    ///
    /// # Safety
    /// `d_in` and `d_out` must be 16-byte aligned device allocations of at
    /// least `n * batch` complex values (they may be equal), valid until the
    /// work on `stream` completes.
    pub unsafe fn fft_device(
        &self,
        d_in: *const Complex<f32>,
        d_out: *mut Complex<f32>,
        n: usize,
        batch: usize,
        inverse: bool,
        stream: *mut c_void,
    ) -> Result<(), Box<dyn Error>> {
        let dir = if inverse { sys::WGFFT_INVERSE } else { sys::WGFFT_FORWARD };
        check(sys::wgfft_exec_device(self.h, d_in as *const f32, d_out as *mut f32, n as u64, batch as u64, dir, stream))
    }

    /// GPU-only timing with the reference's `benchmark_gpu_only` protocol
    /// (`src/lib.rs:109-143`): the batch is uploaded once (not timed), then
    /// `warmup_iters` synchronised transforms and `bench_iters` back-to-back
    /// transforms between HIP events on the handle's stream.  Returns seconds
    /// per transform of the whole batch.
    pub fn benchmark_gpu_only(
        &self,
        n: usize,
        batch_size: usize,
        warmup_iters: usize,
        bench_iters: usize,
    ) -> Result<f64, Box<dyn Error>> {
        let flat: Vec<Complex<f32>> = benchmark::deterministic_inputs(n, batch_size).concat();
        let bytes = flat.len() * std::mem::size_of::<Complex<f32>>();
        let d_in = hip::DeviceBuffer::new(bytes)?;
        let d_out = hip::DeviceBuffer::new(bytes)?;
        d_in.upload(&flat)?;
        let mut sec = 0.0f64;
        check(unsafe {
            sys::wgfft_bench_device(
                self.h,
                d_in.ptr as *const f32,
                d_out.ptr as *mut f32,
                n as u64,
                batch_size as u64,
                sys::WGFFT_FORWARD,
                warmup_iters as c_int,
                bench_iters as c_int,
                &mut sec,
            )
        })?;
        Ok(sec)
    }

    /// Time the compiled kernel variants for length `n` on a scratch batch
    /// and keep the fastest for this executor; returns its description.
    pub fn autotune(&self, n: usize, batch: usize) -> Result<String, Box<dyn Error>> {
        let mut buf = [0u8; 512];
        check(unsafe {
            sys::wgfft_autotune(self.h, n as u64, batch as u64, buf.as_mut_ptr() as *mut c_char, buf.len() as u64)
        })?;
        Ok(desc_string(&buf))
    }

    /// The kernel plan this executor uses for length `n`.
    pub fn describe(&self, n: usize) -> Result<String, Box<dyn Error>> {
        let mut buf = [0u8; 512];
        check(unsafe { sys::wgfft_plan_describe(self.h, n as u64, buf.as_mut_ptr() as *mut c_char, buf.len() as u64) })?;
        Ok(desc_string(&buf))
    }

    /// Plan transforms of lengths `2^log2n[i]` (2^6 .. 2^16) packed at
    /// complex offsets `offsets[i]` of one buffer, run in one persistent
    /// launch.  The plan borrows the executor, so it cannot outlive it.
    pub fn ragged_plan(&self, offsets: &[u64], log2n: &[u32]) -> Result<RaggedPlan<'_>, Box<dyn Error>> {
        if offsets.len() != log2n.len() {
            return Err("offsets and log2n differ in length".into());
        }
        let mut p: *mut sys::wgfft_ragged_t = std::ptr::null_mut();
        check(unsafe {
            sys::wgfft_ragged_create(self.h, offsets.as_ptr(), log2n.as_ptr(), offsets.len() as u64, &mut p)
        })?;
        Ok(RaggedPlan { p, _owner: PhantomData })
    }
}

impl Drop for GpuFft {
    fn drop(&mut self) {
        unsafe { sys::wgfft_destroy(self.h) }
    }
}

impl Default for GpuFft {
    /// Panics without a GPU, as the reference's `Default` does
    /// (`src/lib.rs:946-950`).
    fn default() -> Self {
        Self::new().expect("No GPU available for default GpuFft instance")
    }
}

impl FftExecutor for GpuFft {
    fn name(&self) -> &str {
        &self.name
    }
    fn fft(&self, inputs: &[Vec<Complex<f32>>]) -> Result<Vec<Vec<Complex<f32>>>, Box<dyn Error>> {
        GpuFft::fft(self, inputs)
    }
    fn ifft(&self, inputs: &[Vec<Complex<f32>>]) -> Result<Vec<Vec<Complex<f32>>>, Box<dyn Error>> {
        GpuFft::ifft(self, inputs)
    }
    fn as_any(&self) -> &dyn Any {
        self
    }
}

/// Mixed-length batch: one persistent launch over transforms of uneven
/// power-of-two lengths (no counterpart in the reference, which needs one
/// `fft` call per length, `src/lib.rs:527-531`).
pub struct RaggedPlan<'a> {
    p: *mut sys::wgfft_ragged_t,
    _owner: PhantomData<&'a GpuFft>,
}

impl RaggedPlan<'_> {
    /// Total complex samples the plan covers.
    pub fn samples(&self) -> u64 {
        unsafe { sys::wgfft_ragged_samples(self.p) }
    }

    /// Transform every signal of the plan; `mode` 0 is the persistent
    /// launch, 1 its hipGraph replay, 2 one launch per length.
    ///
    /// # Safety
    /// `d_in` and `d_out` must be 16-byte aligned device buffers covering
    /// every signal of the plan, valid until the work on `stream` completes.
    pub unsafe fn exec_device(
        &self,
        d_in: *const Complex<f32>,
        d_out: *mut Complex<f32>,
        inverse: bool,
        mode: i32,
        stream: *mut c_void,
    ) -> Result<(), Box<dyn Error>> {
        let dir = if inverse { sys::WGFFT_INVERSE } else { sys::WGFFT_FORWARD };
        check(sys::wgfft_ragged_exec_device(self.p, d_in as *const f32, d_out as *mut f32, dir, mode as c_int, stream))
    }
}

impl Drop for RaggedPlan<'_> {
    fn drop(&mut self) {
        unsafe { sys::wgfft_ragged_destroy(self.p) }
    }
}

/// The few HIP runtime calls the host side needs (device buffers for the
/// GPU-only benchmark); everything else goes through libwgfft.so.
mod hip {
    use std::error::Error;
    use std::ffi::{c_int, c_void};

    extern "C" {
        fn hipMalloc(ptr: *mut *mut c_void, size: usize) -> c_int;
        fn hipFree(ptr: *mut c_void) -> c_int;
        fn hipMemcpy(dst: *mut c_void, src: *const c_void, size: usize, kind: c_int) -> c_int;
    }

    const HIP_SUCCESS: c_int = 0;
    const HIP_MEMCPY_HOST_TO_DEVICE: c_int = 1;

    pub struct DeviceBuffer {
        pub ptr: *mut c_void,
        bytes: usize,
    }

    impl DeviceBuffer {
        pub fn new(bytes: usize) -> Result<Self, Box<dyn Error>> {
            let mut ptr: *mut c_void = std::ptr::null_mut();
            if unsafe { hipMalloc(&mut ptr, bytes.max(16)) } != HIP_SUCCESS {
                return Err(format!("hipMalloc({bytes}) failed").into());
            }
            Ok(Self { ptr, bytes })
        }

        pub fn upload<T: Copy>(&self, data: &[T]) -> Result<(), Box<dyn Error>> {
            let size = std::mem::size_of_val(data);
            if size > self.bytes {
                return Err("upload larger than the device buffer".into());
            }
            let rc = unsafe { hipMemcpy(self.ptr, data.as_ptr() as *const c_void, size, HIP_MEMCPY_HOST_TO_DEVICE) };
            if rc != HIP_SUCCESS {
                return Err("hipMemcpy (host to device) failed".into());
            }
            Ok(())
        }
    }

    impl Drop for DeviceBuffer {
        fn drop(&mut self) {
            unsafe {
                hipFree(self.ptr);
            }
        }
    }
}
