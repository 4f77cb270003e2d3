//! Raw bindings of include/wgfft.h, one `extern "C"` declaration per C
//! prototype (same order, same argument types: `wgfft_t*` is
//! `*mut wgfft_t`, `uint64_t` is `u64`, `int` is `c_int`, `const float*` is
//! `*const f32`, a `hipStream_t` is passed as `*mut c_void`).
#![allow(non_camel_case_types)]

use std::ffi::{c_char, c_int, c_void};

/// Opaque library handle (`struct wgfft_handle`).
#[repr(C)]
pub struct wgfft_t {
    _private: [u8; 0],
}

/// Opaque mixed-length plan (`struct wgfft_ragged_plan`).
#[repr(C)]
pub struct wgfft_ragged_t {
    _private: [u8; 0],
}

pub const WGFFT_OK: c_int = 0;
pub const WGFFT_ERR_INVALID_LENGTH: c_int = -1;
pub const WGFFT_ERR_RAGGED_BATCH: c_int = -2;
pub const WGFFT_ERR_NO_DEVICE: c_int = -3;
pub const WGFFT_ERR_OUT_OF_MEMORY: c_int = -4;
pub const WGFFT_ERR_HIP: c_int = -5;
pub const WGFFT_ERR_INVALID_ARGUMENT: c_int = -6;

pub const WGFFT_FORWARD: c_int = -1;
pub const WGFFT_INVERSE: c_int = 1;

extern "C" {
    pub fn wgfft_device_available() -> c_int;
    pub fn wgfft_create(out: *mut *mut wgfft_t, device: c_int) -> c_int;
    pub fn wgfft_destroy(h: *mut wgfft_t);
    pub fn wgfft_last_error() -> *const c_char;
    pub fn wgfft_name() -> *const c_char;
    pub fn wgfft_validate_lengths(lens: *const u64, count: u64, n_out: *mut u64) -> c_int;
    pub fn wgfft_exec_host(h: *mut wgfft_t, input: *const f32, out: *mut f32, n: u64, batch: u64, dir: c_int) -> c_int;
    pub fn wgfft_exec_host_rows(
        h: *mut wgfft_t,
        in_rows: *const *const f32,
        out_rows: *const *mut f32,
        n: u64,
        batch: u64,
        dir: c_int,
    ) -> c_int;
    pub fn wgfft_exec_device(
        h: *mut wgfft_t,
        d_in: *const f32,
        d_out: *mut f32,
        n: u64,
        batch: u64,
        dir: c_int,
        stream: *mut c_void,
    ) -> c_int;
    pub fn wgfft_exec_device_noalloc(
        h: *mut wgfft_t,
        d_in: *const f32,
        d_out: *mut f32,
        n: u64,
        batch: u64,
        dir: c_int,
        stream: *mut c_void,
    ) -> c_int;
    pub fn wgfft_exec_strided_device(
        h: *mut wgfft_t,
        d_in: *const f32,
        d_out: *mut f32,
        n: u64,
        batch: u64,
        istride: u64,
        idist: u64,
        ostride: u64,
        odist: u64,
        dir: c_int,
        stream: *mut c_void,
    ) -> c_int;
    pub fn wgfft_exec_windowed_host(h: *mut wgfft_t, x: *const f32, w: *const f32, y: *mut f32, n: u64, batch: u64)
        -> c_int;
    pub fn wgfft_exec_windowed_device(
        h: *mut wgfft_t,
        d_x: *const f32,
        d_w: *const f32,
        d_y: *mut f32,
        n: u64,
        batch: u64,
        stream: *mut c_void,
    ) -> c_int;
    pub fn wgfft_vjp_windowed_device(
        h: *mut wgfft_t,
        d_x: *const f32,
        d_w: *const f32,
        d_gy: *const f32,
        d_gx: *mut f32,
        d_gw: *mut f32,
        n: u64,
        batch: u64,
        stream: *mut c_void,
    ) -> c_int;
    pub fn wgfft_fwd_adjoint_masked_device(
        h: *mut wgfft_t,
        d_x: *const f32,
        d_w: *const f32,
        d_mask: *const f32,
        d_y: *mut f32,
        d_gx: *mut f32,
        d_gw: *mut f32,
        d_loss: *mut f32,
        n: u64,
        batch: u64,
        stream: *mut c_void,
    ) -> c_int;
    pub fn wgfft_filter_device(
        h: *mut wgfft_t,
        d_x: *const f32,
        d_s: *const f32,
        d_y: *mut f32,
        n: u64,
        batch: u64,
        seed: u32,
        sigma: f32,
        stream: *mut c_void,
    ) -> c_int;
    pub fn wgfft_add_noise_device(
        h: *mut wgfft_t,
        d_x: *mut f32,
        n: u64,
        batch: u64,
        seed: u32,
        sigma: f32,
        stream: *mut c_void,
    ) -> c_int;
    pub fn wgfft_ragged_create(
        h: *mut wgfft_t,
        offsets: *const u64,
        log2n: *const u32,
        count: u64,
        out: *mut *mut wgfft_ragged_t,
    ) -> c_int;
    pub fn wgfft_ragged_destroy(p: *mut wgfft_ragged_t);
    pub fn wgfft_ragged_exec_device(
        p: *mut wgfft_ragged_t,
        d_in: *const f32,
        d_out: *mut f32,
        dir: c_int,
        mode: c_int,
        stream: *mut c_void,
    ) -> c_int;
    pub fn wgfft_ragged_samples(p: *const wgfft_ragged_t) -> u64;
    pub fn wgfft_ragged_describe(p: *const wgfft_ragged_t, desc: *mut c_char, desc_len: u64) -> c_int;
    pub fn wgfft_bench_device(
        h: *mut wgfft_t,
        d_in: *const f32,
        d_out: *mut f32,
        n: u64,
        batch: u64,
        dir: c_int,
        warmup: c_int,
        iters: c_int,
        sec_per_iter: *mut f64,
    ) -> c_int;
    pub fn wgfft_autotune(h: *mut wgfft_t, n: u64, batch: u64, desc: *mut c_char, desc_len: u64) -> c_int;
    pub fn wgfft_plan_describe(h: *mut wgfft_t, n: u64, desc: *mut c_char, desc_len: u64) -> c_int;
    pub fn wgfft_variant_count(n: u64) -> c_int;
    pub fn wgfft_select_variant(h: *mut wgfft_t, n: u64, index: c_int, desc: *mut c_char, desc_len: u64) -> c_int;
    pub fn wgfft_pass_count(n: u64) -> c_int;
    pub fn wgfft_pass_variant_count(n: u64, pass: c_int) -> c_int;
    pub fn wgfft_select_pass_variant(
        h: *mut wgfft_t,
        n: u64,
        pass: c_int,
        index: c_int,
        desc: *mut c_char,
        desc_len: u64,
    ) -> c_int;
    pub fn wgfft_select_split(
        h: *mut wgfft_t,
        n: u64,
        log2r: *const c_int,
        passes: c_int,
        desc: *mut c_char,
        desc_len: u64,
    ) -> c_int;
    pub fn wgfft_adjoint_variant_count(n: u64) -> c_int;
    pub fn wgfft_select_adjoint_variant(h: *mut wgfft_t, n: u64, index: c_int, desc: *mut c_char, desc_len: u64)
        -> c_int;
    pub fn wgfft_synth_uniform(out: *mut f32, nfloats: u64, seed: u64);
    pub fn wgfft_set_alloc_cap(bytes: u64);
}
