//! Links libwgfft.so (built by wgls-rs-fft_amd/build.py into
//! wgls-rs-fft_amd/wgfft/_lib) and the HIP runtime it uses.
//!
//!   WGFFT_LIB_DIR  directory of libwgfft.so (default: ../wgfft/_lib beside
//!                  this crate)
//!   ROCM_PATH      ROCm installation (default /opt/rocm), for libamdhip64
use std::env;
use std::path::PathBuf;

fn main() {
    let manifest = PathBuf::from(env::var("CARGO_MANIFEST_DIR").expect("CARGO_MANIFEST_DIR"));
    let lib_dir = env::var("WGFFT_LIB_DIR")
        .map(PathBuf::from)
        .unwrap_or_else(|_| manifest.join("..").join("wgfft").join("_lib"));
    let rocm = PathBuf::from(env::var("ROCM_PATH").unwrap_or_else(|_| "/opt/rocm".to_string()));
    if !lib_dir.join("libwgfft.so").exists() {
        panic!(
            "libwgfft.so not found in {}: build it with `python wgls-rs-fft_amd/build.py` or set WGFFT_LIB_DIR",
            lib_dir.display()
        );
    }
    println!("cargo:rustc-link-search=native={}", lib_dir.display());
    println!("cargo:rustc-link-lib=dylib=wgfft");
    println!("cargo:rustc-link-search=native={}", rocm.join("lib").display());
    println!("cargo:rustc-link-lib=dylib=amdhip64");
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", lib_dir.display());
    println!("cargo:rustc-link-arg=-Wl,-rpath,{}", rocm.join("lib").display());
    println!("cargo:rerun-if-env-changed=WGFFT_LIB_DIR");
    println!("cargo:rerun-if-env-changed=ROCM_PATH");
    println!("cargo:rerun-if-changed={}", lib_dir.join("libwgfft.so").display());
}
