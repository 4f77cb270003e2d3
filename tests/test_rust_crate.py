"""The Rust crate (wgls-rs-fft_amd/rust) cannot be compiled here (no Rust
toolchain in this image), so its C boundary is checked as text: every
`extern "C"` declaration in src/sys.rs matches the prototype of the same
name in include/wgfft.h argument by argument (types mapped C -> Rust), every
header function is bound, and the crate keeps the reference's public surface
(GpuFft, FftExecutor, Default, Send, the benchmark module)."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CRATE = os.path.join(ROOT, "wgls-rs-fft_amd", "rust")
HEADER = os.path.join(ROOT, "include", "wgfft.h")

C_TO_RUST = {
    "int": "c_int", "void": "()", "uint64_t": "u64", "uint32_t": "u32", "float": "f32", "double": "f64",
    "char*": "*mut c_char", "const char*": "*const c_char", "void*": "*mut c_void",
    "float*": "*mut f32", "const float*": "*const f32", "double*": "*mut f64",
    "uint64_t*": "*mut u64", "const uint64_t*": "*const u64", "const uint32_t*": "*const u32",
    "const int*": "*const c_int", "wgfft_t*": "*mut wgfft_t", "wgfft_t**": "*mut *mut wgfft_t",
    "wgfft_ragged_t*": "*mut wgfft_ragged_t", "const wgfft_ragged_t*": "*const wgfft_ragged_t",
    "wgfft_ragged_t**": "*mut *mut wgfft_ragged_t",
    "const float* const*": "*const *const f32", "float* const*": "*const *mut f32",
}


def _norm_c(t):
    t = " ".join(t.replace("*", " * ").split())
    return t.replace(" *", "*").replace("* *", "**").replace("**", "**")


def header_prototypes():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"^([\w\s\*]+?)\b(wgfft_\w+)\s*\(([^)]*)\)\s*;", text, flags=re.M):
        ret = _norm_c(m.group(1).strip())
        args = []
        for a in m.group(3).split(","):
            a = a.strip()
            if a == "void" or not a:
                continue
            typ = _norm_c(re.sub(r"\b\w+$", "", a).strip())
            args.append(typ)
        out[m.group(2)] = (ret, args)
    return out


def rust_externs():
    text = open(os.path.join(CRATE, "src", "sys.rs")).read()
    block = text[text.index('extern "C" {'):]
    out = {}
    for m in re.finditer(r"pub fn (wgfft_\w+)\s*\(([^)]*)\)\s*(->\s*([^;]+))?;", block, flags=re.S):
        args = []
        for a in m.group(2).split(","):
            a = " ".join(a.split())
            if not a:
                continue
            args.append(a.split(":", 1)[1].strip())
        ret = " ".join(m.group(4).split()) if m.group(4) else "()"
        out[m.group(1)] = (ret, args)
    return out


def test_every_header_function_bound_with_matching_types():
    hdr, rs = header_prototypes(), rust_externs()
    assert len(hdr) >= 30
    assert sorted(hdr) == sorted(rs), set(hdr) ^ set(rs)
    for name, (cret, cargs) in hdr.items():
        rret, rargs = rs[name]
        assert C_TO_RUST[cret] == rret, (name, cret, rret)
        assert [C_TO_RUST[a] for a in cargs] == rargs, (name, cargs, rargs)


def test_error_codes_and_directions_match_header():
    text = open(HEADER).read()
    sysrs = open(os.path.join(CRATE, "src", "sys.rs")).read()
    for name, val in re.findall(r"(WGFFT_(?:OK|ERR_\w+))\s*=\s*(-?\d+)", text):
        assert re.search(rf"pub const {name}: c_int = {val};", sysrs), name
    assert "pub const WGFFT_FORWARD: c_int = -1;" in sysrs and "pub const WGFFT_INVERSE: c_int = 1;" in sysrs


def test_public_surface_of_the_reference_crate():
    lib = open(os.path.join(CRATE, "src", "lib.rs")).read()
    bench = open(os.path.join(CRATE, "src", "benchmark.rs")).read()
    cargo = open(os.path.join(CRATE, "Cargo.toml")).read()
    assert 'name = "wgls_rs_fft"' in cargo and 'name = "wgls-rs-fft"' in cargo
    for item in ("pub trait FftExecutor", "fn name(&self) -> &str;", "fn as_any(&self) -> &dyn Any;",
                 "pub fn new() -> Result<Self, Box<dyn Error>>", "pub fn is_gpu_available() -> bool",
                 "pub fn fft(&self, inputs: &[Vec<Complex<f32>>]) -> Result<Vec<Vec<Complex<f32>>>, Box<dyn Error>>",
                 "pub fn ifft(&self, inputs: &[Vec<Complex<f32>>]) -> Result<Vec<Vec<Complex<f32>>>, Box<dyn Error>>",
                 "impl Default for GpuFft", "unsafe impl Send for GpuFft {}", "impl FftExecutor for GpuFft",
                 "impl Drop for GpuFft", "pub fn fft_windowed(", "No GPU available for default GpuFft instance"):
        assert item in lib, item
    for item in ("pub enum ValidationOutcome", "pub const VALIDATION_TOLERANCE: f32 = 1e-3;",
                 "pub const MAX_TOTAL_SAMPLES: usize = 16 * 1024 * 1024;", "pub fn validate_rival(",
                 "pub fn benchmark_rival(", "pub fn sweep_rival(", "pub fn benchmark_gpu_pipeline(",
                 "pub fn benchmark_gpu_only(", "pub fn choose_batch_size(n: usize) -> usize"):
        assert item in bench, item
    # the plan borrows its executor, so Rust callers cannot free them out of order
    assert "pub struct RaggedPlan<'a>" in lib and "PhantomData<&'a GpuFft>" in lib
