"""CPU checks of the Rust FFI crate (algebra_linear_amd/rust), which cannot be compiled here (no
Rust toolchain): every function its sys.rs declares exists in include/algebra_linear_hip.h with the
same number of parameters, and every host-slice entry point of the header is bound."""
import os
import re

from algebra_linear_amd import _lib

RUST = os.path.join(os.path.dirname(_lib.HERE), "algebra_linear_amd", "rust")
RAW = open(_lib.HEADER).read()
HEADER = re.sub(r"/\*.*?\*/", "", RAW, flags=re.S)


def decls(text, pattern):
    out = {}
    for m in re.finditer(pattern, text, flags=re.S):
        params = m.group(2).strip()
        out[m.group(1)] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def test_sys_rs_matches_the_header():
    sys_rs = open(os.path.join(RUST, "src", "sys.rs")).read()
    rust = decls(sys_rs, r"pub fn (al_\w+)\((.*?)\)")
    c = decls(HEADER, r"\b(al_\w+)\s*\(([^;{]*?)\)\s*;")
    assert len(rust) >= 25
    for name, nparams in rust.items():
        assert name in c, name
        assert c[name] == nparams, (name, c[name], nparams)
    host_part = re.sub(r"/\*.*?\*/", "", RAW.split("Device-pointer API")[0], flags=re.S)
    host = decls(host_part, r"\b(al_\w+)\s*\(([^;{]*?)\)\s*;")
    assert set(host) <= set(rust), sorted(set(host) - set(rust))


def test_safe_wrappers_cover_the_crate_signatures():
    lib_rs = open(os.path.join(RUST, "src", "lib.rs")).read()
    for sig in ["pub fn dgemm(alpha: f64, a: &[f64], b: &[f64], beta: f64, c: &mut [f64], m: usize, n: usize, k: usize)",
                "pub fn dtrsm(alpha: f64, a: &[f64], b: &mut [f64], m: usize, n: usize, left: bool, lower: bool, unit: bool)",
                "pub fn dsyrk(alpha: f64, a: &[f64], beta: f64, c: &mut [f64], n: usize, k: usize, upper: bool)",
                "pub fn dgetrf(a: &mut [f64], n: usize) -> Result<(Vec<usize>, f64), LapackError>",
                "pub fn dgetrf_rect(a: &mut [f64], nrows: usize, ncols: usize, ipiv: &mut [usize]) -> Result<(), LapackError>",
                "pub fn dgesv(a: &mut [f64], b: &mut [f64]) -> Result<(), LapackError>",
                "pub fn dpotrf(a: &mut [f64], n: usize) -> Result<(), LapackError>",
                "pub fn dgetri(lu: &mut [f64], n: usize, ipiv: &[usize]) -> Result<(), LapackError>",
                "pub fn zgemm(alpha: Complex64, a: &[Complex64], b: &[Complex64], beta: Complex64, c: &mut [Complex64], m: usize, n: usize, k: usize)",
                "pub fn zgetrf(a: &mut [Complex64], n: usize) -> Result<Vec<usize>, LapackError>",
                "pub fn zgetrf_rect(a: &mut [Complex64], nrows: usize, ncols: usize, ipiv: &mut [usize]) -> Result<(), LapackError>",
                "pub fn zgesv(a: &mut [Complex64], b: &mut [Complex64]) -> Result<(), LapackError>",
                "pub fn zpotrf(a: &mut [Complex64], n: usize, upper: bool) -> Result<(), LapackError>"]:
        assert sig in " ".join(lib_rs.split()), sig
