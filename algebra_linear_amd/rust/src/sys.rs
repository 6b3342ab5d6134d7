//! Raw declarations of include/algebra_linear_hip.h (host-slice entry points and the
//! mixed-precision solve).  Every function returns an `AL_*` status.
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_int};

pub const AL_OK: c_int = 0;
pub const AL_SINGULAR: c_int = 1;
pub const AL_BAD_SHAPE: c_int = 2;
pub const AL_CONVERGENCE: c_int = 3;
pub const AL_ERR_DEVICE: c_int = -1;
pub const AL_ERR_NO_DEVICE: c_int = -2;
pub const AL_ERR_OOM: c_int = -3;
pub const AL_ERR_ARG: c_int = -4;

pub const AL_FACTOR_FP64: c_int = 0;
pub const AL_FACTOR_FP32: c_int = 1;
pub const AL_FACTOR_BF16: c_int = 2;

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct al_ir_options {
    pub max_iter: c_int,
    pub allow_fallback: c_int,
    pub tol: f64,
    pub stop_tol: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct al_ir_stats {
    pub iterations: c_int,
    pub fell_back: c_int,
    pub factor_breakdown: c_int,
    pub converged: c_int,
    pub backward_error: f64,
    pub backward_error_low: f64,
}

/// `al_complex`: layout-identical to `num_complex::Complex64` (`#[repr(C)]` {re, im}).
pub type al_complex = num_complex::Complex64;

extern "C" {
    pub fn al_last_error() -> *const c_char;
    pub fn al_version() -> *const c_char;
    pub fn al_device_count() -> c_int;

    pub fn al_dgemm(alpha: f64, a: *const f64, a_len: usize, b: *const f64, b_len: usize, beta: f64, c: *mut f64,
                    c_len: usize, m: usize, n: usize, k: usize) -> c_int;
    pub fn al_dgemm_tn(alpha: f64, a: *const f64, a_len: usize, b: *const f64, b_len: usize, beta: f64, c: *mut f64,
                       c_len: usize, m: usize, n: usize, k: usize) -> c_int;
    pub fn al_dgemm_nt(alpha: f64, a: *const f64, a_len: usize, b: *const f64, b_len: usize, beta: f64, c: *mut f64,
                       c_len: usize, m: usize, n: usize, k: usize) -> c_int;
    pub fn al_dgemm_tt(alpha: f64, a: *const f64, a_len: usize, b: *const f64, b_len: usize, beta: f64, c: *mut f64,
                       c_len: usize, m: usize, n: usize, k: usize) -> c_int;
    pub fn al_dtrsm(alpha: f64, a: *const f64, a_len: usize, b: *mut f64, b_len: usize, m: usize, n: usize,
                    left: c_int, lower: c_int, unit: c_int) -> c_int;
    pub fn al_dsyrk(alpha: f64, a: *const f64, a_len: usize, beta: f64, c: *mut f64, c_len: usize, n: usize, k: usize,
                    upper: c_int) -> c_int;
    pub fn al_dgemv(alpha: f64, a: *const f64, a_len: usize, x: *const f64, x_len: usize, beta: f64, y: *mut f64,
                    y_len: usize, m: usize, n: usize) -> c_int;
    pub fn al_dgemv_trans(alpha: f64, a: *const f64, a_len: usize, x: *const f64, x_len: usize, beta: f64,
                          y: *mut f64, y_len: usize, m: usize, n: usize) -> c_int;
    pub fn al_dlange(norm: c_char, a: *const f64, a_len: usize, m: usize, n: usize, out: *mut f64) -> c_int;
    pub fn al_dtrsv(a: *const f64, a_len: usize, x: *mut f64, x_len: usize, n: usize, lower: c_int,
                    unit: c_int) -> c_int;
    pub fn al_dtrsv_transpose(a: *const f64, a_len: usize, x: *mut f64, x_len: usize, n: usize, lower: c_int,
                              unit: c_int) -> c_int;

    pub fn al_dgetrf(a: *mut f64, a_len: usize, n: usize, ipiv: *mut usize, ipiv_len: usize, det: *mut f64) -> c_int;
    pub fn al_dgetrf_rect(a: *mut f64, a_len: usize, nrows: usize, ncols: usize, ipiv: *mut usize,
                          ipiv_len: usize) -> c_int;
    pub fn al_dgetrs(lu: *const f64, lu_len: usize, ipiv: *const usize, ipiv_len: usize, b: *mut f64,
                     b_len: usize) -> c_int;
    pub fn al_dgesv(a: *mut f64, a_len: usize, b: *mut f64, b_len: usize) -> c_int;
    pub fn al_dpotrf(a: *mut f64, a_len: usize, n: usize) -> c_int;
    pub fn al_dgetri(lu: *mut f64, lu_len: usize, n: usize, ipiv: *const usize, ipiv_len: usize) -> c_int;
    pub fn al_dsyr2k(upper: c_int, m: usize, k: usize, alpha: f64, v: *const f64, v_len: usize, w: *const f64,
                     w_len: usize, beta: f64, c: *mut f64, c_len: usize) -> c_int;
    pub fn al_dgesv_nrhs(a: *mut f64, a_len: usize, b: *mut f64, b_len: usize, n: usize, nrhs: usize,
                         ipiv: *mut usize) -> c_int;
    pub fn al_dgetrs_nrhs(lu: *const f64, lu_len: usize, ipiv: *const usize, ipiv_len: usize, b: *mut f64,
                          b_len: usize, n: usize, nrhs: usize) -> c_int;

    pub fn al_zgemm(alpha: al_complex, a: *const al_complex, a_len: usize, b: *const al_complex, b_len: usize,
                    beta: al_complex, c: *mut al_complex, c_len: usize, m: usize, n: usize, k: usize) -> c_int;
    pub fn al_zgemm_tn(alpha: al_complex, a: *const al_complex, a_len: usize, b: *const al_complex, b_len: usize,
                       beta: al_complex, c: *mut al_complex, c_len: usize, m: usize, n: usize, k: usize) -> c_int;
    pub fn al_zgemm_nt(alpha: al_complex, a: *const al_complex, a_len: usize, b: *const al_complex, b_len: usize,
                       beta: al_complex, c: *mut al_complex, c_len: usize, m: usize, n: usize, k: usize) -> c_int;
    pub fn al_zgemm_tt(alpha: al_complex, a: *const al_complex, a_len: usize, b: *const al_complex, b_len: usize,
                       beta: al_complex, c: *mut al_complex, c_len: usize, m: usize, n: usize, k: usize) -> c_int;
    pub fn al_zgetrf(a: *mut al_complex, a_len: usize, n: usize, ipiv: *mut usize, ipiv_len: usize) -> c_int;
    pub fn al_zgetrf_rect(a: *mut al_complex, a_len: usize, nrows: usize, ncols: usize, ipiv: *mut usize,
                          ipiv_len: usize) -> c_int;
    pub fn al_zgesv(a: *mut al_complex, a_len: usize, b: *mut al_complex, b_len: usize) -> c_int;
    pub fn al_zpotrf(a: *mut al_complex, a_len: usize, n: usize, upper: c_int) -> c_int;

    pub fn al_ir_default_options(opt: *mut al_ir_options);
    pub fn al_dgesv_mixed(a: *const f64, a_len: usize, b: *mut f64, b_len: usize, n: usize, nrhs: usize,
                          factor: c_int, opt: *const al_ir_options, stats: *mut al_ir_stats) -> c_int;
}
