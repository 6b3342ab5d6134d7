//! `algebra_linear_hip`: the dense BLAS-3 / LAPACK solve path of the `algebra_linear` crate on
//! one MI355X (gfx950), as safe wrappers over `libalgebra_linear_hip.so`.
//!
//! Every function keeps the crate's signature, argument meaning and error behaviour
//! (`src/blas.rs`, `src/lapack/mod.rs`): column-major slices, 0-based pivots, no workspace
//! arguments; BLAS shape mismatches panic with the crate's own `assert_eq!` message, and the
//! LAPACK drivers return `Err(LapackError::Singular | BadShape)` where the crate does.  A caller
//! switches from the CPU path by changing `use algebra_linear::{blas, lapack}` to
//! `use algebra_linear_hip::{blas, lapack}`.
//!
//! GPU failures (no device, out of memory, a device-side timeout) have no counterpart in the
//! crate, which runs on the CPU only; they panic with the library's message.

pub mod sys;

use std::ffi::CStr;
use std::os::raw::c_int;

/// The message of the last failing call on this thread.
pub fn last_error() -> String {
    unsafe {
        let p = sys::al_last_error();
        if p.is_null() {
            String::new()
        } else {
            CStr::from_ptr(p).to_string_lossy().into_owned()
        }
    }
}

/// The library's version string.
pub fn version() -> String {
    unsafe { CStr::from_ptr(sys::al_version()).to_string_lossy().into_owned() }
}

/// Number of visible HIP devices.
pub fn device_count() -> usize {
    let n = unsafe { sys::al_device_count() };
    if n < 0 { 0 } else { n as usize }
}

/// A BLAS call: `AL_BAD_SHAPE` carries the crate's panic text; anything else that is not
/// `AL_OK` is a device failure.
fn blas_status(st: c_int) {
    match st {
        sys::AL_OK => {}
        sys::AL_BAD_SHAPE => panic!("{}", last_error()),
        _ => panic!("algebra_linear_hip: device error {}: {}", st, last_error()),
    }
}

pub mod blas {
    //! `algebra_linear::blas` routines of the solve path (src/blas.rs).
    use super::{blas_status, sys, Complex64};
    use std::os::raw::c_int;

    /// `C := alpha A B + beta C`; `a` m x k, `b` k x n, `c` m x n (src/blas.rs:201-206).
    pub fn dgemm(alpha: f64, a: &[f64], b: &[f64], beta: f64, c: &mut [f64], m: usize, n: usize, k: usize) {
        blas_status(unsafe {
            sys::al_dgemm(alpha, a.as_ptr(), a.len(), b.as_ptr(), b.len(), beta, c.as_mut_ptr(), c.len(), m, n, k)
        });
    }

    /// `C := alpha A^T B + beta C`; `a` k x m (src/blas.rs:220-225).
    pub fn dgemm_tn(alpha: f64, a: &[f64], b: &[f64], beta: f64, c: &mut [f64], m: usize, n: usize, k: usize) {
        blas_status(unsafe {
            sys::al_dgemm_tn(alpha, a.as_ptr(), a.len(), b.as_ptr(), b.len(), beta, c.as_mut_ptr(), c.len(), m, n, k)
        });
    }

    /// `C := alpha A B^T + beta C`; `b` n x k (src/blas.rs:237-242).
    pub fn dgemm_nt(alpha: f64, a: &[f64], b: &[f64], beta: f64, c: &mut [f64], m: usize, n: usize, k: usize) {
        blas_status(unsafe {
            sys::al_dgemm_nt(alpha, a.as_ptr(), a.len(), b.as_ptr(), b.len(), beta, c.as_mut_ptr(), c.len(), m, n, k)
        });
    }

    /// `C := alpha A^T B^T + beta C` (src/blas.rs:252-257).
    pub fn dgemm_tt(alpha: f64, a: &[f64], b: &[f64], beta: f64, c: &mut [f64], m: usize, n: usize, k: usize) {
        blas_status(unsafe {
            sys::al_dgemm_tt(alpha, a.as_ptr(), a.len(), b.as_ptr(), b.len(), beta, c.as_mut_ptr(), c.len(), m, n, k)
        });
    }

    /// `B := alpha A^-1 B` (left, `a` m x m) or `alpha B A^-1` (right, `a` n x n) (src/blas.rs:1384-1405).
    #[allow(clippy::too_many_arguments)]
    pub fn dtrsm(alpha: f64, a: &[f64], b: &mut [f64], m: usize, n: usize, left: bool, lower: bool, unit: bool) {
        blas_status(unsafe {
            sys::al_dtrsm(alpha, a.as_ptr(), a.len(), b.as_mut_ptr(), b.len(), m, n, left as c_int, lower as c_int,
                          unit as c_int)
        });
    }

    /// One triangle of `C := alpha A A^T + beta C`; `a` n x k (src/blas.rs:609-626).
    pub fn dsyrk(alpha: f64, a: &[f64], beta: f64, c: &mut [f64], n: usize, k: usize, upper: bool) {
        blas_status(unsafe {
            sys::al_dsyrk(alpha, a.as_ptr(), a.len(), beta, c.as_mut_ptr(), c.len(), n, k, upper as c_int)
        });
    }

    /// `y := alpha A x + beta y`; `a` m x n (src/blas.rs:102-130), bit for bit the crate's result.
    pub fn dgemv(alpha: f64, a: &[f64], x: &[f64], beta: f64, y: &mut [f64], m: usize, n: usize) {
        blas_status(unsafe {
            sys::al_dgemv(alpha, a.as_ptr(), a.len(), x.as_ptr(), x.len(), beta, y.as_mut_ptr(), y.len(), m, n)
        });
    }

    /// `y := alpha A^T x + beta y`; `a` m x n, `x` m, `y` n (src/blas.rs:144-158).
    pub fn dgemv_trans(alpha: f64, a: &[f64], x: &[f64], beta: f64, y: &mut [f64], m: usize, n: usize) {
        blas_status(unsafe {
            sys::al_dgemv_trans(alpha, a.as_ptr(), a.len(), x.as_ptr(), x.len(), beta, y.as_mut_ptr(), y.len(), m, n)
        });
    }

    /// `|A|` for `norm` in `'1'`/`'O'`, `'I'`, `'F'`/`'E'`, `'M'` (src/blas.rs:638-657); an
    /// unrecognized norm panics with the crate's message.
    pub fn dlange(norm: char, a: &[f64], m: usize, n: usize) -> f64 {
        let mut out = 0.0f64;
        let code = if norm.is_ascii() { norm as u8 as std::os::raw::c_char } else { b'?' as std::os::raw::c_char };
        let st = unsafe { sys::al_dlange(code, a.as_ptr(), a.len(), m, n, &mut out) };
        if st == sys::AL_ERR_ARG {
            // the crate panics on an unknown norm character; the library reports it as an argument error
            panic!("{}", super::last_error());
        }
        blas_status(st);
        out
    }

    /// Solves `A x = b` in place for a triangular `a` (src/blas.rs:1355-1359).
    pub fn dtrsv(a: &[f64], x: &mut [f64], n: usize, lower: bool, unit: bool) {
        blas_status(unsafe {
            sys::al_dtrsv(a.as_ptr(), a.len(), x.as_mut_ptr(), x.len(), n, lower as c_int, unit as c_int)
        });
    }

    /// Solves `A^T x = b` in place (src/blas.rs:1366-1370).
    pub fn dtrsv_transpose(a: &[f64], x: &mut [f64], n: usize, lower: bool, unit: bool) {
        blas_status(unsafe {
            sys::al_dtrsv_transpose(a.as_ptr(), a.len(), x.as_mut_ptr(), x.len(), n, lower as c_int, unit as c_int)
        });
    }

    /// `C := alpha A B + beta C` for complex data; `a` m x k, `b` k x n (src/blas.rs:1023-1028).
    pub fn zgemm(alpha: Complex64, a: &[Complex64], b: &[Complex64], beta: Complex64, c: &mut [Complex64], m: usize,
                 n: usize, k: usize) {
        blas_status(unsafe {
            sys::al_zgemm(alpha, a.as_ptr(), a.len(), b.as_ptr(), b.len(), beta, c.as_mut_ptr(), c.len(), m, n, k)
        });
    }

    /// `C := alpha A^T B + beta C`, not conjugated; `a` k x m (src/blas.rs:1042-1047).
    pub fn zgemm_tn(alpha: Complex64, a: &[Complex64], b: &[Complex64], beta: Complex64, c: &mut [Complex64],
                    m: usize, n: usize, k: usize) {
        blas_status(unsafe {
            sys::al_zgemm_tn(alpha, a.as_ptr(), a.len(), b.as_ptr(), b.len(), beta, c.as_mut_ptr(), c.len(), m, n, k)
        });
    }

    /// `C := alpha A B^T + beta C`, not conjugated; `b` n x k (src/blas.rs:1059-1064).
    pub fn zgemm_nt(alpha: Complex64, a: &[Complex64], b: &[Complex64], beta: Complex64, c: &mut [Complex64],
                    m: usize, n: usize, k: usize) {
        blas_status(unsafe {
            sys::al_zgemm_nt(alpha, a.as_ptr(), a.len(), b.as_ptr(), b.len(), beta, c.as_mut_ptr(), c.len(), m, n, k)
        });
    }

    /// `C := alpha A^T B^T + beta C`; `a` k x m, `b` n x k (src/blas.rs:1071-1076).
    pub fn zgemm_tt(alpha: Complex64, a: &[Complex64], b: &[Complex64], beta: Complex64, c: &mut [Complex64],
                    m: usize, n: usize, k: usize) {
        blas_status(unsafe {
            sys::al_zgemm_tt(alpha, a.as_ptr(), a.len(), b.as_ptr(), b.len(), beta, c.as_mut_ptr(), c.len(), m, n, k)
        });
    }
}

/// The crate's complex scalar (`algebra_linear::Complex64`, num-complex 0.4).
pub use num_complex::Complex64;

pub mod lapack {
    //! `algebra_linear::lapack` drivers of the solve path (src/lapack/mod.rs, src/lapack/complex.rs).
    use super::{last_error, sys, Complex64};
    use std::os::raw::c_int;

    /// The crate's error kinds (src/lapack/mod.rs:77-96).
    #[derive(Clone, Copy, Debug, PartialEq, Eq)]
    pub enum LapackError {
        /// The matrix is singular (or not positive definite) to working precision.
        Singular,
        /// A shape or length was inconsistent with another.
        BadShape,
        /// An iterative routine (here: mixed-precision refinement with the fallback disabled)
        /// did not converge.
        Convergence,
        /// The routine exists in the interface but has no implementation (never returned here).
        NotImplemented(&'static str),
    }

    /// Value below which a pivot counts as zero (src/lapack/mod.rs:65-73): exactly zero.
    pub const ZERO_DETERMINANT: f64 = 0.0;

    fn result(st: c_int) -> Result<(), LapackError> {
        match st {
            sys::AL_OK => Ok(()),
            sys::AL_SINGULAR => Err(LapackError::Singular),
            sys::AL_BAD_SHAPE => Err(LapackError::BadShape),
            sys::AL_CONVERGENCE => Err(LapackError::Convergence),
            _ => panic!("algebra_linear_hip: device error {}: {}", st, last_error()),
        }
    }

    /// LU with partial pivoting in place; returns the 0-based pivots and the determinant
    /// (src/lapack/mod.rs:181-185).  On `Singular`, `a` is left partly factored.
    pub fn dgetrf(a: &mut [f64], n: usize) -> Result<(Vec<usize>, f64), LapackError> {
        let mut ipiv = vec![0usize; n];
        let mut det = 1.0f64;
        result(unsafe { sys::al_dgetrf(a.as_mut_ptr(), a.len(), n, ipiv.as_mut_ptr(), ipiv.len(), &mut det) })?;
        Ok((ipiv, det))
    }

    /// Panel LU of an nrows x ncols block, nrows >= ncols (src/lapack/mod.rs:202-238).
    pub fn dgetrf_rect(a: &mut [f64], nrows: usize, ncols: usize, ipiv: &mut [usize]) -> Result<(), LapackError> {
        result(unsafe { sys::al_dgetrf_rect(a.as_mut_ptr(), a.len(), nrows, ncols, ipiv.as_mut_ptr(), ipiv.len()) })
    }

    /// Solve with a `dgetrf` factorization, `n = b.len()` (src/lapack/mod.rs:330-354; crate-private
    /// there, public here).
    pub fn dgetrs(lu: &[f64], ipiv: &[usize], b: &mut [f64]) {
        let st = unsafe { sys::al_dgetrs(lu.as_ptr(), lu.len(), ipiv.as_ptr(), ipiv.len(), b.as_mut_ptr(), b.len()) };
        if st != sys::AL_OK {
            // the crate's dgetrs has only debug assertions; a bad call here is a caller bug
            panic!("dgetrs: {}", last_error());
        }
    }

    /// `A x = b` for one right-hand side, `n = b.len()`; `a` is destroyed (src/lapack/mod.rs:309-323).
    pub fn dgesv(a: &mut [f64], b: &mut [f64]) -> Result<(), LapackError> {
        result(unsafe { sys::al_dgesv(a.as_mut_ptr(), a.len(), b.as_mut_ptr(), b.len()) })
    }

    /// Lower Cholesky in place; the upper triangle is neither read nor written (src/lapack/mod.rs:368-399).
    pub fn dpotrf(a: &mut [f64], n: usize) -> Result<(), LapackError> {
        result(unsafe { sys::al_dpotrf(a.as_mut_ptr(), a.len(), n) })
    }

    /// Inverse from a `dgetrf` factorization, in place (src/lapack/mod.rs:249-298).
    pub fn dgetri(lu: &mut [f64], n: usize, ipiv: &[usize]) -> Result<(), LapackError> {
        result(unsafe { sys::al_dgetri(lu.as_mut_ptr(), lu.len(), n, ipiv.as_ptr(), ipiv.len()) })
    }

    /// `dsytrd`'s symmetric rank-2k update (src/lapack/sym.rs:582-668) on packed operands: one
    /// triangle of `C := alpha V W^T + alpha W V^T + beta C`, `v`/`w` m x k, `c` m x m.
    #[allow(clippy::too_many_arguments)]
    pub fn dsyr2k(upper: bool, m: usize, k: usize, alpha: f64, v: &[f64], w: &[f64], beta: f64, c: &mut [f64]) {
        let st = unsafe {
            sys::al_dsyr2k(upper as c_int, m, k, alpha, v.as_ptr(), v.len(), w.as_ptr(), w.len(), beta,
                           c.as_mut_ptr(), c.len())
        };
        super::blas_status(st);
    }

    /// Extension: `A X = B` with `B` n x nrhs column-major; returns the pivots.
    pub fn dgesv_nrhs(a: &mut [f64], b: &mut [f64], n: usize, nrhs: usize) -> Result<Vec<usize>, LapackError> {
        let mut ipiv = vec![0usize; n];
        result(unsafe { sys::al_dgesv_nrhs(a.as_mut_ptr(), a.len(), b.as_mut_ptr(), b.len(), n, nrhs, ipiv.as_mut_ptr()) })?;
        Ok(ipiv)
    }

    /// Extension: solve with a `dgetrf` factorization for `B` n x nrhs.
    pub fn dgetrs_nrhs(lu: &[f64], ipiv: &[usize], b: &mut [f64], n: usize, nrhs: usize) -> Result<(), LapackError> {
        result(unsafe {
            sys::al_dgetrs_nrhs(lu.as_ptr(), lu.len(), ipiv.as_ptr(), ipiv.len(), b.as_mut_ptr(), b.len(), n, nrhs)
        })
    }

    /// Precision of the mixed-precision factorization.
    #[derive(Clone, Copy, Debug, PartialEq, Eq)]
    pub enum Factor {
        /// The FP64 pivoted LU only (the path of `dgesv`).
        Fp64,
        /// FP32 LU (tournament pivoting, FP32 matrix-core updates) + FP64 refinement.
        Fp32,
        /// FP32 panels with split-bf16 matrix-core updates + FP64 refinement.
        Bf16,
    }

    /// What the mixed-precision solve did.
    pub type IrStats = sys::al_ir_stats;

    /// Extension: solves `A X = B` (`b` n x nrhs, overwritten by `X`) with a low-precision LU and
    /// FP64 iterative refinement to an FP64 backward error, falling back to the FP64 LU when
    /// refinement cannot converge; `a` is not modified.
    pub fn dgesv_mixed(a: &[f64], b: &mut [f64], n: usize, nrhs: usize, factor: Factor) -> Result<IrStats, LapackError> {
        let code = match factor {
            Factor::Fp64 => sys::AL_FACTOR_FP64,
            Factor::Fp32 => sys::AL_FACTOR_FP32,
            Factor::Bf16 => sys::AL_FACTOR_BF16,
        };
        let mut stats = IrStats::default();
        result(unsafe {
            sys::al_dgesv_mixed(a.as_ptr(), a.len(), b.as_mut_ptr(), b.len(), n, nrhs, code, std::ptr::null(),
                                &mut stats)
        })?;
        Ok(stats)
    }

    /// Complex LU with |Re| + |Im| partial pivoting in place; returns the 0-based pivots
    /// (src/lapack/complex.rs:67-121).
    pub fn zgetrf(a: &mut [Complex64], n: usize) -> Result<Vec<usize>, LapackError> {
        if a.len() != n * n {
            return Err(LapackError::BadShape);
        }
        let mut ipiv = vec![0usize; n];
        result(unsafe { sys::al_zgetrf(a.as_mut_ptr(), a.len(), n, ipiv.as_mut_ptr(), ipiv.len()) })?;
        Ok(ipiv)
    }

    /// Complex panel LU, nrows >= ncols (src/lapack/complex.rs:169-208).
    pub fn zgetrf_rect(a: &mut [Complex64], nrows: usize, ncols: usize, ipiv: &mut [usize]) -> Result<(), LapackError> {
        result(unsafe { sys::al_zgetrf_rect(a.as_mut_ptr(), a.len(), nrows, ncols, ipiv.as_mut_ptr(), ipiv.len()) })
    }

    /// `A x = b` for one complex right-hand side, `n = b.len()`; `a` is destroyed
    /// (src/lapack/complex.rs:279-320).
    pub fn zgesv(a: &mut [Complex64], b: &mut [Complex64]) -> Result<(), LapackError> {
        result(unsafe { sys::al_zgesv(a.as_mut_ptr(), a.len(), b.as_mut_ptr(), b.len()) })
    }

    /// Complex Cholesky, `A = L L^H` (`upper == false`) or `A = U^H U`; only that triangle is read
    /// or written (src/lapack/complex.rs:344-400).
    pub fn zpotrf(a: &mut [Complex64], n: usize, upper: bool) -> Result<(), LapackError> {
        result(unsafe { sys::al_zpotrf(a.as_mut_ptr(), a.len(), n, upper as c_int) })
    }
}
