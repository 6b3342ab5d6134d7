/*
 * CPU oracle for the dense solve path: a loop-for-loop C restatement of the
 * algebra_linear crate's own CPU routines, used ONLY as test infrastructure
 * (the checker in tests/, smoke() and bench.py's cpu_baseline leg).  Nothing in
 * the product library links or calls this code.
 *
 * Every function cites the reference file:line it follows.  The matrices are
 * column-major with the leading dimension equal to the row count, indices are
 * 0-based, as in the crate.
 *
 * Build: oracle/Makefile (gcc -O2 -mavx2 -mfma -ffp-contract=off).  Explicit FMA
 * appears only in the AVX2 4x4 micro-kernel, as in src/blas.rs:549-598; every
 * other loop is unfused, as rustc leaves it.
 *
 * Status codes of the drivers: 0 Ok, 1 Singular, 2 BadShape (src/lapack/mod.rs:77-96).
 */
#ifndef AL_ORACLE_H
#define AL_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- level 1/2 (src/blas.rs:16-180, :628-657) ---- */
size_t or_idamax(const double* x, size_t n);
double or_dnrm2(const double* x, size_t n);
int or_dgemv(double alpha, const double* a, const double* x, double beta, double* y, size_t m, size_t n);
int or_dgemv_trans(double alpha, const double* a, const double* x, double beta, double* y, size_t m, size_t n);
double or_dlange(char norm, const double* a, size_t m, size_t n);

/* ---- level 3 (src/blas.rs:182-626, :1224-1251) ----
 * trans flags: 0 = as stored, 1 = transposed.  use_avx: -1 = detect, 0/1 = force. */
void or_set_avx(int use_avx);
int or_gemm(int trans_a, int trans_b, double alpha, const double* a, const double* b, double beta, double* c,
            size_t m, size_t n, size_t k);
int or_dsyrk(double alpha, const double* a, double beta, double* c, size_t n, size_t k, int upper);
int or_dsyrk_rank(double alpha, const double* a, double beta, double* c, size_t n, size_t k, int upper,
                  int transposed);

/* src/lapack/sym.rs:582-668 (dsytrd's private dsyr2k): one triangle of
 * C := alpha V W^T + alpha W V^T + beta C, V and W m x k. */
void or_dsyr2k(int upper, size_t m, size_t k, double alpha, const double* v, size_t lv, const double* w,
               size_t lw, double beta, double* c, size_t lc);

/* ---- triangular solves (src/blas.rs:1344-1438) ---- */
int or_dtrsv(const double* a, double* x, size_t n, int lower, int unit);
int or_dtrsv_transpose(const double* a, double* x, size_t n, int lower, int unit);
int or_dtrsm(double alpha, const double* a, double* b, size_t m, size_t n, int left, int lower, int unit);

/* ---- LAPACK drivers (src/lapack/mod.rs:112-399) ---- */
int or_dgetrf(double* a, size_t n, size_t* ipiv, double* det);
int or_dgetrf_rect(double* a, size_t nrows, size_t ncols, size_t* ipiv);
int or_dgetri(double* lu, size_t n, const size_t* ipiv);
void or_dgetrs(const double* lu, const size_t* ipiv, double* b, size_t n);
int or_dgesv(double* a, double* b, size_t n);
int or_dpotrf(double* a, size_t n);

/* ---- TESTING/LIN checkers (src/lapack/checkers.rs:17-168) ---- */
double or_get01(const double* a, const double* lu, const size_t* ipiv, size_t m, size_t n);
double or_get03(const double* a, const double* ainv, size_t n, double* rcond);
double or_get04(const double* x, const double* xact, size_t n, size_t nrhs, double rcond);
double or_pot03(const double* a, const double* ainv, size_t n, double* rcond);

/* ---- TESTING/MATGEN generators (src/lapack/families.rs:166-380, :903-1165) ----
 * dist: 1 uniform(0,1), 2 uniform(-1,1), 3 normal(0,1) (DLARNV's IDIST). */
void or_laruv(uint32_t seed[4], size_t n, double* x);
void or_larnv(int dist, uint32_t seed[4], size_t n, double* x);
double or_laran(uint32_t seed[4]);
int or_latm1(int mode, double cond, int irsign, int dist, uint32_t seed[4], double* d, size_t n);
/* sym: 0 NonSymmetric, 1 PositiveSymmetric, 2 Symmetric; pack: 0 Full, 1 Upper, 2 Lower.
 * Returns DLATMS's INFO (0 on success). */
int or_dlatms(int dist, int sym, int mode, double cond, double dmax, int pack, size_t m, size_t n, double* a,
              double* d, uint32_t seed[4]);

/* ---- compensated-sum backward error used by the tests ----
 * max over rhs of ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf), all FP64 with
 * Neumaier summation in the residual (not a reference routine). */
double or_backward_error(const double* a, const double* x, const double* b, size_t n, size_t nrhs);

/* ---- complex twins (src/blas.rs:1023-1118, src/lapack/complex.rs:67-400); complex data are
 * interleaved (re, im) double pairs, lengths and indices count complex elements ---- */
int or_zgemm(int trans_a, int trans_b, const double* alpha, const double* a, const double* b, const double* beta,
             double* c, size_t m, size_t n, size_t k);
int or_zgetrf(double* a, size_t n, size_t* ipiv);
int or_zgetrf_rect(double* a, size_t nrows, size_t ncols, size_t* ipiv);
int or_zgesv(double* a, double* b, size_t n);
int or_zpotrf(double* a, size_t n, int upper);

#ifdef __cplusplus
}
#endif
#endif
