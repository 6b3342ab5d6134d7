/*
 * algebra_linear_hip: the dense BLAS-3 / LAPACK solve path of the algebra_linear crate
 * on one MI355X (gfx950), behind a C ABI whose entry points mirror the crate's Rust
 * signatures one for one.
 *
 * Conventions kept from the crate (README.md:25-38, AGENTS.md:21-52):
 *   - column-major, element (i, j) at [i + j*nrow], leading dimension = row count;
 *   - slices cross as (pointer, length); lengths are checked exactly as the crate's
 *     assert_eq!/BadShape checks do;
 *   - pivots are 0-based size_t: ipiv[k] = the row interchanged with row k at step k;
 *   - no workspace arguments: the library allocates device memory internally;
 *   - a routine that writes one triangle leaves the other untouched (and does not read it).
 *
 * Status codes: AL_OK, AL_SINGULAR, AL_BAD_SHAPE, AL_CONVERGENCE map to the crate's
 * Ok / LapackError::{Singular, BadShape, Convergence} (src/lapack/mod.rs:77-96).  The
 * crate's BLAS routines panic on a shape mismatch; here they return AL_BAD_SHAPE and
 * al_last_error() returns the crate's panic message (e.g. "dgemm: a must be m*k"), so a
 * binding can re-raise the same panic.  Negative codes are device/runtime failures.
 *
 * Thread safety: every entry point may be called from many threads at once; each thread
 * gets its own HIP stream and device buffers (the crate's functions are reentrant too).
 */
#ifndef ALGEBRA_LINEAR_HIP_H
#define ALGEBRA_LINEAR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    AL_OK = 0,
    AL_SINGULAR = 1,     /* LapackError::Singular */
    AL_BAD_SHAPE = 2,    /* LapackError::BadShape, or a BLAS assert_eq! failure */
    AL_CONVERGENCE = 3,  /* LapackError::Convergence (mixed-precision refinement stalled) */
    AL_ERR_DEVICE = -1,  /* HIP runtime error or a device-side timeout */
    AL_ERR_NO_DEVICE = -2,
    AL_ERR_OOM = -3,
    AL_ERR_ARG = -4      /* invalid argument outside the crate's contract (e.g. null pointer) */
};

/* Message of the last failing call on this thread ("" if none). */
const char* al_last_error(void);
/* Library version string. */
const char* al_version(void);
/* Number of visible HIP devices (0 if none); does not create a context. */
int al_device_count(void);

/* ------------------------------------------------------------------------------------
 * Host-slice API: drop-in replacements of the crate's functions.  Inputs are copied to
 * HBM, computed on the GPU, and results copied back before return.
 * ------------------------------------------------------------------------------------ */

/* blas::dgemm, src/blas.rs:201-206: C := alpha*A*B + beta*C; a m*k, b k*n, c m*n. */
int al_dgemm(double alpha, const double* a, size_t a_len, const double* b, size_t b_len, double beta, double* c,
             size_t c_len, size_t m, size_t n, size_t k);
/* blas::dgemm_tn, src/blas.rs:220-225: C := alpha*A^T*B + beta*C; a k*m, b k*n. */
int al_dgemm_tn(double alpha, const double* a, size_t a_len, const double* b, size_t b_len, double beta, double* c,
                size_t c_len, size_t m, size_t n, size_t k);
/* blas::dgemm_nt, src/blas.rs:237-242: C := alpha*A*B^T + beta*C; a m*k, b n*k. */
int al_dgemm_nt(double alpha, const double* a, size_t a_len, const double* b, size_t b_len, double beta, double* c,
                size_t c_len, size_t m, size_t n, size_t k);
/* blas::dgemm_tt, src/blas.rs:252-257: C := alpha*A^T*B^T + beta*C; a k*m, b n*k. */
int al_dgemm_tt(double alpha, const double* a, size_t a_len, const double* b, size_t b_len, double beta, double* c,
                size_t c_len, size_t m, size_t n, size_t k);

/* blas::dtrsm, src/blas.rs:1384-1405: B := alpha*A^-1*B (left, a m*m) or alpha*B*A^-1
 * (right, a n*n); b m*n; lower/unit as in the crate (unit: the diagonal is not read). */
int al_dtrsm(double alpha, const double* a, size_t a_len, double* b, size_t b_len, size_t m, size_t n, int left,
             int lower, int unit);

/* blas::dsyrk, src/blas.rs:609-626: one triangle of C := alpha*A*A^T + beta*C; a n*k, c n*n. */
int al_dsyrk(double alpha, const double* a, size_t a_len, double beta, double* c, size_t c_len, size_t n, size_t k,
             int upper);

/* lapack::dgetrf, src/lapack/mod.rs:181-185 (dgetrf_with_pivots :112-163): in-place P*A = L*U;
 * writes ipiv (len n) and det (sign-flipped product of U's diagonal, in the crate's order). */
int al_dgetrf(double* a, size_t a_len, size_t n, size_t* ipiv, size_t ipiv_len, double* det);

/* lapack::dgetrf_rect, src/lapack/mod.rs:202-238: panel LU of nrows x ncols, nrows >= ncols;
 * swaps touch only the panel's columns. */
int al_dgetrf_rect(double* a, size_t a_len, size_t nrows, size_t ncols, size_t* ipiv, size_t ipiv_len);

/* lapack::dgetrs (crate-private), src/lapack/mod.rs:330-354: n = b_len, one right-hand side. */
int al_dgetrs(const double* lu, size_t lu_len, const size_t* ipiv, size_t ipiv_len, double* b, size_t b_len);

/* lapack::dgesv, src/lapack/mod.rs:309-323: n = b_len, one right-hand side; a is destroyed. */
int al_dgesv(double* a, size_t a_len, double* b, size_t b_len);

/* lapack::dpotrf, src/lapack/mod.rs:368-399: lower Cholesky in place, upper triangle untouched. */
int al_dpotrf(double* a, size_t a_len, size_t n);

/* Extension (no counterpart in the crate, whose dgesv takes one right-hand side): B is
 * n x nrhs column-major; ipiv (nullable, len n) receives the pivots. */
int al_dgesv_nrhs(double* a, size_t a_len, double* b, size_t b_len, size_t n, size_t nrhs, size_t* ipiv);

/* Extension: multi-RHS dgetrs with the same conventions as al_dgesv_nrhs. */
int al_dgetrs_nrhs(const double* lu, size_t lu_len, const size_t* ipiv, size_t ipiv_len, double* b, size_t b_len,
                   size_t n, size_t nrhs);

/* Support routines of the crate around the solve path. */
/* blas::dgemv, src/blas.rs:102-130: y := alpha*A*x + beta*y; a m*n, x n, y m.  The crate's
 * summation order and unfused arithmetic, so the result is the crate's bit for bit. */
int al_dgemv(double alpha, const double* a, size_t a_len, const double* x, size_t x_len, double beta, double* y,
             size_t y_len, size_t m, size_t n);
/* blas::dgemv_trans, src/blas.rs:145-158: y := alpha*A^T*x + beta*y; a m*n, x m, y n; bit for bit. */
int al_dgemv_trans(double alpha, const double* a, size_t a_len, const double* x, size_t x_len, double beta, double* y,
                   size_t y_len, size_t m, size_t n);
/* blas::dlange, src/blas.rs:638-657: *out := 'M' max |a_ij|, '1'/'O' max column sum, 'I' max row sum
 * (bit for bit), 'F'/'E' Frobenius (the crate's scaled dnrm2, within m*n*eps relative).  An
 * unrecognized norm is AL_ERR_ARG with the crate's panic text. */
int al_dlange(char norm, const double* a, size_t a_len, size_t m, size_t n, double* out);
/* blas::dtrsv / dtrsv_transpose, src/blas.rs:1355-1370: solves A x = b (A^T x = b) in place for a
 * triangular a (n*n, only the named triangle read; unit: the diagonal is not read). */
int al_dtrsv(const double* a, size_t a_len, double* x, size_t x_len, size_t n, int lower, int unit);
int al_dtrsv_transpose(const double* a, size_t a_len, double* x, size_t x_len, size_t n, int lower, int unit);
/* lapack::dgetri, src/lapack/mod.rs:249-298: lu (dgetrf's factors, n*n) := A^{-1}; AL_SINGULAR
 * when U has a zero on its diagonal, lu then unchanged. */
int al_dgetri(double* lu, size_t lu_len, size_t n, const size_t* ipiv, size_t ipiv_len);
/* dsytrd's dsyr2k (src/lapack/sym.rs:582-668, crate-private): one triangle of
 * C := alpha*V*W^T + alpha*W*V^T + beta*C; v, w m*k, c m*m; the other triangle untouched.  With
 * k = 0 only alpha = 0 touches C (beta*C on the triangle), as in the crate. */
int al_dsyr2k(int upper, size_t m, size_t k, double alpha, const double* v, size_t v_len, const double* w,
              size_t w_len, double beta, double* c, size_t c_len);

/* ------------------------------------------------------------------------------------
 * Device-pointer API (extension): the same operations on HBM-resident operands with
 * explicit leading dimensions, enqueued on `stream` (a hipStream_t; NULL = HIP's null
 * stream).  Results are final once the stream is synchronized; the
 * status of a factorization is written to the device int `*d_info` (0 ok, k > 0: first
 * zero / non-positive pivot at 1-based column k, < 0 device error).  Pivots are int32
 * 0-based global row indices on the device.
 * ------------------------------------------------------------------------------------ */
int al_dev_dgemm(int trans_a, int trans_b, size_t m, size_t n, size_t k, double alpha, const double* A, size_t lda,
                 const double* B, size_t ldb, double beta, double* C, size_t ldc, void* stream);
int al_dev_dsyrk(int upper, size_t n, size_t k, double alpha, const double* A, size_t lda, double beta, double* C,
                 size_t ldc, void* stream);
int al_dev_dtrsm(int left, int lower, int unit, size_t m, size_t n, double alpha, const double* A, size_t lda,
                 double* B, size_t ldb, void* stream);
int al_dev_dgetrf(size_t m, size_t n, double* A, size_t lda, int32_t* d_ipiv, int32_t* d_info, void* stream);
int al_dev_dgetrs(size_t n, size_t nrhs, const double* LU, size_t lda, const int32_t* d_ipiv, double* B,
                  size_t ldb, void* stream);
int al_dev_dpotrf(size_t n, double* A, size_t lda, int32_t* d_info, void* stream);
/* Row interchanges d_ipiv[k0..k1) applied in order to the ncols columns of A (the crate's
 * full-row swap, src/lapack/mod.rs:141-146; LAPACK's dlaswp with incx = 1). */
int al_dev_dlaswp(size_t ncols, double* A, size_t lda, size_t k0, size_t k1, const int32_t* d_ipiv, void* stream);

/* Support routines on the device (the host-slice forms above; vectors have unit stride):
 *   al_dev_dgemv:  y := alpha op(A) x + beta y (trans: op(A) = A^T), bit for bit the crate's;
 *   al_dev_dlange: *d_out (device) := the norm of the m x n matrix A;
 *   al_dev_dtrsv:  op(T) x = b in place (trans selects T^T);
 *   al_dev_dgetri: LU := A^{-1} from al_dev_dgetrf's factors and pivots; *d_info = 1-based index
 *                  of a zero on U's diagonal (LU then untouched), else 0;
 *   al_dev_dsyr2k: one triangle of C := alpha V W^T + alpha W V^T + beta C (V, W m x k; V and W
 *                  must not overlap C).
 * A one-launch triangular solve (al_dev_dtrsv, al_dev_dgetrs) whose bounded block waits give up
 * leaves NaN in the rows it could not solve. */
int al_dev_dgemv(int trans, size_t m, size_t n, double alpha, const double* A, size_t lda, const double* x, double beta,
                 double* y, void* stream);
int al_dev_dlange(char norm, size_t m, size_t n, const double* A, size_t lda, double* d_out, void* stream);
int al_dev_dtrsv(int lower, int trans, int unit, size_t n, const double* A, size_t lda, double* x, void* stream);
int al_dev_dgetri(size_t n, double* LU, size_t lda, const int32_t* d_ipiv, int32_t* d_info, void* stream);
int al_dev_dsyr2k(int upper, size_t m, size_t k, double alpha, const double* V, size_t lv, const double* W, size_t lw,
                  double beta, double* C, size_t ldc, void* stream);

/* Low-precision GEMMs (extension; the crate's GEMM is FP64 only, src/blas.rs:201-257), the
 * building blocks of the mixed-precision factorization, column-major FP32 on the device:
 *   al_dev_sgemm:     C := alpha op(A) op(B) + beta C on the FP32 matrix cores;
 *   al_dev_gemm_bf16: C := alpha A B + beta C with A (m x k) and B (k x n) rounded to bfloat16
 *                     (nearest even) and the products accumulated in FP32. */
int al_dev_sgemm(int trans_a, int trans_b, size_t m, size_t n, size_t k, float alpha, const float* A, size_t lda,
                 const float* B, size_t ldb, float beta, float* C, size_t ldc, void* stream);
int al_dev_gemm_bf16(size_t m, size_t n, size_t k, float alpha, const float* A, size_t lda, const float* B, size_t ldb,
                     float beta, float* C, size_t ldc, void* stream);

/* ------------------------------------------------------------------------------------
 * Mixed-precision solve (extension; the crate has no counterpart: its lapack::dgesv,
 * src/lapack/mod.rs:309-323, factors and solves in FP64 only).
 *
 * A x = b is solved by factoring A in low precision and refining the solution in FP64:
 *   1. F = A rounded to FP32 (padded with a scaled identity to a multiple of 64);
 *   2. P F = L U with partial pivoting in FP32; the O(n^3) trailing updates run on the FP32
 *      matrix cores (AL_FACTOR_FP32) or with bf16 operands and FP32 accumulation
 *      (AL_FACTOR_BF16);
 *   3. refinement: r = b - A x in FP64 (A itself, never rounded), d = U^{-1} L^{-1} P r
 *      with FP32 factors and FP64 vectors, x += d, until the normwise backward error
 *      max_j |r_j|_inf / (|A|_inf |x_j|_inf + |b_j|_inf) reaches stop_tol or stops
 *      halving;
 *   4. if the factorization met a zero pivot or the backward error is above tol, the
 *      system is solved again with the FP64 pivoted LU (the path of al_dgesv) when
 *      allow_fallback.
 * AL_FACTOR_FP64 skips 1-3 and solves with the FP64 pivoted LU directly.  A is never
 * modified.  Returns AL_OK when the final backward error is <= tol, AL_CONVERGENCE when it
 * is not (refinement failed and fallback disabled), AL_SINGULAR when the FP64 LU meets an
 * exactly zero pivot.  The refinement loop reads norms back to the host every step, so
 * these calls synchronize their stream.
 * ------------------------------------------------------------------------------------ */
enum { AL_FACTOR_FP64 = 0, AL_FACTOR_FP32 = 1, AL_FACTOR_BF16 = 2 };

typedef struct al_ir_options {
    int max_iter;        /* refinement steps (0: 30) */
    int allow_fallback;  /* nonzero: fall back to FP64 factorizations when refinement fails */
    double tol;          /* accepted backward error (0: 1e-14) */
    double stop_tol;     /* refinement stops once every column is below it (0: 1e-15) */
} al_ir_options;

typedef struct al_ir_stats {
    int iterations;         /* refinement steps taken (the first solve counts as one) */
    int fell_back;          /* 0: the low-precision factor's refinement produced the result; 1: an
                               FP64 LU in that factor's row order (no pivot search) and FP64
                               refinement; 2: the FP64 partially pivoted LU */
    int factor_breakdown;   /* 1-based column of an exactly zero pivot of the FP32 LU, else 0 */
    int converged;          /* 1 if the returned solution meets tol */
    double backward_error;  /* max over columns of the normwise backward error of the result */
    double backward_error_low;  /* the same after the low-precision refinement (before any fallback) */
} al_ir_stats;

/* Fills *opt with the defaults (max_iter 30, allow_fallback 1, tol 1e-14, stop_tol 1e-15). */
void al_ir_default_options(al_ir_options* opt);

/* Host-slice form: a (n*n, read only), b (n*nrhs) is overwritten by x.  opt and stats may
 * be NULL. */
int al_dgesv_mixed(const double* a, size_t a_len, double* b, size_t b_len, size_t n, size_t nrhs, int factor,
                   const al_ir_options* opt, al_ir_stats* stats);

/* Device-pointer form on `stream` (NULL = null stream): X (ldx >= n) receives the solution
 * of A X = B; A and B are not modified. */
int al_dev_dgesv_mixed(size_t n, size_t nrhs, const double* A, size_t lda, const double* B, size_t ldb, double* X,
                       size_t ldx, int factor, const al_ir_options* opt, al_ir_stats* stats, void* stream);

/* ------------------------------------------------------------------------------------
 * Batched solves (extension; the crate solves one system per call): many independent
 * systems of orders 1..512 in one call (size-bucketed persistent kernels; systems of order
 * <= 64 in FP64 mode are solved bit for bit as the crate's dgesv / dpotrf + dtrsv would), each a dgesv (LU with partial pivoting,
 * lapack::dgesv, src/lapack/mod.rs:309-323) or a dposv (lower Cholesky, lapack::dpotrf,
 * :368-399, then the two triangular solves).  Matrices live column-major in one device pool
 * (entry i's n x n matrix at A_pool + a_off), right-hand sides in another (n x nrhs at
 * B_pool + b_off, overwritten by the solutions); entries are a device array, processed in the
 * order given (decreasing size balances best).  factor AL_FACTOR_FP64 factors in place in
 * FP64; AL_FACTOR_FP32 factors an FP32 copy, refines in FP64 to a backward error of 4 eps and
 * falls back to FP64 (in place) per system when refinement fails -- A is then left intact
 * unless that system fell back.  d_info[i]: 0, the 1-based column of a zero pivot /
 * non-positive diagonal (the crate's Singular), or -4 for an order above 512.  d_steps[i]
 * (nullable): refinement steps taken, 0 in FP64 mode, -1 if the system fell back.  The call
 * allocates nothing after its first use with a given pool size and makes no host
 * synchronization, so it can be captured in a HIP graph.
 * ------------------------------------------------------------------------------------ */
enum { AL_BATCH_GESV = 0, AL_BATCH_POSV = 1 };

typedef struct al_batch_entry {
    size_t a_off;  /* element offset of the n x n matrix in A_pool */
    size_t b_off;  /* element offset of the n x nrhs right-hand sides in B_pool */
    int32_t n;
    int32_t nrhs;
    int32_t kind;  /* AL_BATCH_GESV or AL_BATCH_POSV */
    int32_t reserved;
} al_batch_entry;

int al_dev_batch_solve(size_t count, const al_batch_entry* d_entries, double* A_pool, size_t a_pool_len,
                       double* B_pool, int32_t* d_info, int32_t* d_steps, int factor, void* stream);

/* No counterpart in the crate: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The FP64 form of
 * al_dev_batch_solve (same kernels, same results bit for bit) that calls no hipMalloc/hipFree
 * and touches no per-thread workspace: its scratch -- the size buckets' counters and lists --
 * is a static device array of 1 MiB + 256 B per GPU, held until the process ends, so count is
 * at most 65536 (AL_ERR_ARG beyond).  Every launch runs on `stream`.  Each call clears the
 * counters it needs; calls that overlap on one GPU (other streams or threads) share the array
 * and race, so serialize them. */
int al_dev_batch_solve_noalloc(size_t count, const al_batch_entry* d_entries, double* A_pool, double* B_pool,
                               int32_t* d_info, int32_t* d_steps, void* stream);

/* The batched solver's launch table (extension): the systems are split on the device into
 * size buckets -- 0: n <= 64 (one wave per system), 1: n <= 128, 2: n <= 256, 3: n <= 512 (one
 * workgroup per system) -- each served by its own persistent kernel.  Per bucket, kb is the
 * panel width (16 or 32; bucket 3: 16; ignored for bucket 0) and wgs_per_cu the persistent
 * workgroups launched per CU (capped by what the bucket's LDS allows).  The table is read at
 * first use from batch_tuning.json beside the library (written by algebra_linear_amd/
 * tune_batch.py); al_batch_tune_set overrides one entry for the process. */
int al_batch_tune_set(int bucket, int kb, int wgs_per_cu);
int al_batch_tune_get(int bucket, int* kb, int* wgs_per_cu);

/* ------------------------------------------------------------------------------------
 * Solve server (extension): continuous batching of independent requests on one device.
 * Requests (dgesv or dposv, FP64 or FP32 + FP64 refinement, any order up to 32704, any
 * number of right-hand sides) are admitted in arrival order while their HBM footprint fits
 * the budget, run concurrently on `slots` worker streams, and small ones (order <= 512, no
 * resident factorization) waiting together are solved by one batched launch.  A request may
 * keep its factorization resident (its footprint stays reserved until released) and later
 * requests may solve new right-hand sides with it.  Host buffers a, b, x must stay valid until
 * al_server_wait returns for the request.  Results are bit-for-bit the same alone or in a mix.
 * Do not release a factorization while requests that use it are in flight.
 * ------------------------------------------------------------------------------------ */
typedef struct al_server al_server;

typedef struct al_request {
    int32_t kind;       /* AL_BATCH_GESV or AL_BATCH_POSV */
    int32_t factor;     /* AL_FACTOR_FP64, or AL_FACTOR_FP32 (FP32 LU + FP64 refinement, both kinds) */
    size_t n;
    size_t nrhs;
    const double* a;    /* n x n column-major (unused when factor_id >= 0) */
    const double* b;    /* n x nrhs */
    double* x;          /* n x nrhs: receives the solution */
    int64_t factor_id;  /* >= 0: solve with this resident factorization; -1: factor a */
    int32_t keep;       /* nonzero: keep this request's factorization resident */
    int32_t reserved;
} al_request;

typedef struct al_result {
    int32_t status;     /* AL_OK, AL_SINGULAR, or a negative error */
    int32_t steps;      /* refinement steps (FP32 mode), -1 if it fell back to FP64, 0 in FP64 mode */
    int64_t factor_id;  /* the factorization this request kept resident, or -1 */
    int32_t batch_size; /* requests solved by the same launch as this one */
    int32_t reserved;
    double queue_ms;    /* submission to start of service */
    double service_ms;  /* start to end of service */
} al_result;

al_server* al_server_create(size_t hbm_budget_bytes, int slots);
int al_server_submit(al_server* server, const al_request* request, uint64_t* ticket);
int al_server_wait(al_server* server, uint64_t ticket, al_result* result);
int al_server_release_factor(al_server* server, int64_t factor_id);
void al_server_destroy(al_server* server);

/* ------------------------------------------------------------------------------------
 * Complex twins.  al_complex is the crate's Complex64 (num-complex 0.4 Complex<f64>, a
 * #[repr(C)] pair {re, im}), passed by value for scalars; complex slices are arrays of it and
 * their lengths count complex elements.  Semantics, shape checks and messages are the crate's.
 * ------------------------------------------------------------------------------------ */
typedef struct al_complex {
    double re;
    double im;
} al_complex;

/* blas::zgemm, src/blas.rs:1023-1028: C := alpha*A*B + beta*C; a m*k, b k*n, c m*n.  Every C
 * entry becomes alpha*sum + beta*c, beta*c formed even for beta = 0 (src/blas.rs:1115-1116). */
int al_zgemm(al_complex alpha, const al_complex* a, size_t a_len, const al_complex* b, size_t b_len, al_complex beta,
             al_complex* c, size_t c_len, size_t m, size_t n, size_t k);
/* blas::zgemm_tn, src/blas.rs:1042-1047: C := alpha*A^T*B + beta*C (not conjugated); a k*m. */
int al_zgemm_tn(al_complex alpha, const al_complex* a, size_t a_len, const al_complex* b, size_t b_len,
                al_complex beta, al_complex* c, size_t c_len, size_t m, size_t n, size_t k);
/* blas::zgemm_nt, src/blas.rs:1059-1064: C := alpha*A*B^T + beta*C (not conjugated); b n*k. */
int al_zgemm_nt(al_complex alpha, const al_complex* a, size_t a_len, const al_complex* b, size_t b_len,
                al_complex beta, al_complex* c, size_t c_len, size_t m, size_t n, size_t k);
/* blas::zgemm_tt, src/blas.rs:1071-1076: C := alpha*A^T*B^T + beta*C; a k*m, b n*k. */
int al_zgemm_tt(al_complex alpha, const al_complex* a, size_t a_len, const al_complex* b, size_t b_len,
                al_complex beta, al_complex* c, size_t c_len, size_t m, size_t n, size_t k);

/* lapack::zgetrf, src/lapack/complex.rs:67-121: in-place P*A = L*U with |Re|+|Im| pivoting;
 * the crate returns the pivots, here they are written to ipiv (len n).  AL_SINGULAR on a zero
 * pivot (a then holds the partial factorization). */
int al_zgetrf(al_complex* a, size_t a_len, size_t n, size_t* ipiv, size_t ipiv_len);
/* lapack::zgetrf_rect, src/lapack/complex.rs:169-208: the nrows x ncols panel (nrows >= ncols). */
int al_zgetrf_rect(al_complex* a, size_t a_len, size_t nrows, size_t ncols, size_t* ipiv, size_t ipiv_len);
/* lapack::zgesv, src/lapack/complex.rs:279-320: solves A x = b for one right-hand side
 * (n = b_len); a is destroyed (holds the LU factors), b receives x. */
int al_zgesv(al_complex* a, size_t a_len, al_complex* b, size_t b_len);
/* lapack::zpotrf, src/lapack/complex.rs:344-400: A = L L^H (upper = 0) or A = U^H U (upper = 1);
 * only that triangle is read and written.  On AL_SINGULAR the failing pivot value (ajj, 0) is
 * on the diagonal, as the crate leaves it. */
int al_zpotrf(al_complex* a, size_t a_len, size_t n, int upper);

/* Device-pointer forms (leading dimensions in complex elements, asynchronous on `stream`):
 *   al_dev_zgemm: C := alpha op(A) op(B) + beta C, op = transpose without conjugation;
 *   al_dev_zgetrf: m x n (m >= n) LU, int32 0-based pivots, *d_info as in al_dev_dgetrf;
 *   al_dev_zgetrs: B (n x nrhs) := A^-1 B from al_dev_zgetrf's factors;
 *   al_dev_zpotrf: Cholesky of one triangle, *d_info = 1-based failing column or 0. */
int al_dev_zgemm(int trans_a, int trans_b, size_t m, size_t n, size_t k, al_complex alpha, const al_complex* A,
                 size_t lda, const al_complex* B, size_t ldb, al_complex beta, al_complex* C, size_t ldc,
                 void* stream);
int al_dev_zgetrf(size_t m, size_t n, al_complex* A, size_t lda, int32_t* d_ipiv, int32_t* d_info, void* stream);
int al_dev_zgetrs(size_t n, size_t nrhs, const al_complex* LU, size_t lda, const int32_t* d_ipiv, al_complex* B,
                  size_t ldb, void* stream);
int al_dev_zpotrf(int upper, size_t n, al_complex* A, size_t lda, int32_t* d_info, void* stream);

/* ------------------------------------------------------------------------------------
 * Kernel instrumentation (extension): while enabled, every launch of an instrumented
 * kernel class is bracketed by HIP events on the stream it runs on, with its algorithmic
 * flops and bytes.  Classes: "dgemm_nn", "dgemm_tn", "dgemm_nt", "dgemm_tt",
 * "dsyrk_lower", "dsyrk_upper", "panel_getf2", "zpanel_getf2", "zherk_lower_real_view",
 * "zherk_upper_real_view".  al_prof_query synchronizes the recorded
 * events and returns the totals for one class.
 * ------------------------------------------------------------------------------------ */
void al_prof_enable(int on);
void al_prof_reset(void);
int al_prof_query(const char* kernel_class, int* launches, double* total_ms, double* flops, double* bytes);

/* ------------------------------------------------------------------------------------
 * Test hooks (extension).  Bounded device-side waits -- the cooperative panel's column
 * exchange and the one-launch triangular solves' block flags -- give up after a limit in ticks
 * of the 100 MHz wall clock (default 2 s); al_debug_set_wait_limit changes it for the process
 * (<= 0 restores the default), so a test can make them give up from the host and check that
 * every entry point reports AL_ERR_DEVICE.  al_debug_coop_units_in_use returns the
 * cooperative-panel workgroups currently reserved by unfinished FP64 factorizations: their sum
 * is kept at or below half the device's CU count (a panel may run on one half of the chip), so
 * concurrent factorizations on several streams cannot fill the device with panels that wait on
 * each other.
 * ------------------------------------------------------------------------------------ */
void al_debug_set_wait_limit(long long ticks);
int al_debug_coop_units_in_use(void);
/* The cooperative panels of a workspace (one per calling thread and stream) run on one half of
 * the chip (XCDs 0-3 or 4-7) when a calibration at the workspace's creation finds its exchange
 * words nearer to that half; al_debug_panel_pin_half returns that half for `stream` (creating
 * the workspace if needed), or -1 when its panels run anywhere.  al_debug_set_panel_pinning(0)
 * makes every later panel run anywhere, (1) restores the calibrated choice; results are the
 * same either way. */
int al_debug_panel_pin_half(void* stream);
void al_debug_set_panel_pinning(int on);

#ifdef __cplusplus
}
#endif
#endif /* ALGEBRA_LINEAR_HIP_H */
