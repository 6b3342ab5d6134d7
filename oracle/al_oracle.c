/*
 * CPU oracle (test infrastructure only): a loop-for-loop C restatement of the
 * algebra_linear crate's dense solve path.  Each function names the reference
 * file:line it follows.  Built with -ffp-contract=off so the only fused
 * multiply-adds are the explicit ones in the AVX2 micro-kernel, matching the
 * crate (src/blas.rs:549-598).  Not part of the product library.
 */
#include "al_oracle.h"

#include <immintrin.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(i, j, ld) ((i) + (j) * (ld))

/* ------------------------------------------------------------------------- */
/* level 1/2                                                                 */
/* ------------------------------------------------------------------------- */

/* src/blas.rs:77-88: first index of the largest |x|, strict '>' from -inf. */
size_t or_idamax(const double* x, size_t n) {
    size_t best = 0;
    double best_abs = -INFINITY;
    for (size_t i = 0; i < n; i++) {
        double a = fabs(x[i]);
        if (a > best_abs) {
            best_abs = a;
            best = i;
        }
    }
    return best;
}

/* src/blas.rs:54-71: dnrm2 scaled by the global max. */
double or_dnrm2(const double* x, size_t n) {
    double scale = 0.0;
    for (size_t i = 0; i < n; i++) {
        double a = fabs(x[i]);
        if (a > scale) scale = a;
    }
    if (scale == 0.0 || !isfinite(scale)) {
        double s = 0.0;
        for (size_t i = 0; i < n; i++) s += x[i] * x[i];
        return sqrt(s);
    }
    double inv = 1.0 / scale;
    double ss = 0.0;
    for (size_t i = 0; i < n; i++) ss += (x[i] * inv) * (x[i] * inv);
    return scale * sqrt(ss);
}

/* src/blas.rs:102-130 */
int or_dgemv(double alpha, const double* a, const double* x, double beta, double* y, size_t m, size_t n) {
    if (beta == 0.0) {
        for (size_t i = 0; i < m; i++) y[i] = 0.0;
    } else if (beta != 1.0) {
        for (size_t i = 0; i < m; i++) y[i] *= beta;
    }
    for (size_t j = 0; j < n; j++) {
        double xj = x[j];
        if (xj == 0.0) continue;
        double c = alpha * xj;
        const double* col = a + j * m;
        for (size_t i = 0; i < m; i++) y[i] += c * col[i];
    }
    return 0;
}

/* src/blas.rs:145-158 */
int or_dgemv_trans(double alpha, const double* a, const double* x, double beta, double* y, size_t m, size_t n) {
    for (size_t j = 0; j < n; j++) {
        const double* col = a + j * m;
        double sum = 0.0;
        for (size_t i = 0; i < m; i++) sum += col[i] * x[i];
        y[j] = (beta == 0.0) ? alpha * sum : alpha * sum + beta * y[j];
    }
    return 0;
}

/* src/blas.rs:638-657.  'M', '1'/'O', 'I', 'F'/'E'; NaN-ignoring max like f64::max. */
double or_dlange(char norm, const double* a, size_t m, size_t n) {
    double r = 0.0;
    switch (norm) {
        case 'M': case 'm':
            for (size_t i = 0; i < m * n; i++) r = fmax(r, fabs(a[i]));
            return r;
        case '1': case 'O': case 'o':
            for (size_t j = 0; j < n; j++) {
                double s = 0.0;
                for (size_t i = 0; i < m; i++) s += fabs(a[IDX(i, j, m)]);
                r = fmax(r, s);
            }
            return r;
        case 'I': case 'i':
            for (size_t i = 0; i < m; i++) {
                double s = 0.0;
                for (size_t j = 0; j < n; j++) s += fabs(a[IDX(i, j, m)]);
                r = fmax(r, s);
            }
            return r;
        case 'F': case 'f': case 'E': case 'e':
            return or_dnrm2(a, m * n);
        default:
            return NAN;
    }
}

/* ------------------------------------------------------------------------- */
/* level 3: the Goto-blocked gemm_core of src/blas.rs:283-598                */
/* ------------------------------------------------------------------------- */

static int g_use_avx = -1;

void or_set_avx(int use_avx) { g_use_avx = use_avx; }

static int have_avx2_fma(void) {
    /* src/blas.rs:466-471: runtime CPUID check, cached. */
    if (g_use_avx < 0) {
        __builtin_cpu_init();
        g_use_avx = __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
    }
    return g_use_avx;
}

/* src/blas.rs:427-457: the direct small-product kernel; beta already applied. */
static void gemm_direct(double alpha, const double* a, const double* b, double* c, size_t m, size_t n, size_t k,
                        size_t ada, size_t adb, int ta, int tb) {
    for (size_t j = 0; j < n; j++) {
        double* ccol = c + j * m;
        for (size_t p = 0; p < k; p++) {
            double v = tb ? b[j + p * adb] : b[p + j * adb];
            double av = alpha * v;
            for (size_t i = 0; i < m; i++) {
                double u = ta ? a[p + i * ada] : a[i + p * ada];
                ccol[i] += av * u;
            }
        }
    }
}

/* src/blas.rs:549-598: 4x4 AVX2+FMA micro-kernel, C read-modify-written once. */
__attribute__((target("avx2,fma"))) static void gemm_micro_4x4(const double* ap, size_t lda, const double* bp,
                                                                size_t ldb, double* c, size_t ldc, size_t kc) {
    __m256d c0 = _mm256_setzero_pd(), c1 = _mm256_setzero_pd();
    __m256d c2 = _mm256_setzero_pd(), c3 = _mm256_setzero_pd();
    const double* b0 = bp;
    const double* b1 = bp + ldb;
    const double* b2 = bp + 2 * ldb;
    const double* b3 = bp + 3 * ldb;
    for (size_t p = 0; p < kc; p++) {
        __m256d av = _mm256_loadu_pd(ap + p * lda);
        c0 = _mm256_fmadd_pd(av, _mm256_broadcast_sd(b0 + p), c0);
        c1 = _mm256_fmadd_pd(av, _mm256_broadcast_sd(b1 + p), c1);
        c2 = _mm256_fmadd_pd(av, _mm256_broadcast_sd(b2 + p), c2);
        c3 = _mm256_fmadd_pd(av, _mm256_broadcast_sd(b3 + p), c3);
    }
    _mm256_storeu_pd(c, _mm256_add_pd(_mm256_loadu_pd(c), c0));
    _mm256_storeu_pd(c + ldc, _mm256_add_pd(_mm256_loadu_pd(c + ldc), c1));
    _mm256_storeu_pd(c + 2 * ldc, _mm256_add_pd(_mm256_loadu_pd(c + 2 * ldc), c2));
    _mm256_storeu_pd(c + 3 * ldc, _mm256_add_pd(_mm256_loadu_pd(c + 3 * ldc), c3));
}

/* src/blas.rs:479-533: packed panel product with scalar non-FMA remainders. */
static void gemm_panel(const double* ap, const double* bp, double* c, size_t ldc, size_t mc, size_t nc, size_t kc,
                       size_t lda, size_t ldb, int use_avx) {
    size_t j = 0;
    while (j < nc) {
        size_t i = 0;
        if (j + 4 <= nc && use_avx) {
            while (i + 4 <= mc) {
                gemm_micro_4x4(ap + i, lda, bp + j * ldb, ldb, c + i + j * ldc, ldc, kc);
                i += 4;
            }
            while (i < mc) {
                for (size_t jj = j; jj < j + 4; jj++) {
                    double sum = 0.0;
                    for (size_t p = 0; p < kc; p++) sum += ap[i + p * lda] * bp[p + jj * ldb];
                    c[i + jj * ldc] += sum;
                }
                i++;
            }
            j += 4;
        } else {
            for (size_t jj = j; jj < nc; jj++) {
                for (size_t ii = 0; ii < mc; ii++) {
                    double sum = 0.0;
                    for (size_t p = 0; p < kc; p++) sum += ap[ii + p * lda] * bp[p + jj * ldb];
                    c[ii + jj * ldc] += sum;
                }
            }
            j = nc;
        }
    }
}

/* src/blas.rs:283-420 */
int or_gemm(int ta, int tb, double alpha, const double* a, const double* b, double beta, double* c, size_t m,
            size_t n, size_t k) {
    if (m == 0 || n == 0) return 0;
    if (beta == 0.0) {
        for (size_t i = 0; i < m * n; i++) c[i] = 0.0;
    } else if (beta != 1.0) {
        for (size_t i = 0; i < m * n; i++) c[i] *= beta;
    }
    if (k == 0 || alpha == 0.0) return 0;
    size_t ada = ta ? k : m;
    size_t adb = tb ? n : k;
    if (m * n * k <= 1024) {
        gemm_direct(alpha, a, b, c, m, n, k, ada, adb, ta, tb);
        return 0;
    }
    enum { MC = 128, KC = 256, NC = 256 };
    size_t kc_max = k < KC ? k : KC;
    size_t nc_max = n < NC ? n : NC;
    double* ap = (double*)calloc((size_t)MC * kc_max, sizeof(double));
    double* bp = (double*)calloc(kc_max * nc_max, sizeof(double));
    if (!ap || !bp) {
        free(ap);
        free(bp);
        return -1;
    }
    int use_avx = have_avx2_fma();
    for (size_t jc = 0; jc < n; jc += NC) {
        size_t nc = (n - jc) < NC ? (n - jc) : NC;
        for (size_t pc = 0; pc < k; pc += KC) {
            size_t kc = (k - pc) < KC ? (k - pc) : KC;
            if (!tb) {
                for (size_t j = 0; j < nc; j++) memcpy(bp + j * kc, b + pc + (jc + j) * adb, kc * sizeof(double));
            } else {
                for (size_t p = 0; p < kc; p++)
                    for (size_t j = 0; j < nc; j++) bp[j * kc + p] = b[jc + j + (pc + p) * adb];
            }
            for (size_t ic = 0; ic < m; ic += MC) {
                size_t mc = (m - ic) < MC ? (m - ic) : MC;
                if (!ta) {
                    for (size_t p = 0; p < kc; p++) {
                        const double* acol = a + ic + (pc + p) * ada;
                        double* dst = ap + p * MC;
                        if (alpha == 1.0) {
                            memcpy(dst, acol, mc * sizeof(double));
                        } else {
                            for (size_t i = 0; i < mc; i++) dst[i] = alpha * acol[i];
                        }
                    }
                } else {
                    for (size_t i = 0; i < mc; i++)
                        for (size_t p = 0; p < kc; p++) {
                            double v = a[pc + p + (ic + i) * ada];
                            ap[p * MC + i] = (alpha == 1.0) ? v : alpha * v;
                        }
                }
                gemm_panel(ap, bp, c + ic + jc * m, m, mc, nc, kc, MC, kc, use_avx);
            }
        }
    }
    free(ap);
    free(bp);
    return 0;
}

/* src/blas.rs:609-626 */
int or_dsyrk(double alpha, const double* a, double beta, double* c, size_t n, size_t k, int upper) {
    for (size_t j = 0; j < n; j++) {
        size_t i0 = upper ? 0 : j, i1 = upper ? j + 1 : n;
        for (size_t i = i0; i < i1; i++) {
            double sum = 0.0;
            for (size_t l = 0; l < k; l++) sum += a[i + l * n] * a[j + l * n];
            size_t idx = i + j * n;
            c[idx] = (beta == 0.0) ? alpha * sum : alpha * sum + beta * c[idx];
        }
    }
    return 0;
}

/* src/blas.rs:1224-1251 */
int or_dsyrk_rank(double alpha, const double* a, double beta, double* c, size_t n, size_t k, int upper,
                  int transposed) {
    if (!transposed) return or_dsyrk(alpha, a, beta, c, n, k, upper);
    for (size_t j = 0; j < n; j++) {
        size_t i0 = upper ? 0 : j, i1 = upper ? j + 1 : n;
        for (size_t i = i0; i < i1; i++) {
            double sum = 0.0;
            for (size_t l = 0; l < k; l++) sum += a[l + i * k] * a[l + j * k];
            size_t idx = i + j * n;
            c[idx] = (beta == 0.0) ? alpha * sum : alpha * sum + beta * c[idx];
        }
    }
    return 0;
}

/* src/lapack/sym.rs:582-668: dsytrd's dsyr2k, C := alpha V W^T + alpha W V^T + beta C on one
 * triangle (V, W m x k with leading dimensions lv, lw; C m x m with lc).  The crate gathers
 * compact copies of V, W and their transposes, forms the full product t with two plain dgemm
 * calls (t = alpha V W^T, then t = alpha W V^T + t), and folds t into the named triangle
 * (beta == 0 writes t).  m == 0 or k == 0: only alpha == 0 scales the triangle by beta. */
void or_dsyr2k(int upper, size_t m, size_t k, double alpha, const double* v, size_t lv, const double* w,
               size_t lw, double beta, double* c, size_t lc) {
    if (m == 0 || k == 0) {
        if (alpha == 0.0 && beta != 1.0) {
            for (size_t j = 0; j < m; j++) {
                size_t i0 = upper ? 0 : j, i1 = upper ? j + 1 : m;
                for (size_t i = i0; i < i1; i++) c[i + j * lc] = (beta == 0.0) ? 0.0 : c[i + j * lc] * beta;
            }
        }
        return;
    }
    double* vc = calloc(m * k, sizeof(double));
    double* wc = calloc(m * k, sizeof(double));
    double* vt = calloc(k * m, sizeof(double));
    double* wt = calloc(k * m, sizeof(double));
    double* t = calloc(m * m, sizeof(double));
    for (size_t l = 0; l < k; l++)
        for (size_t i = 0; i < m; i++) {
            double x = v[i + l * lv], y = w[i + l * lw];
            vc[i + l * m] = x;
            wc[i + l * m] = y;
            vt[l + i * k] = x;
            wt[l + i * k] = y;
        }
    or_gemm(0, 0, alpha, vc, wt, 0.0, t, m, m, k);
    or_gemm(0, 0, alpha, wc, vt, 1.0, t, m, m, k);
    for (size_t j = 0; j < m; j++) {
        size_t i0 = upper ? 0 : j, i1 = upper ? j + 1 : m;
        for (size_t i = i0; i < i1; i++) {
            double tv = t[i + j * m];
            c[i + j * lc] = (beta == 0.0) ? tv : beta * c[i + j * lc] + tv;
        }
    }
    free(vc);
    free(wc);
    free(vt);
    free(wt);
    free(t);
}

/* ------------------------------------------------------------------------- */
/* triangular solves                                                         */
/* ------------------------------------------------------------------------- */

/* src/blas.rs:1410-1438: x read/written with `stride`; trans swaps the triangle. */
static void tri_solve(const double* a, double* x, size_t n, size_t stride, int lower, int unit, int trans) {
    int eff_lower = trans ? !lower : lower;
    if (eff_lower) {
        for (size_t i = 0; i < n; i++) {
            double sum = x[i * stride];
            for (size_t k = 0; k < i; k++) {
                double aik = trans ? a[k + i * n] : a[i + k * n];
                sum -= aik * x[k * stride];
            }
            if (!unit) sum /= a[i + i * n];
            x[i * stride] = sum;
        }
    } else {
        for (size_t ii = n; ii-- > 0;) {
            double sum = x[ii * stride];
            for (size_t k = ii + 1; k < n; k++) {
                double aik = trans ? a[k + ii * n] : a[ii + k * n];
                sum -= aik * x[k * stride];
            }
            if (!unit) sum /= a[ii + ii * n];
            x[ii * stride] = sum;
        }
    }
}

/* src/blas.rs:1355-1358 */
int or_dtrsv(const double* a, double* x, size_t n, int lower, int unit) {
    tri_solve(a, x, n, 1, lower, unit, 0);
    return 0;
}

/* src/blas.rs:1366-1370 */
int or_dtrsv_transpose(const double* a, double* x, size_t n, int lower, int unit) {
    tri_solve(a, x, n, 1, lower, unit, 1);
    return 0;
}

/* src/blas.rs:1384-1405 */
int or_dtrsm(double alpha, const double* a, double* b, size_t m, size_t n, int left, int lower, int unit) {
    if (left) {
        for (size_t j = 0; j < n; j++) {
            double* x = b + j * m;
            for (size_t i = 0; i < m; i++) x[i] *= alpha;
            tri_solve(a, x, m, 1, lower, unit, 0);
        }
    } else {
        for (size_t i = 0; i < m; i++) {
            double* x = b + i;
            for (size_t k = 0; k < n; k++) x[k * m] *= alpha;
            tri_solve(a, x, n, m, lower, unit, 1);
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* LAPACK drivers                                                            */
/* ------------------------------------------------------------------------- */

/* src/lapack/mod.rs:112-163 (and :181-185): unblocked right-looking LU, full-row
 * swaps, strict '>' pivot scan from the diagonal, exact-zero singular test. */
int or_dgetrf(double* a, size_t n, size_t* ipiv, double* det_out) {
    double det = 1.0;
    for (size_t k = 0; k < n; k++) {
        size_t pivot = k;
        double best = fabs(a[k + k * n]);
        for (size_t i = k + 1; i < n; i++) {
            double cand = fabs(a[i + k * n]);
            if (cand > best) {
                best = cand;
                pivot = i;
            }
        }
        if (best <= 0.0) return 1;
        if (pivot != k) {
            for (size_t j = 0; j < n; j++) {
                double t = a[j * n + k];
                a[j * n + k] = a[j * n + pivot];
                a[j * n + pivot] = t;
            }
            det = -det;
        }
        ipiv[k] = pivot;
        double d = a[k + k * n];
        det *= d;
        for (size_t i = k + 1; i < n; i++) {
            double factor = a[i + k * n] / d;
            a[i + k * n] = factor;
            for (size_t j = k + 1; j < n; j++) a[i + j * n] -= factor * a[k + j * n];
        }
    }
    if (det_out) *det_out = det;
    return 0;
}

/* src/lapack/mod.rs:202-238 */
int or_dgetrf_rect(double* a, size_t nrows, size_t ncols, size_t* ipiv) {
    if (nrows < ncols) return 2;
    for (size_t k = 0; k < ncols; k++) {
        size_t pivot = k;
        double best = fabs(a[k + k * nrows]);
        for (size_t i = k + 1; i < nrows; i++) {
            double cand = fabs(a[i + k * nrows]);
            if (cand > best) {
                best = cand;
                pivot = i;
            }
        }
        if (best <= 0.0) return 1;
        if (pivot != k) {
            for (size_t j = 0; j < ncols; j++) {
                double t = a[j * nrows + k];
                a[j * nrows + k] = a[j * nrows + pivot];
                a[j * nrows + pivot] = t;
            }
        }
        ipiv[k] = pivot;
        double d = a[k + k * nrows];
        for (size_t i = k + 1; i < nrows; i++) {
            double factor = a[i + k * nrows] / d;
            a[i + k * nrows] = factor;
            for (size_t j = k + 1; j < ncols; j++) a[i + j * nrows] -= factor * a[k + j * nrows];
        }
    }
    return 0;
}

/* src/lapack/mod.rs:249-298 */
int or_dgetri(double* lu, size_t n, const size_t* ipiv) {
    if (n == 0) return 0;
    double* inv = (double*)calloc(n * n, sizeof(double));
    double* b = (double*)calloc(n, sizeof(double));
    if (!inv || !b) {
        free(inv);
        free(b);
        return -1;
    }
    for (size_t col = 0; col < n; col++) {
        memset(b, 0, n * sizeof(double));
        b[col] = 1.0;
        for (size_t k = 0; k < n; k++) {
            double t = b[k];
            b[k] = b[ipiv[k]];
            b[ipiv[k]] = t;
        }
        for (size_t i = 0; i < n; i++) {
            double sum = b[i];
            for (size_t j = 0; j < i; j++) sum -= lu[i + j * n] * b[j];
            b[i] = sum;
        }
        for (size_t ii = n; ii-- > 0;) {
            double sum = b[ii];
            for (size_t j = ii + 1; j < n; j++) sum -= lu[ii + j * n] * b[j];
            double d = lu[ii + ii * n];
            if (d == 0.0) {
                free(inv);
                free(b);
                return 1;
            }
            b[ii] = sum / d;
        }
        for (size_t i = 0; i < n; i++) inv[i + col * n] = b[i];
    }
    memcpy(lu, inv, n * n * sizeof(double));
    free(inv);
    free(b);
    return 0;
}

/* src/lapack/mod.rs:330-354: swaps, unit-lower forward, upper back substitution. */
void or_dgetrs(const double* lu, const size_t* ipiv, double* b, size_t n) {
    for (size_t k = 0; k < n; k++) {
        double t = b[k];
        b[k] = b[ipiv[k]];
        b[ipiv[k]] = t;
    }
    for (size_t i = 0; i < n; i++) {
        double sum = b[i];
        for (size_t j = 0; j < i; j++) sum -= lu[i + j * n] * b[j];
        b[i] = sum;
    }
    for (size_t ii = n; ii-- > 0;) {
        double sum = b[ii];
        for (size_t j = ii + 1; j < n; j++) sum -= lu[ii + j * n] * b[j];
        b[ii] = sum / lu[ii + ii * n];
    }
}

/* src/lapack/mod.rs:309-323 */
int or_dgesv(double* a, double* b, size_t n) {
    if (n == 0) return 0;
    size_t* ipiv = (size_t*)calloc(n, sizeof(size_t));
    if (!ipiv) return -1;
    double det;
    int st = or_dgetrf(a, n, ipiv, &det);
    if (st == 0) or_dgetrs(a, ipiv, b, n);
    free(ipiv);
    return st;
}

/* src/lapack/mod.rs:368-399: left-looking dot-product Cholesky, lower only. */
int or_dpotrf(double* a, size_t n) {
    for (size_t j = 0; j < n; j++) {
        double sum = a[j + j * n];
        for (size_t k = 0; k < j; k++) {
            double ljk = a[j + k * n];
            sum -= ljk * ljk;
        }
        if (sum <= 0.0) return 1;
        double ljj = sqrt(sum);
        a[j + j * n] = ljj;
        for (size_t i = j + 1; i < n; i++) {
            double s = a[i + j * n];
            for (size_t k = 0; k < j; k++) s -= a[i + k * n] * a[j + k * n];
            a[i + j * n] = s / ljj;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* TESTING/LIN checkers                                                      */
/* ------------------------------------------------------------------------- */

static const double EPS = 2.220446049250313e-16; /* f64::EPSILON */

/* src/lapack/checkers.rs:17-63 (DGET01) */
double or_get01(const double* a, const double* lu, const size_t* ipiv, size_t m, size_t n) {
    if (m == 0 || n == 0) return 0.0;
    double anorm = or_dlange('1', a, m, n);
    size_t kk = m < n ? m : n;
    double* prod = (double*)calloc(m * n, sizeof(double));
    if (!prod) return NAN;
    for (size_t j = 0; j < n; j++)
        for (size_t i = 0; i < m; i++) {
            double sum = 0.0;
            for (size_t p = 0; p < kk; p++) {
                if (p > i || p > j) continue;
                double l = (p == i) ? 1.0 : lu[i + p * m];
                sum += l * lu[p + j * m];
            }
            prod[i + j * m] = sum;
        }
    for (size_t ii = kk; ii-- > 0;) {
        size_t p = ipiv[ii];
        if (p != ii)
            for (size_t j = 0; j < n; j++) {
                double t = prod[ii + j * m];
                prod[ii + j * m] = prod[p + j * m];
                prod[p + j * m] = t;
            }
    }
    for (size_t idx = 0; idx < m * n; idx++) prod[idx] -= a[idx];
    double resid = or_dlange('1', prod, m, n);
    free(prod);
    if (anorm <= 0.0) return resid != 0.0 ? 1.0 / EPS : 0.0;
    return (resid / (double)n / anorm) / EPS;
}

/* src/lapack/checkers.rs:71-99 (DGET03): W = -AINV*A + I */
double or_get03(const double* a, const double* ainv, size_t n, double* rcond_out) {
    if (n == 0) {
        if (rcond_out) *rcond_out = 1.0;
        return 0.0;
    }
    double anorm = or_dlange('1', a, n, n);
    double ainvnm = or_dlange('1', ainv, n, n);
    if (anorm <= 0.0 || ainvnm <= 0.0) {
        if (rcond_out) *rcond_out = 0.0;
        return 1.0 / EPS;
    }
    double rcond = (1.0 / anorm) / ainvnm;
    double* w = (double*)calloc(n * n, sizeof(double));
    if (!w) return NAN;
    for (size_t j = 0; j < n; j++)
        for (size_t i = 0; i < n; i++) {
            double sum = 0.0;
            for (size_t p = 0; p < n; p++) sum += ainv[i + p * n] * a[p + j * n];
            w[i + j * n] = -sum;
        }
    for (size_t i = 0; i < n; i++) w[i + i * n] += 1.0;
    double resid = or_dlange('1', w, n, n);
    free(w);
    if (rcond_out) *rcond_out = rcond;
    return resid * rcond / EPS / (double)n;
}

/* src/lapack/checkers.rs:108-134 (DGET04) */
double or_get04(const double* x, const double* xact, size_t n, size_t nrhs, double rcond) {
    if (n == 0 || nrhs == 0) return 0.0;
    if (rcond < 0.0) return 1.0 / EPS;
    double resid = 0.0;
    for (size_t j = 0; j < nrhs; j++) {
        double xnorm = 0.0, diffnm = 0.0;
        for (size_t i = 0; i < n; i++) xnorm = fmax(xnorm, fabs(xact[i + j * n]));
        for (size_t i = 0; i < n; i++) diffnm = fmax(diffnm, fabs(x[i + j * n] - xact[i + j * n]));
        if (xnorm <= 0.0) {
            if (diffnm > 0.0) resid = 1.0 / EPS;
        } else {
            resid = fmax(resid, (diffnm / xnorm) * rcond);
        }
    }
    return (resid * EPS < 1.0) ? resid / EPS : resid;
}

/* src/lapack/checkers.rs:142-168 (DPOT03): W = -A*AINV + I */
double or_pot03(const double* a, const double* ainv, size_t n, double* rcond_out) {
    if (n == 0) {
        if (rcond_out) *rcond_out = 1.0;
        return 0.0;
    }
    double anorm = or_dlange('1', a, n, n);
    double ainvnm = or_dlange('1', ainv, n, n);
    if (anorm <= 0.0 || ainvnm <= 0.0) {
        if (rcond_out) *rcond_out = 0.0;
        return 1.0 / EPS;
    }
    double rcond = (1.0 / anorm) / ainvnm;
    double* w = (double*)calloc(n * n, sizeof(double));
    if (!w) return NAN;
    for (size_t j = 0; j < n; j++)
        for (size_t i = 0; i < n; i++) {
            double sum = 0.0;
            for (size_t p = 0; p < n; p++) sum += a[i + p * n] * ainv[p + j * n];
            w[i + j * n] = -sum;
        }
    for (size_t i = 0; i < n; i++) w[i + i * n] += 1.0;
    double resid = or_dlange('1', w, n, n);
    free(w);
    if (rcond_out) *rcond_out = rcond;
    return resid * rcond / EPS / (double)n;
}

/* ------------------------------------------------------------------------- */
/* TESTING/MATGEN generators                                                 */
/* ------------------------------------------------------------------------- */

/* DLARUV's multiplier table (src/lapack/families.rs:22-153) holds the i-th power
 * (i = 1..128) of 33952834046453 modulo 2**48, split into four 12-bit words; it is
 * generated here from that definition rather than transcribed. */
static uint32_t g_mm[128][4];
static int g_mm_ready = 0;

static void mm_init(void) {
    if (g_mm_ready) return;
    const uint64_t mult = 33952834046453ULL, mask = (1ULL << 48) - 1;
    uint64_t v = 1;
    for (int i = 0; i < 128; i++) {
        v = (uint64_t)(((unsigned __int128)v * mult) & mask);
        g_mm[i][0] = (uint32_t)((v >> 36) & 4095);
        g_mm[i][1] = (uint32_t)((v >> 24) & 4095);
        g_mm[i][2] = (uint32_t)((v >> 12) & 4095);
        g_mm[i][3] = (uint32_t)(v & 4095);
    }
    g_mm_ready = 1;
}

/* src/lapack/families.rs:166-219 (DLARUV): 12-bit-word 48-bit LCG, n <= 128. */
void or_laruv(uint32_t seed[4], size_t n, double* x) {
    const uint32_t IPW2 = 4096;
    const double R = 1.0 / 4096.0;
    mm_init();
    if (n < 1) return;
    if (n > 128) n = 128;
    uint32_t i1 = seed[0], i2 = seed[1], i3 = seed[2], i4 = seed[3];
    uint32_t it1 = 0, it2 = 0, it3 = 0, it4 = 0;
    for (size_t i = 0; i < n; i++) {
        uint32_t m1 = g_mm[i][0], m2 = g_mm[i][1], m3 = g_mm[i][2], m4 = g_mm[i][3];
        for (;;) {
            it4 = i4 * m4;
            it3 = it4 / IPW2;
            it4 -= IPW2 * it3;
            it3 += i3 * m4 + i4 * m3;
            it2 = it3 / IPW2;
            it3 -= IPW2 * it2;
            it2 += i2 * m4 + i3 * m3 + i4 * m2;
            it1 = it2 / IPW2;
            it2 -= IPW2 * it1;
            it1 += i1 * m4 + i2 * m3 + i3 * m2 + i4 * m1;
            it1 %= IPW2;
            double v = R * ((double)it1 + R * ((double)it2 + R * ((double)it3 + R * (double)it4)));
            if (v == 1.0) {
                i1 += 2;
                i2 += 2;
                i3 += 2;
                i4 += 2;
                continue;
            }
            x[i] = v;
            break;
        }
    }
    seed[0] = it1;
    seed[1] = it2;
    seed[2] = it3;
    seed[3] = it4;
}

/* src/lapack/families.rs:230-249 (DLARNV): batches of 64, Box-Muller for normal. */
void or_larnv(int dist, uint32_t seed[4], size_t n, double* x) {
    const double TAU = 6.283185307179586476925286766559;
    double u[128];
    size_t iv = 0;
    while (iv < n) {
        size_t il = (n - iv) < 64 ? (n - iv) : 64;
        size_t il2 = (dist == 3) ? 2 * il : il;
        or_laruv(seed, il2, u);
        for (size_t i = 0; i < il; i++) {
            if (dist == 1)
                x[iv + i] = u[i];
            else if (dist == 2)
                x[iv + i] = 2.0 * u[i] - 1.0;
            else
                x[iv + i] = sqrt(-2.0 * log(u[2 * i])) * cos(TAU * u[2 * i + 1]);
        }
        iv += 64;
    }
}

/* src/lapack/families.rs:256-260 (DLARAN) */
double or_laran(uint32_t seed[4]) {
    double x;
    or_laruv(seed, 1, &x);
    return x;
}

/* Rust's f64::powi is binary exponentiation (square-and-multiply from the low bit,
 * reciprocal at the end for a negative exponent), not C's pow. */
static double powi_binary(double a, int b) {
    int recip = b < 0;
    double r = 1.0;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

/* src/lapack/families.rs:280-356 (DLATM1) */
int or_latm1(int mode, double cond, int irsign, int dist, uint32_t seed[4], double* d, size_t n) {
    int ordinary = mode != -6 && mode != 0 && mode != 6;
    int info = 0;
    if (mode < -6 || mode > 6)
        info = -1;
    else if (ordinary && irsign != 0 && irsign != 1)
        info = -2;
    else if (ordinary && cond < 1.0)
        info = -3;
    if (info != 0) return info;
    if (n == 0 || mode == 0) return 0;
    switch (abs(mode)) {
        case 1:
            for (size_t i = 0; i < n; i++) d[i] = 1.0 / cond;
            d[0] = 1.0;
            break;
        case 2:
            for (size_t i = 0; i < n; i++) d[i] = 1.0;
            d[n - 1] = 1.0 / cond;
            break;
        case 3:
            d[0] = 1.0;
            if (n > 1) {
                double alpha = pow(cond, -1.0 / (double)(n - 1));
                for (size_t i = 1; i < n; i++) d[i] = powi_binary(alpha, (int)i);
            }
            break;
        case 4:
            d[0] = 1.0;
            if (n > 1) {
                double temp = 1.0 / cond;
                double alpha = (1.0 - temp) / (double)(n - 1);
                for (size_t i = 1; i < n; i++) d[i] = (double)(n - i - 1) * alpha + temp;
            }
            break;
        case 5: {
            double alpha = log(1.0 / cond);
            for (size_t i = 0; i < n; i++) d[i] = exp(alpha * or_laran(seed));
            break;
        }
        default:
            or_larnv(dist, seed, n, d);
            break;
    }
    if (ordinary && irsign == 1)
        for (size_t i = 0; i < n; i++)
            if (or_laran(seed) > 0.5) d[i] = -d[i];
    if (mode < 0)
        for (size_t i = 0; i < n / 2; i++) {
            double t = d[i];
            d[i] = d[n - 1 - i];
            d[n - 1 - i] = t;
        }
    return 0;
}

/* src/lapack/families.rs:363-380: the reference BLAS DNRM2 with a running scale. */
static double dnrm2_reference(const double* x, size_t n) {
    double scale = 0.0, ssq = 1.0;
    for (size_t i = 0; i < n; i++) {
        double absxi = fabs(x[i]);
        if (absxi > 0.0) {
            if (scale < absxi) {
                double r = scale / absxi;
                ssq = 1.0 + ssq * r * r;
                scale = absxi;
            } else {
                double r = absxi / scale;
                ssq += r * r;
            }
        }
    }
    return scale * sqrt(ssq);
}

/* One Householder vector from Normal draws, in place; returns tau (shared by
 * DLAGGE and DLAGSY, src/lapack/families.rs:1009-1023). */
static double householder_from_normal(double* w, size_t len, uint32_t seed[4]) {
    or_larnv(3, seed, len, w);
    double wn = dnrm2_reference(w, len);
    double wa = (w[0] < 0.0) ? -wn : wn;
    if (wn == 0.0) return 0.0;
    double wb = w[0] + wa;
    double inv = 1.0 / wb;
    for (size_t k = 1; k < len; k++) w[k] *= inv;
    w[0] = 1.0;
    return wb / wa;
}

/* src/lapack/families.rs:994-1079 (DLAGGE, dense case): A = U D V. */
static void dlagge(size_t m, size_t n, const double* d, double* a, uint32_t seed[4]) {
    size_t mn = m < n ? m : n;
    size_t lw = m > n ? m : n;
    memset(a, 0, m * n * sizeof(double));
    for (size_t i = 0; i < mn; i++) a[i + i * m] = d[i];
    double* w = (double*)calloc(lw, sizeof(double));
    double* y = (double*)calloc(lw, sizeof(double));
    for (size_t i = mn; i-- > 0;) {
        if (i + 1 < m) {
            size_t len = m - i;
            double tau = householder_from_normal(w, len, seed);
            if (tau != 0.0) {
                size_t ncols = n - i;
                for (size_t j = 0; j < ncols; j++) {
                    double sum = 0.0;
                    for (size_t k = 0; k < len; k++) sum += a[(i + k) + (i + j) * m] * w[k];
                    y[j] = sum;
                }
                for (size_t j = 0; j < ncols; j++)
                    for (size_t k = 0; k < len; k++) a[(i + k) + (i + j) * m] -= tau * w[k] * y[j];
            }
        }
        if (i + 1 < n) {
            size_t len = n - i;
            double tau = householder_from_normal(w, len, seed);
            if (tau != 0.0) {
                size_t nrows = m - i;
                for (size_t k = 0; k < nrows; k++) {
                    double sum = 0.0;
                    for (size_t j = 0; j < len; j++) sum += a[(i + k) + (i + j) * m] * w[j];
                    y[k] = sum;
                }
                for (size_t j = 0; j < len; j++)
                    for (size_t k = 0; k < nrows; k++) a[(i + k) + (i + j) * m] -= tau * y[k] * w[j];
            }
        }
    }
    free(w);
    free(y);
}

/* src/lapack/families.rs:1087-1165 (DLAGSY, dense case): A = U D U^T. */
static void dlagsy(size_t n, const double* d, double* a, uint32_t seed[4]) {
    memset(a, 0, n * n * sizeof(double));
    for (size_t i = 0; i < n; i++) a[i + i * n] = d[i];
    double* w = (double*)calloc(n ? n : 1, sizeof(double));
    double* y = (double*)calloc(n ? n : 1, sizeof(double));
    for (size_t i = (n > 0 ? n - 1 : 0); i-- > 0;) {
        size_t len = n - i;
        double tau = householder_from_normal(w, len, seed);
        if (tau == 0.0) continue;
        for (size_t k = 0; k < len; k++) y[k] = 0.0;
        for (size_t j = 0; j < len; j++) {
            double t1 = tau * w[j];
            double t2 = 0.0;
            y[j] += t1 * a[(i + j) + (i + j) * n];
            for (size_t k = j + 1; k < len; k++) {
                double akj = a[(i + k) + (i + j) * n];
                y[k] += t1 * akj;
                t2 += akj * w[k];
            }
            y[j] += tau * t2;
        }
        double dot = 0.0;
        for (size_t k = 0; k < len; k++) dot += y[k] * w[k];
        double alpha = -0.5 * tau * dot;
        for (size_t k = 0; k < len; k++) y[k] += alpha * w[k];
        for (size_t j = 0; j < len; j++) {
            if (w[j] != 0.0 || y[j] != 0.0) {
                double t1 = -y[j];
                double t2 = -w[j];
                for (size_t k = j; k < len; k++) a[(i + k) + (i + j) * n] += w[k] * t1 + y[k] * t2;
            }
        }
    }
    for (size_t j = 0; j < n; j++)
        for (size_t i = j + 1; i < n; i++) a[j + i * n] = a[i + j * n];
    free(w);
    free(y);
}

/* src/lapack/families.rs:903-986 (DLATMS, dense Method A only). */
int or_dlatms(int dist, int sym, int mode, double cond, double dmax, int pack, size_t m, size_t n, double* a,
              double* d, uint32_t seed[4]) {
    size_t mn = m < n ? m : n;
    if (m == 0 || n == 0) return 0;
    if (sym != 0 && m != n) return -1;
    if (abs(mode) > 6) return -7;
    if (mode != 0 && abs(mode) != 6 && cond < 1.0) return -8;
    if (sym == 0 && pack != 0) return -12;
    for (int s = 0; s < 4; s++) seed[s] %= 4096;
    if (seed[3] % 2 != 1) seed[3] += 1;
    int irsign = (sym == 2) ? 1 : 0;
    if (or_latm1(mode, cond, irsign, dist, seed, d, mn) != 0) return 1;
    if (mode != 0 && abs(mode) != 6) {
        double temp = 0.0;
        for (size_t i = 0; i < mn; i++) temp = fmax(temp, fabs(d[i]));
        if (temp > 0.0) {
            double alpha = dmax / temp;
            for (size_t i = 0; i < mn; i++) d[i] *= alpha;
        } else {
            return 2;
        }
    }
    if (m == 1 && n == 1) {
        memset(a, 0, sizeof(double));
        a[0] = d[0];
    } else if (sym == 0) {
        dlagge(m, n, d, a, seed);
    } else {
        dlagsy(m, d, a, seed);
    }
    if (pack == 1) {
        for (size_t j = 0; j < n; j++)
            for (size_t i = j + 1; i < n; i++) a[i + j * n] = 0.0;
    } else if (pack == 2) {
        for (size_t j = 1; j < n; j++)
            for (size_t i = 0; i < j; i++) a[i + j * n] = 0.0;
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* normwise backward error with compensated residual sums                    */
/* ------------------------------------------------------------------------- */

double or_backward_error(const double* a, const double* x, const double* b, size_t n, size_t nrhs) {
    double anorm = or_dlange('I', a, n, n);
    double worst = 0.0;
    double* r = (double*)calloc(n ? n : 1, sizeof(double));
    double* comp = (double*)calloc(n ? n : 1, sizeof(double));
    for (size_t c = 0; c < nrhs; c++) {
        const double* xc = x + c * n;
        const double* bc = b + c * n;
        for (size_t i = 0; i < n; i++) {
            r[i] = bc[i];
            comp[i] = 0.0;
        }
        for (size_t j = 0; j < n; j++) {
            double xj = xc[j];
            const double* col = a + j * n;
            for (size_t i = 0; i < n; i++) {
                /* Neumaier: r[i] += -(a_ij x_j), with the product's own error kept. */
                double p = col[i] * xj;
                double perr = fma(col[i], xj, -p);
                double t = r[i] - p;
                if (fabs(r[i]) >= fabs(p))
                    comp[i] += (r[i] - t) - p;
                else
                    comp[i] += (-p - t) + r[i];
                comp[i] -= perr;
                r[i] = t;
            }
        }
        double rn = 0.0, xn = 0.0, bn = 0.0;
        for (size_t i = 0; i < n; i++) {
            rn = fmax(rn, fabs(r[i] + comp[i]));
            xn = fmax(xn, fabs(xc[i]));
            bn = fmax(bn, fabs(bc[i]));
        }
        double den = anorm * xn + bn;
        double be = den > 0.0 ? rn / den : (rn > 0.0 ? INFINITY : 0.0);
        if (be > worst || isnan(be)) worst = be;
    }
    free(r);
    free(comp);
    return worst;
}

/* ------------------------------------------------------------------------- */
/* complex twins                                                             */
/* ------------------------------------------------------------------------- */

/*
 * Complex64 is num-complex 0.4's Complex<f64> (Cargo.toml:32-33; the crate is not vendored
 * under the reference): two doubles {re, im}, so a complex slice is interleaved re, im pairs.
 * Its published operators, restated here unfused (rustc does not contract):
 *   a * b = (a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re)
 *   a / b = ((a.re*b.re + a.im*b.im) / q, (a.im*b.re - a.re*b.im) / q), q = b.re*b.re + b.im*b.im
 *   a / r = (a.re / r, a.im / r) for a real r;  a +- b and a -= b componentwise;
 *   norm_sqr(a) = a.re*a.re + a.im*a.im;  conj(a) = (a.re, -a.im).
 * MulAssign and DivAssign expand to the same roundings as * and /.
 */
typedef struct {
    double re, im;
} zc;

static inline zc z_ld(const double* p, size_t i) {
    zc r = {p[2 * i], p[2 * i + 1]};
    return r;
}
static inline void z_st(double* p, size_t i, zc v) {
    p[2 * i] = v.re;
    p[2 * i + 1] = v.im;
}
static inline zc z_mul(zc a, zc b) {
    zc r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}
static inline zc z_div(zc a, zc b) {
    double q = b.re * b.re + b.im * b.im;
    zc r = {(a.re * b.re + a.im * b.im) / q, (a.im * b.re - a.re * b.im) / q};
    return r;
}
static inline zc z_add(zc a, zc b) {
    zc r = {a.re + b.re, a.im + b.im};
    return r;
}
static inline zc z_sub(zc a, zc b) {
    zc r = {a.re - b.re, a.im - b.im};
    return r;
}
static inline zc z_conj(zc a) {
    zc r = {a.re, -a.im};
    return r;
}
/* complex.rs:52-54: IZAMAX's measure |Re| + |Im| */
static inline double z_abs1(zc a) { return fabs(a.re) + fabs(a.im); }

/* src/blas.rs:1023-1118 (zgemm, zgemm_tn, zgemm_nt, zgemm_tt -> zgemm_core): the plain triple
 * loop, op(A) element (i, l) at a[i + l*m] or (transposed, A stored k x m) a[l + i*k], op(B)
 * element (l, j) at b[l + j*k] or (B stored n x k) b[j + l*n]; never conjugated.  Every C
 * element becomes alpha*sum + beta*c, beta*c evaluated even when beta is zero. */
int or_zgemm(int trans_a, int trans_b, const double* alpha, const double* a, const double* b, const double* beta,
             double* c, size_t m, size_t n, size_t k) {
    zc al = {alpha[0], alpha[1]}, be = {beta[0], beta[1]};
    size_t ada = trans_a ? k : m, adb = trans_b ? n : k;
    for (size_t j = 0; j < n; j++) {
        for (size_t i = 0; i < m; i++) {
            zc sum = {0.0, 0.0};
            for (size_t l = 0; l < k; l++) {
                zc av = trans_a ? z_ld(a, l + i * ada) : z_ld(a, i + l * ada);
                zc bv = trans_b ? z_ld(b, j + l * adb) : z_ld(b, l + j * adb);
                sum = z_add(sum, z_mul(av, bv));
            }
            size_t idx = i + j * m;
            z_st(c, idx, z_add(z_mul(al, sum), z_mul(be, z_ld(c, idx))));
        }
    }
    return 0;
}

/* src/lapack/complex.rs:67-121 (zgetf2): first maximum of |Re|+|Im| from the diagonal down,
 * Singular when it is <= 0 (ZERO_DETERMINANT, src/lapack/mod.rs:73), whole-row swap, the column
 * divided by the pivot, then the rank-1 update column by column.  ipiv 0-based. */
int or_zgetrf(double* a, size_t n, size_t* ipiv) {
    for (size_t j = 0; j < n; j++) {
        double best = z_abs1(z_ld(a, j + j * n));
        size_t jp = j;
        for (size_t i = j + 1; i < n; i++) {
            double cand = z_abs1(z_ld(a, i + j * n));
            if (cand > best) {
                best = cand;
                jp = i;
            }
        }
        ipiv[j] = jp;
        if (best <= 0.0) return 1;
        if (jp != j) {
            for (size_t k = 0; k < n; k++) {
                zc t = z_ld(a, k * n + j);
                z_st(a, k * n + j, z_ld(a, k * n + jp));
                z_st(a, k * n + jp, t);
            }
        }
        zc d = z_ld(a, j + j * n);
        for (size_t i = j + 1; i < n; i++) z_st(a, i + j * n, z_div(z_ld(a, i + j * n), d));
        for (size_t k = j + 1; k < n; k++) {
            zc ajk = z_ld(a, j + k * n);
            for (size_t i = j + 1; i < n; i++)
                z_st(a, i + k * n, z_sub(z_ld(a, i + k * n), z_mul(z_ld(a, i + j * n), ajk)));
        }
    }
    return 0;
}

/* src/lapack/complex.rs:169-208: the rectangular panel (nrows >= ncols), factor = a/d then
 * a[i][j] -= factor * a[k][j] row by row. */
int or_zgetrf_rect(double* a, size_t nrows, size_t ncols, size_t* ipiv) {
    if (nrows < ncols) return 2;
    for (size_t k = 0; k < ncols; k++) {
        size_t pivot = k;
        double best = z_abs1(z_ld(a, k + k * nrows));
        for (size_t i = k + 1; i < nrows; i++) {
            double cand = z_abs1(z_ld(a, i + k * nrows));
            if (cand > best) {
                best = cand;
                pivot = i;
            }
        }
        if (best <= 0.0) return 1;
        if (pivot != k) {
            for (size_t j = 0; j < ncols; j++) {
                zc t = z_ld(a, j * nrows + k);
                z_st(a, j * nrows + k, z_ld(a, j * nrows + pivot));
                z_st(a, j * nrows + pivot, t);
            }
        }
        ipiv[k] = pivot;
        zc d = z_ld(a, k + k * nrows);
        for (size_t i = k + 1; i < nrows; i++) {
            zc factor = z_div(z_ld(a, i + k * nrows), d);
            z_st(a, i + k * nrows, factor);
            for (size_t j = k + 1; j < ncols; j++)
                z_st(a, i + j * nrows, z_sub(z_ld(a, i + j * nrows), z_mul(factor, z_ld(a, k + j * nrows))));
        }
    }
    return 0;
}

/* src/lapack/complex.rs:279-320: zgetrf, the interchanges applied to b in order, the unit-lower
 * forward sweep and the upper back substitution (one right-hand side). */
int or_zgesv(double* a, double* b, size_t n) {
    if (n == 0) return 0;
    size_t* ipiv = (size_t*)calloc(n, sizeof(size_t));
    if (!ipiv) return -1;
    int st = or_zgetrf(a, n, ipiv);
    if (st == 0) {
        for (size_t k = 0; k < n; k++) {
            zc t = z_ld(b, k);
            z_st(b, k, z_ld(b, ipiv[k]));
            z_st(b, ipiv[k], t);
        }
        for (size_t i = 0; i < n; i++) {
            zc sum = z_ld(b, i);
            for (size_t j = 0; j < i; j++) sum = z_sub(sum, z_mul(z_ld(a, i + j * n), z_ld(b, j)));
            z_st(b, i, sum);
        }
        for (size_t ii = n; ii-- > 0;) {
            zc sum = z_ld(b, ii);
            for (size_t j = ii + 1; j < n; j++) sum = z_sub(sum, z_mul(z_ld(a, ii + j * n), z_ld(b, j)));
            z_st(b, ii, z_div(sum, z_ld(a, ii + ii * n)));
        }
    }
    free(ipiv);
    return st;
}

/* src/lapack/complex.rs:344-400 (zpotf2): upper -> A = U^H U reading and writing i <= j only;
 * lower -> A = L L^H reading and writing i >= j only.  The diagonal is Re(a_jj) minus the
 * squared moduli; a non-positive or NaN value is written to the diagonal (imaginary part 0)
 * before Singular is returned. */
int or_zpotrf(double* a, size_t n, int upper) {
    for (size_t j = 0; j < n; j++) {
        double ajj = a[2 * (j + j * n)];
        for (size_t q = 0; q < j; q++) {
            zc v = upper ? z_ld(a, q + j * n) : z_ld(a, j + q * n);
            ajj -= v.re * v.re + v.im * v.im;
        }
        if (ajj <= 0.0 || isnan(ajj)) {
            zc bad = {ajj, 0.0};
            z_st(a, j + j * n, bad);
            return 1;
        }
        ajj = sqrt(ajj);
        zc dg = {ajj, 0.0};
        z_st(a, j + j * n, dg);
        for (size_t o = j + 1; o < n; o++) {
            if (upper) {
                zc sum = z_ld(a, j + o * n);
                for (size_t i = 0; i < j; i++) sum = z_sub(sum, z_mul(z_ld(a, i + o * n), z_conj(z_ld(a, i + j * n))));
                zc r = {sum.re / ajj, sum.im / ajj};
                z_st(a, j + o * n, r);
            } else {
                zc sum = z_ld(a, o + j * n);
                for (size_t q = 0; q < j; q++) sum = z_sub(sum, z_mul(z_ld(a, o + q * n), z_conj(z_ld(a, j + q * n))));
                zc r = {sum.re / ajj, sum.im / ajj};
                z_st(a, o + j * n, r);
            }
        }
    }
    return 0;
}
