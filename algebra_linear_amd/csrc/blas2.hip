// Level-2 BLAS, matrix norms, the inverse from an LU factorization and dsytrd's symmetric
// rank-2k update: the crate's support routines around the solve path.
//
//   dgemv / dgemv_trans (src/blas.rs:102-158): y := alpha op(A) x + beta y.  Both kernels keep
//     the crate's summation order and its unfused multiply-then-add, so the results are bit for
//     bit the crate's: dgemv gives each row to one thread, which walks the columns in order
//     (skipping a column whose x_j is zero, as the crate does); dgemv_trans gives each column to
//     one thread, which sums down its column in order from a 64 x 64 tile staged in LDS (the
//     matrix is read once, in 512-byte column runs).  Both are HBM-bound: 8 m n bytes.
//   dlange (src/blas.rs:638-657): 'M' (largest |a_ij|, NaN ignored), '1'/'O' (largest column
//     sum), 'I' (largest row sum) with the crate's sums, bit for bit; 'F'/'E' (dnrm2 of the whole
//     array, src/blas.rs:54-71, scaled by the largest magnitude) with a fixed-order parallel sum,
//     equal to the crate's within (m n) eps relative.
//   dgetri (src/lapack/mod.rs:249-298): X := A^{-1} from A's LU as n solves against one factor:
//     X = P I built from the row origins, then the unit-lower and upper triangular solves (FP64
//     MFMA blocked), then copied over the factor.  A zero on U's diagonal is the crate's Singular
//     and leaves the factor untouched.
//   dsyr2k (src/lapack/sym.rs:582-668): one triangle of C := alpha V W^T + alpha W V^T + beta C
//     as two triangle-masked FP64 MFMA products; the crate forms the full m x m product t first
//     and adds it to beta C, so the two agree within the GEMM rounding bound.
#include "al_internal.hpp"

#include <cmath>

namespace al {
namespace {

#pragma clang fp contract(off)

constexpr int GV_UNROLL = 16;

// y_i := beta y_i (0 when beta == 0, y unread), then y_i += (alpha x_j) a_ij for j in order,
// columns with x_j == 0 skipped (src/blas.rs:102-130).  ABS: |a_ij| with alpha = x_j = 1 (the
// row sums of dlange 'I'), written to y.
template <bool ABS>
__global__ __launch_bounds__(64) void gemv_n_k(int m, int n, double alpha, const double* __restrict__ A, size_t lda,
                                               const double* __restrict__ x, double beta, double* __restrict__ y) {
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i >= m) return;
    double yi = ABS ? 0.0 : (beta == 0.0 ? 0.0 : (beta == 1.0 ? y[i] : y[i] * beta));
    const double* a = A + i;
    int j = 0;
    for (; j + GV_UNROLL <= n; j += GV_UNROLL) {
        double av[GV_UNROLL], xv[GV_UNROLL];
#pragma unroll
        for (int q = 0; q < GV_UNROLL; ++q) {
            av[q] = a[(size_t)(j + q) * lda];
            xv[q] = ABS ? 1.0 : x[j + q];
        }
#pragma unroll
        for (int q = 0; q < GV_UNROLL; ++q) {
            if (ABS) {
                yi = yi + fabs(av[q]);
            } else if (xv[q] != 0.0) {
                const double c = alpha * xv[q];
                yi = yi + c * av[q];
            }
        }
    }
    for (; j < n; ++j) {
        const double aij = a[(size_t)j * lda];
        if (ABS) {
            yi = yi + fabs(aij);
        } else {
            const double xj = x[j];
            if (xj != 0.0) {
                const double c = alpha * xj;
                yi = yi + c * aij;
            }
        }
    }
    y[i] = yi;
}

constexpr int GT_TILE = 64;

// sum_j = sum over i in order of a_ij x_i (ABS: |a_ij|), then y_j := alpha sum_j (+ beta y_j
// when beta != 0) (src/blas.rs:145-158).  256 threads per 64 columns: the four waves load the
// 64 x 64 tile (wave w: columns 16w..16w+15, a lane per row) and x's 64 entries into LDS, the
// first wave sums, one thread per column.  The next tile is requested before the current one is
// summed.
template <bool ABS>
__global__ __launch_bounds__(256) void gemv_t_k(int m, int n, double alpha, const double* __restrict__ A, size_t lda,
                                                const double* __restrict__ x, double beta, double* __restrict__ y) {
    __shared__ double T[GT_TILE][GT_TILE + 1];  // T[c][r] = a(i0 + r, j0 + c)
    __shared__ double xs[GT_TILE];
    const int t = threadIdx.x, r = t & 63, cg = t >> 6;
    const int j0 = blockIdx.x * GT_TILE;
    const int col = j0 + t;  // the column summed by thread t < 64
    double sum = 0.0;
    double nxt[16];
    double xn = 0.0;
    auto load = [&](int i0) {
        const int i = i0 + r;
        const bool rok = i < m;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int c = j0 + cg * 16 + q;
            nxt[q] = (rok && c < n) ? A[(size_t)i + (size_t)c * lda] : 0.0;
        }
        if (cg == 0) xn = (!ABS && rok) ? x[i] : 0.0;
    };
    if (m > 0) load(0);
    for (int i0 = 0; i0 < m; i0 += GT_TILE) {
        __syncthreads();  // the previous tile has been summed
#pragma unroll
        for (int q = 0; q < 16; ++q) T[cg * 16 + q][r] = nxt[q];
        if (cg == 0) xs[r] = xn;
        __syncthreads();
        if (i0 + GT_TILE < m) load(i0 + GT_TILE);
        if (t < GT_TILE) {
            const int rmax = (m - i0) < GT_TILE ? (m - i0) : GT_TILE;
            for (int q = 0; q < rmax; ++q) sum = ABS ? sum + fabs(T[t][q]) : sum + T[t][q] * xs[q];
        }
    }
    if (t < GT_TILE && col < n) y[col] = ABS ? sum : (beta == 0.0 ? alpha * sum : alpha * sum + beta * y[col]);
}

// ---- dlange helpers: a non-negative double's bits order like the value, so a maximum of
// non-negative values (NaN skipped, as f64::max does) is an atomic maximum of the bits.
__device__ __forceinline__ void max_bits(unsigned long long* out, double v) {
    if (v == v) __hip_atomic_fetch_max(out, (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double wave_fmax(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// largest |a_ij| over the m x n matrix (fmax ignores NaN) into *out (bits; zeroed beforehand)
__global__ __launch_bounds__(256) void absmax_k(int m, int n, const double* __restrict__ A, size_t lda,
                                                unsigned long long* __restrict__ out) {
    __shared__ double red[4];
    const size_t total = (size_t)m * (size_t)n;
    double v = 0.0;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
        const size_t j = e / (size_t)m, i = e - j * (size_t)m;
        v = fmax(v, fabs(A[i + j * lda]));
    }
    v = wave_fmax(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) max_bits(out, fmax(fmax(red[0], red[1]), fmax(red[2], red[3])));
}

// largest of the k values in v (NaN skipped) into *out (bits)
__global__ __launch_bounds__(256) void vmax_k(int k, const double* __restrict__ v, unsigned long long* __restrict__ out) {
    __shared__ double red[4];
    double a = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < k; i += gridDim.x * 256) a = fmax(a, v[i]);
    a = wave_fmax(a);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) max_bits(out, fmax(fmax(red[0], red[1]), fmax(red[2], red[3])));
}

constexpr int SQ_BLOCKS = 512;

// Frobenius norm, pass 2: each workgroup sums the squares (scaled by 1/scale when the largest
// magnitude is finite and nonzero, as dnrm2 does) of a fixed contiguous range of the linear
// index e = i + j m, threads strided, then a fixed tree; partial[blockIdx.x] = the sum.
__global__ __launch_bounds__(256) void sumsq_k(int m, int n, const double* __restrict__ A, size_t lda,
                                               const unsigned long long* __restrict__ scale_bits,
                                               double* __restrict__ partial) {
    __shared__ double red[256];
    const double scale = __longlong_as_double((long long)*scale_bits);
    const bool scaled = scale != 0.0 && isfinite(scale);
    const double inv = scaled ? 1.0 / scale : 1.0;
    const size_t total = (size_t)m * (size_t)n;
    const size_t per = (total + gridDim.x - 1) / gridDim.x;
    const size_t e0 = (size_t)blockIdx.x * per, e1 = e0 + per < total ? e0 + per : total;
    double s = 0.0;
    for (size_t e = e0 + threadIdx.x; e < e1; e += 256) {
        const size_t j = e / (size_t)m, i = e - j * (size_t)m;
        const double v = A[i + j * lda];
        const double w = scaled ? v * inv : v;
        s = s + w * w;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h >= 1; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Frobenius norm, pass 3: the partial sums in order, then scale * sqrt(sum) (or sqrt(sum) of the
// unscaled squares when the largest magnitude is 0 or not finite).
__global__ void fro_finish_k(int nparts, const double* __restrict__ partial,
                             const unsigned long long* __restrict__ scale_bits, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double scale = __longlong_as_double((long long)*scale_bits);
    double ss = 0.0;
    for (int p = 0; p < nparts; ++p) ss = ss + partial[p];
    *out = (scale != 0.0 && isfinite(scale)) ? scale * sqrt(ss) : sqrt(ss);
}

__global__ void bits_to_double_k(const unsigned long long* __restrict__ bits, double* __restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = __longlong_as_double((long long)*bits);
}

// ---- dgetri helpers
// info := 1-based index of the last exactly zero diagonal entry of U (the one the crate's first
// back substitution meets first), 0 if none; info starts at 0
__global__ __launch_bounds__(256) void diag_zero_k(int n, const double* __restrict__ LU, size_t lda, int* info) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        if (LU[(size_t)i + (size_t)i * lda] == 0.0) atomicMax(info, i + 1);
}

// X (n x n, leading dimension n) := P I: X[i][c] = 1 where row i of the permuted identity came
// from row c (perm[i] == c), else 0; skipped once info != 0
__global__ __launch_bounds__(256) void perm_eye_k(int n, const double* __restrict__ perm, double* __restrict__ X,
                                                  const int* __restrict__ info) {
    if (*info != 0) return;
    const size_t total = (size_t)n * (size_t)n;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
        const size_t c = e / (size_t)n, i = e - c * (size_t)n;
        X[e] = ((size_t)perm[i] == c) ? 1.0 : 0.0;
    }
}

// LU (leading dimension lda) := X (leading dimension n) when info == 0
__global__ __launch_bounds__(256) void copy_if_ok_k(int n, const double* __restrict__ X, double* __restrict__ LU,
                                                    size_t lda, const int* __restrict__ info) {
    if (*info != 0) return;
    const size_t total = (size_t)n * (size_t)n;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
        const size_t c = e / (size_t)n, i = e - c * (size_t)n;
        LU[i + c * lda] = X[e];
    }
}

// ---- dsyr2k's alpha == 0, k == 0 branch: C := beta C on one triangle (0 when beta == 0)
__global__ __launch_bounds__(256) void scale_tri_k(int m, bool upper, double beta, double* __restrict__ C, size_t ldc) {
    const size_t total = (size_t)m * (size_t)m;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
        const size_t j = e / (size_t)m, i = e - j * (size_t)m;
        if (upper ? (i <= j) : (i >= j)) C[i + j * ldc] = (beta == 0.0) ? 0.0 : C[i + j * ldc] * beta;
    }
}

#pragma clang fp contract(on)

unsigned grid_for(size_t total, unsigned cap = 2048) {
    const size_t g = (total + 255) / 256;
    return (unsigned)(g == 0 ? 1 : (g > cap ? cap : g));
}

}  // namespace

int gemv_f64(hipStream_t s, bool trans, size_t m, size_t n, double alpha, const double* A, size_t lda,
             const double* x, double beta, double* y) {
    if (m > INT32_MAX || n > INT32_MAX) {
        set_error("dgemv: dimension beyond 2^31");
        return AL_ERR_ARG;
    }
    const int ticket = prof_begin(s);
    if (!trans) {
        if (m == 0) return AL_OK;
        hipLaunchKernelGGL(gemv_n_k<false>, dim3(ceil_div(m, 64)), dim3(64), 0, s, (int)m, (int)n, alpha, A, lda, x,
                           beta, y);
    } else {
        if (n == 0) return AL_OK;
        hipLaunchKernelGGL(gemv_t_k<false>, dim3(ceil_div(n, GT_TILE)), dim3(256), 0, s, (int)m, (int)n, alpha, A,
                           lda, x, beta, y);
    }
    AL_HIP(hipGetLastError());
    if (ticket >= 0) {
        const double dm = (double)m, dn = (double)n;
        prof_end(ticket, s, trans ? "dgemv_t" : "dgemv_n", 2.0 * dm * dn,
                 8.0 * (dm * dn + (trans ? dm + 2.0 * dn : dn + 2.0 * dm)));
    }
    return AL_OK;
}

size_t lange_scratch_doubles(size_t m, size_t n) {
    size_t w = m > n ? m : n;
    return 8 + (w > (size_t)SQ_BLOCKS ? w : (size_t)SQ_BLOCKS);
}

int lange_f64(hipStream_t s, char norm, size_t m, size_t n, const double* A, size_t lda, double* out,
              double* scratch) {
    if (m > INT32_MAX || n > INT32_MAX) {
        set_error("dlange: dimension beyond 2^31");
        return AL_ERR_ARG;
    }
    unsigned long long* mx = (unsigned long long*)scratch;  // scratch[0]: running maximum (bits)
    double* partial = scratch + 8;                           // partial sums, or one sum per line
    const size_t total = m * n;
    switch (norm) {
        case 'M': case 'm':
        case '1': case 'O': case 'o':
        case 'I': case 'i':
        case 'F': case 'f': case 'E': case 'e':
            break;
        default:
            set_error(std::string("dlange: unrecognized norm '") + norm + "'; expected one of 1, O, I, F, E, M");
            return AL_ERR_ARG;
    }
    AL_CHECK(fill_u32(s, (unsigned*)mx, 2, 0u));  // +0.0: the crate's fold start
    const int ticket = prof_begin(s);
    if (total == 0) {
        // every norm of an empty matrix is 0
    } else if (norm == 'M' || norm == 'm') {
        hipLaunchKernelGGL(absmax_k, dim3(grid_for(total)), dim3(256), 0, s, (int)m, (int)n, A, lda, mx);
    } else if (norm == 'I' || norm == 'i') {
        hipLaunchKernelGGL(gemv_n_k<true>, dim3(ceil_div(m, 64)), dim3(64), 0, s, (int)m, (int)n, 1.0, A, lda,
                           (const double*)nullptr, 0.0, partial);
        hipLaunchKernelGGL(vmax_k, dim3(grid_for(m, 64)), dim3(256), 0, s, (int)m, partial, mx);
    } else if (norm == 'F' || norm == 'f' || norm == 'E' || norm == 'e') {
        hipLaunchKernelGGL(absmax_k, dim3(grid_for(total)), dim3(256), 0, s, (int)m, (int)n, A, lda, mx);
        const unsigned nb = grid_for(total, SQ_BLOCKS);
        hipLaunchKernelGGL(sumsq_k, dim3(nb), dim3(256), 0, s, (int)m, (int)n, A, lda, mx, partial);
        hipLaunchKernelGGL(fro_finish_k, dim3(1), dim3(64), 0, s, (int)nb, partial, mx, out);
        AL_HIP(hipGetLastError());
        if (ticket >= 0) prof_end(ticket, s, "dlange", 4.0 * (double)total, 16.0 * (double)total);
        return AL_OK;
    } else {  // one-norm: the column sums
        hipLaunchKernelGGL(gemv_t_k<true>, dim3(ceil_div(n, GT_TILE)), dim3(256), 0, s, (int)m, (int)n, 1.0, A, lda,
                           (const double*)nullptr, 0.0, partial);
        hipLaunchKernelGGL(vmax_k, dim3(grid_for(n, 64)), dim3(256), 0, s, (int)n, partial, mx);
    }
    hipLaunchKernelGGL(bits_to_double_k, dim3(1), dim3(64), 0, s, mx, out);
    AL_HIP(hipGetLastError());
    if (ticket >= 0) prof_end(ticket, s, "dlange", 2.0 * (double)total, 8.0 * (double)total);
    return AL_OK;
}

int getri_f64(hipStream_t s, Workspace* ws, size_t n, double* LU, size_t lda, const int* ipiv, int* info) {
    if (n == 0) return fill_i32(s, info, 1, 0);
    if (n > INT32_MAX) {
        set_error("dgetri: order beyond 2^31");
        return AL_ERR_ARG;
    }
    AL_CHECK(ws_reserve(ws, n, n * n));
    double* X = ws->wbuf;
    double* perm = ws->perm;
    AL_CHECK(perm_from_ipiv(s, ipiv, n, perm));  // the interchanges composed, in dgetrf's order
    AL_CHECK(fill_i32(s, info, 1, 0));
    hipLaunchKernelGGL(diag_zero_k, dim3(grid_for(n, 64)), dim3(256), 0, s, (int)n, LU, lda, info);
    AL_HIP(hipGetLastError());
    hipLaunchKernelGGL(perm_eye_k, dim3(grid_for(n * n)), dim3(256), 0, s, (int)n, perm, X, info);
    AL_HIP(hipGetLastError());
    // L^{-1} (unit lower) then U^{-1}: every flop of the n x n solves on the FP64 matrix cores
    AL_CHECK(trsm_f64(s, true, true, false, true, n, n, 1.0, LU, lda, X, n, info));
    AL_CHECK(trsm_f64(s, true, false, false, false, n, n, 1.0, LU, lda, X, n, info));
    hipLaunchKernelGGL(copy_if_ok_k, dim3(grid_for(n * n)), dim3(256), 0, s, (int)n, X, LU, lda, info);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

int syr2k_f64(hipStream_t s, bool upper, size_t m, size_t k, double alpha, const double* V, size_t lv,
              const double* W, size_t lw, double beta, double* C, size_t lc) {
    if (m == 0) return AL_OK;
    if (m > INT32_MAX || k > INT32_MAX) {
        set_error("dsyr2k: dimension beyond 2^31");
        return AL_ERR_ARG;
    }
    if (k == 0) {
        // the crate touches C here only when alpha == 0 (src/lapack/sym.rs:605-610)
        if (alpha == 0.0 && beta != 1.0) {
            hipLaunchKernelGGL(scale_tri_k, dim3(grid_for(m * m)), dim3(256), 0, s, (int)m, upper, beta, C, lc);
            AL_HIP(hipGetLastError());
        }
        return AL_OK;
    }
    const int tri = upper ? 2 : 1;
    AL_CHECK(gemm_f64(s, false, true, m, m, k, alpha, V, lv, W, lw, beta, C, lc, tri));
    return gemm_f64(s, false, true, m, m, k, alpha, W, lw, V, lv, 1.0, C, lc, tri);
}

}  // namespace al
