// Complex twins of the dense path: blas::zgemm{,_tn,_nt,_tt} (src/blas.rs:1023-1118),
// lapack::zgetrf / zgetrf_rect (src/lapack/complex.rs:67-121, :169-208), zgesv (:279-320) and
// zpotrf (:344-400).
//
// The crate runs these as naive loops on Complex64.  On MI355X the flops go to the FP64 matrix
// cores through the real form of a complex matrix: a complex m x n matrix Z stored interleaved
// (re, im) with leading dimension ld is, read as doubles, the 2m x n real matrix V(Z) (leading
// dimension 2 ld, even rows real parts, odd rows imaginary parts), and the 2m x 2n matrix R(Z)
// that replaces every entry x + iy by the block [[x, -y], [y, x]] satisfies
//     V(Z W) = R(Z) V(W),   R(Z W) = R(Z) R(W),   R(Z^H) = R(Z)^T.
// So a complex product is ONE real FP64 MFMA GEMM of depth 2k whose left operand R(op(A)) is
// staged (2x the storage of A) while B and C are used in place through V -- 8mnk real flops,
// the complex product's own count, with no separate real and imaginary passes.  The same
// identities turn the triangular solves into real ones: R of a unit-lower complex factor is
// unit-lower, R of a Cholesky factor (real diagonal) is triangular, and an upper factor U is
// solved as the unit-upper diag(U)^-1 U after dividing the right-hand side by diag(U).
//
// Pivoting stays complex: the LU panel is the cooperative panel kernel instantiated for `zd`
// (|Re| + |Im| keys, the crate's complex division and multiply-subtract, 32 columns), and the
// Cholesky diagonal blocks are factored by `zpotf2_k` in the crate's own order and arithmetic,
// so a matrix that fits one panel (LU: n <= 32) or one diagonal block (Cholesky: n <= 64)
// factors bit for bit as the crate factors it.  Larger matrices agree to rounding: the blocked
// updates sum in another order.
#include <algorithm>
#include <cmath>

#include "al_internal.hpp"

namespace al {
namespace {

constexpr int ZT = 32;        // tile of the staging kernels
constexpr int ZNBO = 128;     // LU outer block: the trailing GEMM has real depth 256
constexpr int ZPB = 64;       // Cholesky diagonal block (one wave)
constexpr int ZPBO = 256;     // Cholesky outer block: trailing depth 512

enum Part : int {
    PART_FULL = 0,
    PART_UNIT_LOWER = 1,  // strictly lower entries, identity blocks on the diagonal, zeros above
    PART_LOWER = 2,       // lower triangle with the diagonal, zeros above
    PART_UPPER = 3,       // upper triangle with the diagonal, zeros below
    PART_LU_SPLIT = 4     // strictly lower L, identity diagonal, strictly upper diag(U)^-1 U
};

__device__ __forceinline__ bool part_reads(int part, int i, int l) {
    switch (part) {
        case PART_UNIT_LOWER: return l < i;
        case PART_LOWER: return l <= i;
        case PART_UPPER: return l >= i;
        case PART_LU_SPLIT: return l != i;
        default: return true;
    }
}

// R (2m x 2k, leading dimension ldr) := R(f(op(A))) restricted to `part`, where op(A)(i, l) is
// A[i + l*lda] or (TRANS) A[l + i*lda], conjugated if conj_a, and f multiplies by alpha
// (scale 2), negates (scale 1, exact) or is the identity (scale 0).  32 x 32 tiles through LDS
// so that both the reads of A (either orientation) and the 16-byte writes of R are coalesced.
template <bool TRANS>
__global__ __launch_bounds__(256) void expand_k(int m, int k, const zd* __restrict__ A, size_t lda, int conj_a,
                                                int scale, zd alpha, int part, double* __restrict__ R, size_t ldr) {
    __shared__ zd tile[ZT][ZT + 1];
    const int i0 = blockIdx.x * ZT, l0 = blockIdx.y * ZT;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int yy = ty + 8 * r;
        if (TRANS) {
            const int i = i0 + yy, l = l0 + tx;
            zd v(0.0);
            if (i < m && l < k && part_reads(part, i, l)) v = A[(size_t)l + (size_t)i * lda];
            tile[yy][tx] = v;
        } else {
            const int i = i0 + tx, l = l0 + yy;
            zd v(0.0);
            if (i < m && l < k && part_reads(part, i, l)) v = A[(size_t)i + (size_t)l * lda];
            tile[tx][yy] = v;
        }
    }
    __syncthreads();
    const int i = i0 + tx;
    if (i >= m) return;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int l = l0 + ty + 8 * r;
        if (l >= k) continue;
        zd v = tile[tx][ty + 8 * r];
        if (conj_a) v = conj(v);
        if (part == PART_LU_SPLIT && l > i) {
            v = v / A[(size_t)i + (size_t)i * lda];  // diag(U)^-1 U: U(i, i) of the untransposed factor
        }
        if (scale == 1) v = zd(-v.re, -v.im);
        else if (scale == 2) v = alpha * v;
        if (l == i && (part == PART_UNIT_LOWER || part == PART_LU_SPLIT)) v = zd(1.0, 0.0);
        double* c0 = R + 2 * (size_t)i + 2 * (size_t)l * ldr;
        double* c1 = c0 + ldr;
        c0[0] = v.re;
        c0[1] = v.im;
        c1[0] = -v.im;
        c1[1] = v.re;
    }
}

// dst (cols x rows, leading dimension ldd) := src^T or src^H, src rows x cols (lds).
__global__ __launch_bounds__(256) void ztrans_k(int rows, int cols, const zd* __restrict__ src, size_t lds,
                                                int conj_a, zd* __restrict__ dst, size_t ldd) {
    __shared__ zd tile[ZT][ZT + 1];
    const int r0 = blockIdx.x * ZT, c0 = blockIdx.y * ZT;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int r = r0 + tx, c = c0 + ty + 8 * q;
        if (r < rows && c < cols) tile[ty + 8 * q][tx] = src[(size_t)r + (size_t)c * lds];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = c0 + tx, r = r0 + ty + 8 * q;
        if (r < rows && c < cols) {
            zd v = tile[tx][ty + 8 * q];
            dst[(size_t)c + (size_t)r * ldd] = conj_a ? conj(v) : v;
        }
    }
}

// C := beta * C, evaluated for every entry (the crate's alpha*sum + beta*c always forms beta*c).
__global__ void zscale_k(int m, int n, zd beta, zd* __restrict__ C, size_t ldc) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)m * n) return;
    const size_t i = idx % m, j = idx / m;
    zd* p = C + i + j * ldc;
    *p = beta * *p;
}

__global__ void zgather_k(int n, int nrhs, const zd* __restrict__ B, size_t ldb, const double* __restrict__ perm,
                          zd* __restrict__ W) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)n * nrhs) return;
    const size_t i = idx % n, c = idx / n;
    W[idx] = B[(size_t)perm[i] + c * ldb];
}

__global__ void zcopy_out_k(int n, int nrhs, const zd* __restrict__ W, zd* __restrict__ B, size_t ldb) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)n * nrhs) return;
    const size_t i = idx % n, c = idx / n;
    B[i + c * ldb] = W[idx];
}

// W(i, c) := W(i, c) / LU(i, i), the complex division of the crate's back substitution.
__global__ void zdiv_diag_k(int n, int nrhs, const zd* __restrict__ LU, size_t lda, zd* __restrict__ W) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)n * nrhs) return;
    const size_t i = idx % n;
    W[idx] = W[idx] / LU[i + i * lda];
}

// zpotf2 (src/lapack/complex.rs:344-400) on an nb x nb (nb <= 64) diagonal block, in the
// crate's order: column (row, for UPPER) j takes ajj = Re(a_jj) minus the squared moduli of
// the finished part, fails on ajj <= 0 or NaN (writing (ajj, 0) to the diagonal, info =
// col0 + j + 1), else the square root goes on the diagonal and the other entries of the
// column (row) subtract a_ik conj(a_jk) (upper: a_ik conj(a_ij)) in k order and divide by it.
// The block sits in LDS; only the requested triangle is read or written.
template <bool UPPER>
__global__ __launch_bounds__(64) void zpotf2_k(zd* A, size_t lda, int nb, int col0, int* info) {
#pragma clang fp contract(off)
    if (*info != 0) return;
    __shared__ zd L[ZPB][ZPB + 1];
    __shared__ double djj;
    __shared__ int bad;
    const int t = threadIdx.x;
    for (int idx = t; idx < nb * nb; idx += 64) {
        const int i = idx % nb, j = idx / nb;
        if (UPPER ? (i <= j) : (i >= j)) L[i][j] = A[(size_t)i + (size_t)j * lda];
    }
    if (t == 0) bad = 0;
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
        if (t == 0) {
            double ajj = L[j][j].re;
            for (int q = 0; q < j; ++q) ajj -= norm_sqr(UPPER ? L[q][j] : L[j][q]);
            if (ajj <= 0.0 || ajj != ajj) {
                L[j][j] = zd(ajj, 0.0);
                bad = 1;
                *info = col0 + j + 1;
            } else {
                const double r = sqrt(ajj);
                L[j][j] = zd(r, 0.0);
                djj = r;
            }
        }
        __syncthreads();
        if (bad) break;
        const double d = djj;
        for (int o = j + 1 + t; o < nb; o += 64) {
            if (UPPER) {
                zd sum = L[j][o];
                for (int i = 0; i < j; ++i) sum -= L[i][o] * conj(L[i][j]);
                L[j][o] = sum / d;
            } else {
                zd sum = L[o][j];
                for (int q = 0; q < j; ++q) sum -= L[o][q] * conj(L[j][q]);
                L[o][j] = sum / d;
            }
        }
        __syncthreads();
    }
    __syncthreads();
    for (int idx = t; idx < nb * nb; idx += 64) {
        const int i = idx % nb, j = idx / nb;
        if (UPPER ? (i <= j) : (i >= j)) A[(size_t)i + (size_t)j * lda] = L[i][j];
    }
}

// ---------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------

int zbuf(Workspace* ws, size_t bytes, void** out) {
    if (ws->zbuf_bytes < bytes) {
        if (ws->zbuf) AL_HIP(hipFree(ws->zbuf));
        ws->zbuf = nullptr;
        ws->zbuf_bytes = 0;
        if (hipMalloc(&ws->zbuf, bytes) != hipSuccess) {
            (void)hipGetLastError();
            ws->zbuf = nullptr;
            set_error("device allocation failed (complex staging)");
            return AL_ERR_OOM;
        }
        ws->zbuf_bytes = bytes;
    }
    *out = ws->zbuf;
    return AL_OK;
}

inline size_t round256(size_t b) { return (b + 255) / 256 * 256; }

int expand(hipStream_t s, bool trans, size_t m, size_t k, const zd* A, size_t lda, bool conj_a, int scale, zd alpha,
           int part, double* R, size_t ldr) {
    if (m == 0 || k == 0) return AL_OK;
    const dim3 grid(ceil_div(m, ZT), ceil_div(k, ZT));
    if (trans)
        hipLaunchKernelGGL(expand_k<true>, grid, dim3(256), 0, s, (int)m, (int)k, A, lda, conj_a ? 1 : 0, scale, alpha,
                           part, R, ldr);
    else
        hipLaunchKernelGGL(expand_k<false>, grid, dim3(256), 0, s, (int)m, (int)k, A, lda, conj_a ? 1 : 0, scale,
                           alpha, part, R, ldr);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

int ztrans(hipStream_t s, size_t rows, size_t cols, const zd* src, size_t lds, bool conj_a, zd* dst, size_t ldd) {
    if (rows == 0 || cols == 0) return AL_OK;
    hipLaunchKernelGGL(ztrans_k, dim3(ceil_div(rows, ZT), ceil_div(cols, ZT)), dim3(256), 0, s, (int)rows, (int)cols,
                       src, lds, conj_a ? 1 : 0, dst, ldd);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

// C (m x n) += R(-L) V(U): the complex update C -= L U as one real GEMM; R holds R(-L)
// (2m x 2kb, leading dimension 2m), U is kb x n at (U, ldu), C at (C, ldc).
int minus_lu(hipStream_t s, size_t m, size_t n, size_t kb, const double* Rneg, const zd* U, size_t ldu, zd* C,
             size_t ldc) {
    if (m == 0 || n == 0 || kb == 0) return AL_OK;
    return gemm_f64(s, false, false, 2 * m, n, 2 * kb, 1.0, Rneg, 2 * m, (const double*)U, 2 * ldu, 1.0, (double*)C,
                    2 * ldc);
}

// U12 := L11^-1 U12 with L11 unit lower (jb x jb at L, R11 scratch of (2jb)^2 doubles).
int unit_lower_solve(hipStream_t s, size_t jb, const zd* L, size_t lda, zd* B, size_t ldb, size_t ncols, double* R11,
                     const int* info) {
    if (jb == 0 || ncols == 0) return AL_OK;
    AL_CHECK(expand(s, false, jb, jb, L, lda, false, 0, zd(1.0), PART_UNIT_LOWER, R11, 2 * jb));
    return trsm_f64(s, true, true, false, true, 2 * jb, ncols, 1.0, R11, 2 * jb, (double*)B, 2 * ldb, info);
}

// One outer block of the complex LU: 32-wide cooperative panels over columns [k0, k0 + n),
// swaps, U12 solves and updates inside the block only.
int zpanel_blocked(hipStream_t s, Workspace* ws, zd* A, size_t lda, size_t M, size_t k0, size_t n, int* ipiv,
                   int* info, double* R11, double* R21) {
    const size_t PW = (size_t)panel_z_width();
    for (size_t kk = k0; kk < k0 + n; kk += PW) {
        const size_t ib = std::min(PW, k0 + n - kk);
        AL_CHECK(panel_z(s, ws, A + kk + kk * lda, lda, kk, M - kk, ib, ipiv, info));
        const size_t rest = k0 + n - kk - ib;
        AL_CHECK(swap_block_z(s, ws, ipiv, kk, ib, A + k0 * lda, kk - k0, A + (kk + ib) * lda, rest, lda, info));
        if (rest > 0) {
            zd* A12 = A + kk + (kk + ib) * lda;
            AL_CHECK(unit_lower_solve(s, ib, A + kk + kk * lda, lda, A12, lda, rest, R11, info));
            const size_t mb = M - kk - ib;
            if (mb > 0) {
                AL_CHECK(expand(s, false, mb, ib, A + (kk + ib) + kk * lda, lda, false, 1, zd(1.0), PART_FULL, R21,
                                2 * mb));
                AL_CHECK(minus_lu(s, mb, rest, ib, R21, A12, lda, A + (kk + ib) + (kk + ib) * lda, lda));
            }
        }
    }
    return AL_OK;
}

int zgetrf_enqueue(hipStream_t s, Workspace* ws, size_t M, size_t N, zd* A, size_t lda, int* ipiv, int* info) {
    const size_t kmax = std::min(M, N);
    const size_t r11 = round256((size_t)(2 * ZNBO) * (2 * ZNBO) * sizeof(double));
    const size_t r21 = round256(2 * M * (2 * ZNBO) * sizeof(double));
    void* buf;
    AL_CHECK(zbuf(ws, r11 + r21, &buf));
    double* R11 = (double*)buf;
    double* R21 = (double*)((char*)buf + r11);
    AL_CHECK(fill_i32(s, info, 1, 0));
    AL_CHECK(fill_u32(s, ws->abort_flag, 1, 0u));
    for (size_t k = 0; k < kmax; k += ZNBO) {
        const size_t jb = std::min((size_t)ZNBO, kmax - k);
        AL_CHECK(zpanel_blocked(s, ws, A, lda, M, k, jb, ipiv, info, R11, R21));
        const size_t right = (k + jb < N) ? N - k - jb : 0;
        AL_CHECK(swap_block_z(s, ws, ipiv, k, jb, A, k, A + (k + jb) * lda, right, lda, info));
        if (right > 0) {
            zd* A12 = A + k + (k + jb) * lda;
            AL_CHECK(unit_lower_solve(s, jb, A + k + k * lda, lda, A12, lda, right, R11, info));
            const size_t mb = M - k - jb;
            if (mb > 0) {
                AL_CHECK(expand(s, false, mb, jb, A + (k + jb) + k * lda, lda, false, 1, zd(1.0), PART_FULL, R21,
                                2 * mb));
                AL_CHECK(minus_lu(s, mb, right, jb, R21, A12, lda, A + (k + jb) + (k + jb) * lda, lda));
            }
        }
    }
    return AL_OK;
}

// Blocked complex Cholesky of the n x n block at A (global column offset col0), diagonal
// blocks of nb (<= ZPB: one zpotf2_k; larger: one level down with ZPB).
//   lower: L21 := A21 L11^-H solved as L11 W = A21^H (W = L21^H staged, R(L11) lower
//          triangular since the diagonal is real), A21 := W^H, then the lower triangle of
//          A22 -= L21 L21^H as R(-L21) V(W), masked to complex rows >= columns;
//   upper: U12 := U11^-H A12 in place as R(U11)^T V(U12) = V(A12), then the upper triangle of
//          A22 -= U12^H U12 as R(-U12)^T V(U12).
int zpotrf_level(hipStream_t s, size_t n, zd* A, size_t lda, int* info, size_t col0, size_t nb, bool upper,
                 double* R11, zd* W, double* R21) {
    for (size_t k = 0; k < n; k += nb) {
        const size_t jb = std::min(nb, n - k);
        zd* Akk = A + k + k * lda;
        if (jb <= (size_t)ZPB) {
            if (upper)
                hipLaunchKernelGGL(zpotf2_k<true>, dim3(1), dim3(64), 0, s, Akk, lda, (int)jb, (int)(col0 + k), info);
            else
                hipLaunchKernelGGL(zpotf2_k<false>, dim3(1), dim3(64), 0, s, Akk, lda, (int)jb, (int)(col0 + k), info);
            AL_HIP(hipGetLastError());
        } else {
            AL_CHECK(zpotrf_level(s, jb, Akk, lda, info, col0 + k, ZPB, upper, R11, W, R21));
        }
        if (k + jb >= n) break;
        const size_t mr = n - k - jb;
        zd* A22 = Akk + jb + jb * lda;
        if (!upper) {
            zd* A21 = Akk + jb;
            AL_CHECK(ztrans(s, mr, jb, A21, lda, true, W, jb));  // W = A21^H (jb x mr)
            AL_CHECK(expand(s, false, jb, jb, Akk, lda, false, 0, zd(1.0), PART_LOWER, R11, 2 * jb));
            AL_CHECK(trsm_f64(s, true, true, false, false, 2 * jb, mr, 1.0, R11, 2 * jb, (double*)W, 2 * jb, info));
            AL_CHECK(ztrans(s, jb, mr, W, jb, true, A21, lda));  // A21 = W^H = L21
            AL_CHECK(expand(s, false, mr, jb, A21, lda, false, 1, zd(1.0), PART_FULL, R21, 2 * mr));
            AL_CHECK(gemm_f64_ctri(s, false, 2 * mr, mr, 2 * jb, 1.0, R21, 2 * mr, (const double*)W, 2 * jb, 1.0,
                                   (double*)A22, 2 * lda));
        } else {
            zd* A12 = Akk + jb * lda;
            AL_CHECK(expand(s, false, jb, jb, Akk, lda, false, 0, zd(1.0), PART_UPPER, R11, 2 * jb));
            AL_CHECK(trsm_f64(s, true, false, true, false, 2 * jb, mr, 1.0, R11, 2 * jb, (double*)A12, 2 * lda, info));
            AL_CHECK(expand(s, false, jb, mr, A12, lda, false, 1, zd(1.0), PART_FULL, R21, 2 * jb));
            AL_CHECK(gemm_f64_ctri(s, true, 2 * mr, mr, 2 * jb, 1.0, R21, 2 * jb, (const double*)A12, 2 * lda, 1.0,
                                   (double*)A22, 2 * lda));
        }
    }
    return AL_OK;
}

}  // namespace

int zgemm_dev(hipStream_t s, Workspace* ws, bool ta, bool tb, size_t m, size_t n, size_t k, zd alpha, const zd* A,
              size_t lda, const zd* B, size_t ldb, zd beta, zd* C, size_t ldc) {
    if (m == 0 || n == 0) return AL_OK;
    if (2 * m > INT32_MAX || 2 * k > INT32_MAX || n > INT32_MAX) {
        set_error("zgemm: dimensions beyond 2^30 are not supported");
        return AL_ERR_ARG;
    }
    // beta*C for every entry first, then C += R(alpha op(A)) V(op(B)) on the FP64 matrix cores
    hipLaunchKernelGGL(zscale_k, dim3(ceil_div(m * n, 256)), dim3(256), 0, s, (int)m, (int)n, beta, C, ldc);
    AL_HIP(hipGetLastError());
    if (k == 0) return AL_OK;
    const size_t rbytes = round256(4 * m * k * sizeof(double));
    const size_t bbytes = tb ? k * n * sizeof(zd) : 0;
    void* buf;
    AL_CHECK(zbuf(ws, rbytes + bbytes, &buf));
    double* R = (double*)buf;
    const bool one = alpha.re == 1.0 && alpha.im == 0.0;
    AL_CHECK(expand(s, ta, m, k, A, lda, false, one ? 0 : 2, alpha, PART_FULL, R, 2 * m));
    const zd* Bop = B;
    size_t ldbop = ldb;
    if (tb) {
        zd* Bt = (zd*)((char*)buf + rbytes);
        AL_CHECK(ztrans(s, n, k, B, ldb, false, Bt, k));  // op(B) = B^T, k x n
        Bop = Bt;
        ldbop = k;
    }
    return gemm_f64(s, false, false, 2 * m, n, 2 * k, 1.0, R, 2 * m, (const double*)Bop, 2 * ldbop, 1.0, (double*)C,
                    2 * ldc);
}

int zgetrf_dev(hipStream_t s, Workspace* ws, size_t M, size_t N, zd* A, size_t lda, int* ipiv, int* info) {
    if (M > (size_t)32768 || N > INT32_MAX) {
        set_error("zgetrf: more rows than the cooperative panel holds (32768)");
        return AL_ERR_ARG;
    }
    int units = 0;
    AL_CHECK(coop_reserve(s, M, &units));
    const int st = zgetrf_enqueue(s, ws, M, N, A, lda, ipiv, info);
    if (st != AL_OK) {
        (void)hipStreamSynchronize(s);
        coop_release(units);
        return st;
    }
    return coop_hold(s, units);
}

int zgetrs_dev(hipStream_t s, Workspace* ws, size_t n, size_t nrhs, const zd* LU, size_t lda, const int* ipiv,
               zd* B, size_t ldb, const int* info) {
    if (n == 0 || nrhs == 0) return AL_OK;
    if (2 * n > INT32_MAX) {
        set_error("zgetrs: order beyond 2^30");
        return AL_ERR_ARG;
    }
    AL_CHECK(ws_reserve(ws, n, 0));
    AL_CHECK(fill_u32(s, ws->abort_flag, 1, 0u));
    const size_t ebytes = round256(4 * n * n * sizeof(double));
    void* buf;
    AL_CHECK(zbuf(ws, ebytes + n * nrhs * sizeof(zd), &buf));
    double* E = (double*)buf;
    zd* W = (zd*)((char*)buf + ebytes);
    AL_CHECK(perm_from_ipiv(s, ipiv, n, ws->perm));
    const size_t tot = n * nrhs;
    hipLaunchKernelGGL(zgather_k, dim3(ceil_div(tot, 256)), dim3(256), 0, s, (int)n, (int)nrhs, B, ldb, ws->perm, W);
    AL_HIP(hipGetLastError());
    // E: R(L) strictly below, identity 2x2 blocks on the diagonal, R(diag(U)^-1 U) above
    AL_CHECK(expand(s, false, n, n, LU, lda, false, 0, zd(1.0), PART_LU_SPLIT, E, 2 * n));
    const size_t n2 = 2 * n;
    if (trsv_fits(nrhs)) {
        AL_CHECK(trsv_solve<double>(s, ws, true, true, n2, nrhs, E, n2, (double*)W, n2));
    } else {
        AL_CHECK(trsm_f64(s, true, true, false, true, n2, nrhs, 1.0, E, n2, (double*)W, n2, info));
    }
    hipLaunchKernelGGL(zdiv_diag_k, dim3(ceil_div(tot, 256)), dim3(256), 0, s, (int)n, (int)nrhs, LU, lda, W);
    AL_HIP(hipGetLastError());
    if (trsv_fits(nrhs)) {
        AL_CHECK(trsv_solve<double>(s, ws, false, true, n2, nrhs, E, n2, (double*)W, n2));
    } else {
        AL_CHECK(trsm_f64(s, true, false, false, true, n2, nrhs, 1.0, E, n2, (double*)W, n2, info));
    }
    hipLaunchKernelGGL(zcopy_out_k, dim3(ceil_div(tot, 256)), dim3(256), 0, s, (int)n, (int)nrhs, W, B, ldb);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

int zpotrf_dev(hipStream_t s, Workspace* ws, size_t n, zd* A, size_t lda, bool upper, int* info) {
    AL_CHECK(fill_i32(s, info, 1, 0));
    if (n == 0) return AL_OK;
    if (2 * n > INT32_MAX) {
        set_error("zpotrf: order beyond 2^30");
        return AL_ERR_ARG;
    }
    const size_t nb = n > (size_t)ZPB ? (size_t)ZPBO : (size_t)ZPB;
    const size_t r11 = round256((2 * nb) * (2 * nb) * sizeof(double));
    const size_t wb = round256(nb * n * sizeof(zd));
    const size_t r21 = round256(2 * n * 2 * nb * sizeof(double));
    void* buf;
    AL_CHECK(zbuf(ws, r11 + wb + r21, &buf));
    double* R11 = (double*)buf;
    zd* W = (zd*)((char*)buf + r11);
    double* R21 = (double*)((char*)buf + r11 + wb);
    return zpotrf_level(s, n, A, lda, info, 0, nb, upper, R11, W, R21);
}

}  // namespace al
