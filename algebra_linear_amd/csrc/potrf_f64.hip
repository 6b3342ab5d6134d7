// Lower Cholesky, lapack::dpotrf (src/lapack/mod.rs:368-399).
//
// The crate runs a left-looking dot-product loop over the whole matrix.  Here it is a
// right-looking blocked factorization with 256-wide outer blocks:
//   * an order <= 64 matrix is one block factored by one wave with its rows in registers
//     (`potf2_wave`, the crate's own dot-product order per element, unfused, so it factors
//     bit-for-bit as the crate does; `potf2_lds` is the same arithmetic with the block in LDS,
//     kept as the reference form);
//   * inside a larger factorization the 64 x 64 diagonal blocks after the first use
//     `potf2_rl`, the same one-wave register layout with a right-looking (outer-product) order,
//     whose chain per column is one square root, one division and independent fused updates
//     (the first block, untouched by any update, keeps the crate's order);
//   * a 256-wide outer block's diagonal part is factored as 64-wide blocks with their small
//     solves and products; the panel below it, A21 := A21 L11^{-T}, is one FP64 MFMA product with
//     the explicit inverse of L11 (`trtri_lower_k`: one wave per column of the inverse, L11's
//     columns streamed through LDS 16 at a time);
//   * the trailing matrix gets a lower-triangle-only FP64 MFMA rank-256 update.
// Look-ahead on two streams: the side stream updates the next outer block's columns with the
// current panel and factors that block (diagonal part and panel) while the caller's stream
// applies the trailing update to every column right of it.  Only the lower triangle is read or
// written anywhere.
#include "al_internal.hpp"

namespace al {
namespace {

constexpr int PB = 64;
constexpr int PBO = 256;

#pragma clang fp contract(off)
__device__ __forceinline__ double lane_d(double v, int l) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | lo);
}

// The 64 x 64 diagonal block in one wave, lane i holding row i of the lower triangle in
// registers.  Column j: every lane runs s = a[i][j] - sum_k L[i][k] L[j][k] over k < j in
// order, unfused, L[j][k] taken from lane j by v_readlane -- in lane j that is the diagonal's
// own sum -- then lane j takes the square root and the lanes below divide by it.  Each element
// sees exactly the operations of the crate's dot-product order (as potf2_lds below), with no
// LDS or barrier on the chain.  On a non-positive pivot the block is left as potf2_lds leaves
// it: columns before j factored, column j and those after it as given.
__global__ __launch_bounds__(64) void potf2_wave(double* A, size_t lda, int nb, int col0, int* info) {
    if (*info != 0) return;
    const int lane = threadIdx.x;
    const bool in = lane < nb;
    double x[PB];
#pragma unroll
    for (int k = 0; k < PB; ++k) x[k] = (in && k <= lane && k < nb) ? A[(size_t)lane + (size_t)k * lda] : 0.0;
    int done = nb;     // columns factored
    bool alive = true;  // no non-positive pivot yet (uniform)
#pragma unroll
    for (int j = 0; j < PB; ++j) {
        const bool col = alive && j < nb;
        double sacc = x[j];
#pragma unroll
        for (int k = 0; k < j; ++k) sacc -= x[k] * lane_d(x[k], j);
        const double sum = lane_d(sacc, j);  // the diagonal's sum, from lane j
        if (col && sum <= 0.0) {
            if (lane == 0) *info = col0 + j + 1;
            done = j;
            alive = false;
        } else if (col) {
            const double ljj = sqrt(sum);
            x[j] = lane == j ? ljj : (lane > j ? sacc / ljj : x[j]);
        }
    }
    if (in) {
#pragma unroll
        for (int k = 0; k < PB; ++k)
            if (k <= lane && k < done) A[(size_t)lane + (size_t)k * lda] = x[k];
    }
}

__global__ __launch_bounds__(256) void potf2_lds(double* A, size_t lda, int nb, int col0, int* info) {
    if (*info != 0) return;
    __shared__ double L[PB][PB + 1];
    __shared__ int bad;
    const int t = threadIdx.x;
    for (int idx = t; idx < nb * nb; idx += 256) {
        const int i = idx % nb, j = idx / nb;
        if (i >= j) L[i][j] = A[(size_t)i + (size_t)j * lda];
    }
    if (t == 0) bad = 0;
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
        if (t == 0) {
            double sum = L[j][j];
            for (int k = 0; k < j; ++k) {
                const double ljk = L[j][k];
                sum -= ljk * ljk;
            }
            if (sum <= 0.0) {
                bad = 1;
                *info = col0 + j + 1;
            } else {
                L[j][j] = sqrt(sum);
            }
        }
        __syncthreads();
        if (bad) break;
        const double ljj = L[j][j];
        for (int i = j + 1 + t; i < nb; i += 256) {
            double s = L[i][j];
            for (int k = 0; k < j; ++k) s -= L[i][k] * L[j][k];
            L[i][j] = s / ljj;
        }
        __syncthreads();
    }
    // write back the lower triangle (partially factored on failure, like the crate)
    for (int idx = t; idx < nb * nb; idx += 256) {
        const int i = idx % nb, j = idx / nb;
        if (i >= j) A[(size_t)i + (size_t)j * lda] = L[i][j];
    }
}
// A 64 x 64 diagonal block inside a larger factorization, right-looking in one wave: lane i
// holds row i of the lower triangle in registers; at column j the diagonal comes from lane j
// by v_readlane, lane j takes its square root and the lanes below divide by it, then each later
// column k takes L[k][j] from lane k and every lane updates its entry (i, k) with one fused
// multiply-add (entries above the diagonal are never written back).  On a non-positive pivot
// the columns before j are written back factored and the rest left as given.
__global__ __launch_bounds__(64) void potf2_rl(double* A, size_t lda, int nb, int col0, int* info) {
    if (*info != 0) return;
    const int lane = threadIdx.x;
    const bool in = lane < nb;
    double x[PB];
#pragma unroll
    for (int k = 0; k < PB; ++k) x[k] = (in && k <= lane && k < nb) ? A[(size_t)lane + (size_t)k * lda] : 0.0;
    int done = nb;
#pragma unroll
    for (int j = 0; j < PB; ++j) {
        if (j < done) {  // uniform
            const double djj = lane_d(x[j], j);
            if (djj <= 0.0) {
                if (lane == 0) *info = col0 + j + 1;
                done = j;
            } else {
                const double ljj = sqrt(djj);
                x[j] = lane == j ? ljj : (lane > j ? x[j] / ljj : x[j]);
#pragma unroll
                for (int k = j + 1; k < PB; ++k) x[k] = fma(-x[j], lane_d(x[j], k), x[k]);
            }
        }
    }
    if (in) {
#pragma unroll
        for (int k = 0; k < PB; ++k)
            if (k <= lane && k < done) A[(size_t)lane + (size_t)k * lda] = x[k];
    }
}

// W := L^{-1} for the jb x jb (jb <= 256) lower triangle L with a non-zero diagonal; W is
// written in full (zeros above its diagonal).  Wave q of workgroup b computes column
// j = 4b + q of W by forward substitution, L w = e_j: lane l holds w_i for i = l + 64r
// (r < 4); at step k, w_k = (delta_jk - acc_k) / L[k][k] comes from its lane by v_readlane and
// every lane adds L[i][k] w_k to acc_i for its rows i > k.  The workgroup's four columns share
// one pass over k (a column's steps before j contribute nothing), with L's columns staged in
// LDS 16 at a time behind one barrier.
constexpr int TRTRI_MAX = 256;
constexpr int TRTRI_CH = 16;
__global__ __launch_bounds__(256) void trtri_lower_k(int jb, const double* __restrict__ L, size_t ldl,
                                                     double* __restrict__ W, size_t ldw, const int* __restrict__ info) {
    if (*info != 0) return;
    __shared__ double Ls[TRTRI_CH][TRTRI_MAX];
    __shared__ double rdiag[TRTRI_MAX];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int j0 = blockIdx.x * 4;
    const int j = j0 + wave;
    for (int i = t; i < jb; i += 256) rdiag[i] = 1.0 / L[(size_t)i * (ldl + 1)];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    double w[4] = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = j0; k0 < jb; k0 += TRTRI_CH) {
        __syncthreads();  // the previous chunk is no longer read (and rdiag is ready)
        for (int idx = t; idx < TRTRI_CH * TRTRI_MAX; idx += 256) {
            const int c = idx / TRTRI_MAX, i = idx % TRTRI_MAX;
            const int k = k0 + c;
            Ls[c][i] = (k < jb && i > k && i < jb) ? L[(size_t)i + (size_t)k * ldl] : 0.0;
        }
        __syncthreads();
        const int kend = (jb - k0) < TRTRI_CH ? (jb - k0) : TRTRI_CH;
        for (int c = 0; c < kend; ++c) {
            const int k = k0 + c;
            const int rk = k >> 6, lk = k & 63;
            double a = 0.0;
#pragma unroll
            for (int r = 0; r < 4; ++r)
                if (r == rk) a = acc[r];
            const double cand = ((k == j) ? 1.0 : 0.0) - a;
            const double wk = (k < j) ? 0.0 : lane_d(cand, lk) * rdiag[k];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                if (r == rk && lane == lk) w[r] = wk;
                acc[r] = fma(Ls[c][lane + 64 * r], wk, acc[r]);
            }
        }
    }
    if (j < jb) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = lane + 64 * r;
            if (i < jb) W[(size_t)i + (size_t)j * ldw] = w[r];
        }
    }
}

#pragma clang fp contract(on)

// Blocked factorization of the n x n block at A (columns col0.. of the whole matrix), blocks of
// nb; 64-wide blocks are the leaves: potf2_rl, or the crate's order (potf2_wave) when bit_exact
// and for the whole matrix's first block, which no update has touched (so it matches the
// crate's factor of the leading 64 x 64 block bit for bit).
int potrf_level(hipStream_t s, size_t n, double* A, size_t lda, int* info, size_t col0, size_t nb, bool bit_exact) {
    for (size_t k = 0; k < n; k += nb) {
        const size_t jb = (n - k) < nb ? (n - k) : nb;
        double* Akk = A + k + k * lda;
        if (jb <= (size_t)PB) {
            if (bit_exact || col0 + k == 0)
                hipLaunchKernelGGL(potf2_wave, dim3(1), dim3(64), 0, s, Akk, lda, (int)jb, (int)(col0 + k), info);
            else
                hipLaunchKernelGGL(potf2_rl, dim3(1), dim3(64), 0, s, Akk, lda, (int)jb, (int)(col0 + k), info);
            AL_HIP(hipGetLastError());
        } else {
            AL_CHECK(potrf_level(s, jb, Akk, lda, info, col0 + k, PB, bit_exact));
        }
        if (k + jb < n) {
            double* A21 = Akk + jb;
            const size_t mr = n - k - jb;
            // A21 := A21 * L11^{-T}
            AL_CHECK(trsm_f64(s, false, true, true, false, mr, jb, 1.0, Akk, lda, A21, lda, info));
            // A22 := A22 - A21 * A21^T, lower triangle only
            AL_CHECK(gemm_f64(s, false, true, mr, mr, jb, -1.0, A21, lda, A21, lda, 1.0, A21 + jb * lda, lda, 1));
        }
    }
    return AL_OK;
}

__global__ void copy_cols_k(int m, int n, const double* __restrict__ S, size_t lds, double* __restrict__ D,
                            size_t ldd, const int* __restrict__ info) {
    if (*info != 0) return;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)m * n) return;
    const int i = (int)(idx % m), c = (int)(idx / m);
    D[(size_t)i + (size_t)c * ldd] = S[(size_t)i + (size_t)c * lds];
}

// One outer block [k, k + jb) of the whole n x n factorization: its diagonal part, then the
// panel below it as A21 := T L11^{-T} with T a copy of A21 and L11^{-1} formed explicitly.
int potrf_block(hipStream_t s, size_t n, double* A, size_t lda, int* info, size_t k, size_t jb, double* W,
                double* T) {
    double* Akk = A + k + k * lda;
    AL_CHECK(potrf_level(s, jb, Akk, lda, info, k, PB, false));
    const size_t mr = n - k - jb;
    if (mr == 0) return AL_OK;
    hipLaunchKernelGGL(trtri_lower_k, dim3((unsigned)ceil_div(jb, (size_t)4)), dim3(256), 0, s, (int)jb, Akk, lda, W,
                       (size_t)PBO, info);
    AL_HIP(hipGetLastError());
    double* A21 = Akk + jb;
    hipLaunchKernelGGL(copy_cols_k, dim3((unsigned)ceil_div(mr * jb, (size_t)256)), dim3(256), 0, s, (int)mr, (int)jb,
                       A21, lda, T, mr, info);
    AL_HIP(hipGetLastError());
    return gemm_f64(s, false, true, mr, jb, jb, 1.0, T, mr, W, PBO, 0.0, A21, lda);
}

}  // namespace

int potrf_f64(hipStream_t s, Workspace* ws, size_t n, double* A, size_t lda, int* info) {
    AL_CHECK(fill_i32(s, info, 1, 0));
    if (n == 0) return AL_OK;
    if (n <= (size_t)PB) return potrf_level(s, n, A, lda, info, 0, PB, true);
    if (n <= (size_t)PBO) return potrf_level(s, n, A, lda, info, 0, PB, false);
    // the inverse of a diagonal block (PBO x PBO) and the copy of a panel (n x PBO)
    const size_t wbytes = (size_t)PBO * PBO * sizeof(double);
    const size_t need = wbytes + n * (size_t)PBO * sizeof(double);
    if (ws->cbuf_bytes < need) {
        if (ws->cbuf) AL_HIP(hipFree(ws->cbuf));
        ws->cbuf = nullptr;
        ws->cbuf_bytes = 0;
        if (hipMalloc(&ws->cbuf, need) != hipSuccess) {
            ws->cbuf = nullptr;
            set_error("device allocation failed");
            return AL_ERR_OOM;
        }
        ws->cbuf_bytes = need;
    }
    double* W = (double*)ws->cbuf;
    double* T = (double*)((char*)ws->cbuf + wbytes);
    AL_CHECK(ensure_side_stream(ws));
    hipStream_t side = ws->side;
    AL_HIP(hipEventRecord(ws->ev_main, s));
    AL_HIP(hipStreamWaitEvent(side, ws->ev_main, 0));
    AL_CHECK(potrf_block(side, n, A, lda, info, 0, PBO, W, T));
    for (size_t k = 0; k < n; k += PBO) {
        const size_t ke = k + PBO;  // every block but the last is full
        if (ke >= n) break;
        const size_t ke2 = (n - ke) < (size_t)PBO ? n : ke + PBO;
        double* L21 = A + ke + k * lda;  // the panel of block k, rows ke..n
        // side: the next block's columns, once the main stream's last update of them is done
        AL_HIP(hipStreamWaitEvent(side, ws->ev_main, 0));
        AL_CHECK(gemm_f64(side, false, true, n - ke, ke2 - ke, PBO, -1.0, L21, lda, L21, lda, 1.0, A + ke + ke * lda,
                          lda, 1));
        AL_HIP(hipEventRecord(ws->ev_side, side));
        AL_CHECK(potrf_block(side, n, A, lda, info, ke, ke2 - ke, W, T));
        // main: the trailing matrix right of the next block
        AL_HIP(hipStreamWaitEvent(s, ws->ev_side, 0));
        if (ke2 < n) {
            const double* Lr = A + ke2 + k * lda;
            AL_CHECK(gemm_f64(s, false, true, n - ke2, n - ke2, PBO, -1.0, Lr, lda, Lr, lda, 1.0, A + ke2 + ke2 * lda,
                              lda, 1));
        }
        AL_HIP(hipEventRecord(ws->ev_main, s));
    }
    // the caller's stream ends after the side stream's last block
    AL_HIP(hipEventRecord(ws->ev_side, side));
    AL_HIP(hipStreamWaitEvent(s, ws->ev_side, 0));
    return AL_OK;
}

}  // namespace al
