// Lower Cholesky, lapack::dpotrf (src/lapack/mod.rs:368-399).
//
// The crate runs a left-looking dot-product loop over the whole matrix.  Here it is a
// right-looking blocked factorization: a 64x64 diagonal block is factored by one wave with its
// rows in registers (`potf2_wave`, the crate's own dot-product order per element, unfused, so a
// matrix of order <= 64 factors bit-for-bit as the crate factors it; `potf2_lds` is the same
// arithmetic with the block in LDS, kept as the reference form), the panel below it is a
// triangular solve against the block's transpose, and the trailing matrix gets a
// lower-triangle-only FP64 MFMA rank-k update.  Two levels (64 inside 256) keep the trailing
// update at depth 256.  Only the lower triangle is read or written anywhere.
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
#pragma clang fp contract(on)

int potrf_level(hipStream_t s, size_t n, double* A, size_t lda, int* info, size_t col0, size_t nb) {
    for (size_t k = 0; k < n; k += nb) {
        const size_t jb = (n - k) < nb ? (n - k) : nb;
        double* Akk = A + k + k * lda;
        if (jb <= (size_t)PB) {
            hipLaunchKernelGGL(potf2_wave, dim3(1), dim3(64), 0, s, Akk, lda, (int)jb, (int)(col0 + k), info);
            AL_HIP(hipGetLastError());
        } else {
            AL_CHECK(potrf_level(s, jb, Akk, lda, info, col0 + k, PB));
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

}  // namespace

int potrf_f64(hipStream_t s, Workspace* ws, size_t n, double* A, size_t lda, int* info) {
    (void)ws;
    AL_CHECK(fill_i32(s, info, 1, 0));
    if (n == 0) return AL_OK;
    return potrf_level(s, n, A, lda, info, 0, n > (size_t)PB ? (size_t)PBO : (size_t)PB);
}

}  // namespace al
