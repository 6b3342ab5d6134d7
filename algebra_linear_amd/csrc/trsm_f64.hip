// Triangular solves and row interchanges.
//
// blas::dtrsm (src/blas.rs:1384-1438) solves one vector at a time with a scalar
// substitution.  Here a triangular solve is blocked in 64-wide diagonal blocks: each
// diagonal block is solved by substitution (`trsm_fwd_wave` for lower triangles, one wave per
// four right-hand-side vectors; `trsm_diag` for upper ones, one thread per vector), and the
// off-diagonal coupling is an FP64 MFMA GEMM, so almost all flops of a large solve run on the
// matrix cores.
#include <algorithm>

#include "al_internal.hpp"

namespace al {
namespace {

constexpr int DB = 64;   // diagonal block order

// Solves M x = x in place for nvec vectors, M the nb x nb (nb <= 64) triangle whose element
// (i, k) is T[i*t_rs + k*t_cs]; vector v's element i is X[i*x_es + v*x_vs].  One thread per
// vector, the vector in registers, the triangle staged in LDS (padded to 64 with the
// identity).  Each x_i is the crate's row sweep, its terms subtracted in the crate's order
// (src/blas.rs:1410-1438).  Only the triangle (and the diagonal unless `unit`) is read.
__global__ __launch_bounds__(256) void trsm_diag(int nb, int nvec, const double* __restrict__ T, size_t t_rs,
                                                 size_t t_cs, double* __restrict__ X, size_t x_es, size_t x_vs,
                                                 int forward, int unit, const int* __restrict__ info) {
    if (info && *info != 0) return;
    __shared__ double Ts[DB][DB + 1];
    const int t = threadIdx.x;
    for (int idx = t; idx < DB * DB; idx += 256) {
        const int i = idx % DB, k = idx / DB;
        double v = 0.0;
        const bool in_tri = forward ? (k < i) : (k > i);
        if (i < nb && k < nb) {
            if (in_tri) v = T[(size_t)i * t_rs + (size_t)k * t_cs];
            else if (i == k) v = unit ? 1.0 : T[(size_t)i * t_rs + (size_t)i * t_cs];
        } else if (i == k) {
            v = 1.0;
        }
        Ts[i][k] = v;
    }
    __syncthreads();
    const int vec = blockIdx.x * 256 + t;
    if (vec >= nvec) return;
    double x[DB];
    double* xp = X + (size_t)vec * x_vs;
#pragma unroll
    for (int i = 0; i < DB; ++i) x[i] = (i < nb) ? xp[(size_t)i * x_es] : 0.0;
    if (forward) {
#pragma unroll
        for (int i = 0; i < DB; ++i) {
            double sacc = x[i];
#pragma unroll
            for (int k = 0; k < i; ++k) sacc -= Ts[i][k] * x[k];
            x[i] = unit ? sacc : sacc / Ts[i][i];
        }
    } else {
#pragma unroll
        for (int i = DB - 1; i >= 0; --i) {
            double sacc = x[i];
#pragma unroll
            for (int k = i + 1; k < DB; ++k) sacc -= Ts[i][k] * x[k];
            x[i] = unit ? sacc : sacc / Ts[i][i];
        }
    }
#pragma unroll
    for (int i = 0; i < DB; ++i)
        if (i < nb) xp[(size_t)i * x_es] = x[i];
}

__device__ __forceinline__ double lane_d(double v, int l) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | lo);
}

// The forward case of trsm_diag (M lower triangular) with the same arithmetic, organized for
// latency: one wave solves four vectors at a time, lane i holding element i of each.  Lane i
// keeps row i of M left of the diagonal in registers (loaded once per wave; nothing on or above
// the diagonal is read, except the diagonal itself when not unit); step j finishes x_j in lane j
// (the division by M[j][j] of the fully accumulated value, as the row sweep does) and hands it
// to every lane with v_readlane, and the lanes below subtract M[i][j] x_j.  Each x_i thus sees
// exactly the row sweep's FMAs, in the same order (j ascending), with no LDS on the chain.
// Waves stride over the groups of four vectors.
template <bool UNIT>
__global__ __launch_bounds__(256) void trsm_fwd_wave(int nb, int nvec, const double* __restrict__ T, size_t t_rs,
                                                     size_t t_cs, double* __restrict__ X, size_t x_es, size_t x_vs,
                                                     const int* __restrict__ info) {
    if (info && *info != 0) return;
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6), nwaves = gridDim.x * 4;
    const bool in = lane < nb;
    double trow[DB];
#pragma unroll
    for (int j = 0; j < DB; ++j) trow[j] = (in && j < lane) ? T[(size_t)lane * t_rs + (size_t)j * t_cs] : 0.0;
    const double dg = (!UNIT && in) ? T[(size_t)lane * (t_rs + t_cs)] : 1.0;
    for (int g = wave * 4; g < nvec; g += nwaves * 4) {
        double x[4];
        double* xp[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            xp[q] = X + (size_t)min(g + q, nvec - 1) * x_vs;  // a short last group repeats, unstored
            x[q] = in ? xp[q][(size_t)lane * x_es] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < DB; ++j) {
            const double dj = UNIT ? 1.0 : lane_d(dg, j);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                double xj = lane_d(x[q], j);
                if (!UNIT) {
                    xj = xj / dj;
                    x[q] = lane == j ? xj : x[q];
                }
                const double v = fma(-trow[j], xj, x[q]);
                x[q] = lane > j ? v : x[q];
            }
        }
        if (in) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (g + q < nvec) xp[q][(size_t)lane * x_es] = x[q];
        }
    }
}

__global__ void scale_block(int m, int n, double alpha, double* B, size_t ldb, const int* info) {
    if (info && *info != 0) return;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)m * n) return;
    const int i = (int)(idx % m), j = (int)(idx / m);
    B[(size_t)i + (size_t)j * ldb] *= alpha;
}

__global__ void fill_i32_k(int* p, size_t n, int v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void iota_i32_k(int* p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (int)i;
}

__global__ void fill_u32_k(unsigned* p, size_t n, unsigned v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Element (r, c) of op(T) for a sub-block of op(T) at rows r0.., cols c0..: pointer and
// transposition flag for the GEMM that consumes it.
inline const double* opT_block(const double* T, size_t ldt, bool trans, size_t r0, size_t c0) {
    return trans ? T + c0 + r0 * ldt : T + r0 + c0 * ldt;
}

int diag_solve(hipStream_t s, int nb, int nvec, const double* T, size_t t_rs, size_t t_cs, double* X, size_t x_es,
               size_t x_vs, bool forward, bool unit, const int* info) {
    if (nvec <= 0 || nb <= 0) return AL_OK;
    // one wave per four vectors, except for many vectors strided across columns (the rows of a
    // tall right-hand side): those go one per thread, which keeps the loads coalesced across
    // threads and still fills the chip
    if (forward && (x_es == 1 || nvec < 4096)) {
        const unsigned wgs = (unsigned)std::min<size_t>(ceil_div((size_t)nvec, (size_t)16), (size_t)512);
        if (unit)
            hipLaunchKernelGGL(trsm_fwd_wave<true>, dim3(wgs), dim3(256), 0, s, nb, nvec, T, t_rs, t_cs, X, x_es, x_vs,
                               info);
        else
            hipLaunchKernelGGL(trsm_fwd_wave<false>, dim3(wgs), dim3(256), 0, s, nb, nvec, T, t_rs, t_cs, X, x_es,
                               x_vs, info);
        AL_HIP(hipGetLastError());
        return AL_OK;
    }
    hipLaunchKernelGGL(trsm_diag, dim3(ceil_div((size_t)nvec, 256)), dim3(256), 0, s, nb, nvec, T, t_rs, t_cs, X,
                       x_es, x_vs, forward ? 1 : 0, unit ? 1 : 0, info);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

}  // namespace

int fill_i32(hipStream_t s, int* p, size_t n, int v) {
    if (n == 0) return AL_OK;
    hipLaunchKernelGGL(fill_i32_k, dim3(ceil_div(n, 256)), dim3(256), 0, s, p, n, v);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

int iota_i32(hipStream_t s, int* p, size_t n) {
    if (n == 0) return AL_OK;
    hipLaunchKernelGGL(iota_i32_k, dim3(ceil_div(n, 256)), dim3(256), 0, s, p, n);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

int fill_u32(hipStream_t s, unsigned* p, size_t n, unsigned v) {
    if (n == 0) return AL_OK;
    hipLaunchKernelGGL(fill_u32_k, dim3(ceil_div(n, 256)), dim3(256), 0, s, p, n, v);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

int trsm_f64(hipStream_t s, bool left, bool lower, bool trans, bool unit, size_t m, size_t n, double alpha,
             const double* T, size_t ldt, double* B, size_t ldb, const int* info) {
    if (m == 0 || n == 0) return AL_OK;
    if (alpha != 1.0) {
        // the crate scales the right-hand side first, then solves (src/blas.rs:1388-1403)
        hipLaunchKernelGGL(scale_block, dim3(ceil_div(m * n, 256)), dim3(256), 0, s, (int)m, (int)n, alpha, B, ldb,
                           info);
        AL_HIP(hipGetLastError());
    }
    // op(T) is lower triangular iff lower != trans
    const bool op_lower = (lower != trans);
    // element (i, k) of op(T)'s diagonal block at (d, d): op(T)[i][k] = trans ? T[k][i] : T[i][k]
    const size_t opt_rs = trans ? ldt : 1, opt_cs = trans ? 1 : ldt;
    if (left) {
        // op(T) X = B, T is m x m; vectors are the columns of B
        const size_t nblk = (m + DB - 1) / DB;
        for (size_t bi = 0; bi < nblk; ++bi) {
            const size_t b = op_lower ? bi : (nblk - 1 - bi);
            const size_t d = b * DB, nb = (m - d) < (size_t)DB ? (m - d) : (size_t)DB;
            const double* Tdd = T + d + d * ldt;
            AL_CHECK(diag_solve(s, (int)nb, (int)n, Tdd, opt_rs, opt_cs, B + d, 1, ldb, op_lower, unit, info));
            if (op_lower && d + nb < m) {
                // B[d+nb:, :] -= op(T)[d+nb:, d:d+nb] * X_b
                AL_CHECK(gemm_f64(s, trans, false, m - d - nb, n, nb, -1.0, opT_block(T, ldt, trans, d + nb, d), ldt,
                                  B + d, ldb, 1.0, B + d + nb, ldb));
            } else if (!op_lower && d > 0) {
                // B[:d, :] -= op(T)[:d, d:d+nb] * X_b
                AL_CHECK(gemm_f64(s, trans, false, d, n, nb, -1.0, opT_block(T, ldt, trans, 0, d), ldt, B + d, ldb,
                                  1.0, B, ldb));
            }
        }
    } else {
        // X op(T) = B, T is n x n; vectors are the rows of B.  Row x^T of a column block solves
        // op(T)_bb^T x = b, whose matrix element (i, k) is op(T)[k][i].
        const size_t mt_rs = opt_cs, mt_cs = opt_rs;
        const size_t nblk = (n + DB - 1) / DB;
        for (size_t bi = 0; bi < nblk; ++bi) {
            // op(T) upper -> forward over column blocks; lower -> backward
            const size_t b = op_lower ? (nblk - 1 - bi) : bi;
            const size_t d = b * DB, nb = (n - d) < (size_t)DB ? (n - d) : (size_t)DB;
            const double* Tdd = T + d + d * ldt;
            AL_CHECK(diag_solve(s, (int)nb, (int)m, Tdd, mt_rs, mt_cs, B + d * ldb, ldb, 1, !op_lower, unit, info));
            if (!op_lower && d + nb < n) {
                // B[:, d+nb:] -= X_b * op(T)[d:d+nb, d+nb:]
                AL_CHECK(gemm_f64(s, false, trans, m, n - d - nb, nb, -1.0, B + d * ldb, ldb,
                                  opT_block(T, ldt, trans, d, d + nb), ldt, 1.0, B + (d + nb) * ldb, ldb));
            } else if (op_lower && d > 0) {
                // B[:, :d] -= X_b * op(T)[d:d+nb, :d]
                AL_CHECK(gemm_f64(s, false, trans, m, d, nb, -1.0, B + d * ldb, ldb, opT_block(T, ldt, trans, d, 0),
                                  ldt, 1.0, B, ldb));
            }
        }
    }
    return AL_OK;
}

}  // namespace al
