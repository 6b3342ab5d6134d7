// Batched solves of many small independent systems (orders 1..512) in one persistent launch:
// dgesv (LU with partial pivoting) or dposv (lower Cholesky) per system, in plain FP64 or as
// an FP32 factorization refined in FP64 (with an FP64 refactorization if refinement fails).
//
// The crate solves one system per call (lapack::dgesv, src/lapack/mod.rs:309-323; dpotrf,
// :368-399); launching the GPU path once per small system leaves the device idle between
// launches.  Here one 256-thread workgroup owns one system at a time and takes the next from an
// atomic queue counter when it is done, so a batch of uneven sizes is one launch and every
// workgroup exits when the queue is empty.  Systems are taken in the order given: ordering
// them by decreasing size balances the tail.
//
// Per system, in the workgroup:
//   * blocked right-looking factorization with 16-column panels: the panel (rows k..n) is held
//     in LDS for its elimination (the pivot search is a DPP wave maximum plus one LDS exchange
//     between the four waves), the row interchanges and the U12 solve are done by one thread
//     per column, and the trailing update is a register-tiled (4 x 4 per thread) product of
//     the panel and 64-column stripes of U12 staged in LDS;
//   * the triangular solves run 64 rows at a time: the diagonal block by one wave with the
//     solved value broadcast from lane to lane, the rows below by every thread;
//   * mixed mode: A is copied to an FP32 work matrix and factored there; x = solve(b), then up
//     to max_iter steps of r = b - A x (FP64, A itself), x += solve(r) until the normwise
//     backward error of every column is at most 4 eps; a system that does not get there is
//     factored again in FP64 in place (A is then overwritten, as in the FP64 mode) and its
//     remaining columns solved, and its step count says so.
#include <cfloat>

#include "al_internal.hpp"

namespace al {
namespace {

constexpr int NT = 256;
constexpr int NMAX = 512;     // largest order handled here
constexpr int PNB = 16;       // panel width
constexpr int RHS = 8;        // right-hand sides per pass
constexpr int TILE = 64;      // trailing-update stripe

struct BatchLds {
    double panel[NMAX * PNB];     // panel, column-major with leading dimension m = n - k (T = double)
    double stripe[PNB * TILE];    // 16 x 64 stripe of U12 (or L21^T)
    double y[NMAX * RHS];         // right-hand sides / solution
    int ipiv[NMAX];
    int perm[NMAX];
    float wmag[4];
    int widx[4];
    int job;
    double norm[RHS * 3];
    double anorm;
};

// ------------------------------------------------------------------------------------------
// workgroup-wide helpers
// ------------------------------------------------------------------------------------------

__device__ __forceinline__ float dpp_max(float v) {
    const int id = __float_as_int(-2.f);
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(id, __float_as_int(v), 0x111, 0xf, 0xf, false)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(id, __float_as_int(v), 0x112, 0xf, 0xf, false)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(id, __float_as_int(v), 0x114, 0xf, 0xf, false)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(id, __float_as_int(v), 0x118, 0xf, 0xf, false)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(id, __float_as_int(v), 0x142, 0xa, 0xf, false)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(id, __float_as_int(v), 0x143, 0xc, 0xf, false)));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

__device__ __forceinline__ int wave_min_int(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const int o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    return v;
}

__device__ __forceinline__ double bcast(double v, int src) {
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(bits & 0xffffffffLL), src);
    const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// First maximum of |P[i]| over i in [lo, m) (a NaN never wins, as in the crate's scan; with
// no finite candidate the first row is chosen).  Magnitudes are compared in FP32 first and the
// exact FP64 comparison decides between the rows that tie in FP32.
template <typename T>
__device__ int block_argmax(const T* P, int lo, int m, BatchLds& L) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    double best = -1.0;
    int bi = 0x7fffffff;
    for (int i = lo + t; i < m; i += NT) {
        const double v = fabs((double)P[i]);
        if (v > best) {  // first maximum among this thread's rows (NaN never compares greater)
            best = v;
            bi = i;
        }
    }
    // candidates above the FP32 rounding of the maximum, then the exact first maximum among them
    const float fb = (float)best;
    const float wm = dpp_max(fb);
    const float gm_w = wm;
    if (lane == 0) L.wmag[wave] = gm_w;
    __syncthreads();
    float gm = L.wmag[0];
#pragma unroll
    for (int w = 1; w < 4; ++w) gm = fmaxf(gm, L.wmag[w]);
    __syncthreads();
    // every row whose FP32 magnitude reaches gm is a candidate; the exact maximum among the
    // candidates, ties to the lowest index, wins
    double cbest = (fb == gm && bi != 0x7fffffff) ? best : -1.0;
    // exact max among candidates: compare in FP64 through the wave, then across waves
    double wmax = cbest;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(wmax, off, 64);
        wmax = o > wmax ? o : wmax;
    }
    int cand = (cbest == wmax && cbest >= 0.0) ? bi : 0x7fffffff;
    cand = wave_min_int(cand);
    if (lane == 0) {
        L.widx[wave] = cand;
        L.norm[wave] = wmax;
    }
    __syncthreads();
    double gbest = L.norm[0];
    int gi = L.widx[0];
#pragma unroll
    for (int w = 1; w < 4; ++w) {
        const double bw = L.norm[w];
        const int iw = L.widx[w];
        if (bw > gbest || (bw == gbest && iw < gi)) {
            gbest = bw;
            gi = iw;
        }
    }
    __syncthreads();
    return gi == 0x7fffffff ? lo : gi;
}

// ------------------------------------------------------------------------------------------
// factorizations of the n x n matrix M (leading dimension n) by one workgroup
// ------------------------------------------------------------------------------------------

// Trailing update C -= Lp * U over rows [r0, n) and columns [c0, n) (lower: only c <= r),
// Lp the panel in LDS (rows from k0, leading dimension m, kb columns), U kb x (n - c0) read
// from M rows k0..k0+kb (or, for Cholesky, the transpose of the panel).
template <typename T, bool CHOL>
__device__ void trailing_update(T* M, int n, int k0, int kb, const T* Pl, int m, BatchLds& L) {
    const int t = threadIdx.x;
    const int r0 = k0 + kb;
    const int tr = t & 15, tc = t >> 4;
    T* stripe = (T*)L.stripe;
    for (int cc = r0; cc < n; cc += TILE) {
        const int cn = min(TILE, n - cc);
        __syncthreads();
        for (int idx = t; idx < kb * TILE; idx += NT) {
            const int k = idx % kb, c = idx / kb;
            T v = (T)0;
            if (c < cn) v = CHOL ? Pl[(cc + c - k0) + k * m] : M[(size_t)(k0 + k) + (size_t)(cc + c) * n];
            stripe[k + c * PNB] = v;
        }
        __syncthreads();
        for (int rc = CHOL ? cc : r0; rc < n; rc += TILE) {
            T acc[4][4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = (T)0;
            for (int k = 0; k < kb; ++k) {
                T a[4], b[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int r = rc + tr + 16 * i;
                    a[i] = r < n ? Pl[(r - k0) + k * m] : (T)0;
                }
#pragma unroll
                for (int j = 0; j < 4; ++j) b[j] = stripe[k + (tc + 16 * j) * PNB];
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int c = cc + tc + 16 * j;
                if (c >= n) continue;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int r = rc + tr + 16 * i;
                    if (r >= n || (CHOL && r < c)) continue;
                    T* p = M + (size_t)r + (size_t)c * n;
                    *p = *p - acc[i][j];
                }
            }
        }
    }
    __syncthreads();
}

// LU with partial pivoting; pivots (0-based) in L.ipiv; returns the first zero pivot's
// 1-based column or 0.
template <typename T>
__device__ int lu_factor(T* M, int n, BatchLds& L) {
    const int t = threadIdx.x;
    T* P = (T*)L.panel;
    for (int k0 = 0; k0 < n; k0 += PNB) {
        const int kb = min(PNB, n - k0);
        const int m = n - k0;
        for (int idx = t; idx < m * kb; idx += NT) {
            const int i = idx % m, c = idx / m;
            P[i + c * m] = M[(size_t)(k0 + i) + (size_t)(k0 + c) * n];
        }
        __syncthreads();
        for (int j = 0; j < kb; ++j) {
            const int p = block_argmax(P + j * m, j, m, L);
            if (t == 0) L.ipiv[k0 + j] = k0 + p;
            if (p != j && t < kb) {
                const T a = P[j + t * m];
                P[j + t * m] = P[p + t * m];
                P[p + t * m] = a;
            }
            __syncthreads();
            const T piv = P[j + j * m];
            if (piv == (T)0) {
                // the crate stops at an exactly zero pivot (src/lapack/mod.rs:134-139)
                __syncthreads();
                return k0 + j + 1;
            }
            for (int i = j + 1 + t; i < m; i += NT) {
                // the crate's own operations inside a panel: divide, then a separate multiply
                // and subtract (no FMA), so an exactly singular matrix meets its exact zero
#pragma clang fp contract(off)
                const T l = P[i + j * m] / piv;
                P[i + j * m] = l;
                for (int c = j + 1; c < kb; ++c) P[i + c * m] -= l * P[j + c * m];
            }
            __syncthreads();
        }
        for (int idx = t; idx < m * kb; idx += NT) {
            const int i = idx % m, c = idx / m;
            M[(size_t)(k0 + i) + (size_t)(k0 + c) * n] = P[i + c * m];
        }
        // the panel's interchanges on the other columns, one thread per column
        for (int c = t; c < n; c += NT) {
            if (c >= k0 && c < k0 + kb) continue;
            for (int s = 0; s < kb; ++s) {
                const int r2 = L.ipiv[k0 + s];
                if (r2 != k0 + s) {
                    T* pc = M + (size_t)c * n;
                    const T a = pc[k0 + s];
                    pc[k0 + s] = pc[r2];
                    pc[r2] = a;
                }
            }
        }
        __syncthreads();
        if (k0 + kb < n) {
            // U12 := L11^{-1} A12, one thread per column, in place (rolled loops: a register
            // copy of the 16-element column plus the hoisted L11 loads would spill)
            for (int c = k0 + kb + t; c < n; c += NT) {
                T* pc = M + (size_t)c * n + k0;
                for (int r = 1; r < kb; ++r) {
                    T v = pc[r];
                    for (int i = 0; i < r; ++i) v -= P[r + i * m] * pc[i];
                    pc[r] = v;
                }
            }
            __syncthreads();
            trailing_update<T, false>(M, n, k0, kb, P, m, L);
        }
    }
    return 0;
}

// Lower Cholesky (only the lower triangle is read or written); returns the first column whose
// diagonal is not positive (1-based) or 0, as lapack::dpotrf (src/lapack/mod.rs:368-399).
template <typename T>
__device__ int chol_factor(T* M, int n, BatchLds& L) {
    const int t = threadIdx.x;
    T* P = (T*)L.panel;
    for (int k0 = 0; k0 < n; k0 += PNB) {
        const int kb = min(PNB, n - k0);
        const int m = n - k0;
        for (int idx = t; idx < m * kb; idx += NT) {
            const int i = idx % m, c = idx / m;
            P[i + c * m] = i >= c ? M[(size_t)(k0 + i) + (size_t)(k0 + c) * n] : (T)0;
        }
        __syncthreads();
        for (int j = 0; j < kb; ++j) {
            const T d = P[j + j * m];
            if (!(d > (T)0)) {
                __syncthreads();
                return k0 + j + 1;
            }
            const T ljj = sqrt(d);
            __syncthreads();
            for (int i = j + t; i < m; i += NT) P[i + j * m] = i == j ? ljj : P[i + j * m] / ljj;
            __syncthreads();
            for (int i = j + 1 + t; i < m; i += NT) {
                const T lij = P[i + j * m];
                for (int c = j + 1; c < kb && c <= i; ++c) P[i + c * m] -= lij * P[c + j * m];
            }
            __syncthreads();
        }
        for (int idx = t; idx < m * kb; idx += NT) {
            const int i = idx % m, c = idx / m;
            if (i >= c) M[(size_t)(k0 + i) + (size_t)(k0 + c) * n] = P[i + c * m];
        }
        __syncthreads();
        if (k0 + kb < n) trailing_update<T, true>(M, n, k0, kb, P, m, L);
    }
    return 0;
}

// ------------------------------------------------------------------------------------------
// triangular solves on L.y (n x nr, leading dimension NMAX), FP64 vectors
// ------------------------------------------------------------------------------------------

// op = L (unit or not) forward, or U / L^T backward.  TRANS: the factor is read transposed
// (L^T x = y for Cholesky).
template <typename T, bool LOWER, bool UNIT, bool TRANS>
__device__ void tri_solve(const T* M, int n, int nr, BatchLds& L) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    double* y = L.y;
    auto el = [&](int i, int j) -> double {  // element (i, j) of op(factor)
        return TRANS ? (double)M[(size_t)j + (size_t)i * n] : (double)M[(size_t)i + (size_t)j * n];
    };
    const int nblk = (n + 63) / 64;
    for (int bb = 0; bb < nblk; ++bb) {
        const int b = LOWER ? bb : nblk - 1 - bb;
        const int r0 = b * 64, nb = min(64, n - r0);
        if (wave == 0) {
            const int i = r0 + lane;
            double v[RHS];
#pragma unroll
            for (int k = 0; k < RHS; ++k) v[k] = (lane < nb && k < nr) ? y[i + k * NMAX] : 0.0;
            for (int s = 0; s < nb; ++s) {
                const int jj = LOWER ? s : nb - 1 - s;
                const int j = r0 + jj;
                const double ljj = UNIT ? 1.0 : el(j, j);
                const double lij = (lane < nb) ? el(i, j) : 0.0;
                const bool below = LOWER ? (lane > jj) : (lane < jj);
#pragma unroll
                for (int k = 0; k < RHS; ++k) {
                    if (k < nr) {
                        if (!UNIT && lane == jj) v[k] = v[k] / ljj;
                        const double xj = bcast(v[k], jj);
                        if (below && lane < nb) v[k] -= lij * xj;
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < RHS; ++k)
                if (lane < nb && k < nr) y[i + k * NMAX] = v[k];
        }
        __syncthreads();
        // the rows not yet solved, in solve order
        const int lo = LOWER ? r0 + nb : 0, hi = LOWER ? n : r0;
        for (int i = lo + t; i < hi; i += NT) {
            double acc[RHS];
#pragma unroll
            for (int k = 0; k < RHS; ++k) acc[k] = y[i + k * NMAX];
            for (int j = r0; j < r0 + nb; ++j) {
                const double a = el(i, j);
#pragma unroll
                for (int k = 0; k < RHS; ++k)
                    if (k < nr) acc[k] -= a * y[j + k * NMAX];
            }
#pragma unroll
            for (int k = 0; k < RHS; ++k)
                if (k < nr) y[i + k * NMAX] = acc[k];
        }
        __syncthreads();
    }
}

// L.perm := the row origins of the pivot sequence in L.ipiv (each row traced back through the
// interchanges), so that P y is a gather.
__device__ void compute_perm(int n, BatchLds& L) {
    for (int i = threadIdx.x; i < n; i += NT) {
        int x = i;
        for (int s = n - 1; s >= 0; --s) {
            const int p = L.ipiv[s];
            x = x == s ? p : (x == p ? s : x);
        }
        L.perm[i] = x;
    }
    __syncthreads();
}

// L.y := (factored M)^{-1} L.y (for LU, L.perm must hold the pivots' row origins)
template <typename T>
__device__ void factor_solve(const T* M, int n, int nr, bool chol, BatchLds& L) {
    const int t = threadIdx.x;
    if (chol) {
        tri_solve<T, true, false, false>(M, n, nr, L);
        tri_solve<T, false, false, true>(M, n, nr, L);
    } else {
        double tmp[2][RHS];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int i = t + q * NT;
#pragma unroll
            for (int k = 0; k < RHS; ++k) tmp[q][k] = (i < n && k < nr) ? L.y[L.perm[i] + k * NMAX] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int i = t + q * NT;
#pragma unroll
            for (int k = 0; k < RHS; ++k)
                if (i < n && k < nr) L.y[i + k * NMAX] = tmp[q][k];
        }
        __syncthreads();
        tri_solve<T, true, true, false>(M, n, nr, L);
        tri_solve<T, false, false, false>(M, n, nr, L);
    }
}

// ------------------------------------------------------------------------------------------
// the persistent kernel
// ------------------------------------------------------------------------------------------

// Normwise backward error test of the iterate in registers x (rows t and t + 256) against
// b = Bc: r = b - A x is formed in FP64 from A itself (for Cholesky the upper triangle is read
// from the lower); returns true when every column has |r|_inf <= 4 eps (|A|_inf |x|_inf +
// |b|_inf).  r is left in L.y.
__device__ bool residual_small(const double* A, const double* Bc, int n, int nr, bool chol, const double (&x)[2][RHS],
                               BatchLds& L) {
    const int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int k = 0; k < RHS; ++k) {
            const int i = t + q * NT;
            if (i < n && k < nr) L.y[i + k * NMAX] = x[q][k];
        }
    if (t < RHS * 3) L.norm[t] = 0.0;
    if (t == 0) L.anorm = 0.0;
    __syncthreads();
    double r[2][RHS], an[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int i = t + q * NT;
        an[q] = 0.0;
#pragma unroll
        for (int k = 0; k < RHS; ++k) r[q][k] = 0.0;
        if (i < n) {
            for (int j = 0; j < n; ++j) {
                const double a = (!chol || i >= j) ? A[(size_t)i + (size_t)j * n] : A[(size_t)j + (size_t)i * n];
                an[q] += fabs(a);
#pragma unroll
                for (int k = 0; k < RHS; ++k)
                    if (k < nr) r[q][k] = fma(a, L.y[j + k * NMAX], r[q][k]);
            }
#pragma unroll
            for (int k = 0; k < RHS; ++k)
                if (k < nr) r[q][k] = Bc[i + (size_t)k * n] - r[q][k];
        }
    }
    double amax = fmax(an[0], an[1]);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) amax = fmax(amax, __shfl_xor(amax, off, 64));
    for (int k = 0; k < nr; ++k) {
        double rm = 0.0, xm = 0.0, bm = 0.0;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int i = t + q * NT;
            if (i < n) {
                rm = fmax(rm, fabs(r[q][k]));
                xm = fmax(xm, fabs(x[q][k]));
                bm = fmax(bm, fabs(Bc[i + (size_t)k * n]));
            }
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            rm = fmax(rm, __shfl_xor(rm, off, 64));
            xm = fmax(xm, __shfl_xor(xm, off, 64));
            bm = fmax(bm, __shfl_xor(bm, off, 64));
        }
        if ((t & 63) == 0) {
            atomicMax((unsigned long long*)&L.norm[3 * k], (unsigned long long)__double_as_longlong(rm));
            atomicMax((unsigned long long*)&L.norm[3 * k + 1], (unsigned long long)__double_as_longlong(xm));
            atomicMax((unsigned long long*)&L.norm[3 * k + 2], (unsigned long long)__double_as_longlong(bm));
        }
    }
    if ((t & 63) == 0) atomicMax((unsigned long long*)&L.anorm, (unsigned long long)__double_as_longlong(amax));
    __syncthreads();  // all of x read from L.y; norms complete
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int k = 0; k < RHS; ++k) {
            const int i = t + q * NT;
            if (i < n && k < nr) L.y[i + k * NMAX] = r[q][k];
        }
    bool all = true;
    for (int k = 0; k < nr; ++k) {
        const double rm = L.norm[3 * k], xm = L.norm[3 * k + 1], bm = L.norm[3 * k + 2];
        all = all && (rm == 0.0 || rm <= 4.0 * DBL_EPSILON * (L.anorm * xm + bm));
    }
    __syncthreads();
    return all;
}

// MIXED: factor an FP32 copy (in W) and refine in FP64, falling back to FP64; otherwise factor
// A in place in FP64.  info[job]: 0, or the 1-based column of a zero pivot / non-positive
// diagonal, or -4 for an order this kernel does not take; steps[job] (nullable): refinement
// steps of a refined system, 0 for a system solved in FP64, -1 if the mixed mode fell back.
template <bool MIXED>
__global__ __launch_bounds__(NT, 1) void batch_k(const al_batch_entry* __restrict__ entries, int count,
                                                 double* __restrict__ A_pool, double* __restrict__ B_pool,
                                                 float* __restrict__ W_pool, int* __restrict__ info,
                                                 int* __restrict__ steps, unsigned* __restrict__ queue, int max_iter) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    BatchLds& L = *(BatchLds*)lds_raw;
    const int t = threadIdx.x;
    for (;;) {
        if (t == 0) L.job = (int)atomicAdd(queue, 1u);
        __syncthreads();
        const int job = L.job;
        __syncthreads();
        if (job >= count) return;
        const al_batch_entry e = entries[job];
        const int n = e.n, nrhs = e.nrhs;
        const bool chol = e.kind == AL_BATCH_POSV;
        double* A = A_pool + e.a_off;
        double* B = B_pool + e.b_off;
        if (n <= 0 || n > NMAX || nrhs < 0) {
            if (t == 0) {
                info[job] = n == 0 ? 0 : -4;
                if (steps) steps[job] = 0;
            }
            continue;
        }
        int first_col = 0;  // right-hand sides [0, first_col) are solved
        int iters = 0;
        if (MIXED) {
            float* W = W_pool + e.a_off;
            for (size_t idx = t; idx < (size_t)n * n; idx += NT) W[idx] = (float)A[idx];
            __syncthreads();
            const int st = chol ? chol_factor<float>(W, n, L) : lu_factor<float>(W, n, L);
            __syncthreads();
            if (st == 0) {
                if (!chol) compute_perm(n, L);
                for (int k0 = 0; k0 < nrhs; k0 += RHS) {
                    const int nr = min(RHS, nrhs - k0);
                    double* Bc = B + (size_t)k0 * n;
                    for (int i = t; i < n; i += NT)
                        for (int k = 0; k < nr; ++k) L.y[i + k * NMAX] = Bc[i + (size_t)k * n];
                    __syncthreads();
                    factor_solve<float>(W, n, nr, chol, L);
                    double x[2][RHS];
#pragma unroll
                    for (int q = 0; q < 2; ++q)
#pragma unroll
                        for (int k = 0; k < RHS; ++k) {
                            const int i = t + q * NT;
                            x[q][k] = (i < n && k < nr) ? L.y[i + k * NMAX] : 0.0;
                        }
                    __syncthreads();
                    bool conv = false;
                    int it = 0;
                    for (; it <= max_iter; ++it) {
                        if (residual_small(A, Bc, n, nr, chol, x, L)) {
                            conv = true;
                            break;
                        }
                        if (it == max_iter) break;
                        factor_solve<float>(W, n, nr, chol, L);  // d = solve(r), r in L.y
#pragma unroll
                        for (int q = 0; q < 2; ++q)
#pragma unroll
                            for (int k = 0; k < RHS; ++k) {
                                const int i = t + q * NT;
                                if (i < n && k < nr) x[q][k] += L.y[i + k * NMAX];
                            }
                        __syncthreads();
                    }
                    if (!conv) break;
                    iters = max(iters, it + 1);
#pragma unroll
                    for (int q = 0; q < 2; ++q)
#pragma unroll
                        for (int k = 0; k < RHS; ++k) {
                            const int i = t + q * NT;
                            if (i < n && k < nr) Bc[i + (size_t)k * n] = x[q][k];
                        }
                    __syncthreads();
                    first_col = k0 + nr;
                }
            }
            if (first_col >= nrhs) {
                if (t == 0) {
                    info[job] = 0;
                    if (steps) steps[job] = iters;
                }
                __syncthreads();
                continue;
            }
        }
        // plain FP64 (or the fallback of the mixed mode for the columns not yet solved)
        const int st = chol ? chol_factor<double>(A, n, L) : lu_factor<double>(A, n, L);
        __syncthreads();
        if (st == 0) {
            if (!chol) compute_perm(n, L);
            for (int k0 = first_col; k0 < nrhs; k0 += RHS) {
                const int nr = min(RHS, nrhs - k0);
                double* Bc = B + (size_t)k0 * n;
                for (int i = t; i < n; i += NT)
                    for (int k = 0; k < nr; ++k) L.y[i + k * NMAX] = Bc[i + (size_t)k * n];
                __syncthreads();
                factor_solve<double>(A, n, nr, chol, L);
                for (int i = t; i < n; i += NT)
                    for (int k = 0; k < nr; ++k) Bc[i + (size_t)k * n] = L.y[i + k * NMAX];
                __syncthreads();
            }
        }
        if (t == 0) {
            info[job] = st;
            if (steps) steps[job] = MIXED ? -1 : 0;
        }
        __syncthreads();
    }
}

}  // namespace

size_t batch_lds_bytes() { return sizeof(BatchLds); }

int batch_solve(hipStream_t s, size_t count, const al_batch_entry* d_entries, double* A_pool, double* B_pool,
                float* W_pool, int* d_info, int* d_steps, unsigned* queue, bool mixed, int max_iter) {
    if (count == 0) return AL_OK;
    int dev = 0, cus = 0;
    AL_HIP(hipGetDevice(&dev));
    AL_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    static bool attr_set = false;
    if (!attr_set) {
        AL_HIP(hipFuncSetAttribute((const void*)batch_k<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)sizeof(BatchLds)));
        AL_HIP(hipFuncSetAttribute((const void*)batch_k<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)sizeof(BatchLds)));
        attr_set = true;
    }
    AL_CHECK(fill_u32(s, queue, 1, 0u));
    const unsigned grid = (unsigned)std::min<size_t>(count, (size_t)cus);
    const int ticket = prof_begin(s);
    if (mixed)
        hipLaunchKernelGGL(batch_k<true>, dim3(grid), dim3(NT), sizeof(BatchLds), s, d_entries, (int)count, A_pool,
                           B_pool, W_pool, d_info, d_steps, queue, max_iter);
    else
        hipLaunchKernelGGL(batch_k<false>, dim3(grid), dim3(NT), sizeof(BatchLds), s, d_entries, (int)count, A_pool,
                           B_pool, W_pool, d_info, d_steps, queue, max_iter);
    AL_HIP(hipGetLastError());
    if (ticket >= 0) prof_end(ticket, s, mixed ? "batch_mixed" : "batch_fp64", 0.0, 0.0);
    return AL_OK;
}

}  // namespace al
