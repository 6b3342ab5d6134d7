// Batched solves of many small independent systems (orders 1..512) in one call: dgesv (LU with
// partial pivoting) or dposv (lower Cholesky) per system, in plain FP64 or as an FP32
// factorization refined in FP64 (with an FP64 refactorization if refinement fails).
//
// The crate solves one system per call (lapack::dgesv, src/lapack/mod.rs:309-323; dpotrf,
// :368-399); launching the GPU path once per small system leaves the device idle between
// launches.  Here a batch of uneven sizes is split on the device into size buckets, and each
// bucket is served by a persistent kernel shaped for its sizes:
//
//   bucket 0, n <= 64:   one wave per system (a 64-thread workgroup), the whole matrix in LDS,
//                        lane i owning row i.  The elimination, the pivot search (a wave
//                        reduction: largest magnitude, ties to the first row, a NaN only on the
//                        diagonal), the Cholesky dot products and both substitutions follow the
//                        crate's own loops in its order and without FMA, so an FP64 system is
//                        solved bit for bit as lapack::dgesv / dpotrf + dtrsv would solve it;
//   buckets 1-3, n in (64, 128], (128, 256], (256, 512]: one 256-thread workgroup per system,
//                        blocked right-looking with KB-column panels held in LDS (the pivot
//                        search over the panel column, the crate's arithmetic inside a panel),
//                        and the trailing update on the matrix cores (v_mfma_f64_16x16x4_f64, or
//                        the FP32 one for the mixed mode's factor): each wave owns 16 x 64 blocks
//                        of the trailing matrix, its operands the panel (LDS) and a 64-column
//                        stripe of U12 (LDS).  LDS is sized per bucket, so smaller orders fit more
//                        workgroups per CU.
//
// A pre-pass kernel partitions the entries into the buckets stably (each bucket keeps the
// caller's order; decreasing size balances best).  The largest bucket runs on the caller's
// stream and the three others on a second stream of the same workspace, joined by events, so
// their tails overlap; every launch reads its bucket's count from device memory, so the call
// needs no host synchronization and can be captured in a HIP graph.  The panel width and the
// workgroups per CU of each bucket come from a table written by the project's tuner
// (batch_tuning.json beside the library, see tune_batch.py), or from al_batch_tune_set.
//
// Mixed mode: A is rounded to FP32 (bucket 0: in LDS; buckets 1-3: in the caller-sized work
// pool) and factored there; x = solve(b), then up to max_iter steps of r = b - A x (FP64, A
// itself), x += solve(r) until the normwise backward error of every column is at most 4 eps; a
// system that does not get there is factored again in FP64 and its remaining columns solved,
// and its step count says so.
#include <cfloat>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dlfcn.h>
#include <mutex>
#include <string>

#include "al_internal.hpp"

namespace al {
namespace {

constexpr int NBUCKETS = 4;
constexpr int BUCKET_MAX[NBUCKETS] = {64, 128, 256, 512};
constexpr int NT = 256;
constexpr int RHS = 8;        // right-hand sides per pass (workgroup kernels)
constexpr int TILE = 64;      // trailing-update stripe width
constexpr int SMAX = 64;      // bucket 0: largest order
constexpr int SLD = SMAX + 1; // bucket 0: LDS leading dimension
constexpr int SRB = 8;        // bucket 0: right-hand sides per pass (one lane each)

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------------------------------
// shared helpers
// ------------------------------------------------------------------------------------------

__device__ __forceinline__ double bcast(double v, int src) {
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(bits & 0xffffffffLL), src);
    const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ int wave_min_int(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const int o = __shfl_xor(v, off, 64);
        v = o < v ? o : v;
    }
    return v;
}

__device__ __forceinline__ double wave_max_double(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

// The pivot key of candidate row i (current position) of column k, as the crate's scan sees it
// (src/lapack/mod.rs:126-134): the scan starts from the diagonal's magnitude and takes a later
// row only when strictly larger, so a NaN on the diagonal keeps the diagonal and a later NaN
// never wins.  Keys: |v|, +inf for a NaN on the diagonal, -1 for a later NaN.
__device__ __forceinline__ double pivot_key(double v, bool diag) {
    if (v != v) return diag ? (double)INFINITY : -1.0;
    return fabs(v);
}

// Per-element-type MFMA (16x16x4): FP64 for the FP64 factorizations, FP32 for the mixed
// mode's factor.  Fed as (B^T, A^T), accumulator r of lane l holds C[row = l & 15][col =
// drow(l >> 4, r)] of a 16 x 16 block.
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef d4 acc_t;
    static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int kq, int r) { return kq + 4 * r; }
};
template <>
struct Mfma<float> {
    typedef f4 acc_t;
    static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int drow(int kq, int r) { return 4 * kq + r; }
};

// ------------------------------------------------------------------------------------------
// pre-pass: stable partition of the entries into the size buckets
// ------------------------------------------------------------------------------------------

__device__ __forceinline__ int bucket_of(int n) {
    if (n <= BUCKET_MAX[0] || n > BUCKET_MAX[NBUCKETS - 1]) return 0;  // bucket 0 also reports bad orders
    if (n <= BUCKET_MAX[1]) return 1;
    if (n <= BUCKET_MAX[2]) return 2;
    return 3;
}

// One 1024-thread workgroup: chunks of 1024 entries; per chunk and bucket a ballot prefix count
// inside each wave and a scan over the 16 waves give every entry its place.  lists: NBUCKETS
// arrays of `count` ints; counts[b]: entries in bucket b.
__global__ __launch_bounds__(1024) void bucket_k(const al_batch_entry* __restrict__ entries, int count,
                                                 int* __restrict__ lists, int* __restrict__ counts) {
    __shared__ int wtot[NBUCKETS][16];
    __shared__ int base[NBUCKETS];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (t < NBUCKETS) base[t] = 0;
    __syncthreads();
    for (int c0 = 0; c0 < count; c0 += 1024) {
        const int idx = c0 + t;
        const int b = idx < count ? bucket_of(entries[idx].n) : -1;
        int prefix = 0;
#pragma unroll
        for (int bb = 0; bb < NBUCKETS; ++bb) {
            const unsigned long long m = __ballot(b == bb);
            if (b == bb) prefix = __popcll(m & ((1ull << lane) - 1ull));
            if (lane == 0) wtot[bb][wave] = __popcll(m);
        }
        __syncthreads();
        if (b >= 0) {
            int off = base[b];
            for (int w = 0; w < wave; ++w) off += wtot[b][w];
            lists[(size_t)b * count + off + prefix] = idx;
        }
        __syncthreads();
        if (t < NBUCKETS) {
            int s = 0;
            for (int w = 0; w < 16; ++w) s += wtot[t][w];
            base[t] += s;
        }
        __syncthreads();
    }
    if (t < NBUCKETS) counts[t] = base[t];
}

// ------------------------------------------------------------------------------------------
// bucket 0: one wave per system, the crate's loops
// ------------------------------------------------------------------------------------------

struct SmallLds {
    double S[SMAX * SLD];   // the matrix (FP64, or FP32 in its first half in the mixed mode)
    double Y[SMAX * SRB];   // right-hand sides / solutions, column k at Y + k * SMAX
    double X[SMAX * SRB];   // mixed mode: the iterate
    int piv[SMAX];
    double norm[3 * SRB + 1];
};

#pragma clang fp contract(off)

// In-place LU with partial pivoting of the n x n matrix S (leading dimension SLD), lane i owning
// row i: the crate's dgetrf_with_pivots (src/lapack/mod.rs:112-163) step for step.  Returns the
// 1-based column of an exactly zero pivot (the crate's Singular) or 0.
template <typename T>
__device__ int small_lu(T* S, int n, int* piv) {
    const int i = threadIdx.x;
    for (int k = 0; k < n; ++k) {
        const double key = (i >= k && i < n) ? pivot_key((double)S[i + k * SLD], i == k) : -2.0;
        const double best = wave_max_double(key);
        const int p = wave_min_int((key == best && i >= k && i < n) ? i : 0x7fffffff);
        // |pivot| <= 0 (ZERO_DETERMINANT): Singular; a NaN diagonal (key +inf) is not
        if (!(best > 0.0)) return k + 1;
        if (i == 0) piv[k] = p;
        if (p != k && i < n) {  // swap rows k and p across every column: lane i takes column i
            const T a = S[k + i * SLD];
            S[k + i * SLD] = S[p + i * SLD];
            S[p + i * SLD] = a;
        }
        __syncthreads();
        const T d = S[k + k * SLD];
        if (i > k && i < n) {
            const T f = S[i + k * SLD] / d;
            S[i + k * SLD] = f;
            for (int j = k + 1; j < n; ++j) S[i + j * SLD] -= f * S[k + j * SLD];
        }
        __syncthreads();
    }
    return 0;
}

// Lower Cholesky of S in place (only the lower triangle read or written): the crate's dpotrf
// (src/lapack/mod.rs:368-399), lane i computing row i's entries.  Returns the 1-based column of
// a non-positive diagonal or 0.
template <typename T>
__device__ int small_chol(T* S, int n) {
    const int i = threadIdx.x;
    for (int j = 0; j < n; ++j) {
        T sum = S[j + j * SLD];
        for (int k = 0; k < j; ++k) {
            const T ljk = S[j + k * SLD];
            sum -= ljk * ljk;
        }
        if (!(sum > (T)0)) return j + 1;  // sum <= 0 (or NaN): not positive definite
        const T ljj = sqrt(sum);
        __syncthreads();                  // every lane has read the diagonal entry
        if (i == j) S[j + j * SLD] = ljj;
        if (i > j && i < n) {
            T s = S[i + j * SLD];
            for (int k = 0; k < j; ++k) s -= S[i + k * SLD] * S[j + k * SLD];
            S[i + j * SLD] = s / ljj;
        }
        __syncthreads();
    }
    return 0;
}

// Solves with the factors in S for the nr <= SRB columns of Y (lane k < nr solves column k,
// sequentially, in the crate's order): LU -- the interchanges in order, then the unit-lower
// and the upper sweeps of dgetrs (src/lapack/mod.rs:330-354); Cholesky -- dtrsv with L, then
// dtrsv_transpose with L (src/blas.rs:1355-1370, tri_solve_f64 :1410-1438).  The factor may be
// FP32 (mixed mode); the vectors are FP64.
template <typename T>
__device__ void small_solve(const T* S, int n, int nr, bool chol, const int* piv, double* Y) {
    const int k = threadIdx.x;
    if (k < nr) {
        double* y = Y + k * SMAX;
        if (!chol) {
            for (int s = 0; s < n; ++s) {
                const int p = piv[s];
                const double a = y[s];
                y[s] = y[p];
                y[p] = a;
            }
            for (int r = 0; r < n; ++r) {
                double sum = y[r];
                for (int j = 0; j < r; ++j) sum -= (double)S[r + j * SLD] * y[j];
                y[r] = sum;
            }
            for (int r = n - 1; r >= 0; --r) {
                double sum = y[r];
                for (int j = r + 1; j < n; ++j) sum -= (double)S[r + j * SLD] * y[j];
                y[r] = sum / (double)S[r + r * SLD];
            }
        } else {
            for (int r = 0; r < n; ++r) {
                double sum = y[r];
                for (int j = 0; j < r; ++j) sum -= (double)S[r + j * SLD] * y[j];
                y[r] = sum / (double)S[r + r * SLD];
            }
            for (int r = n - 1; r >= 0; --r) {
                double sum = y[r];
                for (int j = r + 1; j < n; ++j) sum -= (double)S[j + r * SLD] * y[j];
                y[r] = sum / (double)S[r + r * SLD];
            }
        }
    }
    __syncthreads();
}

#pragma clang fp contract(on)

// Mixed mode, bucket 0: r = b - A x for the nr columns (A FP64 from global, the full matrix for
// LU, the lower triangle mirrored for Cholesky); r -> Y; true when every column's normwise
// backward error is at most 4 eps.
__device__ bool small_residual(const double* A, const double* Bc, int n, int nr, bool chol, SmallLds& L) {
    const int i = threadIdx.x;
    double an = 0.0;
    double r[SRB];
#pragma unroll
    for (int k = 0; k < SRB; ++k) r[k] = 0.0;
    if (i < n) {
        for (int j = 0; j < n; ++j) {
            const double a = (!chol || i >= j) ? A[(size_t)i + (size_t)j * n] : A[(size_t)j + (size_t)i * n];
            an += fabs(a);
#pragma unroll
            for (int k = 0; k < SRB; ++k)
                if (k < nr) r[k] = fma(a, L.X[j + k * SMAX], r[k]);
        }
#pragma unroll
        for (int k = 0; k < SRB; ++k)
            if (k < nr) r[k] = Bc[i + (size_t)k * n] - r[k];
    }
    const double anorm = wave_max_double(an);
    bool all = true;
#pragma unroll
    for (int k = 0; k < SRB; ++k) {
        if (k >= nr) continue;
        const double rm = wave_max_double(i < n ? fabs(r[k]) : 0.0);
        const double xm = wave_max_double(i < n ? fabs(L.X[i + k * SMAX]) : 0.0);
        const double bm = wave_max_double(i < n ? fabs(Bc[i + (size_t)k * n]) : 0.0);
        all = all && (rm == 0.0 || rm <= 4.0 * DBL_EPSILON * (anorm * xm + bm));
        if (i < n) L.Y[i + k * SMAX] = r[k];
    }
    __syncthreads();
    return all;
}

template <bool MIXED>
__global__ __launch_bounds__(64) void batch_small_k(const al_batch_entry* __restrict__ entries,
                                                    const int* __restrict__ list, const int* __restrict__ counts,
                                                    double* __restrict__ A_pool, double* __restrict__ B_pool,
                                                    int* __restrict__ info, int* __restrict__ steps,
                                                    unsigned* __restrict__ queue, int max_iter) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    SmallLds& L = *(SmallLds*)lds_raw;
    const int i = threadIdx.x;
    const int count = counts[0];
    for (;;) {
        int slot = 0;
        if (i == 0) slot = (int)atomicAdd(queue, 1u);
        slot = __builtin_amdgcn_readfirstlane(slot);
        if (slot >= count) return;
        const int job = list[slot];
        const al_batch_entry e = entries[job];
        const int n = e.n, nrhs = e.nrhs;
        const bool chol = e.kind == AL_BATCH_POSV;
        double* A = A_pool + e.a_off;
        double* B = B_pool + e.b_off;
        if (n <= 0 || n > SMAX || nrhs < 0) {
            if (i == 0) {
                info[job] = n == 0 ? 0 : -4;  // -4: an order beyond 512
                if (steps) steps[job] = 0;
            }
            continue;
        }
        int first_col = 0;
        int iters = 0;
        if (MIXED) {
            float* Sf = (float*)L.S;
            if (i < n)
                for (int c = 0; c < n; ++c)
                    Sf[i + c * SLD] = (!chol || i >= c) ? (float)A[(size_t)i + (size_t)c * n] : 0.f;
            __syncthreads();
            const int st = chol ? small_chol<float>(Sf, n) : small_lu<float>(Sf, n, L.piv);
            __syncthreads();
            if (st == 0) {
                for (int k0 = 0; k0 < nrhs; k0 += SRB) {
                    const int nr = min(SRB, nrhs - k0);
                    double* Bc = B + (size_t)k0 * n;
                    if (i < n)
                        for (int k = 0; k < nr; ++k) L.Y[i + k * SMAX] = Bc[i + (size_t)k * n];
                    __syncthreads();
                    small_solve<float>(Sf, n, nr, chol, L.piv, L.Y);
                    if (i < n)
                        for (int k = 0; k < nr; ++k) L.X[i + k * SMAX] = L.Y[i + k * SMAX];
                    __syncthreads();
                    bool conv = false;
                    int it = 0;
                    for (; it <= max_iter; ++it) {
                        if (small_residual(A, Bc, n, nr, chol, L)) {
                            conv = true;
                            break;
                        }
                        if (it == max_iter) break;
                        small_solve<float>(Sf, n, nr, chol, L.piv, L.Y);
                        if (i < n)
                            for (int k = 0; k < nr; ++k) L.X[i + k * SMAX] += L.Y[i + k * SMAX];
                        __syncthreads();
                    }
                    if (!conv) break;
                    iters = max(iters, it + 1);
                    if (i < n)
                        for (int k = 0; k < nr; ++k) Bc[i + (size_t)k * n] = L.X[i + k * SMAX];
                    __syncthreads();
                    first_col = k0 + nr;
                }
            }
            if (first_col >= nrhs) {
                if (i == 0) {
                    info[job] = 0;
                    if (steps) steps[job] = iters;
                }
                continue;
            }
        }
        // FP64 (or the mixed mode's fallback for the columns not yet solved): factor in LDS,
        // write the factors back over A
        if (i < n)
            for (int c = 0; c < n; ++c) L.S[i + c * SLD] = A[(size_t)i + (size_t)c * n];
        __syncthreads();
        const int st = chol ? small_chol<double>(L.S, n) : small_lu<double>(L.S, n, L.piv);
        __syncthreads();
        if (i < n)
            for (int c = 0; c < n; ++c)
                if (!chol || i >= c) A[(size_t)i + (size_t)c * n] = L.S[i + c * SLD];
        if (st == 0) {
            for (int k0 = first_col; k0 < nrhs; k0 += SRB) {
                const int nr = min(SRB, nrhs - k0);
                double* Bc = B + (size_t)k0 * n;
                if (i < n)
                    for (int k = 0; k < nr; ++k) L.Y[i + k * SMAX] = Bc[i + (size_t)k * n];
                __syncthreads();
                small_solve<double>(L.S, n, nr, chol, L.piv, L.Y);
                if (i < n)
                    for (int k = 0; k < nr; ++k) Bc[i + (size_t)k * n] = L.Y[i + k * SMAX];
                __syncthreads();
            }
        }
        if (i == 0) {
            info[job] = st;
            if (steps) steps[job] = MIXED ? -1 : 0;
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------
// buckets 1-3: one workgroup per system
// ------------------------------------------------------------------------------------------

constexpr int TB = 32;  // diagonal block of the triangular solves (workgroup kernels)

// The panel is needed only while a system is factored, the right-hand sides and the staged
// diagonal block only while it is solved: they share their LDS, which lets two (order 512)
// to six (order 128) workgroups share a CU and hide each other's memory latency.
template <int NMAX, int KB>
struct BatchLds {
    union {
        double panel[NMAX * KB];          // panel, column-major with leading dimension m = n - k0 (T = double)
        struct {
            double y[NMAX * RHS];         // right-hand sides / solution, column k at y + k * NMAX
            double diag[TB * (TB + 1)];   // the solve's current diagonal block
        } s;
    } u;
    double stripe[KB * (TILE + 1)];       // KB x 64 stripe of U12 (or L21^T), row p at stripe + p * (TILE + 1)
    int ipiv[NMAX];
    int perm[NMAX];
    int swp[1 + 2 * 32];                  // 16 pivots' interchanges composed: count, destinations, sources
    int widx[4];
    int job;
    double norm[RHS * 3];
    double anorm;
};

// The workgroup's pivot from every thread's candidate (its first maximum by pivot_key over its
// rows, -2 if none): the largest key, ties to the lowest row; -1 when no candidate is positive
// (the crate's exactly zero pivot).  One barrier: each wave reduces in registers and writes
// one slot, then every thread reads the four slots (the next call's writes come two barriers
// later).
template <typename LDS>
__device__ int panel_pick(double key, int idx, LDS& L) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const double wmax = wave_max_double(key);
    const int widx = wave_min_int(key == wmax ? idx : 0x7fffffff);
    if (lane == 0) {
        L.norm[wave] = wmax;
        L.widx[wave] = widx;
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
    return (gi == 0x7fffffff || !(gbest > 0.0)) ? -1 : gi;
}

// Trailing update M[r, c] -= sum_k Lp[r, k] U[k, c] over rows [k0 + kb, n) and columns
// [k0 + kb, n) (Cholesky: the lower triangle only) on the matrix cores.  Lp is the panel in LDS
// (rows from k0, leading dimension m, kb columns); U is staged per 64-column stripe from M's
// rows k0..k0+kb (LU) or from the panel transposed (Cholesky).  Each wave owns blocks of
// 16 UR rows x 64 columns (UR = 1: the smallest blocks measured fastest -- 21.5 ms for the
// 257..512 bucket of the 20,000-system batch against 29.9 with UR = 2 and 31.1 with UR = 4);
// out of line, so its accumulators do not compete with the factorization's registers.
constexpr int UR = 1;
template <typename T, bool CHOL, typename LDS, int KB>
__device__ __noinline__ void trailing_update(T* M, int n, int k0, int kb, const T* Pl, int m, LDS& L) {
    typedef typename Mfma<T>::acc_t acc_t;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int col = lane & 15, kq = lane >> 4;
    const int r0 = k0 + kb;
    T* stripe = (T*)L.stripe;
    for (int cc = r0; cc < n; cc += TILE) {
        const int cn = min(TILE, n - cc);
        __syncthreads();
        for (int idx = t; idx < KB * TILE; idx += NT) {
            // LU: consecutive threads read consecutive rows of a U12 column (contiguous in M)
            const int k = CHOL ? idx / TILE : idx % KB, c = CHOL ? idx % TILE : idx / KB;
            T v = (T)0;
            if (c < cn && k < kb) v = CHOL ? Pl[(cc + c - k0) + k * m] : M[(size_t)(k0 + k) + (size_t)(cc + c) * n];
            stripe[c + k * (TILE + 1)] = v;
        }
        __syncthreads();
        const int rstart = CHOL ? cc : r0;
        for (int rc = rstart + wave * (16 * UR); rc < n; rc += 4 * 16 * UR) {
            // the accumulators start as the C block itself (loads issued together), then
            // acc += (-Lp) U: one memory round trip per 64 x 64 block
            acc_t acc[UR][4];
#pragma unroll
            for (int i = 0; i < UR; ++i) {
                const int r = rc + i * 16 + col;
#pragma unroll
                for (int j = 0; j < 4; ++j)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int c = cc + j * 16 + Mfma<T>::drow(kq, q);
                        acc[i][j][q] = (r < n && c < n) ? M[(size_t)r + (size_t)c * n] : (T)0;
                    }
            }
#pragma unroll
            for (int s = 0; s < KB / 4; ++s) {
                const int kk = 4 * s + kq;
                T af[UR], bf[4];
#pragma unroll
                for (int i = 0; i < UR; ++i) {
                    const int r = rc + i * 16 + col;
                    af[i] = (kk < kb && r < n) ? -Pl[(r - k0) + kk * m] : (T)0;
                }
#pragma unroll
                for (int j = 0; j < 4; ++j) bf[j] = stripe[j * 16 + col + kk * (TILE + 1)];
#pragma unroll
                for (int i = 0; i < UR; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[i][j] = Mfma<T>::op(bf[j], af[i], acc[i][j]);
            }
#pragma unroll
            for (int i = 0; i < UR; ++i) {
                const int r = rc + i * 16 + col;
                if (r >= n) continue;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int c = cc + j * 16 + Mfma<T>::drow(kq, q);
                        if (c >= n || (CHOL && r < c)) continue;
                        M[(size_t)r + (size_t)c * n] = acc[i][j][q];
                    }
                }
            }
        }
    }
    __syncthreads();
}

// The panel's interchanges ipiv[k0..k0+kb) on the columns outside the panel, 16 pivots at a
// time (applying the groups in order equals applying the swaps in order): thread 0 composes a
// group's swaps into (destination, source) row pairs -- at most 32 rows move -- then one thread
// per column reads every moving row's value (independent loads) and writes them moved.  Kept
// out of line: its 32 values per thread would otherwise crowd the factorization's registers.
template <typename T, typename LDS>
__device__ __noinline__ void apply_swaps(T* M, int n, int k0, int kb, LDS& L) {
    const int t = threadIdx.x;
    for (int g0 = 0; g0 < kb; g0 += 16) {
        const int g1 = min(kb, g0 + 16);
        if (t == 0) {
            int* dst = L.swp + 1;
            int* src = L.swp + 1 + 32;
            int np = 0;
            for (int s2 = g0; s2 < g1; ++s2) {
                const int a = k0 + s2, b = L.ipiv[k0 + s2];
                if (a == b) continue;
                int ia = -1, ib = -1;
                for (int q = 0; q < np; ++q) {
                    if (dst[q] == a) ia = q;
                    if (dst[q] == b) ib = q;
                }
                if (ia < 0) {
                    dst[np] = a;
                    src[np] = a;
                    ia = np++;
                }
                if (ib < 0) {
                    dst[np] = b;
                    src[np] = b;
                    ib = np++;
                }
                const int tmp = src[ia];
                src[ia] = src[ib];
                src[ib] = tmp;
            }
            L.swp[0] = np;
        }
        __syncthreads();
        const int np = L.swp[0];
        const int* dst = L.swp + 1;
        const int* src = L.swp + 1 + 32;
        if (np > 0) {
            for (int c = t; c < n; c += NT) {
                if (c >= k0 && c < k0 + kb) continue;
                T* pc = M + (size_t)c * n;
                T v[32];
#pragma unroll
                for (int q = 0; q < 32; ++q)
                    if (q < np) v[q] = pc[src[q]];
#pragma unroll
                for (int q = 0; q < 32; ++q)
                    if (q < np) pc[dst[q]] = v[q];
            }
        }
        __syncthreads();
    }
}

// U12 := L11^{-1} A12 for the columns right of the KB-column panel at k0, in place: one thread
// per column with the column's kb entries in registers (row r: its entry minus L11[r, i] times
// the solved entries i < r, in that order), L11 from the panel in LDS (P, leading dimension m).
// Out of line, so its registers do not add to the factorization's.
template <typename T, int KB>
__device__ __noinline__ void panel_u12(T* M, int n, int k0, int kb, const T* P, int m) {
    for (int c = k0 + kb + (int)threadIdx.x; c < n; c += NT) {
        T* pc = M + (size_t)c * n + k0;
        T u[KB];
#pragma unroll
        for (int r = 0; r < KB; ++r) u[r] = r < kb ? pc[r] : (T)0;
#pragma unroll
        for (int r = 1; r < KB; ++r) {
            if (r < kb) {
                T v = u[r];
#pragma unroll
                for (int i = 0; i < r; ++i) v -= P[r + i * m] * u[i];
                u[r] = v;
            }
        }
#pragma unroll
        for (int r = 1; r < KB; ++r)
            if (r < kb) pc[r] = u[r];
    }
}

// LU with partial pivoting of the n x n matrix M (leading dimension n); pivots (0-based) in
// L.ipiv; returns the first zero pivot's 1-based column or 0.
template <typename T, int KB, typename LDS>
__device__ int lu_factor(T* M, int n, LDS& L) {
    const int t = threadIdx.x;
    T* P = (T*)L.u.panel;
    for (int k0 = 0; k0 < n; k0 += KB) {
        const int kb = min(KB, n - k0);
        const int m = n - k0;
        for (int idx = t; idx < m * kb; idx += NT) {
            const int i = idx % m, c = idx / m;
            P[i + c * m] = M[(size_t)(k0 + i) + (size_t)(k0 + c) * n];
        }
        __syncthreads();
        // the first column's candidates: each thread's first maximum over its rows
        double key = -2.0;
        int kidx = 0x7fffffff;
        for (int i = t; i < m; i += NT) {
            const double v = pivot_key((double)P[i], i == 0);
            if (v > key) {
                key = v;
                kidx = i;
            }
        }
        for (int j = 0; j < kb; ++j) {
            const int p = panel_pick(key, kidx, L);
            if (p < 0) {
                // the crate stops at an exactly zero pivot (src/lapack/mod.rs:134-139)
                __syncthreads();
                return k0 + j + 1;
            }
            if (t == 0) L.ipiv[k0 + j] = k0 + p;
            if (p != j && t < kb) {
                const T a = P[j + t * m];
                P[j + t * m] = P[p + t * m];
                P[p + t * m] = a;
            }
            __syncthreads();
            const T piv = P[j + j * m];
            // eliminate column j, updating column j + 1 first so the next candidates come out of
            // the same pass
            key = -2.0;
            kidx = 0x7fffffff;
            for (int i = j + 1 + t; i < m; i += NT) {
                // the crate's own operations inside a panel: divide, then a separate multiply
                // and subtract (no FMA), so an exactly singular matrix meets its exact zero
#pragma clang fp contract(off)
                const T l = P[i + j * m] / piv;
                P[i + j * m] = l;
                for (int c = j + 1; c < kb; ++c) P[i + c * m] -= l * P[j + c * m];
                if (j + 1 < kb) {
                    const double v = pivot_key((double)P[i + (j + 1) * m], i == j + 1);
                    if (v > key) {
                        key = v;
                        kidx = i;
                    }
                }
            }
            __syncthreads();
        }
        for (int idx = t; idx < m * kb; idx += NT) {
            const int i = idx % m, c = idx / m;
            M[(size_t)(k0 + i) + (size_t)(k0 + c) * n] = P[i + c * m];
        }
        apply_swaps<T>(M, n, k0, kb, L);
        if (k0 + kb < n) {
            panel_u12<T, KB>(M, n, k0, kb, P, m);
            __syncthreads();
            trailing_update<T, false, LDS, KB>(M, n, k0, kb, P, m, L);
        }
    }
    return 0;
}

// Lower Cholesky (only the lower triangle is read or written); returns the first column whose
// diagonal is not positive (1-based) or 0, as lapack::dpotrf (src/lapack/mod.rs:368-399).
template <typename T, int KB, typename LDS>
__device__ int chol_factor(T* M, int n, LDS& L) {
    const int t = threadIdx.x;
    T* P = (T*)L.u.panel;
    for (int k0 = 0; k0 < n; k0 += KB) {
        const int kb = min(KB, n - k0);
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
        if (k0 + kb < n) trailing_update<T, true, LDS, KB>(M, n, k0, kb, P, m, L);
    }
    return 0;
}

// Triangular solves on the right-hand sides in LDS (n x nr, leading dimension NMAX), FP64
// vectors.  op = L (unit or not) forward, or U / L^T backward.  TRANS: the factor is read
// transposed (L^T x = y).  Per 32-row diagonal block: the block's triangle is staged in LDS by
// every thread, one wave solves it (lane = row, x_j handed out by readlane), then every thread
// updates the rows not yet solved with the block's columns.
template <typename T, bool LOWER, bool UNIT, bool TRANS, int NMAX, typename LDS>
__device__ void tri_solve(const T* M, int n, int nr, LDS& L) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    double* y = L.u.s.y;
    double* D = L.u.s.diag;
    auto el = [&](int i, int j) -> double {  // element (i, j) of op(factor)
        return TRANS ? (double)M[(size_t)j + (size_t)i * n] : (double)M[(size_t)i + (size_t)j * n];
    };
    const int nblk = (n + TB - 1) / TB;
    for (int bb = 0; bb < nblk; ++bb) {
        const int b = LOWER ? bb : nblk - 1 - bb;
        const int r0 = b * TB, nb = min(TB, n - r0);
        for (int idx = t; idx < TB * TB; idx += NT) {
            const int i = idx % TB, j = idx / TB;
            const bool need = i < nb && j < nb && (LOWER ? (j < i || (!UNIT && j == i)) : (j > i || (!UNIT && j == i)));
            D[i + j * (TB + 1)] = need ? el(r0 + i, r0 + j) : 0.0;
        }
        __syncthreads();
        if (wave == 0) {
            const int i = r0 + lane;
            double v[RHS];
#pragma unroll
            for (int k = 0; k < RHS; ++k) v[k] = (lane < nb && k < nr) ? y[i + k * NMAX] : 0.0;
            for (int s = 0; s < nb; ++s) {
                const int jj = LOWER ? s : nb - 1 - s;
                const double ljj = UNIT ? 1.0 : D[jj + jj * (TB + 1)];
                const double lij = (lane < nb) ? D[lane + jj * (TB + 1)] : 0.0;
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

// L.perm := the row origins of the pivot sequence in L.ipiv, so that P y is a gather.
template <typename LDS>
__device__ void compute_perm(int n, LDS& L) {
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
template <typename T, int NMAX, typename LDS>
__device__ void factor_solve(const T* M, int n, int nr, bool chol, LDS& L) {
    constexpr int Q = (NMAX + NT - 1) / NT;
    const int t = threadIdx.x;
    if (chol) {
        tri_solve<T, true, false, false, NMAX>(M, n, nr, L);
        tri_solve<T, false, false, true, NMAX>(M, n, nr, L);
    } else {
        double tmp[Q][RHS];
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int i = t + q * NT;
#pragma unroll
            for (int k = 0; k < RHS; ++k) tmp[q][k] = (i < n && k < nr) ? L.u.s.y[L.perm[i] + k * NMAX] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int i = t + q * NT;
#pragma unroll
            for (int k = 0; k < RHS; ++k)
                if (i < n && k < nr) L.u.s.y[i + k * NMAX] = tmp[q][k];
        }
        __syncthreads();
        tri_solve<T, true, true, false, NMAX>(M, n, nr, L);
        tri_solve<T, false, false, false, NMAX>(M, n, nr, L);
    }
}

// Normwise backward error test of the iterate x (rows t + q*256) against b = Bc: r = b - A x in
// FP64 from A itself (Cholesky: the upper triangle read from the lower); true when every column
// has |r|_inf <= 4 eps (|A|_inf |x|_inf + |b|_inf).  r is left in L.y.
template <int NMAX, typename LDS, int Q>
__device__ bool residual_small(const double* A, const double* Bc, int n, int nr, bool chol, const double (&x)[Q][RHS],
                               LDS& L) {
    const int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
        for (int k = 0; k < RHS; ++k) {
            const int i = t + q * NT;
            if (i < n && k < nr) L.u.s.y[i + k * NMAX] = x[q][k];
        }
    if (t < RHS * 3) L.norm[t] = 0.0;
    if (t == 0) L.anorm = 0.0;
    __syncthreads();
    double r[Q][RHS], an[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
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
                    if (k < nr) r[q][k] = fma(a, L.u.s.y[j + k * NMAX], r[q][k]);
            }
#pragma unroll
            for (int k = 0; k < RHS; ++k)
                if (k < nr) r[q][k] = Bc[i + (size_t)k * n] - r[q][k];
        }
    }
    double amax = an[0];
#pragma unroll
    for (int q = 1; q < Q; ++q) amax = fmax(amax, an[q]);
    amax = wave_max_double(amax);
    for (int k = 0; k < nr; ++k) {
        double rm = 0.0, xm = 0.0, bm = 0.0;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int i = t + q * NT;
            if (i < n) {
                rm = fmax(rm, fabs(r[q][k]));
                xm = fmax(xm, fabs(x[q][k]));
                bm = fmax(bm, fabs(Bc[i + (size_t)k * n]));
            }
        }
        rm = wave_max_double(rm);
        xm = wave_max_double(xm);
        bm = wave_max_double(bm);
        if ((t & 63) == 0) {
            atomicMax((unsigned long long*)&L.norm[3 * k], (unsigned long long)__double_as_longlong(rm));
            atomicMax((unsigned long long*)&L.norm[3 * k + 1], (unsigned long long)__double_as_longlong(xm));
            atomicMax((unsigned long long*)&L.norm[3 * k + 2], (unsigned long long)__double_as_longlong(bm));
        }
    }
    if ((t & 63) == 0) atomicMax((unsigned long long*)&L.anorm, (unsigned long long)__double_as_longlong(amax));
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
        for (int k = 0; k < RHS; ++k) {
            const int i = t + q * NT;
            if (i < n && k < nr) L.u.s.y[i + k * NMAX] = r[q][k];
        }
    bool all = true;
    for (int k = 0; k < nr; ++k) {
        const double rm = L.norm[3 * k], xm = L.norm[3 * k + 1], bm = L.norm[3 * k + 2];
        all = all && (rm == 0.0 || rm <= 4.0 * DBL_EPSILON * (L.anorm * xm + bm));
    }
    __syncthreads();
    return all;
}

// One workgroup per system of the bucket (orders in (NMAX / 2, NMAX]); see the file comment.
template <int NMAX, int KB, bool MIXED>
__global__ __launch_bounds__(NT, 1) void batch_k(const al_batch_entry* __restrict__ entries, const int* __restrict__ list,
                                                 const int* __restrict__ counts, int bucket,
                                                 double* __restrict__ A_pool, double* __restrict__ B_pool,
                                                 float* __restrict__ W_pool, int* __restrict__ info,
                                                 int* __restrict__ steps, unsigned* __restrict__ queue, int max_iter) {
    typedef BatchLds<NMAX, KB> LDS;
    constexpr int Q = (NMAX + NT - 1) / NT;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    LDS& L = *(LDS*)lds_raw;
    const int t = threadIdx.x;
    const int count = counts[bucket];
    for (;;) {
        if (t == 0) L.job = (int)atomicAdd(queue, 1u);
        __syncthreads();
        const int slot = L.job;
        __syncthreads();
        if (slot >= count) return;
        const int job = list[slot];
        const al_batch_entry e = entries[job];
        const int n = e.n, nrhs = e.nrhs;
        const bool chol = e.kind == AL_BATCH_POSV;
        double* A = A_pool + e.a_off;
        double* B = B_pool + e.b_off;
        if (nrhs < 0) {
            if (t == 0) {
                info[job] = -4;
                if (steps) steps[job] = 0;
            }
            continue;
        }
        int first_col = 0;
        int iters = 0;
        if (MIXED) {
            float* W = W_pool + e.a_off;
            for (size_t idx = t; idx < (size_t)n * n; idx += NT) W[idx] = (float)A[idx];
            __syncthreads();
            const int st = chol ? chol_factor<float, KB>(W, n, L) : lu_factor<float, KB>(W, n, L);
            __syncthreads();
            if (st == 0) {
                if (!chol) compute_perm(n, L);
                for (int k0 = 0; k0 < nrhs; k0 += RHS) {
                    const int nr = min(RHS, nrhs - k0);
                    double* Bc = B + (size_t)k0 * n;
                    for (int i = t; i < n; i += NT)
                        for (int k = 0; k < nr; ++k) L.u.s.y[i + k * NMAX] = Bc[i + (size_t)k * n];
                    __syncthreads();
                    factor_solve<float, NMAX>(W, n, nr, chol, L);
                    double x[Q][RHS];
#pragma unroll
                    for (int q = 0; q < Q; ++q)
#pragma unroll
                        for (int k = 0; k < RHS; ++k) {
                            const int i = t + q * NT;
                            x[q][k] = (i < n && k < nr) ? L.u.s.y[i + k * NMAX] : 0.0;
                        }
                    __syncthreads();
                    bool conv = false;
                    int it = 0;
                    for (; it <= max_iter; ++it) {
                        if (residual_small<NMAX>(A, Bc, n, nr, chol, x, L)) {
                            conv = true;
                            break;
                        }
                        if (it == max_iter) break;
                        factor_solve<float, NMAX>(W, n, nr, chol, L);
#pragma unroll
                        for (int q = 0; q < Q; ++q)
#pragma unroll
                            for (int k = 0; k < RHS; ++k) {
                                const int i = t + q * NT;
                                if (i < n && k < nr) x[q][k] += L.u.s.y[i + k * NMAX];
                            }
                        __syncthreads();
                    }
                    if (!conv) break;
                    iters = max(iters, it + 1);
#pragma unroll
                    for (int q = 0; q < Q; ++q)
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
        const int st = chol ? chol_factor<double, KB>(A, n, L) : lu_factor<double, KB>(A, n, L);
        __syncthreads();
        if (st == 0) {
            if (!chol) compute_perm(n, L);
            for (int k0 = first_col; k0 < nrhs; k0 += RHS) {
                const int nr = min(RHS, nrhs - k0);
                double* Bc = B + (size_t)k0 * n;
                for (int i = t; i < n; i += NT)
                    for (int k = 0; k < nr; ++k) L.u.s.y[i + k * NMAX] = Bc[i + (size_t)k * n];
                __syncthreads();
                factor_solve<double, NMAX>(A, n, nr, chol, L);
                for (int i = t; i < n; i += NT)
                    for (int k = 0; k < nr; ++k) Bc[i + (size_t)k * n] = L.u.s.y[i + k * NMAX];
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

// ------------------------------------------------------------------------------------------
// host side: the tuning table, the launches
// ------------------------------------------------------------------------------------------

struct BucketConfig {
    int kb;          // panel width (buckets 1-3): 16 or 32
    int wgs_per_cu;  // persistent workgroups per CU (the LDS caps what fits)
};

// Defaults, used until batch_tuning.json (or al_batch_tune_set) says otherwise.
BucketConfig g_cfg[NBUCKETS] = {{0, 3}, {16, 6}, {16, 3}, {16, 2}};
std::once_flag g_cfg_once;

// Reads batch_tuning.json from the library's directory: {"table": [{"bucket": .., "max_n": ..,
// "kb": .., "wgs_per_cu": ..}, ...], ...} in bucket order (the tuner's measurements follow the
// table).  A missing or malformed file keeps the defaults.
void load_tuning_table() {
    Dl_info dli;
    if (!dladdr((void*)&load_tuning_table, &dli) || !dli.dli_fname) return;
    std::string path(dli.dli_fname);
    const size_t slash = path.rfind('/');
    path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/batch_tuning.json";
    FILE* f = fopen(path.c_str(), "r");
    if (!f) return;
    std::string text;
    char buf[4096];
    size_t got;
    while ((got = fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, got);
    fclose(f);
    BucketConfig cfg[NBUCKETS];
    for (int b = 0; b < NBUCKETS; ++b) cfg[b] = g_cfg[b];
    size_t pos = text.find("\"table\"");
    if (pos == std::string::npos) return;
    for (int b = 0; b < NBUCKETS; ++b) {
        const size_t kp = text.find("\"kb\"", pos);
        const size_t wp = text.find("\"wgs_per_cu\"", pos);
        if (kp == std::string::npos || wp == std::string::npos) return;
        const int kb = atoi(text.c_str() + text.find(':', kp) + 1);
        const int w = atoi(text.c_str() + text.find(':', wp) + 1);
        if ((b > 0 && kb != 16 && kb != 32) || w < 1 || w > 16) return;
        cfg[b] = BucketConfig{b == 0 ? 0 : kb, w};
        pos = (kp > wp ? kp : wp) + 1;
    }
    for (int b = 0; b < NBUCKETS; ++b) g_cfg[b] = cfg[b];
}

template <typename K>
int set_lds(K kernel, size_t bytes) {
    AL_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return AL_OK;
}

// LDS of a workgroup kernel, and how many fit in a CU's 160 KiB
template <int NMAX, int KB>
constexpr size_t wg_lds() { return sizeof(BatchLds<NMAX, KB>); }

template <int NMAX, int KB>
int launch_wg(hipStream_t s, const al_batch_entry* d_entries, const int* lists, const int* counts, int bucket,
              size_t count, double* A_pool, double* B_pool, float* W_pool, int* d_info, int* d_steps, unsigned* queue,
              bool mixed, int max_iter, int wgs_per_cu) {
    static std::once_flag once;
    int rc = AL_OK;
    std::call_once(once, [&] {
        rc = set_lds(batch_k<NMAX, KB, true>, wg_lds<NMAX, KB>());
        if (rc == AL_OK) rc = set_lds(batch_k<NMAX, KB, false>, wg_lds<NMAX, KB>());
    });
    AL_CHECK(rc);
    const int fit = (int)((160 * 1024) / wg_lds<NMAX, KB>());
    const int per_cu = wgs_per_cu < fit ? wgs_per_cu : (fit > 0 ? fit : 1);
    const size_t g = (size_t)per_cu * (size_t)cu_count();
    const unsigned grid = (unsigned)(g < count ? g : count);
    const int* list = lists + (size_t)bucket * count;
    const size_t lds = wg_lds<NMAX, KB>();
    if (mixed)
        hipLaunchKernelGGL((batch_k<NMAX, KB, true>), dim3(grid), dim3(NT), lds, s, d_entries, list, counts, bucket,
                           A_pool, B_pool, W_pool, d_info, d_steps, queue, max_iter);
    else
        hipLaunchKernelGGL((batch_k<NMAX, KB, false>), dim3(grid), dim3(NT), lds, s, d_entries, list, counts, bucket,
                           A_pool, B_pool, W_pool, d_info, d_steps, queue, max_iter);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

int launch_bucket(int b, hipStream_t s, const al_batch_entry* d_entries, const int* lists, const int* counts,
                  size_t count, double* A_pool, double* B_pool, float* W_pool, int* d_info, int* d_steps,
                  unsigned* queue, bool mixed, int max_iter) {
    const BucketConfig c = g_cfg[b];
    if (b == 0) {
        static std::once_flag once;
        int rc = AL_OK;
        std::call_once(once, [&] {
            rc = set_lds(batch_small_k<true>, sizeof(SmallLds));
            if (rc == AL_OK) rc = set_lds(batch_small_k<false>, sizeof(SmallLds));
        });
        AL_CHECK(rc);
        const int fit = (int)((160 * 1024) / sizeof(SmallLds));
        const int per_cu = c.wgs_per_cu < fit ? c.wgs_per_cu : fit;
        const size_t g = (size_t)per_cu * (size_t)cu_count();
        const unsigned grid = (unsigned)(g < count ? g : count);
        if (mixed)
            hipLaunchKernelGGL(batch_small_k<true>, dim3(grid), dim3(64), sizeof(SmallLds), s, d_entries, lists, counts,
                               A_pool, B_pool, d_info, d_steps, queue, max_iter);
        else
            hipLaunchKernelGGL(batch_small_k<false>, dim3(grid), dim3(64), sizeof(SmallLds), s, d_entries, lists,
                               counts, A_pool, B_pool, d_info, d_steps, queue, max_iter);
        AL_HIP(hipGetLastError());
        return AL_OK;
    }
    const int w = c.wgs_per_cu;
    if (b == 1)
        return c.kb == 32 ? launch_wg<128, 32>(s, d_entries, lists, counts, 1, count, A_pool, B_pool, W_pool, d_info,
                                               d_steps, queue, mixed, max_iter, w)
                          : launch_wg<128, 16>(s, d_entries, lists, counts, 1, count, A_pool, B_pool, W_pool, d_info,
                                               d_steps, queue, mixed, max_iter, w);
    if (b == 2)
        return c.kb == 32 ? launch_wg<256, 32>(s, d_entries, lists, counts, 2, count, A_pool, B_pool, W_pool, d_info,
                                               d_steps, queue, mixed, max_iter, w)
                          : launch_wg<256, 16>(s, d_entries, lists, counts, 2, count, A_pool, B_pool, W_pool, d_info,
                                               d_steps, queue, mixed, max_iter, w);
    return launch_wg<512, 16>(s, d_entries, lists, counts, 3, count, A_pool, B_pool, W_pool, d_info, d_steps, queue,
                              mixed, max_iter, w);
}

}  // namespace

size_t batch_scratch_bytes(size_t count) { return 256 + (size_t)NBUCKETS * count * sizeof(int); }

namespace {

// The launches of one batched call.  work: batch_scratch_bytes(count) device bytes -- [0, 64)
// the buckets' counts, [64, 128) their queue counters, from 256 the bucket lists; the counts and
// counters are cleared here (by a kernel, so the sequence can be captured in a graph).  With a
// side stream (and its two events) the largest bucket runs on s and the others beside it;
// without one every launch is on s.
int batch_launch(hipStream_t s, hipStream_t side, hipEvent_t ev_fork, hipEvent_t ev_join, size_t count,
                 const al_batch_entry* d_entries, double* A_pool, double* B_pool, float* W_pool, int* d_info,
                 int* d_steps, void* work, bool mixed, int max_iter) {
    std::call_once(g_cfg_once, load_tuning_table);
    int* counts = (int*)work;
    unsigned* queues = (unsigned*)((char*)work + 64);
    int* lists = (int*)((char*)work + 256);
    AL_CHECK(fill_u32(s, (unsigned*)work, 64, 0u));
    hipLaunchKernelGGL(bucket_k, dim3(1), dim3(1024), 0, s, d_entries, (int)count, lists, counts);
    AL_HIP(hipGetLastError());
    if (side) {
        AL_HIP(hipEventRecord(ev_fork, s));
        AL_HIP(hipStreamWaitEvent(side, ev_fork, 0));
    }
    AL_CHECK(launch_bucket(3, s, d_entries, lists, counts, count, A_pool, B_pool, W_pool, d_info, d_steps,
                           queues + 3, mixed, max_iter));
    for (int b = 2; b >= 0; --b)
        AL_CHECK(launch_bucket(b, side ? side : s, d_entries, lists, counts, count, A_pool, B_pool, W_pool, d_info,
                               d_steps, queues + b, mixed, max_iter));
    if (side) {
        AL_HIP(hipEventRecord(ev_join, side));
        AL_HIP(hipStreamWaitEvent(s, ev_join, 0));
    }
    return AL_OK;
}

}  // namespace

int batch_solve(hipStream_t s, Workspace* ws, size_t count, const al_batch_entry* d_entries, double* A_pool,
                double* B_pool, float* W_pool, int* d_info, int* d_steps, void* work, bool mixed, int max_iter) {
    if (count == 0) return AL_OK;
    AL_CHECK(ensure_side_stream(ws));
    const int ticket = prof_begin(s);
    AL_CHECK(batch_launch(s, ws->side, ws->ev_main, ws->ev_side, count, d_entries, A_pool, B_pool, W_pool, d_info,
                          d_steps, work, mixed, max_iter));
    if (ticket >= 0) prof_end(ticket, s, mixed ? "batch_mixed" : "batch_fp64", 0.0, 0.0);
    return AL_OK;
}

// ------------------------------------------------------------------------------------------
// the FP64 batched solve that allocates nothing
// ------------------------------------------------------------------------------------------

namespace {

// Scratch of al_dev_batch_solve_noalloc: the buckets' counts and queue counters and the bucket
// lists of up to NOALLOC_MAX_SYSTEMS systems (1 MiB + 256 B), one copy per GPU, held until the
// process ends.  Host code never passes the array's name to a kernel: it is a host-side
// stand-in, not a device address; noalloc_scratch() resolves each device's copy once.
constexpr size_t NOALLOC_MAX_SYSTEMS = 65536;
constexpr size_t NOALLOC_SCRATCH_BYTES = 256 + (size_t)NBUCKETS * NOALLOC_MAX_SYSTEMS * sizeof(int);
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// The current device's copy of `scratch`, resolved with hipGetSymbolAddress on the first call on
// that device (with the device current) and kept.
int noalloc_scratch(void** out) {
    int dev = 0;
    AL_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        set_error("batch_solve_noalloc: device index beyond 63");
        return AL_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fail_hip(e == hipSuccess ? hipErrorInvalidSymbol : e,
                                                  "hipGetSymbolAddress(scratch)");
        g_noalloc_ptr[dev] = p;
    }
    *out = g_noalloc_ptr[dev];
    return AL_OK;
}

// The tuning table is read when the library is loaded, so the no-allocation entry point never
// opens a file.
struct LoadTableAtLoad {
    LoadTableAtLoad() { std::call_once(g_cfg_once, load_tuning_table); }
} g_load_table_at_load;

}  // namespace

int batch_solve_noalloc(hipStream_t s, size_t count, const al_batch_entry* d_entries, double* A_pool, double* B_pool,
                        int* d_info, int* d_steps) {
    if (count > NOALLOC_MAX_SYSTEMS) {
        set_error("batch_solve_noalloc: more than 65536 systems (the static scratch holds the bucket lists of 65536)");
        return AL_ERR_ARG;
    }
    if (count == 0) return AL_OK;
    void* work = nullptr;
    AL_CHECK(noalloc_scratch(&work));
    // every bucket on the caller's stream: no side stream, no events, no workspace
    return batch_launch(s, nullptr, nullptr, nullptr, count, d_entries, A_pool, B_pool, nullptr, d_info, d_steps, work,
                        false, 0);
}

int batch_tune_set(int bucket, int kb, int wgs_per_cu) {
    std::call_once(g_cfg_once, load_tuning_table);
    if (bucket < 0 || bucket >= NBUCKETS || wgs_per_cu < 1 || wgs_per_cu > 16 ||
        (bucket > 0 && kb != 16 && kb != 32) || (bucket == 3 && kb != 16)) {
        set_error("batch_tune_set: bucket 0-3, kb 16 or 32 (bucket 3: 16), 1-16 workgroups per CU");
        return AL_ERR_ARG;
    }
    g_cfg[bucket] = BucketConfig{bucket == 0 ? 0 : kb, wgs_per_cu};
    return AL_OK;
}

int batch_tune_get(int bucket, int* kb, int* wgs_per_cu) {
    std::call_once(g_cfg_once, load_tuning_table);
    if (bucket < 0 || bucket >= NBUCKETS) {
        set_error("batch_tune_get: bucket 0-3");
        return AL_ERR_ARG;
    }
    if (kb) *kb = g_cfg[bucket].kb;
    if (wgs_per_cu) *wgs_per_cu = g_cfg[bucket].wgs_per_cu;
    return AL_OK;
}

}  // namespace al
