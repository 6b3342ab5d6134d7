// Triangular solves with few right-hand sides in one launch (the substitution sweeps of
// lapack::dgetrs, src/lapack/mod.rs:340-353, for nrhs <= 16).
//
// A blocked substitution launched per diagonal block costs two launches per 64 rows; at
// n = 16384 that is 512 launches whose diagonal solves are each a long dependency chain.
// Here one workgroup owns one 64-row block.  It accumulates y = b - sum_c A[b,c] x_c over the
// column blocks c in solve order (matrix reads coalesced down columns, the next block's
// requested while the current one is summed), reduces the four partial sums in a fixed order, solves its diagonal block
// with one wave (lane = row; x_i broadcast by readlane), writes x_b over b and raises the
// block's flag with a release store.  Workgroups are dispatched in increasing id and a block
// waits only on blocks of smaller id, so the smallest unfinished block is always resident:
// the chain always progresses.  Every spin is bounded; a timeout raises the abort flag.
//
// The matrix may be FP64 or FP32 (the low-precision factors of the mixed-precision solver);
// right-hand sides and all accumulation are FP64.
#include <algorithm>
#include <type_traits>
#include <utility>

#include "al_internal.hpp"

namespace al {
namespace {

constexpr int TB = 64;      // block order
constexpr int NRHS_MAX = 16;

// Value of v in lane `src` (wave-uniform) for every lane, through scalar registers.
__device__ __forceinline__ double bcast_lane(double v, int src) {
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(bits & 0xffffffffLL), src);
    const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// NR: right-hand sides the registers are sized for (nrhs <= NR); every loop over them is
// unrolled, so no accumulator is indexed dynamically.
template <typename TA, bool LOWER, bool UNIT, int NR>
__global__ __launch_bounds__(256) void trsv_blocked(int n, int nrhs, const TA* __restrict__ A, size_t lda,
                                                    double* __restrict__ X, size_t ldx, unsigned* __restrict__ flags,
                                                    unsigned gen, unsigned* __restrict__ abort_flag,
                                                    long long timeout) {
    __shared__ double red[4][TB][NR];
    __shared__ double D[TB][TB + 1];
    __shared__ int ok_lds;
    const int nblk = (n + TB - 1) / TB;
    const int b = LOWER ? (int)blockIdx.x : nblk - 1 - (int)blockIdx.x;
    const int r0 = b * TB;
    const int nb = (n - r0) < TB ? (n - r0) : TB;
    const int t = threadIdx.x, row = t & 63, part = t >> 6;

    // stage the diagonal block's triangle
    for (int idx = t; idx < TB * TB; idx += 256) {
        const int i = idx & 63, k = idx >> 6;
        const bool tri = LOWER ? (k <= i) : (k >= i);
        D[i][k] = (i < nb && k < nb && tri && !(UNIT && i == k)) ? (double)A[(size_t)(r0 + i) + (size_t)(r0 + k) * lda]
                                                                  : 0.0;
    }

    double acc[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) acc[k] = 0.0;

    // column blocks this block depends on: [0, b) for lower, (b, nblk) for upper, visited in
    // the order they are solved.  The matrix part of the next block is loaded into registers
    // before waiting for its flag, so only the x_c reads remain after the wait.
    const int ndep = LOWER ? b : nblk - 1 - b;
    const bool rok = (r0 + row) < n;
    const int row_c = rok ? r0 + row : 0;  // rows past n read row 0 and are replaced by zeros
    auto load_a = [&](int q, double (&av)[16]) {
        const int c0 = (LOWER ? q : nblk - 1 - q) * TB + part * 16;
        if (c0 + 16 <= n) {
            const TA* p = A + (size_t)row_c + (size_t)c0 * lda;
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) {
                const double v = (double)*p;
                av[jj] = rok ? v : 0.0;
                p += lda;
            }
        } else {
#pragma unroll
            for (int jj = 0; jj < 16; ++jj) {
                const bool okc = c0 + jj < n;
                const double v = (double)A[(size_t)row_c + (okc ? (size_t)(c0 + jj) : 0) * lda];
                av[jj] = (rok && okc) ? v : 0.0;
            }
        }
    };
    double a_next[16];
    if (ndep > 0) load_a(0, a_next);
    bool ok = true;
    for (int q = 0; q < ndep; ++q) {
        double a_cur[16];
#pragma unroll
        for (int jj = 0; jj < 16; ++jj) a_cur[jj] = a_next[jj];
        const int c = LOWER ? q : nblk - 1 - q;
        if (t == 0) {
            int good = 1;
            const long long t0 = wall_clock64();
            while (__hip_atomic_load(&flags[c * 16], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != gen) {
                if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
                    good = 0;
                    break;
                }
                if (wall_clock64() - t0 > timeout) {
                    __hip_atomic_store(abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    good = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (good) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            ok_lds = good;
        }
        __syncthreads();
        ok = ok_lds != 0;
        __syncthreads();
        if (!ok) break;
        // the next block's matrix part is requested only now: the acquire fence above waits for
        // every load in flight, so a prefetch issued before it would serialize with the wait
        if (q + 1 < ndep) load_a(q + 1, a_next);
        const int c0 = c * TB + part * 16;
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            if (k < nrhs) {
                double loc = 0.0;  // this block's partial sum, then one add into the running total
#pragma unroll
                for (int jj = 0; jj < 16; ++jj)
                    loc += a_cur[jj] * ((c0 + jj < n) ? X[(size_t)(c0 + jj) + (size_t)k * ldx] : 0.0);
                acc[k] += loc;
            }
        }
    }
    if (!ok) {
        // the chain was aborted: poison this block's rows so the partial solve cannot pass for a
        // result (the host entry points also read the raised abort flag and report an error)
        for (int idx = t; idx < nb * nrhs; idx += 256)
            X[(size_t)(r0 + idx % nb) + (size_t)(idx / nb) * ldx] = __builtin_nan("");
        return;
    }

#pragma unroll
    for (int k = 0; k < NR; ++k) red[part][row][k] = acc[k];
    __syncthreads();
    if (part == 0) {
        // y = b - (p0 + p1 + p2 + p3), in this fixed order
        double y[NR];
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const double sum = ((red[0][row][k] + red[1][row][k]) + red[2][row][k]) + red[3][row][k];
            y[k] = (k < nrhs && row < nb) ? X[(size_t)(r0 + row) + (size_t)k * ldx] - sum : 0.0;
        }
        // diagonal block, one lane per row, right-looking: lane i divides, x_i is broadcast with
        // readlane (no LDS round trip), every lane beyond i subtracts its term
        for (int s = 0; s < nb; ++s) {
            const int i = LOWER ? s : nb - 1 - s;
            const double lik = D[row][i];
            const bool below = LOWER ? (row > i) : (row < i);
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                if (k < nrhs) {
                    if (!UNIT && row == i) y[k] = y[k] / D[i][i];
                    const double xi = bcast_lane(y[k], i);
                    if (below) y[k] -= lik * xi;
                }
            }
        }
        if (row < nb)
#pragma unroll
            for (int k = 0; k < NR; ++k)
                if (k < nrhs) X[(size_t)(r0 + row) + (size_t)k * ldx] = y[k];
        if (row == 0) __hip_atomic_store(&flags[b * 16], gen, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// FP32 factors (the correction solves of the mixed-precision solver) take 128-row blocks: half
// as many workgroup-to-workgroup hops on the critical path as 64-row blocks, each hop costing a
// flag round trip through the memory system.  A workgroup thread holds rows r and r + 64 of
// its block for 32 of the 128 columns of each block it depends on (FP64 sums of FP32 entries
// times FP64 x, x_c read once per lane and broadcast by readlane).  The diagonal block is
// solved by one wave, lane l owning rows l and l + 64: the half solved first also updates the
// other half in the same steps, and a nonunit diagonal multiplies by its FP64 reciprocal.
constexpr int TB2 = 128;
constexpr int NR2_MAX = 8;

template <int V>
using ic = std::integral_constant<int, V>;

constexpr int DINV_ROW = TB2 + 2;  // doubles per row of a staged diagonal-block inverse
constexpr int NR_INV_MAX = 4;      // right-hand sides of the inverse-based solve (LDS bound)

template <bool INV, int NR>
constexpr size_t trsv128_lds_bytes() {
    return INV ? (size_t)TB2 * DINV_ROW * sizeof(double) + (size_t)4 * TB2 * NR * sizeof(double) +
                     (size_t)TB2 * NR * sizeof(double) + 16
               : (size_t)TB2 * (TB2 + 1) * sizeof(float) + (size_t)4 * TB2 * NR * sizeof(double) +
                     TB2 * sizeof(double) + 16;
}

// INV: the block's diagonal triangle is applied as its precomputed FP64 inverse Dinv (one
// 128 x 128 column-major block per row block), staged into LDS while the block waits; the
// diagonal step of the critical path is then a matrix-vector product over all four waves
// instead of a 128-step substitution.
template <bool LOWER, bool UNIT, int NR, bool INV>
__global__ __launch_bounds__(256) void trsv128_f32(int n, int nrhs, const float* __restrict__ A, size_t lda,
                                                   double* __restrict__ X, size_t ldx, unsigned* __restrict__ flags,
                                                   unsigned gen, unsigned* __restrict__ abort_flag, long long timeout,
                                                   const double* __restrict__ Dinv) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    double* red = (double*)lds_raw;                       // [4][TB2][NR]
    double* rdiag = red + 4 * TB2 * NR;                   // [TB2] (substitution)
    double* ys = rdiag;                                   // [TB2][NR] (inverse)
    float* D = (float*)(rdiag + TB2);                     // [TB2][TB2 + 1], strictly triangular (substitution)
    double* Di = ys + TB2 * NR;                           // [TB2][DINV_ROW], row-major (inverse)
    int* ok_lds = INV ? (int*)(Di + TB2 * DINV_ROW) : (int*)(D + TB2 * (TB2 + 1));
    const int nblk = (n + TB2 - 1) / TB2;
    // a launch may carry several independent chains, one per group of NR right-hand sides:
    // workgroup ids interleave them block by block, so a block still waits only on lower ids
    const int nch = (int)gridDim.x / nblk, chunk = (int)blockIdx.x % nch, bidx = (int)blockIdx.x / nch;
    X += (size_t)chunk * NR * ldx;
    nrhs = min(NR, nrhs - chunk * NR);
    flags += (size_t)chunk * nblk * 16;
    const int b = LOWER ? bidx : nblk - 1 - bidx;
    const int r0 = b * TB2;
    const int nb = (n - r0) < TB2 ? (n - r0) : TB2;
    const int t = threadIdx.x, lane = t & 63, part = t >> 6;

    if (INV) {
        const double* src = Dinv + (size_t)b * TB2 * TB2;  // column-major: src[k * 128 + i] = Dinv[i][k]
        for (int idx = t; idx < TB2 * TB2; idx += 256) {
            const int i = idx & (TB2 - 1), k = idx >> 7;
            Di[i * DINV_ROW + k] = src[idx];
        }
    } else {
        for (int idx = t; idx < TB2 * TB2; idx += 256) {
            const int i = idx & (TB2 - 1), k = idx >> 7;
            const bool tri = LOWER ? (k < i) : (k > i);
            D[i * (TB2 + 1) + k] = (i < nb && k < nb && tri) ? A[(size_t)(r0 + i) + (size_t)(r0 + k) * lda] : 0.f;
        }
        if (!UNIT && t < TB2) rdiag[t] = t < nb ? 1.0 / (double)A[(size_t)(r0 + t) + (size_t)(r0 + t) * lda] : 0.0;
    }

    double acc0[NR], acc1[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) acc0[k] = acc1[k] = 0.0;

    const int ndep = LOWER ? b : nblk - 1 - b;
    const int ra = r0 + lane, rb = r0 + 64 + lane;
    const bool oka = ra < n, okb = rb < n;
    // this block's right-hand side, read now rather than on the critical path
    double ba[NR], bb[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        ba[k] = (part == 0 && k < nrhs && oka) ? X[(size_t)ra + (size_t)k * ldx] : 0.0;
        bb[k] = (part == 0 && k < nrhs && okb) ? X[(size_t)rb + (size_t)k * ldx] : 0.0;
    }
    // rows past n read row 0 and columns past n (only in an upper solve's partial last block)
    // column 0, and are replaced by zeros: the loads are branch-free and walk the columns with a
    // running pointer
    const int ra_c = oka ? ra : 0, rb_c = okb ? rb : 0;
    auto load_a = [&](int q, float (&av)[2][32]) {
        const int c0 = (LOWER ? q : nblk - 1 - q) * TB2 + part * 32;
        if (c0 + 32 <= n) {
            const float* p0 = A + (size_t)ra_c + (size_t)c0 * lda;
            const float* p1 = A + (size_t)rb_c + (size_t)c0 * lda;
#pragma unroll
            for (int jj = 0; jj < 32; ++jj) {
                const float v0 = *p0, v1 = *p1;
                av[0][jj] = oka ? v0 : 0.f;
                av[1][jj] = okb ? v1 : 0.f;
                p0 += lda;
                p1 += lda;
            }
        } else {
#pragma unroll
            for (int jj = 0; jj < 32; ++jj) {
                const bool okc = c0 + jj < n;
                const size_t col = okc ? (size_t)(c0 + jj) : 0;
                const float v0 = A[(size_t)ra_c + col * lda], v1 = A[(size_t)rb_c + col * lda];
                av[0][jj] = (oka && okc) ? v0 : 0.f;
                av[1][jj] = (okb && okc) ? v1 : 0.f;
            }
        }
    };
    float a_next[2][32];
    if (ndep > 0) load_a(0, a_next);
    bool ok = true;
    for (int q = 0; q < ndep; ++q) {
        float a_cur[2][32];
#pragma unroll
        for (int jj = 0; jj < 32; ++jj) {
            a_cur[0][jj] = a_next[0][jj];
            a_cur[1][jj] = a_next[1][jj];
        }
        const int c = LOWER ? q : nblk - 1 - q;
        if (t == 0) {
            int good = 1;
            const long long t0 = wall_clock64();
            while (__hip_atomic_load(&flags[c * 16], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != gen) {
                if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
                    good = 0;
                    break;
                }
                if (wall_clock64() - t0 > timeout) {
                    __hip_atomic_store(abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    good = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (good) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            *ok_lds = good;
        }
        __syncthreads();
        ok = *ok_lds != 0;
        __syncthreads();
        if (!ok) break;
        if (q + 1 < ndep) load_a(q + 1, a_next);
        const int c0 = c * TB2 + part * 32;
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            if (k < nrhs) {
                const double xv = (lane < 32 && c0 + lane < n) ? X[(size_t)(c0 + lane) + (size_t)k * ldx] : 0.0;
                double l0 = 0.0, l1 = 0.0;
#pragma unroll
                for (int jj = 0; jj < 32; ++jj) {
                    const double xj = bcast_lane(xv, jj);
                    l0 += (double)a_cur[0][jj] * xj;
                    l1 += (double)a_cur[1][jj] * xj;
                }
                acc0[k] += l0;
                acc1[k] += l1;
            }
        }
    }
    if (!ok) {
        for (int idx = t; idx < nb * nrhs; idx += 256)
            X[(size_t)(r0 + idx % nb) + (size_t)(idx / nb) * ldx] = __builtin_nan("");
        return;
    }

#pragma unroll
    for (int k = 0; k < NR; ++k) {
        red[(part * TB2 + lane) * NR + k] = acc0[k];
        red[(part * TB2 + 64 + lane) * NR + k] = acc1[k];
    }
    __syncthreads();
    if (!INV && part != 0) return;
    double ya[NR], yb[NR];
    if (part == 0) {
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const double sa = ((red[lane * NR + k] + red[(TB2 + lane) * NR + k]) + red[(2 * TB2 + lane) * NR + k]) +
                              red[(3 * TB2 + lane) * NR + k];
            const double sb = ((red[(64 + lane) * NR + k] + red[(TB2 + 64 + lane) * NR + k]) +
                               red[(2 * TB2 + 64 + lane) * NR + k]) +
                              red[(3 * TB2 + 64 + lane) * NR + k];
            ya[k] = (k < nrhs && oka) ? ba[k] - sa : 0.0;
            yb[k] = (k < nrhs && okb) ? bb[k] - sb : 0.0;
        }
    }
    if (INV) {
        // x = Dinv y: wave 0 publishes y, every thread sums its two rows over its 32 columns,
        // wave 0 adds the four partial sums in a fixed order
        if (part == 0) {
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                ys[lane * NR + k] = ya[k];
                ys[(64 + lane) * NR + k] = yb[k];
            }
        }
        __syncthreads();
        double pa[NR], pb[NR];
#pragma unroll
        for (int k = 0; k < NR; ++k) pa[k] = pb[k] = 0.0;
        const double* Ra = Di + lane * DINV_ROW + 32 * part;
        const double* Rb = Di + (64 + lane) * DINV_ROW + 32 * part;
#pragma unroll
        for (int kk = 0; kk < 32; ++kk) {
            const double da = Ra[kk], db = Rb[kk];
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                const double yv = ys[(32 * part + kk) * NR + k];
                pa[k] = fma(da, yv, pa[k]);
                pb[k] = fma(db, yv, pb[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            red[(part * TB2 + lane) * NR + k] = pa[k];
            red[(part * TB2 + 64 + lane) * NR + k] = pb[k];
        }
        __syncthreads();
        if (part != 0) return;
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            ya[k] = ((red[lane * NR + k] + red[(TB2 + lane) * NR + k]) + red[(2 * TB2 + lane) * NR + k]) +
                    red[(3 * TB2 + lane) * NR + k];
            yb[k] = ((red[(64 + lane) * NR + k] + red[(TB2 + 64 + lane) * NR + k]) +
                     red[(2 * TB2 + 64 + lane) * NR + k]) +
                    red[(3 * TB2 + 64 + lane) * NR + k];
        }
    } else {
    // The diagonal block in groups of 8 rows, in solve order.  For a group every lane reads the
    // 8 current values by readlane and solves the 8 x 8 triangle itself (a dependency chain of 8
    // short steps instead of 8 lane-to-lane broadcasts), then updates the rows it owns that come
    // later in the solve.  Rows past nb hold y = 0, zero D entries and a zero reciprocal, so
    // running every group of both halves changes nothing for them.  The D reads do not depend on
    // y, so they are issued ahead of the chain.
    const float* Da = D + lane * (TB2 + 1);
    const float* Db = D + (64 + lane) * (TB2 + 1);
    // one group of the half whose rows start at H (its lanes' values in y), later rows of the
    // same half in y, every row of the other half in yo (when OTHER)
    auto group = [&](auto Hc, int G, auto OTHERc, double (&y)[NR], double (&yo)[NR]) {
        constexpr int H = decltype(Hc)::value;
        constexpr bool OTHER = decltype(OTHERc)::value;
        const int R0 = H + 8 * G;  // first row of the group
        float dg[8][8];
#pragma unroll
        for (int r = 0; r < 8; ++r)
#pragma unroll
            for (int c2 = 0; c2 < 8; ++c2) dg[r][c2] = ((LOWER && c2 < r) || (!LOWER && c2 > r)) ? D[(R0 + r) * (TB2 + 1) + R0 + c2] : 0.f;
        double rd[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) rd[r] = UNIT ? 1.0 : rdiag[R0 + r];
        const float* Dy = H == 0 ? Da : Db;   // this lane's row in the group's half
        const float* Do = H == 0 ? Db : Da;   // .. and in the other half
        float dy[8], dother[8];
#pragma unroll
        for (int c2 = 0; c2 < 8; ++c2) {
            dy[c2] = Dy[R0 + c2];
            dother[c2] = OTHER ? Do[R0 + c2] : 0.f;
        }
        const int gl = 8 * G;  // the group's first lane
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            if (k < nrhs) {
                double xs[8];
#pragma unroll
                for (int r = 0; r < 8; ++r) xs[r] = bcast_lane(y[k], gl + r);
                if (LOWER) {
#pragma unroll
                    for (int r = 0; r < 8; ++r) {
#pragma unroll
                        for (int c2 = 0; c2 < r; ++c2) xs[r] = fma(-(double)dg[r][c2], xs[c2], xs[r]);
                        if (!UNIT) xs[r] *= rd[r];
                    }
                } else {
#pragma unroll
                    for (int r = 7; r >= 0; --r) {
#pragma unroll
                        for (int c2 = r + 1; c2 < 8; ++c2) xs[r] = fma(-(double)dg[r][c2], xs[c2], xs[r]);
                        if (!UNIT) xs[r] *= rd[r];
                    }
                }
                double u = y[k], uo = yo[k];
#pragma unroll
                for (int c2 = 0; c2 < 8; ++c2) {
                    u = fma(-(double)dy[c2], xs[c2], u);
                    if (OTHER) uo = fma(-(double)dother[c2], xs[c2], uo);
                }
                const bool later = LOWER ? lane >= gl + 8 : lane < gl;
                double v = later ? u : y[k];
#pragma unroll
                for (int r = 0; r < 8; ++r) v = lane == gl + r ? xs[r] : v;
                y[k] = v;
                if (OTHER) yo[k] = uo;
            }
        }
    };
    // groups in solve order: 0..7 for a lower solve, 7..0 for an upper one
    if (LOWER) {
#pragma unroll
        for (int g = 0; g < 8; ++g) group(ic<0>{}, g, std::true_type{}, ya, yb);
#pragma unroll
        for (int g = 0; g < 8; ++g) group(ic<64>{}, g, std::false_type{}, yb, ya);
    } else {
#pragma unroll
        for (int g = 7; g >= 0; --g) group(ic<64>{}, g, std::true_type{}, yb, ya);
#pragma unroll
        for (int g = 7; g >= 0; --g) group(ic<0>{}, g, std::false_type{}, ya, yb);
    }
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        if (k < nrhs) {
            if (oka) X[(size_t)ra + (size_t)k * ldx] = ya[k];
            if (okb) X[(size_t)rb + (size_t)k * ldx] = yb[k];
        }
    }
    if (lane == 0) __hip_atomic_store(&flags[b * 16], gen, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Inverses of the 128 x 128 diagonal blocks of an LU factor held in one FP32 matrix: Linv of
// the unit lower triangle (blockIdx.y == 0) and Uinv of the upper triangle (blockIdx.y == 1),
// in FP64 from the FP32 entries, column-major 128 x 128 per block (identity / zero past the
// matrix edge).  Column j of a block's inverse is solved by thread j (rows 0..63) and thread
// 128 + j (rows 64..127): the half solved first publishes its values through LDS, the other
// half subtracts the off-diagonal block times them and solves its own triangle.
constexpr int DI_ROW = TB2 + 4;   // floats per staged column of the block
constexpr int DI_ZROW = 66;       // doubles per published half column
constexpr size_t DIAG_INV_LDS = (size_t)TB2 * DI_ROW * sizeof(float) + (size_t)TB2 * DI_ZROW * sizeof(double) +
                                TB2 * sizeof(double);

__global__ __launch_bounds__(256) void diag_inv_k(int n, const float* __restrict__ A, size_t lda,
                                                  double* __restrict__ Linv, double* __restrict__ Uinv) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    float* S = (float*)lds_raw;                 // S[k * DI_ROW + i] = A[r0 + i][r0 + k]
    double* Z = (double*)(S + TB2 * DI_ROW);    // Z[j * DI_ZROW + i]: the first-solved half of column j
    double* rd = Z + TB2 * DI_ZROW;             // reciprocals of U's diagonal
    const int b = blockIdx.x;
    const bool upper = blockIdx.y == 1;
    const int r0 = b * TB2;
    const int nb = (n - r0) < TB2 ? (n - r0) : TB2;
    const int t = threadIdx.x;
    for (int idx = t; idx < TB2 * TB2; idx += 256) {
        const int i = idx & (TB2 - 1), k = idx >> 7;
        S[k * DI_ROW + i] = (i < nb && k < nb) ? A[(size_t)(r0 + i) + (size_t)(r0 + k) * lda] : 0.f;
    }
    if (t < TB2) rd[t] = (upper && t < nb) ? 1.0 / (double)A[(size_t)(r0 + t) + (size_t)(r0 + t) * lda] : 0.0;
    __syncthreads();
    const int j = t & (TB2 - 1);
    const int h = __builtin_amdgcn_readfirstlane(t >> 7);  // which half of the column, uniform per wave
    double z[64];
#pragma unroll
    for (int i = 0; i < 64; ++i) z[i] = (64 * h + i == j) ? 1.0 : 0.0;
    if (!upper) {
        if (h == 0) {  // rows 0..63: L11 z = e_j
#pragma unroll
            for (int k = 0; k < 64; ++k) {
                const double zk = z[k];
#pragma unroll
                for (int i = k + 1; i < 64; ++i) z[i] = fma(-(double)S[k * DI_ROW + i], zk, z[i]);
            }
#pragma unroll
            for (int i = 0; i < 64; ++i) Z[j * DI_ZROW + i] = z[i];
        }
        __syncthreads();
        if (h == 1) {  // rows 64..127: L22 z = e_j - L21 z1
#pragma unroll
            for (int k = 0; k < 64; ++k) {
                const double zk = Z[j * DI_ZROW + k];
#pragma unroll
                for (int i = 0; i < 64; ++i) z[i] = fma(-(double)S[k * DI_ROW + 64 + i], zk, z[i]);
            }
#pragma unroll
            for (int k = 0; k < 64; ++k) {
                const double zk = z[k];
#pragma unroll
                for (int i = k + 1; i < 64; ++i) z[i] = fma(-(double)S[(64 + k) * DI_ROW + 64 + i], zk, z[i]);
            }
        }
    } else {
        if (h == 1) {  // rows 64..127: U22 z = e_j, backward
#pragma unroll
            for (int k = 63; k >= 0; --k) {
                z[k] *= rd[64 + k];
                const double zk = z[k];
#pragma unroll
                for (int i = 0; i < k; ++i) z[i] = fma(-(double)S[(64 + k) * DI_ROW + 64 + i], zk, z[i]);
            }
#pragma unroll
            for (int i = 0; i < 64; ++i) Z[j * DI_ZROW + i] = z[i];
        }
        __syncthreads();
        if (h == 0) {  // rows 0..63: U11 z = e_j - U12 z2
#pragma unroll
            for (int k = 0; k < 64; ++k) {
                const double zk = Z[j * DI_ZROW + k];
#pragma unroll
                for (int i = 0; i < 64; ++i) z[i] = fma(-(double)S[(64 + k) * DI_ROW + i], zk, z[i]);
            }
#pragma unroll
            for (int k = 63; k >= 0; --k) {
                z[k] *= rd[k];
                const double zk = z[k];
#pragma unroll
                for (int i = 0; i < k; ++i) z[i] = fma(-(double)S[k * DI_ROW + i], zk, z[i]);
            }
        }
    }
    double* out = (upper ? Uinv : Linv) + (size_t)b * TB2 * TB2 + (size_t)j * TB2 + 64 * h;
#pragma unroll
    for (int i = 0; i < 64; ++i) out[i] = z[i];
}

template <bool LOWER, bool UNIT, int NR, bool INV>
int launch128_t(hipStream_t s, unsigned nblk, int n, int nrhs, const float* A, size_t lda, double* X, size_t ldx,
                unsigned* flags, unsigned gen, unsigned* abort_flag, long long timeout, const double* Dinv,
                unsigned nch = 1) {
    static bool attr_set = false;  // benign race: every thread sets the same value
    auto fn = trsv128_f32<LOWER, UNIT, NR, INV>;
    constexpr size_t bytes = trsv128_lds_bytes<INV, NR>();
    if (!attr_set) {
        AL_HIP(hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
        attr_set = true;
    }
    hipLaunchKernelGGL(fn, dim3(nblk * nch), dim3(256), bytes, s, n, nrhs, A, lda, X, ldx, flags, gen, abort_flag,
                       timeout, Dinv);
    return AL_OK;
}

template <bool LOWER, bool UNIT, int NR>
int launch128(hipStream_t s, unsigned nblk, int n, int nrhs, const float* A, size_t lda, double* X, size_t ldx,
              unsigned* flags, unsigned gen, unsigned* abort_flag, long long timeout, const double* Dinv) {
    if constexpr (NR <= NR_INV_MAX) {
        if (Dinv)
            return launch128_t<LOWER, UNIT, NR, true>(s, nblk, n, nrhs, A, lda, X, ldx, flags, gen, abort_flag, timeout,
                                                      Dinv);
    }
    return launch128_t<LOWER, UNIT, NR, false>(s, nblk, n, nrhs, A, lda, X, ldx, flags, gen, abort_flag, timeout,
                                               nullptr);
}

template <typename TA, bool LOWER, bool UNIT, int NR>
void launch_nr(hipStream_t s, unsigned nblk, int n, int nrhs, const TA* A, size_t lda, double* X, size_t ldx,
               unsigned* flags, unsigned gen, unsigned* abort_flag, long long timeout) {
    hipLaunchKernelGGL((trsv_blocked<TA, LOWER, UNIT, NR>), dim3(nblk), dim3(256), 0, s, n, nrhs, A, lda, X, ldx, flags,
                       gen, abort_flag, timeout);
}

// nch > 1 (FP32 with Dinv only): nch chains of NR_INV_MAX right-hand sides in one launch.
template <typename TA, bool LOWER, bool UNIT>
int launch_trsv(hipStream_t s, Workspace* ws, size_t n, size_t nrhs, const TA* A, size_t lda, double* X, size_t ldx,
                const double* Dinv, unsigned nch = 1) {
    const size_t nblk = (n + TB - 1) / TB;
    AL_CHECK(ws_reserve_flags(ws, nblk * 16 * nch));
    if (ws->trsv_gen > 0xF0000000u) {  // keep generations unique: restart from a cleared array
        AL_CHECK(fill_u32(s, ws->trsv_flags, ws->trsv_flags_len, 0u));
        ws->trsv_gen = 0;
    }
    const unsigned gen = ++ws->trsv_gen;
    const long long timeout = wait_limit_ticks();
    const int ticket = prof_begin(s);
    const unsigned g = (unsigned)nblk;
    if (nch > 1) {
        const unsigned g2 = (unsigned)((n + TB2 - 1) / TB2);
        AL_CHECK((launch128_t<LOWER, UNIT, NR_INV_MAX, true>(s, g2, (int)n, (int)nrhs, (const float*)A, lda, X, ldx,
                                                             ws->trsv_flags, gen, ws->abort_flag, timeout, Dinv,
                                                             nch)));
    } else if (std::is_same<TA, float>::value && nrhs <= (size_t)NR2_MAX) {
        const unsigned g2 = (unsigned)((n + TB2 - 1) / TB2);
        const float* Af = (const float*)A;
        if (nrhs <= 1)
            AL_CHECK((launch128<LOWER, UNIT, 1>(s, g2, (int)n, (int)nrhs, Af, lda, X, ldx, ws->trsv_flags, gen,
                                                ws->abort_flag, timeout, Dinv)));
        else if (nrhs <= 2)
            AL_CHECK((launch128<LOWER, UNIT, 2>(s, g2, (int)n, (int)nrhs, Af, lda, X, ldx, ws->trsv_flags, gen,
                                                ws->abort_flag, timeout, Dinv)));
        else if (nrhs <= 4)
            AL_CHECK((launch128<LOWER, UNIT, 4>(s, g2, (int)n, (int)nrhs, Af, lda, X, ldx, ws->trsv_flags, gen,
                                                ws->abort_flag, timeout, Dinv)));
        else
            AL_CHECK((launch128<LOWER, UNIT, 8>(s, g2, (int)n, (int)nrhs, Af, lda, X, ldx, ws->trsv_flags, gen,
                                                ws->abort_flag, timeout, Dinv)));
    } else if (nrhs <= 1)
        launch_nr<TA, LOWER, UNIT, 1>(s, g, (int)n, (int)nrhs, A, lda, X, ldx, ws->trsv_flags, gen, ws->abort_flag, timeout);
    else if (nrhs <= 2)
        launch_nr<TA, LOWER, UNIT, 2>(s, g, (int)n, (int)nrhs, A, lda, X, ldx, ws->trsv_flags, gen, ws->abort_flag, timeout);
    else if (nrhs <= 4)
        launch_nr<TA, LOWER, UNIT, 4>(s, g, (int)n, (int)nrhs, A, lda, X, ldx, ws->trsv_flags, gen, ws->abort_flag, timeout);
    else if (nrhs <= 8)
        launch_nr<TA, LOWER, UNIT, 8>(s, g, (int)n, (int)nrhs, A, lda, X, ldx, ws->trsv_flags, gen, ws->abort_flag, timeout);
    else
        launch_nr<TA, LOWER, UNIT, 16>(s, g, (int)n, (int)nrhs, A, lda, X, ldx, ws->trsv_flags, gen, ws->abort_flag, timeout);
    AL_HIP(hipGetLastError());
    if (ticket >= 0) {
        const double dn = (double)n;
        prof_end(ticket, s, "trsv", dn * dn * (double)nrhs, 0.5 * dn * dn * sizeof(TA) + 16.0 * dn * (double)nrhs);
    }
    return AL_OK;
}

}  // namespace

bool trsv_fits(size_t nrhs) { return nrhs >= 1 && nrhs <= (size_t)NRHS_MAX; }

template <typename TA>
int trsv_solve(hipStream_t s, Workspace* ws, bool lower, bool unit, size_t n, size_t nrhs, const TA* A, size_t lda,
               double* X, size_t ldx, const double* Dinv) {
    if (n == 0 || nrhs == 0) return AL_OK;
    if (!trsv_fits(nrhs) || n > INT32_MAX) {
        set_error("trsv: more right-hand sides than the one-launch solve takes");
        return AL_ERR_ARG;
    }
    if (Dinv && std::is_same<TA, float>::value && nrhs > (size_t)NR_INV_MAX) {
        // the inverse-based solve holds at most NR_INV_MAX right-hand sides per chain; each chain
        // is bound by its hop latency, so several run side by side in one launch: up to twice as
        // many workgroups as CUs (one fits per CU, for its LDS; the later ones are dispatched as
        // the first blocks finish, which still beats a second launch: 82.5 against 88.4 ms for 16
        // right-hand sides at n = 16384)
        const size_t nblk2 = (n + TB2 - 1) / TB2;
        const size_t fit = std::max<size_t>(1, (size_t)(2 * cu_count()) / nblk2);
        for (size_t c0 = 0; c0 < nrhs; c0 += NR_INV_MAX * fit) {
            const size_t nr = std::min(nrhs - c0, (size_t)NR_INV_MAX * fit);
            const unsigned nch = (unsigned)((nr + NR_INV_MAX - 1) / NR_INV_MAX);
            if (nch == 1) {
                AL_CHECK(trsv_solve<TA>(s, ws, lower, unit, n, nr, A, lda, X + c0 * ldx, ldx, Dinv));
                continue;
            }
            if (lower)
                AL_CHECK((unit ? launch_trsv<TA, true, true>(s, ws, n, nr, A, lda, X + c0 * ldx, ldx, Dinv, nch)
                               : launch_trsv<TA, true, false>(s, ws, n, nr, A, lda, X + c0 * ldx, ldx, Dinv, nch)));
            else
                AL_CHECK((unit ? launch_trsv<TA, false, true>(s, ws, n, nr, A, lda, X + c0 * ldx, ldx, Dinv, nch)
                               : launch_trsv<TA, false, false>(s, ws, n, nr, A, lda, X + c0 * ldx, ldx, Dinv, nch)));
        }
        return AL_OK;
    }
    if (lower) return unit ? launch_trsv<TA, true, true>(s, ws, n, nrhs, A, lda, X, ldx, Dinv)
                           : launch_trsv<TA, true, false>(s, ws, n, nrhs, A, lda, X, ldx, Dinv);
    return unit ? launch_trsv<TA, false, true>(s, ws, n, nrhs, A, lda, X, ldx, Dinv)
                : launch_trsv<TA, false, false>(s, ws, n, nrhs, A, lda, X, ldx, Dinv);
}

template int trsv_solve<double>(hipStream_t, Workspace*, bool, bool, size_t, size_t, const double*, size_t, double*,
                                size_t, const double*);
template int trsv_solve<float>(hipStream_t, Workspace*, bool, bool, size_t, size_t, const float*, size_t, double*,
                               size_t, const double*);

size_t trsv_diag_inverse_len(size_t n) { return (n + TB2 - 1) / TB2 * (size_t)TB2 * TB2; }

int trsv_diag_inverse(hipStream_t s, size_t n, const float* A, size_t lda, double* Linv, double* Uinv) {
    if (n == 0) return AL_OK;
    static bool attr_set = false;  // benign race: every thread sets the same value
    if (!attr_set) {
        AL_HIP(hipFuncSetAttribute((const void*)diag_inv_k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)DIAG_INV_LDS));
        attr_set = true;
    }
    const unsigned nblk = (unsigned)((n + TB2 - 1) / TB2);
    hipLaunchKernelGGL(diag_inv_k, dim3(nblk, 2), dim3(256), DIAG_INV_LDS, s, (int)n, A, lda, Linv, Uinv);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

}  // namespace al
