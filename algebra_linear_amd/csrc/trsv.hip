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
    auto load_a = [&](int q, double (&av)[16]) {
        const int c0 = (LOWER ? q : nblk - 1 - q) * TB + part * 16;
#pragma unroll
        for (int jj = 0; jj < 16; ++jj)
            av[jj] = (rok && c0 + jj < n) ? (double)A[(size_t)(r0 + row) + (size_t)(c0 + jj) * lda] : 0.0;
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
    if (!ok) return;

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

template <typename TA, bool LOWER, bool UNIT, int NR>
void launch_nr(hipStream_t s, unsigned nblk, int n, int nrhs, const TA* A, size_t lda, double* X, size_t ldx,
               unsigned* flags, unsigned gen, unsigned* abort_flag, long long timeout) {
    hipLaunchKernelGGL((trsv_blocked<TA, LOWER, UNIT, NR>), dim3(nblk), dim3(256), 0, s, n, nrhs, A, lda, X, ldx, flags,
                       gen, abort_flag, timeout);
}

template <typename TA, bool LOWER, bool UNIT>
int launch_trsv(hipStream_t s, Workspace* ws, size_t n, size_t nrhs, const TA* A, size_t lda, double* X, size_t ldx) {
    const size_t nblk = (n + TB - 1) / TB;
    AL_CHECK(ws_reserve_flags(ws, nblk * 16));
    if (ws->trsv_gen > 0xF0000000u) {  // keep generations unique: restart from a cleared array
        AL_CHECK(fill_u32(s, ws->trsv_flags, nblk * 16, 0u));
        ws->trsv_gen = 0;
    }
    const unsigned gen = ++ws->trsv_gen;
    const long long timeout = 200000000LL;  // 2 s at 100 MHz
    const int ticket = prof_begin(s);
    const unsigned g = (unsigned)nblk;
    if (nrhs <= 1)
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
               double* X, size_t ldx) {
    if (n == 0 || nrhs == 0) return AL_OK;
    if (!trsv_fits(nrhs) || n > INT32_MAX) {
        set_error("trsv: more right-hand sides than the one-launch solve takes");
        return AL_ERR_ARG;
    }
    if (lower) return unit ? launch_trsv<TA, true, true>(s, ws, n, nrhs, A, lda, X, ldx)
                           : launch_trsv<TA, true, false>(s, ws, n, nrhs, A, lda, X, ldx);
    return unit ? launch_trsv<TA, false, true>(s, ws, n, nrhs, A, lda, X, ldx)
                : launch_trsv<TA, false, false>(s, ws, n, nrhs, A, lda, X, ldx);
}

template int trsv_solve<double>(hipStream_t, Workspace*, bool, bool, size_t, size_t, const double*, size_t, double*,
                                size_t);
template int trsv_solve<float>(hipStream_t, Workspace*, bool, bool, size_t, size_t, const float*, size_t, double*,
                               size_t);

}  // namespace al
