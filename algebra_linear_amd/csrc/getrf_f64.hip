// LU with partial pivoting (lapack::dgetrf / dgetrf_rect, src/lapack/mod.rs:112-238) and the
// solve with its factors (dgetrs, :330-354).
//
// The crate factors with an unblocked right-looking loop whose elimination walks rows with a
// stride of n.  This build blocks it for the GPU:
//   * panel: a cooperative kernel (`panel_getf2`) spreads the m x 64 panel over up to 64
//     workgroups, each holding its rows in LDS for the whole panel.  Each column costs one
//     grid-wide exchange: every workgroup publishes its best candidate (first maximum, the
//     crate's strict '>' scan) plus that row, one grid barrier, then every workgroup picks the
//     same winner, swaps and eliminates its own rows.  Within the panel the arithmetic is the
//     crate's own (divide, then a separate multiply and subtract per element, no FMA), so a
//     matrix that fits in one panel factors bit-for-bit as the crate factors it;
//   * two-level blocking: 64-wide panels inside 256-wide outer panels, so the trailing update
//     of the matrix is a k = 256 FP64 MFMA GEMM;
//   * row swaps of the columns outside a panel, the U12 triangular solve and the trailing
//     update are separate kernels on the same stream; each checks the status word first so
//     that a zero pivot stops the factorization like the crate's early return.
#include "al_internal.hpp"

#include <cmath>

namespace al {
namespace {

constexpr int PNB = 64;        // panel width
constexpr int PT = 256;        // threads per panel workgroup
constexpr int CHUNK_MAX = 256; // rows per workgroup (LDS: 256 x 64 doubles = 128 KiB)
constexpr int MAX_WG = 128;    // cooperative workgroups per panel (one per CU at most)
constexpr int SLOT = PNB + 2;  // key, index, row values
constexpr int NBO = 256;       // outer block (trailing-update depth)

// Candidate ordering: larger key wins, ties go to the lower row (first maximum).
__device__ __forceinline__ bool better(double k1, int i1, double k2, int i2) {
    return (k1 > k2) || (k1 == k2 && i1 < i2);
}

// Pivot key of a candidate: |v|, except that NaN never wins unless it sits on the diagonal,
// where the crate's scan starts (src/lapack/mod.rs:126-134): a NaN start blocks every later
// candidate, a later NaN is skipped.
__device__ __forceinline__ double pivot_key(double v, int g, int j) {
    if (v != v) return (g == j) ? INFINITY : -1.0;
    return fabs(v);
}

__device__ __forceinline__ void block_argmax(double& key, int& idx, double* red_k, int* red_i) {
    // wave64 butterfly, then across the 4 waves through LDS
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double ok = __shfl_xor(key, off, 64);
        const int oi = __shfl_xor(idx, off, 64);
        if (better(ok, oi, key, idx)) {
            key = ok;
            idx = oi;
        }
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red_k[w] = key;
        red_i[w] = idx;
    }
    __syncthreads();
    key = red_k[0];
    idx = red_i[0];
#pragma unroll
    for (int q = 1; q < PT / 64; ++q)
        if (better(red_k[q], red_i[q], key, idx)) {
            key = red_k[q];
            idx = red_i[q];
        }
    __syncthreads();
}

// Grid barrier over `nwg` co-resident workgroups: arrival counter with release/acquire at
// agent scope.  The spin is bounded: past `timeout` ticks of the 100 MHz wall clock it
// raises the abort flag, so a barrier can never hang the device.  Returns false on abort.
__device__ bool grid_barrier(unsigned* counter, unsigned* abort_flag, unsigned target, long long timeout,
                             int* ok_lds) {
    __syncthreads();
    if (threadIdx.x == 0) {
        int ok = 1;
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        const long long t0 = wall_clock64();
        while (__hip_atomic_load(counter, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
                ok = 0;
                break;
            }
            if (wall_clock64() - t0 > timeout) {
                __hip_atomic_store(abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        *ok_lds = ok;
    }
    __syncthreads();
    return *ok_lds != 0;
}

#pragma clang fp contract(off)
// Factors the m x ncols panel at P (leading dimension lda), whose top-left element is the
// global diagonal element (row0, row0).  Workgroup w owns panel rows [w*chunk, (w+1)*chunk).
__global__ __launch_bounds__(PT) void panel_getf2(double* __restrict__ P, size_t lda, int row0, int m, int ncols,
                                                  int chunk, int* __restrict__ ipiv, int* __restrict__ info,
                                                  unsigned* counter, unsigned* abort_flag,
                                                  double* __restrict__ slots, long long timeout) {
    extern __shared__ double S[];  // chunk x PNB, column-major with leading dimension `chunk`
    __shared__ double prow[PNB];
    __shared__ double rj[PNB];
    __shared__ double red_k[PT / 64];
    __shared__ int red_i[PT / 64];
    __shared__ int ok_lds;

    const int nwg = gridDim.x;
    const int w = blockIdx.x, t = threadIdx.x;
    if (*info != 0) return;  // an earlier panel hit a zero pivot
    const int rb = w * chunk;
    const int nr = (m - rb) < chunk ? (m - rb) : chunk;
    double* rowj_buf = slots + (size_t)2 * nwg * SLOT;  // [2][PNB]

    for (int idx = t; idx < nr * ncols; idx += PT) {
        const int r = idx % nr, c = idx / nr;
        S[c * chunk + r] = P[(size_t)(rb + r) + (size_t)c * lda];
    }
    __syncthreads();

    unsigned bar = 0;
    int stop = 0;
    for (int j = 0; j < ncols; ++j) {
        // 1. this workgroup's candidate for pivot j
        double key = -2.0;
        int idx = 0x7fffffff;
        for (int r = t; r < nr; r += PT) {
            const int g = rb + r;
            if (g < j) continue;
            const double kv = pivot_key(S[j * chunk + r], g, j);
            if (better(kv, g, key, idx)) {
                key = kv;
                idx = g;
            }
        }
        block_argmax(key, idx, red_k, red_i);

        // 2. publish it (with its row) and, if owned here, row j
        const int par = j & 1;
        double* my = slots + ((size_t)par * nwg + w) * SLOT;
        if (key > -2.0) {
            const int lr = idx - rb;
            for (int c = t; c < ncols; c += PT) my[2 + c] = S[c * chunk + lr];
        }
        if (t == 0) {
            my[0] = key;
            my[1] = (double)idx;
        }
        if (j >= rb && j < rb + nr)
            for (int c = t; c < ncols; c += PT) rowj_buf[par * PNB + c] = S[c * chunk + (j - rb)];
        __threadfence();

        // 3. one grid barrier per column
        if (nwg > 1) {
            ++bar;
            if (!grid_barrier(counter, abort_flag, bar * (unsigned)nwg, timeout, &ok_lds)) {
                if (w == 0 && t == 0) *info = -1;
                stop = 1;
                break;
            }
        } else {
            __syncthreads();
        }

        // 4. the same winner everywhere: the best of all published candidates
        double bk = -2.0;
        int bi = 0x7fffffff, bw = 0;
        for (int q = t; q < nwg; q += PT) {
            const double* sq = slots + ((size_t)par * nwg + q) * SLOT;
            const double kq = sq[0];
            const int iq = (int)sq[1];
            if (better(kq, iq, bk, bi)) {
                bk = kq;
                bi = iq;
                bw = q;
            }
        }
        // reduce (key, idx) and recover the owner workgroup from the row index
        block_argmax(bk, bi, red_k, red_i);
        bw = bi / chunk;
        if (!(bk > 0.0)) {
            // |pivot| <= 0: singular (src/lapack/mod.rs:136-138); NaN diagonals carry +inf
            if (w == 0 && t == 0) *info = row0 + j + 1;
            stop = 1;
            break;
        }
        const double* win = slots + ((size_t)par * nwg + bw) * SLOT;
        for (int c = t; c < ncols; c += PT) {
            prow[c] = win[2 + c];
            rj[c] = rowj_buf[par * PNB + c];
        }
        __syncthreads();

        // 5. swap rows j and bi (full panel width), record the pivot
        const int p = bi;
        if (p != j) {
            if (p >= rb && p < rb + nr)
                for (int c = t; c < ncols; c += PT) S[c * chunk + (p - rb)] = rj[c];
            if (j >= rb && j < rb + nr)
                for (int c = t; c < ncols; c += PT) S[c * chunk + (j - rb)] = prow[c];
        }
        if (w == 0 && t == 0) ipiv[row0 + j] = row0 + p;
        __syncthreads();

        // 6. eliminate below the pivot: factor = a/d, then a -= factor*u, unfused
        const double d = prow[j];
        for (int r = t; r < nr; r += PT) {
            const int g = rb + r;
            if (g <= j) continue;
            const double f = S[j * chunk + r] / d;
            S[j * chunk + r] = f;
            for (int c = j + 1; c < ncols; ++c) S[c * chunk + r] -= f * prow[c];
        }
        __syncthreads();
    }
    (void)stop;
    for (int idx = t; idx < nr * ncols; idx += PT) {
        const int r = idx % nr, c = idx / nr;
        P[(size_t)(rb + r) + (size_t)c * lda] = S[c * chunk + r];
    }
}
#pragma clang fp contract(on)

// Factors the m x n block at (k0, k0) of A with 64-wide cooperative panels; swaps are applied
// inside columns [k0, k0 + n) only -- the caller applies them outside.
int panel_blocked(hipStream_t s, Workspace* ws, double* A, size_t lda, size_t M, size_t k0, size_t n, int* ipiv,
                  int* info) {
    static bool attr_set = false;
    if (!attr_set) {
        AL_HIP(hipFuncSetAttribute((const void*)panel_getf2, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)(CHUNK_MAX * PNB * sizeof(double))));
        attr_set = true;
    }
    for (size_t kk = k0; kk < k0 + n; kk += PNB) {
        const size_t ib = (k0 + n - kk) < (size_t)PNB ? (k0 + n - kk) : (size_t)PNB;
        const size_t m = M - kk;
        size_t nwg = (m + CHUNK_MAX - 1) / CHUNK_MAX;
        if (nwg > (size_t)MAX_WG) nwg = MAX_WG;
        const size_t chunk = (m + nwg - 1) / nwg;
        if (chunk > (size_t)CHUNK_MAX) {
            set_error("getrf: panel taller than the cooperative kernel holds");
            return AL_ERR_ARG;
        }
        AL_CHECK(fill_u32(s, ws->counter, 2, 0u));  // counter and abort flag
        const long long timeout = 200000000LL;     // 2 s at 100 MHz
        const int ticket = prof_begin(s);
        hipLaunchKernelGGL(panel_getf2, dim3((unsigned)nwg), dim3(PT), chunk * PNB * sizeof(double), s,
                           A + kk + kk * lda, lda, (int)kk, (int)m, (int)ib, (int)chunk, ipiv, info, ws->counter,
                           ws->counter + 1, ws->slots, timeout);
        AL_HIP(hipGetLastError());
        if (ticket >= 0) {
            // unblocked LU of an m x ib panel: m*ib^2 - ib^3/3 flops; the panel read and written once
            const double dm = (double)m, db = (double)ib;
            prof_end(ticket, s, "panel_getf2", dm * db * db - db * db * db / 3.0, 16.0 * dm * db);
        }
        // swaps on this outer panel's other columns, then U12 and the in-panel trailing update
        const size_t rest = k0 + n - kk - ib;
        AL_CHECK(swap_block_f64(s, ws, ipiv, kk, ib, A + k0 * lda, kk - k0, A + (kk + ib) * lda, rest, lda, nullptr,
                                info));
        if (rest > 0) {
            double* A12 = A + kk + (kk + ib) * lda;
            AL_CHECK(trsm_f64(s, true, true, false, true, ib, rest, 1.0, A + kk + kk * lda, lda, A12, lda, info));
            if (M > kk + ib)
                AL_CHECK(gemm_f64(s, false, false, M - kk - ib, rest, ib, -1.0, A + (kk + ib) + kk * lda, lda, A12,
                                  lda, 1.0, A + (kk + ib) + (kk + ib) * lda, lda));
        }
    }
    return AL_OK;
}

}  // namespace

int getrf_f64(hipStream_t s, Workspace* ws, size_t M, size_t N, double* A, size_t lda, int* ipiv, int* info,
              double* perm) {
    const size_t kmax = M < N ? M : N;
    if (M > (size_t)CHUNK_MAX * MAX_WG || N > INT32_MAX) {
        set_error("getrf: more rows than the cooperative panel holds (32768)");
        return AL_ERR_ARG;
    }
    AL_CHECK(fill_i32(s, info, 1, 0));
    if (perm) AL_CHECK(perm_init(s, perm, M));
    for (size_t k = 0; k < kmax; k += NBO) {
        const size_t jb = (kmax - k) < (size_t)NBO ? (kmax - k) : (size_t)NBO;
        AL_CHECK(panel_blocked(s, ws, A, lda, M, k, jb, ipiv, info));
        const size_t right = (k + jb < N) ? N - k - jb : 0;
        AL_CHECK(swap_block_f64(s, ws, ipiv, k, jb, A, k, A + (k + jb) * lda, right, lda, perm, info));
        if (right > 0) {
            double* A12 = A + k + (k + jb) * lda;
            AL_CHECK(trsm_f64(s, true, true, false, true, jb, right, 1.0, A + k + k * lda, lda, A12, lda, info));
            if (M > k + jb)
                AL_CHECK(gemm_f64(s, false, false, M - k - jb, right, jb, -1.0, A + (k + jb) + k * lda, lda, A12, lda,
                                  1.0, A + (k + jb) + (k + jb) * lda, lda));
        }
    }
    return AL_OK;
}

int getrs_f64(hipStream_t s, Workspace* ws, size_t n, size_t nrhs, const double* LU, size_t lda, const int* ipiv,
              const double* perm, double* B, size_t ldb, const int* info) {
    if (n == 0 || nrhs == 0) return AL_OK;
    AL_CHECK(ws_reserve(ws, n, n * nrhs));
    if (!perm) {
        AL_CHECK(perm_from_ipiv(s, ipiv, n, ws->perm));
        perm = ws->perm;
    }
    // apply the interchanges as one gather into the staging block, solve there, copy back
    AL_CHECK(gather_rows(s, B, ldb, n, nrhs, perm, ws->wbuf));
    AL_CHECK(trsm_f64(s, true, true, false, true, n, nrhs, 1.0, LU, lda, ws->wbuf, n, info));
    AL_CHECK(trsm_f64(s, true, false, false, false, n, nrhs, 1.0, LU, lda, ws->wbuf, n, info));
    AL_CHECK(copy_block(s, ws->wbuf, n, nrhs, B, ldb));
    return AL_OK;
}

size_t panel_slots_bytes() { return ((size_t)2 * MAX_WG * SLOT + 2 * PNB) * sizeof(double); }

}  // namespace al
