// Lower Cholesky, lapack::dpotrf (src/lapack/mod.rs:368-399).
//
// The crate runs a left-looking dot-product loop over the whole matrix.  Here it is a
// right-looking blocked factorization with 128-wide diagonal blocks:
//   * a 128 x 128 diagonal block is factored by one 512-thread workgroup with the block's
//     elements in registers (`potrf_diag_k`); the same sweep forms the block's inverse L11^{-1}.
//     The leading block -- the whole matrix when n <= 128 -- is factored with the crate's own
//     operations in the crate's order, so it matches lapack::dpotrf bit for bit;
//   * the panel below the block, A21 := A21 L11^{-T}, is one FP64 MFMA product with that inverse,
//     in place (a workgroup owns whole rows of A21), and the trailing matrix gets a
//     lower-triangle-only FP64 MFMA update.
// The look-ahead blocks are 384 wide (three diagonal blocks, so the trailing update has depth
// 384; 256 and 512 measured 2% slower at n = 8192).  Look-ahead on two streams: once block k is
// factored, the caller's stream applies its panel to the next block's columns first; the side
// stream then factors that block (diagonal blocks, inverses, panels) while the caller's stream
// applies panel k to every column right of it.  Only the lower triangle is read or written.
#include <algorithm>

#include "al_internal.hpp"

namespace al {
namespace {

#pragma clang fp contract(off)
constexpr int DB = 128;  // diagonal block
constexpr int OB = 384;  // look-ahead block: three diagonal blocks, the depth of the trailing update
constexpr int DT = 512;  // threads of the diagonal-block kernel: 64 row pairs x 8 column groups
constexpr int NS = 16;   // slots per row: the 16 columns of the thread's group

// Lane l's value of v, in every lane (l uniform).
__device__ __forceinline__ double lane_d(double v, int l) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | lo);
}

// 1 / sqrt(d) for a positive normal d: the hardware estimate and two Newton steps
// (r <- r + r (1/2 - d r^2 / 2)), to within an ulp or two.
__device__ __forceinline__ double rsqrt_nr(double d) {
    double r = __builtin_amdgcn_rsq(d);
    const double h = 0.5 * d;
#pragma unroll
    for (int k = 0; k < 2; ++k) r = fma(r, fma(-h * r, r, 0.5), r);
    return r;
}

// The nb x nb (nb <= 128) diagonal block at A (global columns col0..), factored right-looking with
// its elements in registers; with INV the same sweep forms the inverse W = L^{-1} (nb x nb, zeros
// above the diagonal, leading dimension ldw).  Thread (p, g) = (t & 63, t >> 6) owns rows 2p and
// 2p + 1 at the columns s = 16g + m, m < 16.  A slot holds L[i][s] until column s is factored and
// then l_ii W[i][s]: at step j every slot of a row i > j loses l_ij u_s, where u_s is l_sj for s > j
// (the factor's update), 1 / l_jj for s = j (the slot was zeroed when column j was handed over)
// and W[j][s] for s < j (forward substitution of L W = I); rows <= j keep their slots, which at
// the end are scaled by their 1 / l_ii into W.
// Per step the waves exchange only column j of L before the division by l_jj, in one LDS array
// `c` (1 at s = j; the pivot's sum in `dg`); two copies alternate between steps so that handing
// the next column over cannot overwrite what a slower wave still reads, and a step has one
// barrier.  A wave reads the values of its 16 columns once (one lane each) and broadcasts them
// through scalar registers; l_jj W[j][s] it takes from its own lane j / 2.  Without EXACT both
// kinds enter unscaled and the row's multiplier is c_i / l_jj^2 = l_ij / l_jj.  The step loop is
// unrolled by 16 so that the slot handed over and the register array of row j are known at
// compile time.
// With EXACT every element of L gets the crate's operations in the crate's order -- the
// subtractions l_ij l_sj for j = 0, 1, .. as a separate multiply and subtract, then the division by
// l_ss (the diagonal: the same subtractions, then the square root) -- so a block that no update
// has touched factors bit for bit as lapack::dpotrf factors it.  Without, the update is one FMA.
// A non-positive pivot at column j sets info to col0 + j + 1 and leaves A's columns j.. as given
// (those before j factored).
template <bool EXACT, bool INV>
__global__ __launch_bounds__(DT) void potrf_diag_k(double* __restrict__ A, size_t lda, int nb, int col0,
                                                  double* __restrict__ W, size_t ldw, int* __restrict__ info) {
    if (*info != 0) return;
    __shared__ double c[2][DB], dg[2];
    const int t = threadIdx.x, p = t & 63, r0 = 2 * p, r1 = r0 + 1;
    const int g = __builtin_amdgcn_readfirstlane(t >> 6);  // column group, uniform per wave
    const int s0 = 16 * g;
    double x0[NS], x1[NS];  // rows 2p and 2p + 1
#pragma unroll
    for (int m = 0; m < NS; ++m) {
        const int s = s0 + m;
        x0[m] = (r0 < nb && s <= r0) ? A[(size_t)r0 + (size_t)s * lda] : 0.0;
        x1[m] = (r1 < nb && s <= r1) ? A[(size_t)r1 + (size_t)s * lda] : 0.0;
    }
    if (g == 0) {  // hand column 0 over
        c[0][r0] = (r0 == 0) ? 1.0 : x0[0];
        c[0][r1] = x1[0];
        if (r0 == 0) dg[0] = x0[0];
    }
    x0[0] = (g == 0) ? 0.0 : x0[0];
    x1[0] = (g == 0) ? 0.0 : x1[0];
    double q0 = 0.0, q1 = 0.0;  // 1 / l_ii of the two rows, once factored
#pragma nounroll
    for (int j0 = 0; j0 < nb; j0 += 16) {
#pragma unroll
        for (int jj = 0; jj < 16; ++jj) {
            const int j = j0 + jj;
            if (j >= nb) break;  // uniform
            const int b = jj & 1, bn = b ^ 1;
            __syncthreads();
            const double d = dg[b];
            if (d <= 0.0) {  // uniform: every thread read the same value
                if (t == 0) *info = col0 + j + 1;
                return;
            }
            double ljj, r;
            if (EXACT || !(d >= 0x1p-1000 && d <= 0x1p+1000)) {  // uniform; rsqrt_nr wants a normal d
                ljj = sqrt(d);
                r = 1.0 / ljj;
            } else {
                r = rsqrt_nr(d);
                ljj = d * r;
            }
            q0 = (r0 == j) ? r : q0;
            q1 = (r1 == j) ? r : q1;
            if (t < nb && t >= j)  // column j of L is final
                A[(size_t)t + (size_t)j * lda] = (t == j) ? ljj : (EXACT ? c[b][t] / ljj : c[b][t] * r);
            // the rows' multipliers (0 for the rows done, which keep their slots) and u_s
            const double c0 = c[b][r0], c1 = c[b][r1];
            const double cv = c[b][s0 + (p & 15)];  // lane m < 16: column s0 + m of L (1 at j)
            double m0, m1, u[NS];
            const double* xj = (jj & 1) ? x1 : x0;  // row j's slots, in lane j / 2
            const int jl = j >> 1;
            // EXACT: multipliers l_ij, u_s = l_sj and W[j][s] = (l_jj W[j][s]) / l_jj; otherwise
            // multipliers l_ij / l_jj and u_s unscaled
            const double r2 = r * r;
            m0 = (r0 > j) ? (EXACT ? c0 / ljj : c0 * r2) : 0.0;
            m1 = (r1 > j) ? (EXACT ? c1 / ljj : c1 * r2) : 0.0;
            const double sv = EXACT ? cv / ljj : cv;
            if (s0 + NS <= j) {  // uniform: every slot of this group is a W slot
#pragma unroll
                for (int m = 0; m < NS; ++m) u[m] = EXACT ? lane_d(xj[m], jl) * r : lane_d(xj[m], jl);
            } else if (s0 > j) {  // uniform: every slot is an L slot
#pragma unroll
                for (int m = 0; m < NS; ++m) u[m] = lane_d(sv, m);
            } else {  // the group of column j: W slots below jj
#pragma unroll
                for (int m = 0; m < NS; ++m)
                    u[m] = (m < jj) ? (EXACT ? lane_d(xj[m], jl) * r : lane_d(xj[m], jl)) : lane_d(sv, m);
            }
#pragma unroll
            for (int m = 0; m < NS; ++m) {
                x0[m] = EXACT ? x0[m] - m0 * u[m] : fma(-m0, u[m], x0[m]);
                x1[m] = EXACT ? x1[m] - m1 * u[m] : fma(-m1, u[m], x1[m]);
            }
            // hand column j + 1 over: group (j + 1) / 16, slot (jj + 1) % 16, rows > j; keep a zero
            const int jn = j + 1;
            const int mh = (jj + 1) & 15;     // a constant in each unrolled step
            const bool own = g == (jn >> 4);  // uniform
            if (own) {
                if (r0 > j) c[bn][r0] = (r0 == jn) ? 1.0 : x0[mh];
                if (r1 > j) c[bn][r1] = (r1 == jn) ? 1.0 : x1[mh];
                if (r0 == jn) dg[bn] = x0[mh];
                if (r1 == jn) dg[bn] = x1[mh];
            }
            x0[mh] = own ? 0.0 : x0[mh];
            x1[mh] = own ? 0.0 : x1[mh];
        }
    }
    if (!INV) return;
#pragma unroll
    for (int m = 0; m < NS; ++m) {
        const int s = s0 + m;
        if (s >= nb) break;  // uniform
        if (r0 < nb) W[(size_t)r0 + (size_t)s * ldw] = (s < r0) ? x0[m] * q0 : ((s == r0) ? q0 : 0.0);
        if (r1 < nb) W[(size_t)r1 + (size_t)s * ldw] = (s < r1) ? x1[m] * q1 : ((s == r1) ? q1 : 0.0);
    }
}
#pragma clang fp contract(on)

// The panel below a diagonal block in place: A21 (m x 128) := A21 W^T with W = L11^{-1} (128 x 128,
// column-major, leading dimension ldw).  A workgroup owns 64 whole rows: it copies them to LDS,
// and wave w forms output columns [32w, 32w + 32) with v_mfma_f64_16x16x4_f64, W's entries from
// L2.  The operands go in as (W, A21^T) so that each lane's accumulators hold 16 consecutive
// rows of one output column.
constexpr int PR = 64;        // rows per workgroup
constexpr int PLD = DB + 1;   // LDS row stride (odd: a step's 16 rows x 4 columns fall in distinct banks)
typedef double pd4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void panel_trmm_k(int m, double* __restrict__ A, size_t lda,
                                                    const double* __restrict__ W, size_t ldw,
                                                    const int* __restrict__ info) {
    if (*info != 0) return;
    extern __shared__ double As[];  // PR x DB, row-major with stride PLD
    const int t = threadIdx.x, l = t & 63, w = t >> 6;
    const int r0 = blockIdx.x * PR;
    const int rows = min(PR, m - r0);
    const int lr = l & 15, lk = l >> 4;
    // every load issued before the first wait: this wave's W entries, then the rows
    double b[DB / 4][2];
#pragma unroll
    for (int s = 0; s < DB / 4; ++s)
#pragma unroll
        for (int jt = 0; jt < 2; ++jt) b[s][jt] = W[(size_t)(32 * w + 16 * jt + lr) + (size_t)(4 * s + lk) * ldw];
    double v[PR * DB / 256];
    const int i = t & (PR - 1), kq = t / PR;  // row, and the first of the columns kq + 4e
#pragma unroll
    for (int e = 0; e < PR * DB / 256; ++e)
        v[e] = i < rows ? A[(size_t)(r0 + i) + (size_t)(kq + 4 * e) * lda] : 0.0;
#pragma unroll
    for (int e = 0; e < PR * DB / 256; ++e) As[i * PLD + kq + 4 * e] = v[e];
    __syncthreads();
    pd4 acc[4][2];
#pragma unroll
    for (int it = 0; it < 4; ++it)
#pragma unroll
        for (int jt = 0; jt < 2; ++jt) acc[it][jt] = pd4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < DB / 4; ++s) {
        double a[4];
#pragma unroll
        for (int it = 0; it < 4; ++it) a[it] = As[(16 * it + lr) * PLD + 4 * s + lk];  // A21[i][k]
#pragma unroll
        for (int it = 0; it < 4; ++it)
#pragma unroll
            for (int jt = 0; jt < 2; ++jt)
                acc[it][jt] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[s][jt], a[it], acc[it][jt], 0, 0, 0);
    }
    // accumulator q of lane l: output row 16 it + (l & 15), column 32 w + 16 jt + (l >> 4) + 4 q
#pragma unroll
    for (int it = 0; it < 4; ++it) {
        const int i = 16 * it + lr;
        if (i >= rows) continue;
#pragma unroll
        for (int jt = 0; jt < 2; ++jt)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                A[(size_t)(r0 + i) + (size_t)(32 * w + 16 * jt + lk + 4 * q) * lda] = acc[it][jt][q];
    }
}

int set_panel_lds() {
    static const int rc = [] {
        if (hipFuncSetAttribute((const void*)panel_trmm_k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                PR * PLD * (int)sizeof(double)) != hipSuccess) {
            (void)hipGetLastError();
            return (int)AL_ERR_DEVICE;
        }
        return (int)AL_OK;
    }();
    if (rc != AL_OK) set_error("dpotrf: the panel kernel's LDS could not be configured");
    return rc;
}

// One diagonal block [k, k + jb) of the whole n x n factorization: the block (and its inverse W),
// then the panel below it, A21 := A21 W^T.  A block with rows below it is always DB wide (only
// the last block is narrower).  The leading block is factored with the crate's arithmetic
// (nothing has updated it), the others with FMA.
int potrf_block(hipStream_t s, size_t n, double* A, size_t lda, int* info, size_t k, size_t jb, double* W) {
    double* Akk = A + k + k * lda;
    const size_t mr = n - k - jb;
    if (mr > 0 && jb != (size_t)DB) {
        set_error("dpotrf: internal error (a narrow block with a panel)");
        return AL_ERR_DEVICE;
    }
    const bool exact = k == 0;
    if (mr == 0) {
        if (exact)
            hipLaunchKernelGGL((potrf_diag_k<true, false>), dim3(1), dim3(DT), 0, s, Akk, lda, (int)jb, (int)k, W,
                               (size_t)DB, info);
        else
            hipLaunchKernelGGL((potrf_diag_k<false, false>), dim3(1), dim3(DT), 0, s, Akk, lda, (int)jb, (int)k, W,
                               (size_t)DB, info);
        AL_HIP(hipGetLastError());
        return AL_OK;
    }
    if (exact)
        hipLaunchKernelGGL((potrf_diag_k<true, true>), dim3(1), dim3(DT), 0, s, Akk, lda, (int)jb, (int)k, W,
                           (size_t)DB, info);
    else
        hipLaunchKernelGGL((potrf_diag_k<false, true>), dim3(1), dim3(DT), 0, s, Akk, lda, (int)jb, (int)k, W,
                           (size_t)DB, info);
    AL_HIP(hipGetLastError());
    AL_CHECK(set_panel_lds());
    hipLaunchKernelGGL(panel_trmm_k, dim3((unsigned)ceil_div(mr, (size_t)PR)), dim3(256), PR * PLD * sizeof(double), s,
                       (int)mr, Akk + jb, lda, W, (size_t)DB, info);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

// Columns [k, k + ob) (ob <= OB) once every earlier panel has been applied to them: diagonal
// blocks of DB with their panels, each after the panels of the block's earlier ones have been
// applied to its columns.
int potrf_outer(hipStream_t s, size_t n, double* A, size_t lda, int* info, size_t k, size_t ob, double* W) {
    for (size_t k1 = k; k1 < k + ob; k1 += DB) {
        const size_t jb = std::min(k + ob - k1, (size_t)DB);
        if (k1 > k) {
            const double* L = A + k1 + k * lda;  // the panels of columns [k, k1), rows k1..n
            AL_CHECK(gemm_f64(s, false, true, n - k1, jb, k1 - k, -1.0, L, lda, L, lda, 1.0, A + k1 + k1 * lda, lda,
                              1));
        }
        AL_CHECK(potrf_block(s, n, A, lda, info, k1, jb, W));
    }
    return AL_OK;
}

}  // namespace

int potrf_f64(hipStream_t s, Workspace* ws, size_t n, double* A, size_t lda, int* info) {
    AL_CHECK(fill_i32(s, info, 1, 0));
    if (n == 0) return AL_OK;
    if (n <= (size_t)DB) return potrf_block(s, n, A, lda, info, 0, n, nullptr);
    // the inverse of a diagonal block (DB x DB)
    const size_t need = (size_t)DB * DB * sizeof(double);
    if (ws->cbuf_bytes < need) {
        if (ws->cbuf) AL_HIP(hipFree(ws->cbuf));
        ws->cbuf = nullptr;
        ws->cbuf_bytes = 0;
        if (hipMalloc(&ws->cbuf, need) != hipSuccess) {
            ws->cbuf = nullptr;
            (void)hipGetLastError();
            set_error("device allocation failed");
            return AL_ERR_OOM;
        }
        ws->cbuf_bytes = need;
    }
    double* W = (double*)ws->cbuf;
    AL_CHECK(ensure_side_stream(ws));
    hipStream_t side = ws->side;
    AL_HIP(hipEventRecord(ws->ev_main, s));
    AL_HIP(hipStreamWaitEvent(side, ws->ev_main, 0));
    const auto width = [&](size_t k) { return std::min(n - k, (size_t)OB); };  // the look-ahead block at column k
    AL_CHECK(potrf_outer(side, n, A, lda, info, 0, width(0), W));
    AL_HIP(hipEventRecord(ws->ev_side, side));
    for (size_t k = 0, kb = width(0); k + kb < n; k += kb, kb = width(k)) {
        const size_t ke = k + kb;
        const size_t ke2 = ke + width(ke);
        // main, once block k is factored: its panel applied to the next block's columns first
        AL_HIP(hipStreamWaitEvent(s, ws->ev_side, 0));
        const double* L21 = A + ke + k * lda;  // the panel of block k, rows ke..n
        AL_CHECK(gemm_f64(s, false, true, n - ke, ke2 - ke, kb, -1.0, L21, lda, L21, lda, 1.0, A + ke + ke * lda,
                          lda, 1));
        AL_HIP(hipEventRecord(ws->ev_main, s));
        // side: the next block ...
        AL_HIP(hipStreamWaitEvent(side, ws->ev_main, 0));
        AL_CHECK(potrf_outer(side, n, A, lda, info, ke, ke2 - ke, W));
        AL_HIP(hipEventRecord(ws->ev_side, side));
        // ... while the main stream applies panel k to every column right of it
        if (ke2 < n) {
            const double* Lr = A + ke2 + k * lda;
            AL_CHECK(gemm_f64(s, false, true, n - ke2, n - ke2, kb, -1.0, Lr, lda, Lr, lda, 1.0, A + ke2 + ke2 * lda,
                              lda, 1));
        }
    }
    // the caller's stream ends after the side stream's last block
    AL_HIP(hipEventRecord(ws->ev_side, side));
    AL_HIP(hipStreamWaitEvent(s, ws->ev_side, 0));
    return AL_OK;
}

}  // namespace al
