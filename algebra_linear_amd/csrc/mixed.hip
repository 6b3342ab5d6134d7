// Mixed-precision solve of A X = B: pivoted LU in low precision, FP64 iterative
// refinement, FP64 pivoted-LU fallback.  The crate solves in FP64 only (lapack::dgesv,
// src/lapack/mod.rs:309-323); this is the GPU extension beside it, with the same result
// contract measured as normwise backward error.
//
// Pipeline:
//   |A|_inf and |B_j|_inf   row sums / column maxima, FP64
//   F = fl32(A)             A rounded to FP32, padded to a multiple of 64 with |A|_inf/n * I
//   P F = L U               blocked right-looking LU in FP32 with tournament-pivoted 64-column
//                           panels, 64-row U solves and FP32 MFMA updates inside a 256/512-wide
//                           outer block, the outer trailing update on FP32 or bf16 matrix cores
//                           (FP32 accumulation); the panel chain runs one outer panel ahead on
//                           a second stream
//   refinement              r = B - A X (FP64, A streamed once per step, deterministic
//                           blocked sums), d = U^{-1} L^{-1} P r with the FP32 factors (their
//                           128-row diagonal blocks applied as FP64 inverses) and FP64
//                           vectors, X += d
//   fallback                FP64 pivoted LU (getrf_f64/getrs_f64) on a copy of A
//
// Pivoting is not optional here: without it (even after a random butterfly transform) the LU
// of a dense random matrix has |L||U| / |A| of 1e4-1e5 at n = 1024, enough to stall FP32
// refinement at n = 16384 and to make a bf16 update useless.  Tournament pivoting keeps the
// growth of partial pivoting without its per-column grid-wide exchange.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "al_internal.hpp"

namespace al {
namespace {

constexpr int PB = 64;          // panel width and padding granule
constexpr int PAIR_LISTS = 8;     // 64-pivot blocks per outer panel, at most (bf16: 512 columns)
constexpr int PAIR_STRIDE = 320;  // ints per pair list: a count and up to 128 (destination, source) pairs
constexpr int GV_ROWS = 256;    // residual: rows per workgroup
constexpr int GV_COLS = 256;    // residual: columns per partial sum
constexpr int NR_MAX = 16;      // right-hand sides per refinement pass

typedef unsigned long long u64;

// ---------------------------------------------------------------------------------------
// FP32 copy and permuted right-hand sides
// ---------------------------------------------------------------------------------------

// F (np x np, FP32, ld np) := fl32([A 0; 0 pad I]), pad = |A|_inf / n (1 if that is zero or
// not finite).
__global__ __launch_bounds__(256) void to_f32_pad_k(int n, int np, const double* __restrict__ A, size_t lda,
                                                    const u64* __restrict__ norms, float* __restrict__ F,
                                                    size_t idx0, size_t idx1) {
    const size_t idx = idx0 + (size_t)blockIdx.x * 256 + threadIdx.x;  // columns idx0 / np ..
    if (idx >= idx1) return;
    const int i = (int)(idx % np), j = (int)(idx / np);
    float v;
    if (i < n && j < n) {
        v = (float)A[(size_t)i + (size_t)j * lda];
    } else {
        const double anorm = __longlong_as_double((long long)norms[0]);
        const double pad = (anorm > 0.0 && anorm <= DBL_MAX) ? anorm / n : 1.0;
        v = i == j ? (float)pad : 0.f;
    }
    F[idx] = v;
}

// Y (np x nr, ld np) := P [R; 0]: Y[i] = R[perm[i]] (perm the row origins of the factorization).
__global__ __launch_bounds__(256) void gather_perm_k(int n, int np, int nr, const double* __restrict__ perm,
                                                     const double* __restrict__ R, size_t ldr, double* __restrict__ Y) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= np * nr) return;
    const int i = idx % np, k = idx / np;
    const int src = (int)perm[i];
    Y[(size_t)i + (size_t)k * np] = src < n ? R[(size_t)src + (size_t)k * ldr] : 0.0;
}

// Af (np x np, FP64, ld np) := P [A 0; 0 pad I] with P the FP32 factor's row origins (perm)
// and pad as in to_f32_pad_k: the padded system in the low-precision factorization's row order.
// One workgroup per column, so a column's scattered reads stay in one L2.
__global__ __launch_bounds__(256) void gather_pad_k(int n, int np, const double* __restrict__ A, size_t lda,
                                                    const double* __restrict__ perm, const u64* __restrict__ norms,
                                                    double* __restrict__ Af) {
    const int c = blockIdx.x;
    const double anorm = __longlong_as_double((long long)norms[0]);
    const double pad = (anorm > 0.0 && anorm <= DBL_MAX) ? anorm / n : 1.0;
    for (int i = threadIdx.x; i < np; i += 256) {
        const int src = (int)perm[i];
        double v;
        if (c < n)
            v = src < n ? A[(size_t)src + (size_t)c * lda] : 0.0;
        else
            v = src == c ? pad : 0.0;
        Af[(size_t)i + (size_t)c * np] = v;
    }
}

// out[0] := min_k |U[k][k]|, out[1] := max_k |U[k][k]| over the n x n factor at U, as the bit
// patterns of non-negative doubles (which order as integers); one workgroup.
__global__ __launch_bounds__(1024) void diag_minmax_k(int n, const double* __restrict__ U, size_t ldu,
                                                      u64* __restrict__ out) {
    __shared__ u64 lo[16], hi[16];
    u64 mn = ~0ull, mx = 0ull;
    for (int k = threadIdx.x; k < n; k += 1024) {
        const u64 b = (u64)__double_as_longlong(fabs(U[(size_t)k * (ldu + 1)]));
        mn = b < mn ? b : mn;
        mx = b > mx ? b : mx;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const u64 a = (u64)__shfl_xor((long long)mn, off, 64), c = (u64)__shfl_xor((long long)mx, off, 64);
        mn = a < mn ? a : mn;
        mx = c > mx ? c : mx;
    }
    if ((threadIdx.x & 63) == 0) {
        lo[threadIdx.x >> 6] = mn;
        hi[threadIdx.x >> 6] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 16; ++w) {
            mn = lo[w] < mn ? lo[w] : mn;
            mx = hi[w] > mx ? hi[w] : mx;
        }
        out[0] = mn;
        out[1] = mx;
    }
}

// X (n x nr, ld ldx) := (accumulate ? X : 0) + Y[0:n).
__global__ __launch_bounds__(256) void add_k(int n, int np, int nr, const double* __restrict__ Y,
                                             double* __restrict__ X, size_t ldx, int accumulate) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= n * nr) return;
    const int i = idx % n, k = idx / n;
    double* p = X + (size_t)i + (size_t)k * ldx;
    const double y = Y[(size_t)i + (size_t)k * np];
    *p = accumulate ? *p + y : y;
}

// ---------------------------------------------------------------------------------------
// FP64 residual and norms
// ---------------------------------------------------------------------------------------

// part[chunk][row][k] = sum over the chunk's 256 columns c of A[row, c] * X[c, k] (ABS:
// |A[row, c]|, X unused).  Each thread owns one row and walks the chunk in order, so the
// sums are reproducible; A is read once, coalesced down its columns.
template <int NR, bool ABS>
__global__ __launch_bounds__(256) void gemv_part_k(int n, int nr, const double* __restrict__ A, size_t lda,
                                                   const double* __restrict__ X, size_t ldx,
                                                   double* __restrict__ part) {
    __shared__ double xs[ABS ? 1 : GV_COLS][NR];
    const int t = threadIdx.x;
    const int r = blockIdx.x * GV_ROWS + t;
    const int c0 = blockIdx.y * GV_COLS;
    const int cn = (n - c0) < GV_COLS ? (n - c0) : GV_COLS;
    if (!ABS) {
        for (int idx = t; idx < GV_COLS * NR; idx += 256) {
            const int c = idx / NR, k = idx % NR;
            xs[c][k] = (c < cn && k < nr) ? X[(size_t)(c0 + c) + (size_t)k * ldx] : 0.0;
        }
        __syncthreads();
    }
    double acc[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) acc[k] = 0.0;
    if (r >= n) return;
    const double* ap = A + (size_t)r + (size_t)c0 * lda;
    int c = 0;
    for (; c + 8 <= cn; c += 8) {
        double a[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) a[u] = ap[(size_t)(c + u) * lda];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (ABS) {
                acc[0] += fabs(a[u]);
            } else {
#pragma unroll
                for (int k = 0; k < NR; ++k) acc[k] = fma(a[u], xs[c + u][k], acc[k]);
            }
        }
    }
    for (; c < cn; ++c) {
        const double a = ap[(size_t)c * lda];
        if (ABS) {
            acc[0] += fabs(a);
        } else {
#pragma unroll
            for (int k = 0; k < NR; ++k) acc[k] = fma(a, xs[c][k], acc[k]);
        }
    }
    const int w = ABS ? 1 : nr;
    double* out = part + ((size_t)blockIdx.y * n + r) * w;
#pragma unroll
    for (int k = 0; k < NR; ++k)
        if (k < w) out[k] = acc[k];
}

__device__ __forceinline__ u64 abs_bits(double v) { return (u64)__double_as_longlong(fabs(v)); }

// Max over the wave of non-negative doubles as their bit patterns (order-preserving; a NaN
// sorts above +inf, so it survives the reduction).
__device__ __forceinline__ u64 wave_max(u64 v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const u64 o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

// out[0] = max over rows of sum over the chunks of part (|A|_inf from the ABS partial sums).
__global__ __launch_bounds__(256) void rowsum_max_k(int n, int nparts, const double* __restrict__ part,
                                                    u64* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    double s = 0.0;
    if (i < n)
        for (int p = 0; p < nparts; ++p) s += part[(size_t)p * n + i];
    const u64 m = wave_max(abs_bits(s));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// out[k] = max_i |B[i, k]|.
__global__ __launch_bounds__(256) void colmax_k(int n, int nr, const double* __restrict__ B, size_t ldb,
                                                u64* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    for (int k = 0; k < nr; ++k) {
        const double v = i < n ? B[(size_t)i + (size_t)k * ldb] : 0.0;
        const u64 m = wave_max(abs_bits(v));
        if ((threadIdx.x & 63) == 0) atomicMax(&out[k], m);
    }
}

// R = B - (sum of the partial products in chunk order); rmax[k] = max |R[:, k]|,
// xmax[k] = max |X[:, k]|.
__global__ __launch_bounds__(256) void resid_k(int n, int nr, int nparts, const double* __restrict__ part,
                                               const double* __restrict__ B, size_t ldb,
                                               const double* __restrict__ X, size_t ldx, double* __restrict__ R,
                                               u64* __restrict__ rmax, u64* __restrict__ xmax) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    for (int k = 0; k < nr; ++k) {
        double r = 0.0, x = 0.0;
        if (i < n) {
            double s = 0.0;
            for (int p = 0; p < nparts; ++p) s += part[((size_t)p * n + i) * nr + k];
            r = B[(size_t)i + (size_t)k * ldb] - s;
            R[(size_t)i + (size_t)k * n] = r;
            x = X[(size_t)i + (size_t)k * ldx];
        }
        const u64 rm = wave_max(abs_bits(r)), xm = wave_max(abs_bits(x));
        if ((threadIdx.x & 63) == 0) {
            atomicMax(&rmax[k], rm);
            atomicMax(&xmax[k], xm);
        }
    }
}

// Row indices below 2^24 are exact in FP32, so the FP32 factor's interchanges can carry them
// as one more column.
__global__ void iota_f32_k(float* p, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = (float)i;
}

__global__ void widen_k(const float* __restrict__ src, double* __restrict__ dst, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = (double)src[i];
}

__global__ void copy2d_k(int m, int n, const double* __restrict__ S, size_t lds, double* __restrict__ D, size_t ldd) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)m * n) return;
    const int i = (int)(idx % m), j = (int)(idx / m);
    D[(size_t)i + (size_t)j * ldd] = S[(size_t)i + (size_t)j * lds];
}

// ---------------------------------------------------------------------------------------
// LU with partial pivoting, FP32
// ---------------------------------------------------------------------------------------

// Pieces of u12_k on four columns at once (their element `lane` of one 64-row block in x):
// x -= L_ba x_a with L_ba's row `lane` at Lp (stride ld between its columns), and the forward
// solve with a unit lower diagonal block.  Kept out of line so that each holds one 64-register
// tile of L at a time.
__device__ __attribute__((noinline)) float4 u12_off(const float* __restrict__ Lp, size_t ld, float4 x, float4 xa) {
    float lr[64];
#pragma unroll
    for (int j = 0; j < 64; ++j) lr[j] = Lp[(size_t)j * ld];
    float xv[4] = {x.x, x.y, x.z, x.w};
    const float av[4] = {xa.x, xa.y, xa.z, xa.w};
#pragma unroll
    for (int j = 0; j < 64; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q)
            xv[q] = fmaf(-lr[j], __int_as_float(__builtin_amdgcn_readlane(__float_as_int(av[q]), j)), xv[q]);
    return float4{xv[0], xv[1], xv[2], xv[3]};
}

__device__ __attribute__((noinline)) float4 u12_diag(const float* __restrict__ Lp, size_t ld, float4 x, int lane) {
    float lr[64];
#pragma unroll
    for (int j = 0; j < 64; ++j) lr[j] = Lp[(size_t)j * ld];
    float xv[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
    for (int j = 0; j < 63; ++j) {
        const bool below = lane > j;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float xj = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(xv[q]), j));
            const float v = fmaf(-lr[j], xj, xv[q]);
            xv[q] = below ? v : xv[q];
        }
    }
    return float4{xv[0], xv[1], xv[2], xv[3]};
}

// U12 := L11^{-1} A12 for the outer panel's NB * 64 rows at K and columns [c0, c1), L11 the
// unit lower triangle of F[K:K+64NB, K:K+64NB], in one launch.  One wave solves four columns at
// a time, block b's element i of each in lane i: block b first subtracts L(b, a) x_a for every
// a < b (the 64 x 64 block's row i streamed into lane i's registers, x_a's elements handed out
// by v_readlane), then solves with its diagonal block as solve_l64_cols does.
template <int NB>
__global__ __launch_bounds__(256) void u12_k(float* __restrict__ F, size_t ld, int K, int c0, int c1) {
    const int lane = threadIdx.x & 63;
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6), nwaves = gridDim.x * 4;
    const int ncols = c1 - c0;
    const float* Lrow = F + (size_t)(K + lane) + (size_t)K * ld;  // L11(lane, 0)
    for (int g = wave * 4; g < ncols; g += nwaves * 4) {
        float* col[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) col[q] = F + (size_t)K + (size_t)(c0 + min(g + q, ncols - 1)) * ld;  // a short last group repeats
        float4 x[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b)
            x[b] = float4{col[0][64 * b + lane], col[1][64 * b + lane], col[2][64 * b + lane], col[3][64 * b + lane]};
#pragma unroll
        for (int b = 0; b < NB; ++b) {
#pragma unroll
            for (int a = 0; a < b; ++a) x[b] = u12_off(Lrow + 64 * b + (size_t)(64 * a) * ld, ld, x[b], x[a]);
            x[b] = u12_diag(Lrow + 64 * b + (size_t)(64 * b) * ld, ld, x[b], lane);
        }
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const float xv[4] = {x[b].x, x[b].y, x[b].z, x[b].w};
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (g + q < ncols) col[q][64 * b + lane] = xv[q];
        }
    }
}

// ---------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------

struct Arena {
    size_t n = 0, np = 0, nr = 0, nparts = 0;
    float* F = nullptr;       // np x np factor
    int* ipiv = nullptr;      // np pivots of the FP32 factorization
    double* perm = nullptr;   // np row origins
    float* rowidx = nullptr;  // np row origins as they are built: the interchanges applied to 0, 1, ..
    double* R = nullptr;      // n x nr residual
    double* Y = nullptr;      // np x nr permuted residual / correction
    double* part = nullptr;   // nparts x n x nr partial sums
    u64* norms = nullptr;     // [0] |A|_inf, [1..17) |B|, [17..33) |R|, [33..49) |X|
    int* minfo = nullptr;     // [0] low-precision factorization status, [2] fallback status
    __bf16* At = nullptr;     // bf16 staging of the outer block's L and U
    __bf16* Bt = nullptr;
    void* tslu = nullptr;     // tournament panel scratch (nominees, L11 \ U11)
    unsigned* counters = nullptr;  // its arrival counters (zero between panels)
    int* pairs_sub = nullptr;      // [2][PAIR_LISTS] row-interchange pair lists, one per 64-pivot
                                   // block of an outer panel, alternating by outer panel
    double* dinv = nullptr;   // FP64 inverses of the factor's 128-row diagonal blocks: L's, then U's
};

constexpr int NORM_A = 0, NORM_B = 1, NORM_R = 17, NORM_X = 33, NORM_DIAG = 49, NORM_ABORT = 63, NORM_WORDS = 64;

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

int grow(void** p, size_t* have, size_t want) {
    if (*have >= want) return AL_OK;
    if (*p) AL_HIP(hipFree(*p));
    *p = nullptr;
    *have = 0;
    hipError_t e = hipMalloc(p, want);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        set_error("device allocation failed (mixed-precision workspace)");
        return AL_ERR_OOM;
    }
    *have = want;
    return AL_OK;
}

int make_arena(hipStream_t s, Workspace* ws, size_t n, size_t np, size_t nr, size_t nbo, bool lowp, bool bf16,
               bool refine, Arena* a) {
    a->n = n;
    a->np = np;
    a->nr = nr;
    a->nparts = ceil_div(n, GV_COLS);
    size_t off = 4096;  // [0, 4096): the tournament's arrival counters, zeroed by every call
    const size_t oF = off;
    if (lowp) off += align256(np * np * sizeof(float));
    const size_t oV = off;
    if (lowp) off += align256(np * sizeof(int));
    const size_t oW = off;
    if (lowp) off += align256(np * sizeof(double));
    const size_t oX = off;
    if (lowp) off += align256(np * sizeof(float));
    const size_t oR = off;
    off += align256(n * nr * sizeof(double));
    const size_t oY = off;
    off += align256(np * nr * sizeof(double));
    const size_t oP = off;
    off += align256(a->nparts * n * nr * sizeof(double));
    const size_t oN = off;
    off += align256(NORM_WORDS * sizeof(u64));
    const size_t oI = off;
    off += 256;
    const size_t oA = off;
    if (bf16) off += align256(np * 3 * nbo * 2);  // split (hi, hi, lo) rows of the outer block's L21
    const size_t oB = off;
    if (bf16) off += align256(np * 3 * nbo * 2);  // split (hi, lo, hi) columns of its U12
    const size_t oT = off;
    if (lowp) off += align256(tslu_scratch_bytes(np));
    const size_t oS = off;
    if (lowp) off += align256(2 * PAIR_LISTS * PAIR_STRIDE * sizeof(int));
    const size_t oD = off;
    if (refine) off += 2 * align256(trsv_diag_inverse_len(np) * sizeof(double));
    AL_CHECK(grow(&ws->mbuf, &ws->mbuf_bytes, off));
    char* base = (char*)ws->mbuf;
    AL_CHECK(fill_u32(s, (unsigned*)base, 1024, 0u));
    a->tslu = base + oT;
    a->pairs_sub = lowp ? (int*)(base + oS) : nullptr;
    a->dinv = refine ? (double*)(base + oD) : nullptr;
    a->counters = (unsigned*)base;
    a->F = (float*)(base + oF);
    a->ipiv = (int*)(base + oV);
    a->perm = (double*)(base + oW);
    a->rowidx = lowp ? (float*)(base + oX) : nullptr;
    a->R = (double*)(base + oR);
    a->Y = (double*)(base + oY);
    a->part = (double*)(base + oP);
    a->norms = (u64*)(base + oN);
    a->minfo = (int*)(base + oI);
    a->At = bf16 ? (__bf16*)(base + oA) : nullptr;
    a->Bt = bf16 ? (__bf16*)(base + oB) : nullptr;
    return AL_OK;
}


// One outer panel of the FP32 LU, columns [K, KE) and rows K..np, 64 columns at a time:
// tournament pivots, the interchanges applied to the outer panel's columns, L21 of the 64
// columns, then the rest of the outer panel updated (64-row U solve and a k = 64 product).
int factor_panel(hipStream_t q, Workspace* ws, Arena& a, int K, int KE, int* lists) {
    const int np = (int)a.np;
    const size_t ld = a.np;
    float* F = a.F;
    auto at = [&](int i, int j) { return F + (size_t)i + (size_t)j * ld; };
    for (int J = K; J < KE; J += PB) {
        int* pairs = lists + (size_t)((J - K) / PB) * PAIR_STRIDE;  // this block's interchanges
        AL_CHECK(tslu_panel(q, F, ld, J, np - J, a.ipiv, a.minfo, a.tslu, a.counters, pairs));
        AL_CHECK(swap_apply_f32(q, pairs, at(0, K), KE - K, nullptr, 0, ld, a.minfo));
        AL_CHECK(tslu_finish(q, F, ld, J, np - J, a.tslu, J + PB, KE));  // L21, and U12 on [J + 64, KE)
        if (J + PB < KE) {
            AL_CHECK(gemm_f32(q, false, false, np - J - PB, KE - J - PB, PB, -1.f, at(J + PB, J), ld, at(J, J + PB), ld,
                              1.f, at(J + PB, J + PB), ld));
        }
    }
    return AL_OK;
}

// A22[:, c0:c1) -= L21 U12[:, c0:c1) for the outer panel [K, KE): FP32 matrix cores, or the
// bf16 matrix cores with FP32 accumulation (L21 staged by the call with c0 == KE, U12 per
// call).  The bf16 form splits each FP32 operand into hi + lo bf16 parts and forms
// Lhi Uhi + Lhi Ulo + Llo Uhi as one product of depth 3 kb: with single bf16 operands
// (u = 2^-8) the factors are too coarse for FP64 refinement to converge at these orders, while
// the split keeps about 16 bits of every product.
int trailing_update(hipStream_t s, Arena& a, bool bf16, int K, int KE, int c0, int c1) {
    const int np = (int)a.np;
    const size_t ld = a.np;
    const int kb = KE - K, m2 = np - KE;
    float* F = a.F;
    auto at = [&](int i, int j) { return F + (size_t)i + (size_t)j * ld; };
    if (c1 <= c0) return AL_OK;
    if (bf16) {
        if (c0 == KE) AL_CHECK(bf16x3_rows(s, m2, kb, at(KE, K), ld, a.At));
        __bf16* Bt = a.Bt + (size_t)(c0 - KE) * 3 * kb;
        AL_CHECK(bf16x3_cols(s, kb, c1 - c0, at(K, c0), ld, Bt));
        return gemm_bf16(s, m2, c1 - c0, 3 * (size_t)kb, -1.f, a.At, Bt, 1.f, at(KE, c0), ld);
    }
    return gemm_f32(s, false, false, m2, c1 - c0, kb, -1.f, at(KE, K), ld, at(K, c0), ld, 1.f, at(KE, c0), ld);
}

// Blocked right-looking LU with partial pivoting of the np x np FP32 matrix a.F
// (np % 64 == 0); pivots as global row indices in a.ipiv, status in a.minfo[0].
//
// Look-ahead on two streams.  The side stream carries the chain that bounds the run: it
// factors an outer panel, then prepares the next panel's columns (the panel's interchanges,
// U12 and the product with L21 on those columns only) and factors that panel, and so on.  The
// main stream updates every other column behind it: once a panel is factored it applies the
// panel's interchanges to the columns left of it and right of the next panel, solves U12 and
// applies the trailing product there.  The next panel's preparation waits for that update of
// the previous panel (its columns are among those updated); nothing else crosses streams,
// because the two streams touch disjoint columns.
// A (n x n, FP64) rounded into columns [c0, c1) of the padded FP32 factor a.F.
int to_f32_cols(hipStream_t s, Arena& a, const double* A, size_t lda, size_t c0, size_t c1) {
    if (c1 <= c0) return AL_OK;
    const size_t i0 = c0 * a.np, i1 = c1 * a.np;
    hipLaunchKernelGGL(to_f32_pad_k, dim3((unsigned)ceil_div(i1 - i0, (size_t)256)), dim3(256), 0, s, (int)a.n,
                       (int)a.np, A, lda, a.norms, a.F, i0, i1);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

// A, when given, is rounded into F here: the first outer panel's columns before the panel
// stream starts, the rest on the main stream while that panel is factored.
int getrf_lowp(hipStream_t s, Workspace* ws, Arena& a, bool bf16, const double* A = nullptr, size_t lda = 0) {
    const int np = (int)a.np;
    const int nbo = bf16 ? 512 : 256;
    float* F = a.F;
    int* info = a.minfo;
    const size_t ld = a.np;
    auto at = [&](int i, int j) { return F + (size_t)i + (size_t)j * ld; };
    AL_CHECK(ensure_side_stream(ws));
    hipStream_t side = ws->side;
    AL_CHECK(fill_i32(s, info, 1, 0));
    AL_CHECK(fill_u32(s, ws->abort_flag, 1, 0u));
    AL_CHECK(iota_i32(s, a.ipiv, a.np));
    if (A) AL_CHECK(to_f32_cols(s, a, A, lda, 0, std::min<size_t>(a.np, (size_t)nbo)));
    hipLaunchKernelGGL(iota_f32_k, dim3((unsigned)ceil_div(a.np, (size_t)256)), dim3(256), 0, s, a.rowidx, np);
    AL_HIP(hipGetLastError());
    AL_HIP(hipEventRecord(ws->ev_main, s));  // also "the previous update is done" for the first panel
    AL_HIP(hipStreamWaitEvent(side, ws->ev_main, 0));
    // the outer panel [K, KE)'s interchanges (its 64-pivot blocks' pair lists, in order) on
    // columns [c0, c0 + n0) and [c2, c2 + n1)
    auto lists = [&](int K) { return a.pairs_sub + (size_t)((K / nbo) & 1) * PAIR_LISTS * PAIR_STRIDE; };
    auto swaps = [&](hipStream_t q, int K, int KE, int c0, int n0, int c2, int n1) -> int {
        return swap_apply_multi_f32(q, lists(K), PAIR_STRIDE, (KE - K) / PB, at(0, c0), n0,
                                    n1 > 0 ? at(0, c2) : nullptr, n1, ld, info);
    };
    // U12 := L11^{-1} A12 on columns [c0, c1) of a full outer panel [K, K + nbo), one launch
    auto u12 = [&](hipStream_t q, int K, int c0, int c1) -> int {
        if (c1 <= c0) return AL_OK;
        const unsigned wgs = (unsigned)std::min<size_t>(ceil_div((size_t)(c1 - c0), (size_t)16), (size_t)512);
        if (bf16)
            hipLaunchKernelGGL(u12_k<8>, dim3(wgs), dim3(256), 0, q, F, ld, K, c0, c1);
        else
            hipLaunchKernelGGL(u12_k<4>, dim3(wgs), dim3(256), 0, q, F, ld, K, c0, c1);
        AL_HIP(hipGetLastError());
        return AL_OK;
    };
    AL_CHECK(factor_panel(side, ws, a, 0, np < nbo ? np : nbo, lists(0)));
    if (A && np > nbo) {
        // the rest of F while the first panel is factored; the panel stream's next wait covers it
        AL_CHECK(to_f32_cols(s, a, A, lda, (size_t)nbo, a.np));
        AL_HIP(hipEventRecord(ws->ev_main, s));
    }
    for (int K = 0; K < np; K += nbo) {
        const int kb = (np - K) < nbo ? (np - K) : nbo;
        const int KE = K + kb;
        if (KE >= np) {  // the last panel: only the columns left of it remain
            AL_HIP(hipEventRecord(ws->ev_side, side));
            AL_HIP(hipStreamWaitEvent(s, ws->ev_side, 0));
            AL_CHECK(swaps(s, K, KE, 0, K, 0, 0));
            AL_CHECK(swap_apply_multi_f32(s, lists(K), PAIR_STRIDE, (KE - K) / PB, a.rowidx, 1, nullptr, 0, ld, info));
            break;
        }
        const int KE2 = (np - KE) < nbo ? np : KE + nbo;
        // side: the next panel's columns, after the main stream's last update of them
        AL_HIP(hipStreamWaitEvent(side, ws->ev_main, 0));
        AL_CHECK(swaps(side, K, KE, KE, KE2 - KE, 0, 0));
        AL_CHECK(u12(side, K, KE, KE2));  // kb == nbo: only the last panel is shorter
        AL_CHECK(trailing_update(side, a, bf16, K, KE, KE, KE2));
        AL_HIP(hipEventRecord(ws->ev_side, side));
        AL_CHECK(factor_panel(side, ws, a, KE, KE2, lists(KE)));  // the other list set
        // main: every other column
        AL_HIP(hipStreamWaitEvent(s, ws->ev_side, 0));
        AL_CHECK(swaps(s, K, KE, 0, K, KE2, np - KE2));
        AL_CHECK(swap_apply_multi_f32(s, lists(K), PAIR_STRIDE, (KE - K) / PB, a.rowidx, 1, nullptr, 0, ld, info));
        AL_CHECK(u12(s, K, KE2, np));
        AL_CHECK(trailing_update(s, a, bf16, K, KE, KE2, np));
        AL_HIP(hipEventRecord(ws->ev_main, s));
    }
    // the row origins: every outer panel's interchanges were applied to 0, 1, .., np-1 in order
    hipLaunchKernelGGL(widen_k, dim3((unsigned)ceil_div(a.np, (size_t)256)), dim3(256), 0, s, a.rowidx, a.perm, np);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

template <int NR>
int launch_gemv(hipStream_t s, const Arena& a, size_t nr, const double* A, size_t lda, const double* X, size_t ldx) {
    const dim3 grid(ceil_div(a.n, GV_ROWS), (unsigned)a.nparts);
    hipLaunchKernelGGL((gemv_part_k<NR, false>), grid, dim3(256), 0, s, (int)a.n, (int)nr, A, lda, X, ldx, a.part);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

// R = B - A X with the norms of R and X per column (into a.norms).
int residual(hipStream_t s, const Arena& a, size_t nr, const double* A, size_t lda, const double* B, size_t ldb,
             const double* X, size_t ldx) {
    const int ticket = prof_begin(s);
    if (nr <= 1)
        AL_CHECK(launch_gemv<1>(s, a, nr, A, lda, X, ldx));
    else if (nr <= 2)
        AL_CHECK(launch_gemv<2>(s, a, nr, A, lda, X, ldx));
    else if (nr <= 4)
        AL_CHECK(launch_gemv<4>(s, a, nr, A, lda, X, ldx));
    else if (nr <= 8)
        AL_CHECK(launch_gemv<8>(s, a, nr, A, lda, X, ldx));
    else
        AL_CHECK(launch_gemv<16>(s, a, nr, A, lda, X, ldx));
    if (ticket >= 0) {
        const double dn = (double)a.n;
        prof_end(ticket, s, "residual_gemv", 2.0 * dn * dn * nr, 8.0 * dn * dn + 8.0 * (double)a.nparts * dn * nr);
    }
    hipLaunchKernelGGL(resid_k, dim3(ceil_div(a.n, 256)), dim3(256), 0, s, (int)a.n, (int)nr, (int)a.nparts, a.part, B,
                       ldb, X, ldx, a.R, a.norms + NORM_R, a.norms + NORM_X);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

// Reads the norm words back, and the workspace's abort flag into host[NORM_ABORT] (the device
// word there is never written): a raised flag means a solve's block wait gave up.
int read_norms(hipStream_t s, Workspace* ws, const Arena& a, u64* host) {
    AL_HIP(hipMemcpyAsync(host, a.norms, NORM_WORDS * sizeof(u64), hipMemcpyDeviceToHost, s));
    AL_HIP(hipMemcpyAsync(host + NORM_ABORT, ws->abort_flag, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    AL_HIP(hipStreamSynchronize(s));
    if ((unsigned)host[NORM_ABORT] != 0u) {
        set_error("dgesv_mixed: a device-side wait timed out in a triangular solve");
        return AL_ERR_DEVICE;
    }
    return AL_OK;
}

double as_double(u64 b) {
    double d;
    std::memcpy(&d, &b, sizeof d);
    return d;
}

// Normwise backward error of one column; a zero residual counts as exact.
double berr_of(double r, double anorm, double x, double b) {
    if (r == 0.0) return 0.0;
    const double den = anorm * x + b;
    return r / den;  // NaN or inf when the iterate has blown up
}

// FP64 pivoted LU solve on a copy of A: X := A^{-1} B.
int fallback_fp64(hipStream_t s, Workspace* ws, Arena& a, size_t n, size_t nrhs, const double* A, size_t lda,
                  const double* B, size_t ldb, double* X, size_t ldx, int* info_out) {
    const size_t mat = align256(n * n * sizeof(double));
    AL_CHECK(grow(&ws->fbuf, &ws->fbuf_bytes, mat + align256(n * sizeof(int))));
    double* Af = (double*)ws->fbuf;
    int* ipiv = (int*)((char*)ws->fbuf + mat);
    AL_CHECK(ws_reserve(ws, n, n * nrhs));
    hipLaunchKernelGGL(copy2d_k, dim3(ceil_div(n * n, 256)), dim3(256), 0, s, (int)n, (int)n, A, lda, Af, n);
    AL_HIP(hipGetLastError());
    hipLaunchKernelGGL(copy2d_k, dim3(ceil_div(n * nrhs, 256)), dim3(256), 0, s, (int)n, (int)nrhs, B, ldb, X, ldx);
    AL_HIP(hipGetLastError());
    int* info = a.minfo + 2;
    AL_CHECK(fill_i32(s, info, 1, 0));
    AL_CHECK(iota_i32(s, ipiv, n));
    AL_CHECK(getrf_f64(s, ws, n, n, Af, n, ipiv, info, ws->perm));
    AL_CHECK(getrs_f64(s, ws, n, nrhs, Af, n, ipiv, ws->perm, X, ldx, info));
    AL_HIP(hipMemcpyAsync(info_out, info, sizeof(int), hipMemcpyDeviceToHost, s));
    AL_HIP(hipStreamSynchronize(s));
    return AL_OK;
}

// The first fallback: the padded system in the low-precision factor's row order (tournament
// pivots, chosen for A itself) factored in FP64 without any pivot search -- the row order keeps
// the growth as small as the low-precision factorization's, so no column needs a grid-wide
// exchange -- then FP64 refinement with that factor.  *ok reports a result within opt.tol; it is
// false (and X unspecified) when a pivot in this order is zero or refinement stalls, and the
// caller then runs the pivoted FP64 LU.
int reordered_fp64(hipStream_t s, Workspace* ws, Arena& a, size_t n, size_t nrhs, const double* A, size_t lda,
                   const double* B, size_t ldb, double* X, size_t ldx, const al_ir_options& opt, u64* hn, bool* ok,
                   double* berr_out, int* iters_out) {
    *ok = false;
    const size_t np = a.np;
    AL_CHECK(grow(&ws->fbuf, &ws->fbuf_bytes, align256(np * np * sizeof(double)) + align256(np * sizeof(int))));
    double* Af = (double*)ws->fbuf;
    hipLaunchKernelGGL(gather_pad_k, dim3((unsigned)np), dim3(256), 0, s, (int)n, (int)np, A, lda, a.perm, a.norms, Af);
    AL_HIP(hipGetLastError());
    int* info = a.minfo + 3;
    AL_CHECK(getrf_nopiv_f64(s, ws, np, Af, np, info));
    hipLaunchKernelGGL(diag_minmax_k, dim3(1), dim3(1024), 0, s, (int)np, Af, np, a.norms + NORM_DIAG);
    AL_HIP(hipGetLastError());
    int hinfo = 0;
    AL_HIP(hipMemcpyAsync(&hinfo, info, sizeof hinfo, hipMemcpyDeviceToHost, s));
    AL_HIP(hipMemcpyAsync(hn + NORM_DIAG, a.norms + NORM_DIAG, 2 * sizeof(u64), hipMemcpyDeviceToHost, s));
    AL_HIP(hipStreamSynchronize(s));
    if (hinfo != 0) return AL_OK;
    // a pivot at rounding level in this order: the matrix may be singular, which only the
    // pivoted factorization decides exactly (the crate's zero-pivot test)
    if (!(as_double(hn[NORM_DIAG]) > (double)np * DBL_EPSILON * as_double(hn[NORM_DIAG + 1]))) return AL_OK;
    double worst = 0.0;
    int iters = 0;
    bool all_ok = true;
    const int max_it = opt.max_iter < 10 ? opt.max_iter : 10;
    for (size_t c0 = 0; c0 < nrhs; c0 += NR_MAX) {
        const size_t nr = (nrhs - c0) < (size_t)NR_MAX ? (nrhs - c0) : (size_t)NR_MAX;
        const double* Bc = B + c0 * ldb;
        double* Xc = X + c0 * ldx;
        AL_CHECK(fill_u32(s, (unsigned*)(a.norms + NORM_B), 2 * (NORM_WORDS - NORM_B), 0u));
        hipLaunchKernelGGL(colmax_k, dim3(ceil_div(n, 256)), dim3(256), 0, s, (int)n, (int)nr, Bc, ldb, a.norms + NORM_B);
        AL_HIP(hipGetLastError());
        hipLaunchKernelGGL(copy2d_k, dim3(ceil_div(n * nr, 256)), dim3(256), 0, s, (int)n, (int)nr, Bc, ldb, a.R, n);
        AL_HIP(hipGetLastError());
        double prev = INFINITY, berr = INFINITY;
        int it = 0;
        while (it < max_it) {
            ++it;
            hipLaunchKernelGGL(gather_perm_k, dim3(ceil_div(np * nr, 256)), dim3(256), 0, s, (int)n, (int)np, (int)nr,
                               a.perm, a.R, n, a.Y);
            AL_HIP(hipGetLastError());
            AL_CHECK(trsv_solve<double>(s, ws, true, true, np, nr, Af, np, a.Y, np));
            AL_CHECK(trsv_solve<double>(s, ws, false, false, np, nr, Af, np, a.Y, np));
            hipLaunchKernelGGL(add_k, dim3(ceil_div(n * nr, 256)), dim3(256), 0, s, (int)n, (int)np, (int)nr, a.Y, Xc,
                               ldx, it > 1 ? 1 : 0);
            AL_HIP(hipGetLastError());
            AL_CHECK(fill_u32(s, (unsigned*)(a.norms + NORM_R), 2 * (NORM_WORDS - NORM_R), 0u));
            AL_CHECK(residual(s, a, nr, A, lda, Bc, ldb, Xc, ldx));
            AL_CHECK(read_norms(s, ws, a, hn));
            const double anorm = as_double(hn[NORM_A]);
            berr = 0.0;
            for (size_t k = 0; k < nr; ++k) {
                const double e = berr_of(as_double(hn[NORM_R + k]), anorm, as_double(hn[NORM_X + k]),
                                         as_double(hn[NORM_B + k]));
                berr = (e > berr || std::isnan(e)) ? e : berr;
            }
            if (berr <= opt.stop_tol) break;
            if (!(berr < 0.5 * prev)) break;
            prev = berr;
        }
        iters += it;
        if (!(berr <= opt.tol)) all_ok = false;
        worst = (berr > worst || std::isnan(berr)) ? berr : worst;
    }
    *ok = all_ok;
    *berr_out = worst;
    *iters_out = iters;
    return AL_OK;
}

}  // namespace

int gesv_mixed(hipStream_t s, Workspace* ws, size_t n, size_t nrhs, const double* A, size_t lda, const double* B,
               size_t ldb, double* X, size_t ldx, int factor, const al_ir_options* opt_in, al_ir_stats* st,
               const MixedFactor* reuse, MixedFactor* keep) {
    al_ir_options opt;
    al_ir_default_options(&opt);
    if (opt_in) {
        opt = *opt_in;
        if (opt.max_iter <= 0) opt.max_iter = 30;
        if (!(opt.tol > 0.0)) opt.tol = 1e-14;
        if (!(opt.stop_tol > 0.0)) opt.stop_tol = 1e-15;
    }
    al_ir_stats stats{};
    if (n == 0 || nrhs == 0) {
        stats.converged = 1;
        if (st) *st = stats;
        return AL_OK;
    }
    if (n > 32768 - PB) {
        set_error("dgesv_mixed: n beyond 32704 is not supported (cooperative panel height)");
        return AL_ERR_ARG;
    }
    if (factor != AL_FACTOR_FP64 && factor != AL_FACTOR_FP32 && factor != AL_FACTOR_BF16) {
        set_error("dgesv_mixed: unknown factor precision");
        return AL_ERR_ARG;
    }
    const bool bf16 = factor == AL_FACTOR_BF16;
    const size_t np = (n + PB - 1) / PB * PB;
    const size_t nr_max = nrhs < (size_t)NR_MAX ? nrhs : (size_t)NR_MAX;
    if (reuse && (reuse->n != n || reuse->np != np || factor == AL_FACTOR_FP64)) {
        set_error("dgesv_mixed: the resident factor does not match the system");
        return AL_ERR_ARG;
    }
    Arena a;
    AL_CHECK(make_arena(s, ws, n, np, nr_max, bf16 ? 512 : 256, factor != AL_FACTOR_FP64 && !reuse, bf16,
                        factor != AL_FACTOR_FP64, &a));
    if (reuse) {
        a.F = reuse->F;
        a.perm = reuse->perm;
    }
    if (!ws->hnorms) AL_HIP(hipHostMalloc((void**)&ws->hnorms, NORM_WORDS * sizeof(u64), hipHostMallocDefault));
    u64* hn = (u64*)ws->hnorms;
    // a flag raised by an earlier call's timed-out wait must not fail this one's solves
    AL_CHECK(fill_u32(s, ws->abort_flag, 1, 0u));

    // |A|_inf (also the padding scale)
    AL_CHECK(fill_u32(s, (unsigned*)a.norms, 2 * NORM_WORDS, 0u));
    {
        const dim3 grid(ceil_div(n, GV_ROWS), (unsigned)a.nparts);
        hipLaunchKernelGGL((gemv_part_k<1, true>), grid, dim3(256), 0, s, (int)n, 1, A, lda, (const double*)nullptr,
                           (size_t)0, a.part);
        AL_HIP(hipGetLastError());
        hipLaunchKernelGGL(rowsum_max_k, dim3(ceil_div(n, 256)), dim3(256), 0, s, (int)n, (int)a.nparts, a.part,
                           a.norms + NORM_A);
        AL_HIP(hipGetLastError());
    }

    double berr_low = INFINITY;
    bool ok_low = false;
    int iters = 0;
    if (factor != AL_FACTOR_FP64) {
        int minfo = 0;
        if (reuse) {
            minfo = reuse->breakdown;
        } else {
            AL_CHECK(getrf_lowp(s, ws, a, bf16, A, lda));  // rounds A into F as it goes
            AL_HIP(hipMemcpyAsync(&minfo, a.minfo, sizeof minfo, hipMemcpyDeviceToHost, s));
            AL_HIP(hipStreamSynchronize(s));
            if (minfo < 0) {
                set_error("dgesv_mixed: a device-side wait timed out in the FP32 factorization");
                return AL_ERR_DEVICE;
            }
            if (keep && keep->F && keep->perm) {
                AL_HIP(hipMemcpyAsync(keep->F, a.F, np * np * sizeof(float), hipMemcpyDeviceToDevice, s));
                AL_HIP(hipMemcpyAsync(keep->perm, a.perm, np * sizeof(double), hipMemcpyDeviceToDevice, s));
                keep->n = n;
                keep->np = np;
                keep->breakdown = minfo;
            }
        }
        stats.factor_breakdown = minfo;
        if (minfo == 0) {
            // the correction solves apply the factor's 128-row diagonal blocks as FP64 inverses
            AL_CHECK(trsv_diag_inverse(s, np, a.F, np, a.dinv, a.dinv + trsv_diag_inverse_len(np)));
            berr_low = 0.0;
            ok_low = true;
            for (size_t c0 = 0; c0 < nrhs; c0 += NR_MAX) {
                const size_t nr = (nrhs - c0) < (size_t)NR_MAX ? (nrhs - c0) : (size_t)NR_MAX;
                const double* Bc = B + c0 * ldb;
                double* Xc = X + c0 * ldx;
                AL_CHECK(fill_u32(s, (unsigned*)(a.norms + NORM_B), 2 * (NORM_WORDS - NORM_B), 0u));
                hipLaunchKernelGGL(colmax_k, dim3(ceil_div(n, 256)), dim3(256), 0, s, (int)n, (int)nr, Bc, ldb,
                                   a.norms + NORM_B);
                AL_HIP(hipGetLastError());
                hipLaunchKernelGGL(copy2d_k, dim3(ceil_div(n * nr, 256)), dim3(256), 0, s, (int)n, (int)nr, Bc, ldb, a.R,
                                   n);
                AL_HIP(hipGetLastError());
                double prev = INFINITY, berr = INFINITY;
                int it = 0;
                while (it < opt.max_iter) {
                    ++it;
                    hipLaunchKernelGGL(gather_perm_k, dim3(ceil_div(np * nr, 256)), dim3(256), 0, s, (int)n, (int)np,
                                       (int)nr, a.perm, a.R, n, a.Y);
                    AL_HIP(hipGetLastError());
                    AL_CHECK(trsv_solve<float>(s, ws, true, true, np, nr, a.F, np, a.Y, np, a.dinv));
                    AL_CHECK(trsv_solve<float>(s, ws, false, false, np, nr, a.F, np, a.Y, np,
                                               a.dinv + trsv_diag_inverse_len(np)));
                    hipLaunchKernelGGL(add_k, dim3(ceil_div(n * nr, 256)), dim3(256), 0, s, (int)n, (int)np, (int)nr, a.Y,
                                       Xc, ldx, it > 1 ? 1 : 0);
                    AL_HIP(hipGetLastError());
                    AL_CHECK(fill_u32(s, (unsigned*)(a.norms + NORM_R), 2 * (NORM_WORDS - NORM_R), 0u));
                    AL_CHECK(residual(s, a, nr, A, lda, Bc, ldb, Xc, ldx));
                    AL_CHECK(read_norms(s, ws, a, hn));
                    const double anorm = as_double(hn[NORM_A]);
                    berr = 0.0;
                    for (size_t k = 0; k < nr; ++k) {
                        const double e = berr_of(as_double(hn[NORM_R + k]), anorm, as_double(hn[NORM_X + k]),
                                                 as_double(hn[NORM_B + k]));
                        berr = (e > berr || std::isnan(e)) ? e : berr;
                    }
                    if (berr <= opt.stop_tol) break;
                    if (!(berr < 0.5 * prev)) break;  // stalled or diverging (NaN included)
                    prev = berr;
                }
                iters += it;
                if (!(berr <= opt.tol)) ok_low = false;
                berr_low = (berr > berr_low || std::isnan(berr)) ? berr : berr_low;
            }
        }
    }
    stats.iterations = iters;
    stats.backward_error_low = berr_low;
    stats.backward_error = berr_low;
    stats.converged = ok_low ? 1 : 0;
    int status = AL_OK;
    bool reordered_ok = false;
    if (!ok_low && factor != AL_FACTOR_FP64 && opt.allow_fallback && stats.factor_breakdown == 0) {
        // the low-precision factor's row order is complete: FP64 LU in that order first
        double be2 = INFINITY;
        int it2 = 0;
        AL_CHECK(reordered_fp64(s, ws, a, n, nrhs, A, lda, B, ldb, X, ldx, opt, hn, &reordered_ok, &be2, &it2));
        if (reordered_ok) {
            stats.fell_back = 1;
            stats.iterations += it2;
            stats.backward_error = be2;
            stats.converged = 1;
        }
    }
    if (!ok_low && !reordered_ok) {
        if (factor == AL_FACTOR_FP64 || opt.allow_fallback) {
            int info = 0;
            AL_CHECK(fallback_fp64(s, ws, a, n, nrhs, A, lda, B, ldb, X, ldx, &info));
            stats.fell_back = factor == AL_FACTOR_FP64 ? 0 : 2;
            if (info > 0) {
                status = AL_SINGULAR;
                stats.converged = 0;
            } else if (info < 0) {
                set_error("dgesv_mixed: a device-side wait timed out in the FP64 factorization");
                return AL_ERR_DEVICE;
            } else {
                // backward error of the FP64 result, for the record
                double worst = 0.0;
                for (size_t c0 = 0; c0 < nrhs; c0 += NR_MAX) {
                    const size_t nr = (nrhs - c0) < (size_t)NR_MAX ? (nrhs - c0) : (size_t)NR_MAX;
                    AL_CHECK(fill_u32(s, (unsigned*)(a.norms + NORM_B), 2 * (NORM_WORDS - NORM_B), 0u));
                    hipLaunchKernelGGL(colmax_k, dim3(ceil_div(n, 256)), dim3(256), 0, s, (int)n, (int)nr, B + c0 * ldb,
                                       ldb, a.norms + NORM_B);
                    AL_HIP(hipGetLastError());
                    AL_CHECK(residual(s, a, nr, A, lda, B + c0 * ldb, ldb, X + c0 * ldx, ldx));
                    AL_CHECK(read_norms(s, ws, a, hn));
                    const double anorm = as_double(hn[NORM_A]);
                    for (size_t k = 0; k < nr; ++k) {
                        const double e = berr_of(as_double(hn[NORM_R + k]), anorm, as_double(hn[NORM_X + k]),
                                                 as_double(hn[NORM_B + k]));
                        worst = (e > worst || std::isnan(e)) ? e : worst;
                    }
                }
                stats.backward_error = worst;
                stats.converged = worst <= opt.tol ? 1 : 0;
                if (!stats.converged) {
                    status = AL_CONVERGENCE;
                    set_error("dgesv_mixed: the FP64 solve did not reach the requested backward error");
                }
            }
        } else {
            status = AL_CONVERGENCE;
            set_error("dgesv_mixed: refinement did not reach the requested backward error");
        }
    }
    if (st) *st = stats;
    return status;
}

}  // namespace al

extern "C" void al_ir_default_options(al_ir_options* opt) {
    if (!opt) return;
    opt->max_iter = 30;
    opt->allow_fallback = 1;
    opt->tol = 1e-14;
    opt->stop_tol = 1e-15;
}
