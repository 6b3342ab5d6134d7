// bf16 GEMM: C := alpha*A*B + beta*C with the operands rounded to bfloat16 and every product
// accumulated in FP32 on the matrix cores (v_mfma_f32_16x16x32_bf16), C held in FP32.  It
// is the trailing update of the mixed-precision LU (alpha = -1, beta = 1) and, through the
// device API, a GEMM of its own.
//
// Operands arrive k-contiguous: At is m x k row-major (A^T column-major) and Bt is n x k
// row-major (B column-major), both bf16, as written by the conversion kernels below.  One
// 256-thread workgroup computes a 128x128 tile of C (2x2 waves of 64x64, each 4x4 MFMA tiles
// of 16x16); K advances 64 at a time through double-buffered LDS rows padded to 72 elements
// (144 B) so the 16-byte fragment reads of 16 consecutive rows fall on distinct banks.  The
// operands are fed as (B, A) so each lane's accumulators hold C[row][col .. col+3] with 16
// lanes spanning 16 consecutive rows: the FP32 read-modify-write of the column-major C is
// made of contiguous 64-byte runs.
#include "al_internal.hpp"

namespace al {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int TM = 128, TN = 128, TK = 64;
constexpr int LROW = TK + 8;                    // padded LDS row (bf16 elements)
constexpr int TILE_ELEMS = TM * LROW;           // one operand tile
constexpr size_t LDS_BYTES = 4 * TILE_ELEMS * 2;  // A and B, double-buffered: 72 KiB
constexpr int CHUNKS = TM * TK / 8 / 256;        // 16-byte chunks per thread per operand tile

// Loads rows [r0, r0 + 128) x k-range [k0, k0 + 64) of a k-contiguous bf16 operand; rows at or
// beyond r_lim read as zero.
__device__ __forceinline__ void load_op(const __bf16* __restrict__ P, size_t ldk, int r0, int r_lim, int k0,
                                        uint4 (&reg)[CHUNKS], int t) {
#pragma unroll
    for (int s = 0; s < CHUNKS; ++s) {
        const int c = t + 256 * s;
        const int row = c >> 3, kc = c & 7;
        if (r0 + row < r_lim)
            reg[s] = *(const uint4*)(P + (size_t)(r0 + row) * ldk + k0 + 8 * kc);
        else
            reg[s] = uint4{0u, 0u, 0u, 0u};
    }
}

__device__ __forceinline__ void store_op(__bf16* S, const uint4 (&reg)[CHUNKS], int t) {
#pragma unroll
    for (int s = 0; s < CHUNKS; ++s) {
        const int c = t + 256 * s;
        const int row = c >> 3, kc = c & 7;
        *(uint4*)(S + row * LROW + 8 * kc) = reg[s];
    }
}

__global__ __launch_bounds__(256, 2) void gemm_bf16_k(int m, int n, int k, float alpha, const __bf16* __restrict__ At,
                                                      const __bf16* __restrict__ Bt, float beta, float* __restrict__ C,
                                                      size_t ldc, int tiles_m, int tiles_n) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __bf16* lds = (__bf16*)lds_raw;
    __bf16* As[2] = {lds, lds + TILE_ELEMS};
    __bf16* Bs[2] = {lds + 2 * TILE_ELEMS, lds + 3 * TILE_ELEMS};

    int tm, tn;
    tile_coords(tiles_m, tiles_n, tm, tn);
    const int bm = tm * TM, bn = tn * TN;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int fr = lane & 15, fk = 8 * (lane >> 4);

    f4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

    uint4 ra[CHUNKS], rb[CHUNKS];
    const int ktiles = k / TK;  // the caller guarantees k % 64 == 0
    load_op(At, (size_t)k, bm, m, 0, ra, t);
    load_op(Bt, (size_t)k, bn, n, 0, rb, t);
    store_op(As[0], ra, t);
    store_op(Bs[0], rb, t);
    __syncthreads();

    for (int kt = 0; kt < ktiles; ++kt) {
        const bool more = (kt + 1) < ktiles;
        if (more) {
            load_op(At, (size_t)k, bm, m, (kt + 1) * TK, ra, t);
            load_op(Bt, (size_t)k, bn, n, (kt + 1) * TK, rb, t);
        }
        const __bf16* Ac = As[kt & 1];
        const __bf16* Bc = Bs[kt & 1];
#pragma unroll
        for (int kk = 0; kk < TK; kk += 32) {
            bf8 af[4], bfv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) af[i] = *(const bf8*)(Ac + (wm * 64 + i * 16 + fr) * LROW + kk + fk);
#pragma unroll
            for (int j = 0; j < 4; ++j) bfv[j] = *(const bf8*)(Bc + (wn * 64 + j * 16 + fr) * LROW + kk + fk);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfv[j], af[i], acc[i][j], 0, 0, 0);
        }
        if (more) {
            store_op(As[(kt + 1) & 1], ra, t);
            store_op(Bs[(kt + 1) & 1], rb, t);
        }
        __syncthreads();
    }

    // D = (B-tile)(A-tile)^T = C^T: accumulator r of lane l holds D[4 (l >> 4) + r][l & 15],
    // i.e. C[row = .. + (l & 15)][col = .. + 4 (l >> 4) + r]
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int row = bm + wm * 64 + i * 16 + fr;
        if (row >= m) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int cc = bn + wn * 64 + j * 16 + fk / 2 + r;
                if (cc >= n) continue;
                float* cp = C + (size_t)row + (size_t)cc * ldc;
                const float v = alpha * acc[i][j][r];
                *cp = beta == 0.f ? v : v + beta * *cp;
            }
        }
    }
}

// At[r][p] = bf16(L[r + p*ld]) for an m x k column-major FP32 block, rows of kp >= k entries
// (zero beyond k): a 64x64 tile transpose through LDS (reads down columns, writes along
// rows, both coalesced).
__global__ __launch_bounds__(256) void to_bf16_rows_k(int m, int k, int kp, const float* __restrict__ L, size_t ld,
                                                       __bf16* __restrict__ At) {
    __shared__ float tile[64][65];
    const int r0 = blockIdx.x * 64, p0 = blockIdx.y * 64;
    const int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int idx = t + 256 * q;
        const int r = idx & 63, p = idx >> 6;
        tile[p][r] = (r0 + r < m && p0 + p < k) ? L[(size_t)(r0 + r) + (size_t)(p0 + p) * ld] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int idx = t + 256 * q;
        const int p = idx & 63, r = idx >> 6;
        if (r0 + r < m && p0 + p < kp) At[(size_t)(r0 + r) * kp + p0 + p] = (__bf16)tile[p][r];
    }
}

// Bt[c][p] = bf16(U[p + c*ld]) for a k x n column-major FP32 block (already k-contiguous),
// rows of kp >= k entries, zero beyond k.
__global__ __launch_bounds__(256) void to_bf16_cols_k(int k, int kp, int n, const float* __restrict__ U, size_t ld,
                                                       __bf16* __restrict__ Bt) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)kp * n) return;
    const int p = (int)(idx % kp), c = (int)(idx / kp);
    Bt[idx] = p < k ? (__bf16)U[(size_t)p + (size_t)c * ld] : (__bf16)0.f;
}

// Two-term bf16 split of an FP32 value: hi = bf16(x), lo = bf16(x - hi) (x - hi is exact in
// FP32), so hi + lo carries x to about 2^-17 relative.
__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
    hi = (__bf16)x;
    lo = (__bf16)(x - (float)hi);
}

// At[r] = [hi(L[r, 0..k)) | hi(L[r, 0..k)) | lo(L[r, 0..k))], rows of 3k entries: the left
// operand of the split product sum_p At[r][p] Bt[c][p] = Lhi Uhi + Lhi Ulo + Llo Uhi.
__global__ __launch_bounds__(256) void to_bf16x3_rows_k(int m, int k, const float* __restrict__ L, size_t ld,
                                                         __bf16* __restrict__ At) {
    __shared__ float tile[64][65];
    const int r0 = blockIdx.x * 64, p0 = blockIdx.y * 64;
    const int t = threadIdx.x;
    const size_t row_len = 3 * (size_t)k;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int idx = t + 256 * q;
        const int r = idx & 63, p = idx >> 6;
        tile[p][r] = (r0 + r < m && p0 + p < k) ? L[(size_t)(r0 + r) + (size_t)(p0 + p) * ld] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int idx = t + 256 * q;
        const int p = idx & 63, r = idx >> 6;
        if (r0 + r < m && p0 + p < k) {
            __bf16 hi, lo;
            split_bf16(tile[p][r], hi, lo);
            __bf16* row = At + (size_t)(r0 + r) * row_len + p0 + p;
            row[0] = hi;
            row[k] = hi;
            row[2 * (size_t)k] = lo;
        }
    }
}

// Bt[c] = [hi(U[0..k, c]) | lo(U[0..k, c]) | hi(U[0..k, c])], rows of 3k entries.
__global__ __launch_bounds__(256) void to_bf16x3_cols_k(int k, int n, const float* __restrict__ U, size_t ld,
                                                         __bf16* __restrict__ Bt) {
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= (size_t)k * n) return;
    const int p = (int)(idx % k), c = (int)(idx / k);
    __bf16 hi, lo;
    split_bf16(U[(size_t)p + (size_t)c * ld], hi, lo);
    __bf16* row = Bt + (size_t)c * 3 * (size_t)k + p;
    row[0] = hi;
    row[k] = lo;
    row[2 * (size_t)k] = hi;
}

}  // namespace

int bf16x3_rows(hipStream_t s, size_t m, size_t k, const float* L, size_t ld, __bf16* At) {
    if (m == 0 || k == 0) return AL_OK;
    hipLaunchKernelGGL(to_bf16x3_rows_k, dim3(ceil_div(m, 64), ceil_div(k, 64)), dim3(256), 0, s, (int)m, (int)k, L,
                       ld, At);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

int bf16x3_cols(hipStream_t s, size_t k, size_t n, const float* U, size_t ld, __bf16* Bt) {
    if (n == 0 || k == 0) return AL_OK;
    hipLaunchKernelGGL(to_bf16x3_cols_k, dim3(ceil_div(k * n, 256)), dim3(256), 0, s, (int)k, (int)n, U, ld, Bt);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

int bf16_rows(hipStream_t s, size_t m, size_t k, size_t kp, const float* L, size_t ld, __bf16* At) {
    if (m == 0 || kp == 0) return AL_OK;
    hipLaunchKernelGGL(to_bf16_rows_k, dim3(ceil_div(m, 64), ceil_div(kp, 64)), dim3(256), 0, s, (int)m, (int)k, (int)kp,
                       L, ld, At);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

int bf16_cols(hipStream_t s, size_t k, size_t kp, size_t n, const float* U, size_t ld, __bf16* Bt) {
    if (n == 0 || kp == 0) return AL_OK;
    hipLaunchKernelGGL(to_bf16_cols_k, dim3(ceil_div(kp * n, 256)), dim3(256), 0, s, (int)k, (int)kp, (int)n, U, ld,
                       Bt);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

int gemm_bf16(hipStream_t s, size_t m, size_t n, size_t k, float alpha, const __bf16* At, const __bf16* Bt, float beta,
              float* C, size_t ldc) {
    if (m == 0 || n == 0 || k == 0) return AL_OK;
    if (k % TK != 0 || m > INT32_MAX || n > INT32_MAX || k > INT32_MAX) {
        set_error("gemm_bf16: k must be a multiple of 64");
        return AL_ERR_ARG;
    }
    static bool attr_set = false;  // benign race: every thread sets the same value
    if (!attr_set) {
        AL_HIP(hipFuncSetAttribute((const void*)gemm_bf16_k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)LDS_BYTES));
        attr_set = true;
    }
    const int tiles_m = (int)ceil_div(m, TM), tiles_n = (int)ceil_div(n, TN);
    const int ticket = prof_begin(s);
    hipLaunchKernelGGL(gemm_bf16_k, dim3(tiles_m * tiles_n), dim3(256), LDS_BYTES, s, (int)m, (int)n, (int)k, alpha, At,
                       Bt, beta, C, ldc, tiles_m, tiles_n);
    AL_HIP(hipGetLastError());
    if (ticket >= 0) {
        const double dm = m, dn = n, dk = k;
        prof_end(ticket, s, "bgemm", 2.0 * dm * dn * dk, 2.0 * (dm * dk + dk * dn) + (beta != 0.f ? 8.0 : 4.0) * dm * dn);
    }
    return AL_OK;
}

}  // namespace al
