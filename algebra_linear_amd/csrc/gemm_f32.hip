// FP32 C := alpha*A*B + beta*C (neither operand transposed) on v_mfma_f32_16x16x4_f32: the
// trailing and panel updates of the FP32 factorization of the mixed-precision solve (the
// low-precision counterpart of blas::dgemm's no-transpose case, src/blas.rs:201-227).
//
// FP32 matrix cores do a 16x16x4 product in half the time of FP64, so per flop the operand
// path must move twice the bytes; this kernel is built around 16-byte accesses everywhere:
//   * one 256-thread workgroup (4 wave64s, 2x2) computes a 128x128 tile of C, K in steps of
//     32 through double-buffered LDS (69 KiB, two workgroups per CU), one barrier per step;
//   * global loads are float4: four rows of a column of A, four k of a column of B.  A's float4
//     lands in LDS k-major (rows padded to 132 floats); B's stays column-major in LDS (columns
//     padded to 36 floats, an odd number of 16-byte units, so a 16-lane b128 access is
//     conflict-free): both are single 16-byte stores, nothing is transposed;
//   * the MFMA's k index is permuted per lane group: lane group q of a K step covers k = 8q ..
//     8q + 7, so one ds_read_b128 of B gives a lane four consecutive k of one column (four MFMA
//     steps), and one ds_read_b128 of A four consecutive rows (its four 16-row tiles).  Lane
//     (c, q)'s columns are c, c + 16, c + 32, c + 48 of its wave's 64-column block, and with the
//     row permutation accumulator r of tile (i, j) is C[16q + 4r + i][c + 16j], so the epilogue
//     reads and writes C as float4 runs down columns;
//   * workgroup ids are remapped so each XCD walks a contiguous, grouped range of tiles.
// Requirements (checked by the caller, which otherwise uses the general kernel): m % 4 == 0,
// k % 4 == 0, lda, ldb, ldc multiples of 4 and A, B, C 16-byte aligned.
#include "al_internal.hpp"

namespace al {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int BM = 128, BN = 128, BK = 32, NT = 256;
constexpr int LROW = 132;  // floats per k-row of the A tile in LDS
constexpr int LDB = 36;    // floats per column of the B tile in LDS
constexpr int ATILE = BK * LROW, BTILE = BN * LDB;
constexpr size_t LDS_BYTES = 2 * (ATILE + BTILE) * sizeof(float);  // A and B, double-buffered

__global__ __launch_bounds__(NT, 2) void sgemm_nn_k(int m, int n, int k, float alpha, const float* __restrict__ A,
                                                    size_t lda, const float* __restrict__ B, size_t ldb, float beta,
                                                    float* __restrict__ C, size_t ldc, int tiles_m, int tiles_n) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    float* lds = (float*)lds_raw;
    int tm, tn;
    tile_coords(tiles_m, tiles_n, tm, tn);
    const int bm = tm * BM, bn = tn * BN;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int c = lane & 15, q = lane >> 4;

    // staging: float4 number f = t + 256 j of a tile.  A: column f >> 5, rows 4 (f & 31) ..;
    // B: column f >> 3, k 4 (f & 7) ..
    f4 ra[4], rb[4];
    // out-of-range float4s read the operand's first element (always valid) and are replaced
    // by zeros, so the loads carry no branches and issue back to back
    auto load = [&](int k0) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int f = t + NT * j;
            const int p = f >> 5, r = 4 * (f & 31);
            const bool ok = (bm + r) < m && (k0 + p) < k;
            const size_t off = ok ? (size_t)(bm + r) + (size_t)(k0 + p) * lda : 0;
            const f4 v = *(const f4*)(A + off);
            ra[j] = ok ? v : f4{0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int f = t + NT * j;
            const int col = f >> 3, kk = 4 * (f & 7);
            const bool ok = (bn + col) < n && (k0 + kk) < k;
            const size_t off = ok ? (size_t)(k0 + kk) + (size_t)(bn + col) * ldb : 0;
            const f4 v = *(const f4*)(B + off);
            rb[j] = ok ? v : f4{0.f, 0.f, 0.f, 0.f};
        }
    };
    // C (when beta != 0) is requested during the last K step, all sixteen float4s at once
    f4 cv[4][4];
    auto load_c = [&]() {
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int col = bn + wn * 64 + c + 16 * j, row = bm + wm * 64 + 16 * q + 4 * r;
                const bool ok = col < n && row < m;
                const f4 v = *(const f4*)(C + (ok ? (size_t)row + (size_t)col * ldc : 0));
                cv[j][r] = ok ? v : f4{0.f, 0.f, 0.f, 0.f};
            }
    };
    auto store = [&](float* As, float* Bs) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int f = t + NT * j;
            *(f4*)(As + (f >> 5) * LROW + 4 * (f & 31)) = ra[j];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int f = t + NT * j;
            *(f4*)(Bs + (f >> 3) * LDB + 4 * (f & 7)) = rb[j];
        }
    };

    f4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

    auto mfma_step = [&](int kt) {
        const float* As = lds + (kt & 1) * ATILE;
        const float* Bs = lds + 2 * ATILE + (kt & 1) * BTILE;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            f4 bq[4];  // four consecutive k (8q + 4h ..) of each of the lane's four columns
#pragma unroll
            for (int j = 0; j < 4; ++j) bq[j] = *(const f4*)(Bs + (wn * 64 + c + 16 * j) * LDB + 8 * q + 4 * h);
#pragma unroll
            for (int s4 = 0; s4 < 4; ++s4) {
                const f4 av = *(const f4*)(As + (8 * q + 4 * h + s4) * LROW + wm * 64 + 4 * c);
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i], bq[j][s4], acc[i][j], 0, 0, 0);
            }
        }
    };

    const int ktiles = (k + BK - 1) / BK;
    load(0);
    store(lds, lds + 2 * ATILE);
    __syncthreads();
    for (int kt = 0; kt + 1 < ktiles; ++kt) {
        load((kt + 1) * BK);
        mfma_step(kt);
        store(lds + ((kt + 1) & 1) * ATILE, lds + 2 * ATILE + ((kt + 1) & 1) * BTILE);
        __syncthreads();
    }
    // the last K step, peeled so C's registers are live only here
    if (beta != 0.f) load_c();
    mfma_step(ktiles - 1);

#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int col = bn + wn * 64 + c + 16 * j;
        if (col >= n) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = bm + wm * 64 + 16 * q + 4 * r;
            if (row >= m) continue;
            f4 v = f4{acc[0][j][r], acc[1][j][r], acc[2][j][r], acc[3][j][r]} * alpha;
            if (beta != 0.f) v += beta * cv[j][r];
            *(f4*)(C + (size_t)row + (size_t)col * ldc) = v;
        }
    }
}

}  // namespace

bool sgemm_nn_fits(size_t m, size_t k, const float* A, size_t lda, const float* B, size_t ldb, const float* C,
                   size_t ldc) {
    const auto al16 = [](const void* p) { return ((uintptr_t)p & 15u) == 0; };
    return m % 4 == 0 && k % 4 == 0 && lda % 4 == 0 && ldb % 4 == 0 && ldc % 4 == 0 && al16(A) && al16(B) && al16(C);
}

int sgemm_nn(hipStream_t s, int m, int n, int k, float alpha, const float* A, size_t lda, const float* B, size_t ldb,
             float beta, float* C, size_t ldc) {
    static bool attr_set = false;  // benign race: every thread sets the same value
    if (!attr_set) {
        AL_HIP(hipFuncSetAttribute((const void*)sgemm_nn_k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)LDS_BYTES));
        attr_set = true;
    }
    const int tiles_m = (m + BM - 1) / BM, tiles_n = (n + BN - 1) / BN;
    const int ticket = prof_begin(s);
    hipLaunchKernelGGL(sgemm_nn_k, dim3(tiles_m * tiles_n), dim3(NT), LDS_BYTES, s, m, n, k, alpha, A, lda, B, ldb,
                       beta, C, ldc, tiles_m, tiles_n);
    AL_HIP(hipGetLastError());
    if (ticket >= 0) {
        const double dm = m, dn = n, dk = k;
        prof_end(ticket, s, "sgemm_nn", 2.0 * dm * dn * dk,
                 4.0 * (dm * dk + dk * dn + dm * dn * (beta != 0.f ? 2.0 : 1.0)));
    }
    return AL_OK;
}

}  // namespace al
