// FP32 C := alpha*A*B + beta*C (neither operand transposed) on v_mfma_f32_16x16x4_f32: the
// trailing and panel updates of the FP32 factorization of the mixed-precision solve (the
// low-precision counterpart of blas::dgemm's no-transpose case, src/blas.rs:201-227).
//
// FP32 matrix cores do a 16x16x4 product in half the time of FP64, so per flop the operand
// path must move twice the bytes; this kernel is built around 16-byte accesses everywhere:
//   * one 256-thread workgroup (4 wave64s, 2x2) computes a 128x128 tile of C, K in steps of
//     32 through double-buffered LDS (66 KiB, two workgroups per CU), one barrier per step;
//   * global loads are float4: four rows of a column of A, four k of a column of B; A's
//     float4 lands in LDS as one 16-byte store, B's as four scalar stores into its k-major
//     layout (rows padded to 132 floats: 16-byte aligned, at most 2-way conflicts);
//   * the MFMA rows and columns are permuted so a lane's operands for its wave's four 16-row
//     (16-column) MFMA tiles are four consecutive rows (columns): one ds_read_b128 of A and
//     one of B feed 16 MFMAs.  With that permutation, accumulator r of tile (i, j) in lane
//     (c = l & 15, q = l >> 4) is C[16q + 4r + i][4c + j] of the wave's 64x64 block, so the
//     epilogue reads and writes C as float4 runs down columns;
//   * workgroup ids are remapped so each XCD walks a contiguous, grouped range of tiles.
// Requirements (checked by the caller, which otherwise uses the general kernel): m % 4 == 0,
// k % 4 == 0, lda, ldb, ldc multiples of 4 and A, B, C 16-byte aligned.
#include "al_internal.hpp"

namespace al {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int BM = 128, BN = 128, BK = 32, NT = 256;
constexpr int LROW = 132;  // floats per k-row of a tile in LDS
constexpr int TILE = BK * LROW;
constexpr size_t LDS_BYTES = 4 * TILE * sizeof(float);  // A and B, double-buffered

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
                const int col = bn + wn * 64 + 4 * c + j, row = bm + wm * 64 + 16 * q + 4 * r;
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
            float* d = Bs + (4 * (f & 7)) * LROW + (f >> 3);
            d[0] = rb[j].x;
            d[LROW] = rb[j].y;
            d[2 * LROW] = rb[j].z;
            d[3 * LROW] = rb[j].w;
        }
    };

    f4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

    auto mfma_step = [&](int kt) {
        const float* As = lds + (kt & 1) * TILE;
        const float* Bs = lds + 2 * TILE + (kt & 1) * TILE;
#pragma unroll
        for (int ks = 0; ks < BK / 4; ++ks) {
            const int kk = 4 * ks + q;
            const f4 av = *(const f4*)(As + kk * LROW + wm * 64 + 4 * c);
            const f4 bv = *(const f4*)(Bs + kk * LROW + wn * 64 + 4 * c);
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[i], bv[j], acc[i][j], 0, 0, 0);
        }
    };

    const int ktiles = (k + BK - 1) / BK;
    load(0);
    store(lds, lds + 2 * TILE);
    __syncthreads();
    for (int kt = 0; kt + 1 < ktiles; ++kt) {
        load((kt + 1) * BK);
        mfma_step(kt);
        store(lds + ((kt + 1) & 1) * TILE, lds + 2 * TILE + ((kt + 1) & 1) * TILE);
        __syncthreads();
    }
    // the last K step, peeled so C's registers are live only here
    if (beta != 0.f) load_c();
    mfma_step(ktiles - 1);

#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int col = bn + wn * 64 + 4 * c + j;
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
