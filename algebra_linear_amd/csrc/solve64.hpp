// Forward substitution with the unit lower triangle of a 64 x 64 block, applied to columns of a
// column-major matrix: X := L^{-1} X for X = F[J:J+64, c0:c1).  Used by the FP32 panel kernel
// that finishes a tournament (L from its row-major copy of L11 \ U11); L may also be read in
// place from a column-major matrix.
#pragma once

#include <hip/hip_runtime.h>

namespace al {

constexpr int S64_G = 4;  // columns per wave and pass

// One wave solves S64_G columns at a time, lane i holding row i of each: the lane keeps row i of
// L in registers (64 loads per wave, once), and step j takes x_j of every column from lane j with
// v_readlane, so the 63 dependent steps wait on neither memory nor LDS.  The FMAs are those of
// column-oriented forward substitution, in the same order.  Waves `wave` of `nwaves` stride over
// the column groups.  L(i, j) is L[i * 64 + j] when ROWMAJOR, else L[i + j * ldl]; the block
// must not overlap X.
template <bool ROWMAJOR>
__device__ __forceinline__ void solve_l64_cols(float* __restrict__ F, size_t ld, int J, int c0, int c1,
                                               const float* __restrict__ L, size_t ldl, int wave, int nwaves) {
    const int lane = threadIdx.x & 63;
    const int ncols = c1 - c0;
    float lrow[64];
    if constexpr (ROWMAJOR) {
        const float4* p = (const float4*)(L + (size_t)lane * 64);
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const float4 q = p[v];
            lrow[4 * v] = q.x;
            lrow[4 * v + 1] = q.y;
            lrow[4 * v + 2] = q.z;
            lrow[4 * v + 3] = q.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < 64; ++j) lrow[j] = L[(size_t)lane + (size_t)j * ldl];
    }
    for (int g = wave * S64_G; g < ncols; g += nwaves * S64_G) {
        float x[S64_G];
        float* col[S64_G];
#pragma unroll
        for (int q = 0; q < S64_G; ++q) {
            const int c = c0 + min(g + q, ncols - 1);  // a short last group repeats its last column, unstored
            col[q] = F + (size_t)J + (size_t)c * ld;
            x[q] = col[q][lane];
        }
#pragma unroll
        for (int j = 0; j < 63; ++j) {
            const bool below = lane > j;
#pragma unroll
            for (int q = 0; q < S64_G; ++q) {
                const float xj = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x[q]), j));
                const float v = fmaf(-lrow[j], xj, x[q]);
                x[q] = below ? v : x[q];
            }
        }
#pragma unroll
        for (int q = 0; q < S64_G; ++q)
            if (g + q < ncols) col[q][lane] = x[q];
    }
}

}  // namespace al
