// Row interchanges without per-column dependency chains.
//
// The crate applies a pivot block as jb sequential full-row swaps (src/lapack/mod.rs:141-146).
// Done literally on the GPU, one thread per column walks jb dependent load/store pairs with a
// column stride of lda: latency-bound.  Here the block's swaps are first composed into one
// permutation of the touched rows (`piv_pairs`: at most 2*jb (destination, source) pairs,
// computed once, each touched row traced back through the swaps in parallel), and then every column applies that permutation as an
// independent gather (`apply_pairs`: read all sources, then write all destinations).  The
// result is identical to the sequential swaps: each touched row ends up holding exactly the
// row the swap sequence would have moved there.
#include "al_internal.hpp"

namespace al {
namespace {

constexpr int JB_MAX = 256;
constexpr int PAIRS_MAX = 2 * JB_MAX;

// Composes the pivots ipiv[k0..k0+jb) into (destination, source) row pairs: out[0] = count,
// out[1 + 2p] = destination row, out[2 + 2p] = source row.  Every row the block touches --
// the rows k0..k0+jb-1 and the pivot rows below them -- is traced backwards through the swaps
// independently (one thread per row, the swaps broadcast from LDS four at a time): walking the
// sequence in reverse from a destination row yields the original row whose value ends there.
// A pivot row is kept only at the first step that touches it.  The loops have fixed trip
// counts and no early exits, and the pairs are written in thread order (top rows ascending,
// then pivot rows in step order), so the list is the same on every run.
__global__ __launch_bounds__(512) void piv_pairs(const int* __restrict__ ipiv, int k0, int jb, int* __restrict__ out,
                                                 const int* __restrict__ info) {
    __shared__ __align__(16) int piv[JB_MAX];
    __shared__ int wcount[8];
    const int t = threadIdx.x;
    if (info && *info != 0) {
        if (t == 0) out[0] = 0;
        return;
    }
    const int jb4 = (jb + 3) & ~3;
    if (t < jb4) piv[t] = t < jb ? ipiv[k0 + t] : k0 + t;  // padding: a row swapped with itself
    __syncthreads();
    const int hi = k0 + jb;
    int dst = -1;
    if (t < jb) {
        dst = k0 + t;
    } else if (t < 2 * jb) {
        const int s0 = t - jb;
        const int p = piv[s0];
        if (p >= hi) {
            bool dup = false;
            for (int s = 0; s < jb4; s += 4) {
                const int4 q = *(const int4*)&piv[s];
                dup = dup || (s < s0 && q.x == p) || (s + 1 < s0 && q.y == p) || (s + 2 < s0 && q.z == p) ||
                      (s + 3 < s0 && q.w == p);
            }
            dst = dup ? -1 : p;
        }
    }
    int x = dst;
    for (int s = jb4 - 4; s >= 0; s -= 4) {
        const int4 q = *(const int4*)&piv[s];
        x = (x == k0 + s + 3) ? q.w : (x == q.w) ? k0 + s + 3 : x;
        x = (x == k0 + s + 2) ? q.z : (x == q.z) ? k0 + s + 2 : x;
        x = (x == k0 + s + 1) ? q.y : (x == q.y) ? k0 + s + 1 : x;
        x = (x == k0 + s) ? q.x : (x == q.x) ? k0 + s : x;
    }
    const bool emit = dst >= 0 && x != dst;
    const unsigned long long bal = __ballot(emit);
    const int lane = t & 63, wave = t >> 6;
    if (lane == 0) wcount[wave] = __popcll(bal);
    __syncthreads();
    int base = 0;
    for (int w = 0; w < wave; ++w) base += wcount[w];
    if (emit) {
        const int r = base + __popcll(bal & ((1ull << lane) - 1ull));
        out[1 + 2 * r] = dst;
        out[2 + 2 * r] = x;
    }
    if (t == 0) {
        int total = 0;
        for (int w = 0; w < 8; ++w) total += wcount[w];
        out[0] = total;
    }
}

// One wave per column: columns [0, n0) of A0, then [0, n1) of A1, then (if perm) the vector
// perm as one more column (FP64 matrices only).
template <typename T>
__global__ __launch_bounds__(256) void apply_pairs(T* __restrict__ A0, int n0, T* __restrict__ A1, int n1, size_t lda,
                                                   double* __restrict__ perm, const int* __restrict__ pairs,
                                                   const int* __restrict__ info) {
    __shared__ int pr[1 + PAIRS_MAX * 2];
    if (info && *info != 0) return;
    const int t = threadIdx.x;
    const int cnt = pairs[0];
    for (int i = t; i < 2 * cnt; i += 256) pr[1 + i] = pairs[1 + i];
    __syncthreads();
    const int lane = t & 63;
    const int c = blockIdx.x * 4 + (t >> 6);
    const int ntot = n0 + n1 + (perm ? 1 : 0);
    T* col = nullptr;
    if (c < n0)
        col = A0 + (size_t)c * lda;
    else if (c < n0 + n1)
        col = A1 + (size_t)(c - n0) * lda;
    else if constexpr (sizeof(T) == sizeof(double)) {
        if (c < ntot) col = (T*)perm;
    }
    T v[PAIRS_MAX / 64];
#pragma unroll
    for (int q = 0; q < PAIRS_MAX / 64; ++q) {
        const int i = lane + 64 * q;
        v[q] = (col && i < cnt) ? col[pr[2 + 2 * i]] : (T)0;
    }
    __syncthreads();  // every source read before any destination is written
#pragma unroll
    for (int q = 0; q < PAIRS_MAX / 64; ++q) {
        const int i = lane + 64 * q;
        if (col && i < cnt) col[pr[1 + 2 * i]] = v[q];
    }
}

// Several pivot blocks' pair lists applied one after the other, as their swap sequences are:
// lists l = 0..nl-1 at pairs + l * stride, each in piv_pairs' format with at most
// SUB_PAIRS_MAX pairs (one 64-pivot block).  One wave per column, so the lists need no
// synchronization between workgroups: a wave issues a list's reads before its writes, and a
// list's writes before the next list's reads.  A later list may read, in another lane, a row an
// earlier list wrote; that relies on the hardware processing one wave's vector memory
// instructions in order (the vector L1 is write-through).  The result was checked bit for bit
// against applying the blocks' composed permutation instead.
constexpr int SUB_PAIRS_MAX = 128;
constexpr int MULTI_MAX = 8;
__global__ __launch_bounds__(256) void apply_pairs_multi(float* __restrict__ A0, int n0, float* __restrict__ A1,
                                                         int n1, size_t lda, const int* __restrict__ pairs,
                                                         size_t stride, int nl, const int* __restrict__ info) {
    __shared__ int pr[MULTI_MAX][1 + 2 * SUB_PAIRS_MAX];
    if (info && *info != 0) return;
    const int t = threadIdx.x;
    for (int l = 0; l < nl; ++l) {
        const int* p = pairs + (size_t)l * stride;
        const int cnt = min(p[0], SUB_PAIRS_MAX);
        if (t == 0) pr[l][0] = cnt;
        for (int i = t; i < 2 * cnt; i += 256) pr[l][1 + i] = p[1 + i];
    }
    __syncthreads();
    const int lane = t & 63;
    const int c = blockIdx.x * 4 + (t >> 6);
    float* col;
    if (c < n0)
        col = A0 + (size_t)c * lda;
    else if (c < n0 + n1)
        col = A1 + (size_t)(c - n0) * lda;
    else
        return;
    for (int l = 0; l < nl; ++l) {
        const int cnt = pr[l][0];
        float v[SUB_PAIRS_MAX / 64];
#pragma unroll
        for (int q = 0; q < SUB_PAIRS_MAX / 64; ++q) {
            const int i = lane + 64 * q;
            v[q] = i < cnt ? col[pr[l][2 + 2 * i]] : 0.f;
        }
#pragma unroll
        for (int q = 0; q < SUB_PAIRS_MAX / 64; ++q) {
            const int i = lane + 64 * q;
            if (i < cnt) col[pr[l][1 + 2 * i]] = v[q];
        }
    }
}

__global__ void perm_iota(double* perm, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) perm[i] = (double)i;
}

// Row origins of a whole pivot sequence: thread i traces row i backwards through all n swaps
// (the pivots staged through LDS in chunks), ending at the original row that lands at i.  The
// pivots are read eight at a time (two 16-byte LDS loads) and the eight steps applied from
// registers, in the same order, so the chain waits on LDS once per eight swaps.
__device__ __forceinline__ int trace_swap(int x, int a, int b) { return (x == a) ? b : (x == b) ? a : x; }

__global__ __launch_bounds__(256) void perm_from_pivots(const int* __restrict__ ipiv, int n, double* __restrict__ perm) {
    __shared__ __align__(16) int piv[1024];
    const int i = blockIdx.x * 256 + threadIdx.x;
    int x = i;
    const int nchunks = (n + 1023) / 1024;
    for (int ch = nchunks - 1; ch >= 0; --ch) {
        const int base = ch * 1024;
        const int len = (n - base) < 1024 ? (n - base) : 1024;
        __syncthreads();
        for (int q = threadIdx.x; q < len; q += 256) piv[q] = ipiv[base + q];
        __syncthreads();
        const int full = len & ~7;
        for (int s = len - 1; s >= full; --s) x = trace_swap(x, base + s, piv[s]);
        for (int g = full - 8; g >= 0; g -= 8) {
            const int4 lo = *(const int4*)&piv[g], hi = *(const int4*)&piv[g + 4];
            x = trace_swap(x, base + g + 7, hi.w);
            x = trace_swap(x, base + g + 6, hi.z);
            x = trace_swap(x, base + g + 5, hi.y);
            x = trace_swap(x, base + g + 4, hi.x);
            x = trace_swap(x, base + g + 3, lo.w);
            x = trace_swap(x, base + g + 2, lo.z);
            x = trace_swap(x, base + g + 1, lo.y);
            x = trace_swap(x, base + g, lo.x);
        }
    }
    if (i < n) perm[i] = (double)x;
}

__global__ void gather_rows_k(const double* __restrict__ B, size_t ldb, int n, int nrhs, const double* __restrict__ perm,
                              double* __restrict__ W) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)n * nrhs) return;
    const int i = (int)(idx % n), c = (int)(idx / n);
    W[idx] = B[(size_t)perm[i] + (size_t)c * ldb];
}

__global__ void copy_block_k(const double* __restrict__ W, int n, int nrhs, double* __restrict__ B, size_t ldb) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)n * nrhs) return;
    const int i = (int)(idx % n), c = (int)(idx / n);
    B[(size_t)i + (size_t)c * ldb] = W[idx];
}

}  // namespace

template <typename T>
int swap_block_t(hipStream_t s, Workspace* ws, const int* ipiv, size_t k0, size_t jb, T* A0, size_t n0, T* A1,
                 size_t n1, size_t lda, double* perm, const int* info, int* pairs) {
    if (jb == 0 || (n0 + n1 == 0 && !perm)) return AL_OK;
    if (jb > (size_t)JB_MAX) {
        set_error("swap_block: pivot block wider than 256");
        return AL_ERR_ARG;
    }
    if (!pairs) pairs = ws->pairs;
    hipLaunchKernelGGL(piv_pairs, dim3(1), dim3(512), 0, s, ipiv, (int)k0, (int)jb, pairs, info);
    AL_HIP(hipGetLastError());
    const size_t ncols = n0 + n1 + (perm ? 1 : 0);
    hipLaunchKernelGGL(apply_pairs<T>, dim3(ceil_div(ncols, 4)), dim3(256), 0, s, A0, (int)n0, A1, (int)n1, lda, perm,
                       pairs, info);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

int swap_block_f64(hipStream_t s, Workspace* ws, const int* ipiv, size_t k0, size_t jb, double* A0, size_t n0,
                   double* A1, size_t n1, size_t lda, double* perm, const int* info, int* pairs) {
    return swap_block_t<double>(s, ws, ipiv, k0, jb, A0, n0, A1, n1, lda, perm, info, pairs);
}

int swap_block_z(hipStream_t s, Workspace* ws, const int* ipiv, size_t k0, size_t jb, zd* A0, size_t n0, zd* A1,
                 size_t n1, size_t lda, const int* info) {
    return swap_block_t<zd>(s, ws, ipiv, k0, jb, A0, n0, A1, n1, lda, nullptr, info, nullptr);
}

int swap_block_f32(hipStream_t s, Workspace* ws, const int* ipiv, size_t k0, size_t jb, float* A0, size_t n0,
                   float* A1, size_t n1, size_t lda, const int* info, int* pairs) {
    return swap_block_t<float>(s, ws, ipiv, k0, jb, A0, n0, A1, n1, lda, nullptr, info, pairs);
}

int swap_apply_f32(hipStream_t s, const int* pairs, float* A0, size_t n0, float* A1, size_t n1, size_t lda,
                   const int* info) {
    if (n0 + n1 == 0) return AL_OK;
    hipLaunchKernelGGL(apply_pairs<float>, dim3(ceil_div(n0 + n1, 4)), dim3(256), 0, s, A0, (int)n0, A1, (int)n1, lda,
                       (double*)nullptr, pairs, info);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

int swap_apply_multi_f32(hipStream_t s, const int* pairs, size_t stride, int nl, float* A0, size_t n0, float* A1,
                         size_t n1, size_t lda, const int* info) {
    if (n0 + n1 == 0 || nl == 0) return AL_OK;
    if (nl > MULTI_MAX) {
        set_error("swap_apply_multi: more than 8 pivot blocks");
        return AL_ERR_ARG;
    }
    hipLaunchKernelGGL(apply_pairs_multi, dim3(ceil_div(n0 + n1, 4)), dim3(256), 0, s, A0, (int)n0, A1, (int)n1, lda,
                       pairs, stride, nl, info);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

int perm_init(hipStream_t s, double* perm, size_t n) {
    if (n == 0) return AL_OK;
    hipLaunchKernelGGL(perm_iota, dim3(ceil_div(n, 256)), dim3(256), 0, s, perm, n);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

int perm_from_ipiv(hipStream_t s, const int* ipiv, size_t n, double* perm) {
    if (n == 0) return AL_OK;
    hipLaunchKernelGGL(perm_from_pivots, dim3(ceil_div(n, 256)), dim3(256), 0, s, ipiv, (int)n, perm);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

int gather_rows(hipStream_t s, const double* B, size_t ldb, size_t n, size_t nrhs, const double* perm, double* W) {
    if (n == 0 || nrhs == 0) return AL_OK;
    hipLaunchKernelGGL(gather_rows_k, dim3(ceil_div(n * nrhs, 256)), dim3(256), 0, s, B, ldb, (int)n, (int)nrhs, perm,
                       W);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

int copy_block(hipStream_t s, const double* W, size_t n, size_t nrhs, double* B, size_t ldb) {
    if (n == 0 || nrhs == 0) return AL_OK;
    hipLaunchKernelGGL(copy_block_k, dim3(ceil_div(n * nrhs, 256)), dim3(256), 0, s, W, (int)n, (int)nrhs, B, ldb);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

}  // namespace al
