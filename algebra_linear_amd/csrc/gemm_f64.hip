// FP64 GEMM on CDNA4 matrix cores: C := alpha*op(A)*op(B) + beta*C, column-major, with
// the four transposition variants of blas::dgemm{,_tn,_nt,_tt} (src/blas.rs:201-257) and a
// triangle-restricted form for dsyrk (src/blas.rs:609-626).  Two more restrictions serve the
// complex Cholesky: on the real view of a complex matrix (row r holds part r & 1 of complex row
// r >> 1) they keep the complex lower (r >> 1 >= column) or upper (r >> 1 <= column) triangle.
//
// Design (MI355X-first, not a translation of the crate's Goto blocking):
//   * one 256-thread workgroup (4 wave64s, 2x2) computes a 128x128 tile of C; each wave
//     owns 64x64 = 4x4 tiles of v_mfma_f64_16x16x4_f64.  Products with fewer 128-tiles than
//     two per CU use 64x64 workgroup tiles (2x2 MFMA tiles per wave) instead;
//   * K advances in steps of 16 through double-buffered LDS (66 KiB) with at most 256
//     registers per lane, so two workgroups share a CU and one's loads and epilogue overlap
//     the other's MFMAs; the next step's global loads
//     are issued into registers before the current step's MFMAs, so HBM/L2 latency hides
//     behind matrix-core time;
//   * transposition is handled entirely in the global->LDS load pattern: both orientations
//     land in the same k-major LDS layout, so the MFMA loop is shared by all variants;
//   * the MFMA operands are fed as (B^T, A^T) so each lane's four accumulators lie in one
//     C row and 16 lanes cover 16 consecutive rows of a column: stores and beta loads of the
//     column-major C are contiguous 128-byte runs;
//   * workgroup ids are remapped so each of the 8 XCDs walks a contiguous range of tiles,
//     grouped 8 tile-rows high, keeping A row-panels and B column-panels in that XCD's L2.
#include "al_internal.hpp"

namespace al {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// Per-element-type MFMA: FP64 uses v_mfma_f64_16x16x4_f64, FP32 v_mfma_f32_16x16x4_f32.
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef d4 acc_t;
    static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // accumulator r of lane l holds D[(l >> 4) + 4r][l & 15]
    static __device__ __forceinline__ int drow(int kq, int r) { return kq + 4 * r; }
};
template <>
struct Mfma<float> {
    typedef f4 acc_t;
    static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // accumulator r of lane l holds D[4 (l >> 4) + r][l & 15]
    static __device__ __forceinline__ int drow(int kq, int r) { return 4 * kq + r; }
};

constexpr int BK = 16;
constexpr int NT = 256;
// Workgroup tile TM x TM (128, or 64 for products too small to fill the chip with 128-tiles):
// 2 x 2 waves of (TM/2) x (TM/2), i.e. (TM/32) x (TM/32) MFMA blocks of 16 x 16 per wave.
template <int TM>
constexpr int lds_row() { return TM + 1; }  // one-element pad breaks the strided stores' bank conflicts
template <int TM>
constexpr int tile_elems() { return BK * lds_row<TM>(); }
template <typename T, int TM>
constexpr size_t lds_bytes() { return 4 * tile_elems<TM>() * sizeof(T); }  // A and B, double-buffered
template <int TM>
constexpr int lpt() { return TM * BK / NT; }  // operand elements per thread per tile

// Loads the TM x BK operand tile whose element (r, p) is base[r + p*ld] (ROWCONTIG false:
// the tile's columns are contiguous, e.g. a non-transposed A) or base[p + r*ld] (true: the
// tile's rows are contiguous, e.g. a transposed A) into lpt registers per thread, zero-filling
// outside [0, r_lim) x [0, kdim).
template <bool ROWCONTIG, int TM, typename T>
__device__ __forceinline__ void load_tile(const T* __restrict__ base, size_t ld, int r0, int r_lim, int k0,
                                          int kdim, T (&reg)[lpt<TM>()], int t) {
    if (!ROWCONTIG) {
        const int r = t & (TM - 1), p0 = t / TM;
        const bool rok = (r0 + r) < r_lim;
#pragma unroll
        for (int j = 0; j < lpt<TM>(); ++j) {
            const int p = k0 + p0 + (NT / TM) * j;
            reg[j] = (rok && p < kdim) ? base[(size_t)(r0 + r) + (size_t)p * ld] : (T)0;
        }
    } else {
        const int p = t % BK, rr = t / BK;
        const bool pok = (k0 + p) < kdim;
#pragma unroll
        for (int j = 0; j < lpt<TM>(); ++j) {
            const int r = r0 + rr + (NT / BK) * j;
            reg[j] = (pok && r < r_lim) ? base[(size_t)(k0 + p) + (size_t)r * ld] : (T)0;
        }
    }
}

template <bool ROWCONTIG, int TM, typename T>
__device__ __forceinline__ void store_tile(T* __restrict__ S, const T (&reg)[lpt<TM>()], int t) {
    constexpr int LR = lds_row<TM>();
    if (!ROWCONTIG) {
        const int r = t & (TM - 1), p0 = t / TM;
#pragma unroll
        for (int j = 0; j < lpt<TM>(); ++j) S[(p0 + (NT / TM) * j) * LR + r] = reg[j];
    } else {
        const int p = t % BK, rr = t / BK;
#pragma unroll
        for (int j = 0; j < lpt<TM>(); ++j) S[p * LR + rr + (NT / BK) * j] = reg[j];
    }
}

template <typename T, bool TA, bool TB, int TRI, int TM>
__global__ __launch_bounds__(NT, 2) void dgemm_mfma(int m, int n, int k, T alpha, const T* __restrict__ A,
                                                 size_t lda, const T* __restrict__ B, size_t ldb, T beta,
                                                 T* __restrict__ C, size_t ldc, int tiles_m, int tiles_n) {
    constexpr int LR = lds_row<TM>();
    constexpr int TE = tile_elems<TM>();
    constexpr int WT = TM / 32;  // MFMA blocks per wave and dimension
    extern __shared__ __align__(16) unsigned char lds_raw[];
    T* lds = (T*)lds_raw;
    T* As0 = lds;
    T* As1 = lds + TE;
    T* Bs0 = lds + 2 * TE;
    T* Bs1 = lds + 3 * TE;
    typedef typename Mfma<T>::acc_t acc_t;

    int tm, tn;
    tile_coords(tiles_m, tiles_n, tm, tn);
    const int bm = tm * TM, bn = tn * TM;
    if (TRI == 1 && bm + TM - 1 < bn) return;  // tile strictly above the diagonal
    if (TRI == 2 && bm > bn + TM - 1) return;  // tile strictly below the diagonal
    if (TRI == 3 && ((bm + TM - 1) >> 1) < bn) return;  // complex rows strictly above
    if (TRI == 4 && (bm >> 1) > bn + TM - 1) return;    // complex rows strictly below

    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    // op(A) is m x k: non-transposed A has contiguous columns; transposed A contiguous rows.
    // op(B) is k x n; its tile element (r = j, p) is B[p + j*ldb] (contiguous in p) when not
    // transposed and B[j + p*ldb] (contiguous in j) when transposed.
    constexpr bool A_RC = TA;
    constexpr bool B_RC = !TB;

    acc_t acc[WT][WT];
#pragma unroll
    for (int i = 0; i < WT; ++i)
#pragma unroll
        for (int j = 0; j < WT; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

    T ra[lpt<TM>()], rb[lpt<TM>()];
    const int ktiles = (k + BK - 1) / BK;
    load_tile<A_RC, TM>(A, lda, bm, m, 0, k, ra, t);
    load_tile<B_RC, TM>(B, ldb, bn, n, 0, k, rb, t);
    store_tile<A_RC, TM>(As0, ra, t);
    store_tile<B_RC, TM>(Bs0, rb, t);
    __syncthreads();

    const int col = lane & 15, kq = lane >> 4;
    for (int kt = 0; kt < ktiles; ++kt) {
        const bool more = (kt + 1) < ktiles;
        if (more) {
            load_tile<A_RC, TM>(A, lda, bm, m, (kt + 1) * BK, k, ra, t);
            load_tile<B_RC, TM>(B, ldb, bn, n, (kt + 1) * BK, k, rb, t);
        }
        const T* As = (kt & 1) ? As1 : As0;
        const T* Bs = (kt & 1) ? Bs1 : Bs0;
#pragma unroll
        for (int s = 0; s < BK / 4; ++s) {
            const int kk = 4 * s + kq;
            T af[WT], bf[WT];
#pragma unroll
            for (int i = 0; i < WT; ++i) af[i] = As[kk * LR + wm * (TM / 2) + i * 16 + col];
#pragma unroll
            for (int j = 0; j < WT; ++j) bf[j] = Bs[kk * LR + wn * (TM / 2) + j * 16 + col];
#pragma unroll
            for (int i = 0; i < WT; ++i)
#pragma unroll
                for (int j = 0; j < WT; ++j) acc[i][j] = Mfma<T>::op(bf[j], af[i], acc[i][j]);
        }
        if (more) {
            store_tile<A_RC, TM>((kt & 1) ? As0 : As1, ra, t);
            store_tile<B_RC, TM>((kt & 1) ? Bs0 : Bs1, rb, t);
        }
        __syncthreads();
    }

    // with the operands fed as (B^T, A^T), D = C^T: accumulator r of a lane holds C[row][col],
    // row = .. + (lane & 15), col = .. + Mfma<T>::drow(lane >> 4, r)
#pragma unroll
    for (int i = 0; i < WT; ++i) {
        const int row = bm + wm * (TM / 2) + i * 16 + col;
        if (row >= m) continue;
#pragma unroll
        for (int j = 0; j < WT; ++j) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int cc = bn + wn * (TM / 2) + j * 16 + Mfma<T>::drow(kq, r);
                if (cc >= n) continue;
                if (TRI == 1 && row < cc) continue;
                if (TRI == 2 && row > cc) continue;
                if (TRI == 3 && (row >> 1) < cc) continue;
                if (TRI == 4 && (row >> 1) > cc) continue;
                T* cp = C + (size_t)row + (size_t)cc * ldc;
                T v = alpha * acc[i][j][r];
                if (beta != (T)0) v += beta * *cp;
                *cp = v;
            }
        }
    }
}

// C := beta*C (0 when beta == 0) -- the crate's k == 0 / alpha == 0 path (src/blas.rs:297-308).
template <typename T>
__global__ void scale_c(int m, int n, T beta, T* C, size_t ldc, int tri) {
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)m * n) return;
    const int i = (int)(idx % m), j = (int)(idx / m);
    if (tri == 1 && i < j) return;
    if (tri == 2 && i > j) return;
    T* cp = C + (size_t)i + (size_t)j * ldc;
    *cp = (beta == (T)0) ? (T)0 : beta * *cp;
}

// 128-tiles unless that leaves the chip short of workgroups (fewer than two per CU).
inline bool small_tiles(int m, int n, int tri) {
    const long long t128 = (long long)((m + 127) / 128) * ((n + 127) / 128);
    const long long eff = tri ? (t128 + (long long)((m + 127) / 128)) / 2 : t128;
    return eff < 2LL * cu_count();
}

template <typename T, bool TA, bool TB, int TRI, int TM>
int launch_tm(hipStream_t s, int m, int n, int k, T alpha, const T* A, size_t lda, const T* B, size_t ldb, T beta,
              T* C, size_t ldc) {
    static bool attr_set = false;  // benign race: every thread sets the same value
    auto fn = dgemm_mfma<T, TA, TB, TRI, TM>;
    constexpr size_t LDS_BYTES = lds_bytes<T, TM>();
    if (!attr_set) {
        AL_HIP(hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)LDS_BYTES));
        attr_set = true;
    }
    const int tiles_m = (m + TM - 1) / TM, tiles_n = (n + TM - 1) / TM;
    hipLaunchKernelGGL(fn, dim3(tiles_m * tiles_n), dim3(NT), LDS_BYTES, s, m, n, k, alpha, A, lda, B, ldb, beta,
                       C, ldc, tiles_m, tiles_n);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

template <typename T, bool TA, bool TB, int TRI>
int launch(hipStream_t s, int m, int n, int k, T alpha, const T* A, size_t lda, const T* B, size_t ldb, T beta, T* C,
           size_t ldc) {
    const int ticket = prof_begin(s);
    if (small_tiles(m, n, TRI))
        AL_CHECK((launch_tm<T, TA, TB, TRI, 64>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)));
    else
        AL_CHECK((launch_tm<T, TA, TB, TRI, 128>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)));
    if (ticket >= 0) {
        // algorithmic work: 2mnk flops (half, plus the diagonal, for a triangle); bytes: A and B
        // once, C written (and read when beta != 0)
        const double dm = m, dn = n, dk = k;
        const double cfrac = TRI ? (dn + 1.0) / (2.0 * dn) : 1.0;
        const double fl = 2.0 * dm * dn * dk * cfrac;
        const double by = 8.0 * (dm * dk + dk * dn + dm * dn * cfrac * (beta != 0.0 ? 2.0 : 1.0));
        static const char* names[2][3][4] = {
            {{"dgemm_nn", "dgemm_tn", "dgemm_nt", "dgemm_tt"},
             {"dsyrk_lower", "dsyrk_lower", "dsyrk_lower", "dsyrk_lower"},
             {"dsyrk_upper", "dsyrk_upper", "dsyrk_upper", "dsyrk_upper"}},
            {{"sgemm_nn", "sgemm_tn", "sgemm_nt", "sgemm_tt"},
             {"ssyrk_lower", "ssyrk_lower", "ssyrk_lower", "ssyrk_lower"},
             {"ssyrk_upper", "ssyrk_upper", "ssyrk_upper", "ssyrk_upper"}}};
        const double by2 = by * (double)sizeof(T) / 8.0;
        const char* nm = TRI >= 3 ? (TRI == 3 ? "zherk_lower_real_view" : "zherk_upper_real_view")
                                  : names[sizeof(T) == 4 ? 1 : 0][TRI][(TA ? 1 : 0) + (TB ? 2 : 0)];
        prof_end(ticket, s, nm, fl, by2);
    }
    return AL_OK;
}

template <typename T, int TRI>
int dispatch(hipStream_t s, bool ta, bool tb, int m, int n, int k, T alpha, const T* A, size_t lda, const T* B,
             size_t ldb, T beta, T* C, size_t ldc) {
    if (!ta && !tb) return launch<T, false, false, TRI>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    if (ta && !tb) return launch<T, true, false, TRI>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    if (!ta && tb) return launch<T, false, true, TRI>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
    return launch<T, true, true, TRI>(s, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

template <typename T>
int gemm_t(hipStream_t s, bool ta, bool tb, size_t m, size_t n, size_t k, T alpha, const T* A, size_t lda, const T* B,
           size_t ldb, T beta, T* C, size_t ldc, int tri) {
    if (m == 0 || n == 0) return AL_OK;
    if (m > INT32_MAX || n > INT32_MAX || k > INT32_MAX) {
        set_error("gemm: dimensions beyond 2^31 are not supported");
        return AL_ERR_ARG;
    }
    if (k == 0 || alpha == (T)0) {
        if (beta == (T)1) return AL_OK;
        const size_t total = m * n;
        hipLaunchKernelGGL(scale_c<T>, dim3(ceil_div(total, 256)), dim3(256), 0, s, (int)m, (int)n, beta, C, ldc, tri);
        AL_HIP(hipGetLastError());
        return AL_OK;
    }
    switch (tri) {
        case 1: return dispatch<T, 1>(s, ta, tb, (int)m, (int)n, (int)k, alpha, A, lda, B, ldb, beta, C, ldc);
        case 2: return dispatch<T, 2>(s, ta, tb, (int)m, (int)n, (int)k, alpha, A, lda, B, ldb, beta, C, ldc);
        default: return dispatch<T, 0>(s, ta, tb, (int)m, (int)n, (int)k, alpha, A, lda, B, ldb, beta, C, ldc);
    }
}

}  // namespace

int gemm_f64(hipStream_t s, bool ta, bool tb, size_t m, size_t n, size_t k, double alpha, const double* A,
             size_t lda, const double* B, size_t ldb, double beta, double* C, size_t ldc, int tri) {
    return gemm_t<double>(s, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tri);
}

int gemm_f64_ctri(hipStream_t s, bool upper, size_t m, size_t n, size_t k, double alpha, const double* A, size_t lda,
                  const double* B, size_t ldb, double beta, double* C, size_t ldc) {
    if (m == 0 || n == 0 || k == 0) return AL_OK;
    if (m > INT32_MAX || n > INT32_MAX || k > INT32_MAX) {
        set_error("gemm: dimensions beyond 2^31 are not supported");
        return AL_ERR_ARG;
    }
    if (upper) return launch<double, true, false, 4>(s, (int)m, (int)n, (int)k, alpha, A, lda, B, ldb, beta, C, ldc);
    return launch<double, false, false, 3>(s, (int)m, (int)n, (int)k, alpha, A, lda, B, ldb, beta, C, ldc);
}

int gemm_f32(hipStream_t s, bool ta, bool tb, size_t m, size_t n, size_t k, float alpha, const float* A, size_t lda,
             const float* B, size_t ldb, float beta, float* C, size_t ldc, int tri) {
    // the 16-byte-access kernel takes the untransposed product when the shapes allow it
    if (!ta && !tb && tri == 0 && m > 0 && n > 0 && k > 0 && alpha != 0.f && m <= INT32_MAX && n <= INT32_MAX &&
        k <= INT32_MAX && sgemm_nn_fits(m, k, A, lda, B, ldb, C, ldc))
        return sgemm_nn(s, (int)m, (int)n, (int)k, alpha, A, lda, B, ldb, beta, C, ldc);
    return gemm_t<float>(s, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tri);
}

}  // namespace al
