// LU with partial pivoting (lapack::dgetrf / dgetrf_rect, src/lapack/mod.rs:112-238) and the
// solve with its factors (dgetrs, :330-354).
//
// The crate factors with an unblocked right-looking loop whose elimination walks rows with a
// stride of n.  This build blocks it for the GPU:
//   * panel: a cooperative kernel (`panel_getf2`) spreads the m x 64 panel over up to 128
//     workgroups of 256 rows, each holding its rows in LDS for the whole panel.  Each column
//     costs one grid-wide exchange: every workgroup publishes its best candidate (first
//     maximum, the crate's strict '>' scan) and that row with write-through stores, then one
//     64-bit word carrying the column's generation, the candidate row and its key's top bits;
//     every workgroup reads all the words, picks the same winner, swaps and eliminates its own
//     rows.  Within the panel the arithmetic is the crate's own (divide, then a separate
//     multiply and subtract per element, no FMA), so a matrix that fits in one panel factors
//     bit-for-bit as the crate factors it.  The workgroups must be resident together: a
//     process-wide ledger (`coop_ledger.hpp`) keeps concurrent panels within the CU count;
//   * two-level blocking: 64-wide panels inside 256-wide outer panels, so the trailing update
//     of the matrix is a k = 256 FP64 MFMA GEMM;
//   * look-ahead on two streams: the workspace's side stream factors an outer panel and
//     prepares the next one's columns (interchanges, U12, their update) while the caller's
//     stream applies the outer panel to every other column; each kernel checks the status word
//     first so that a zero pivot stops the factorization like the crate's early return.
// The panel kernel is also instantiated for float (the mixed-precision factor) and for the
// complex element `zd` (lapack::zgetrf, src/lapack/complex.rs:67-121: |Re| + |Im| pivot keys,
// complex division and multiply-subtract, 32-column panels).
#include "al_internal.hpp"
#include "coop_ledger.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>
#include <vector>

namespace al {
namespace {

constexpr int PNB = 64;        // panel width of the real factorizations
constexpr int PT = 256;        // rows per panel workgroup, one per thread of waves 1-4
constexpr int PTW = PT + 64;   // threads per panel workgroup: wave 0 exchanges, waves 1-4 hold rows
constexpr int CHUNK_MAX = 256; // rows per workgroup, one per thread (LDS: 256 x 64 doubles = 128 KiB)
constexpr int MAX_WG = 128;    // cooperative workgroups per panel (one per CU at most)
constexpr int NBO = 256;       // outer block (trailing-update depth)
std::atomic<bool> g_pinning{true};  // calibrated pinning of the cooperative panels allowed
// Pinned exchange layout (see panel_exchange_init): the publication words in one 8 KiB block of
// the exchange area and three slots of SLOT_PAD words in each of XUSED - 1 others, the blocks
// measured nearest to the chosen half of the chip among XBLOCKS candidates.
constexpr int XBLK_WORDS = 1024;                  // 8 KiB
constexpr int SLOT_PAD = 260;                     // >= every element type's slot_words
constexpr int XUSED = 1 + (2 * MAX_WG + 2) / 3;   // blocks the pinned layout uses
constexpr int XBLOCKS = 6 * XUSED;                // candidate blocks measured

// Per element type: the pivot key's type and the panel width that keeps a CHUNK_MAX-row
// chunk of the panel at 128 KiB of LDS (64 real columns, 32 complex ones).
template <typename T>
struct PanelElem {
    typedef T key_t;
    static constexpr int NB = PNB;
};
template <>
struct PanelElem<zd> {
    typedef double key_t;
    static constexpr int NB = 32;
};
// cooperative exchange slot (global memory), in 64-bit words that each carry a 32-bit part of a
// value in the low half and the column's generation in the high half: the key's parts, the row
// index, then the candidate row and row j, each as its columns' first parts, then their second
// parts, and so on (a store or load of one part for 64 columns covers 512 contiguous bytes)
template <typename T>
constexpr int slot_row() { return (int)(sizeof(typename PanelElem<T>::key_t) / 4) + 1; }
template <typename T>
constexpr int slot_words() { return slot_row<T>() + 2 * PanelElem<T>::NB * (int)(sizeof(T) / 4); }

// Candidate ordering: larger key wins, ties go to the lower row (first maximum).
template <typename T>
__device__ __forceinline__ bool better(T k1, int i1, T k2, int i2) {
    return (k1 > k2) || (k1 == k2 && i1 < i2);
}

// Pivot key of a candidate: |v|, except that NaN never wins unless it sits on the diagonal,
// where the crate's scan starts (src/lapack/mod.rs:126-134): a NaN start blocks every later
// candidate, a later NaN is skipped.
template <typename T>
__device__ __forceinline__ T pivot_key(T v, int g, int j) {
    if (v != v) return (g == j) ? (T)INFINITY : (T)-1;
    return v < (T)0 ? -v : (v == (T)0 ? (T)0 : v);  // -0.0 -> +0.0: keys compare as bit patterns too
}
// complex entries: |Re| + |Im| (src/lapack/complex.rs:78-90), with the same NaN rule
__device__ __forceinline__ double pivot_key(zd v, int g, int j) {
    const double a = abs1(v);
    if (a != a) return (g == j) ? INFINITY : -1.0;
    return a;
}


// Order-preserving integer image of a pivot key: non-negative keys (|v| or +inf) keep their
// IEEE order as unsigned bit patterns; the sentinels -1 (a skipped NaN) and -2 (no candidate)
// sort below every real key.
__device__ __forceinline__ unsigned long long key_code(double k) {
    return k >= 0.0 ? (unsigned long long)__double_as_longlong(k) + 2ull : (k == -1.0 ? 1ull : 0ull);
}
__device__ __forceinline__ unsigned long long key_code(float k) {
    return k >= 0.0f ? (unsigned long long)__float_as_uint(k) + 2ull : (k == -1.0f ? 1ull : 0ull);
}
template <typename T>
__device__ __forceinline__ T key_value(unsigned long long c);
template <>
__device__ __forceinline__ double key_value<double>(unsigned long long c) {
    return c >= 2ull ? __longlong_as_double((long long)(c - 2ull)) : (c == 1ull ? -1.0 : -2.0);
}
template <>
__device__ __forceinline__ float key_value<float>(unsigned long long c) {
    return c >= 2ull ? __uint_as_float((unsigned)(c - 2ull)) : (c == 1ull ? -1.0f : -2.0f);
}

// Agent-scope coherent element accesses for the cooperative panel's exchange area: relaxed
// atomics, which gfx950 issues as write-through stores and L2-missing loads (the sc1 bit), so
// the exchange needs no L2 writeback or invalidate (a concurrent trailing update's dirty lines
// in L2 never have to be flushed for it).  Complex elements travel as their two doubles.
__device__ __forceinline__ void st_agent(double* p, double v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(float* p, float v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(zd* p, zd v) {
    st_agent(&p->re, v.re);
    st_agent(&p->im, v.im);
}
__device__ __forceinline__ double ld_agent(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float ld_agent(const float* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ zd ld_agent(const zd* p) {
    zd v;
    v.re = ld_agent(&p->re);
    v.im = ld_agent(&p->im);
    return v;
}
__device__ __forceinline__ unsigned long long ld_flag(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// A value as generation-tagged words (one per 32-bit part, `stride` words apart), and back: a
// 64-bit relaxed atomic is single-copy atomic, so a word whose tag is the current column's holds
// this column's part.
template <typename V>
__device__ __forceinline__ void put_tagged(unsigned long long* p, V v, unsigned tag, int stride = 1) {
    unsigned u[sizeof(V) / 4];
    __builtin_memcpy(u, &v, sizeof(V));
#pragma unroll
    for (int h = 0; h < (int)(sizeof(V) / 4); ++h)
        __hip_atomic_store(&p[h * stride], ((unsigned long long)tag << 32) | u[h], __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}
template <typename V>
__device__ __forceinline__ bool get_tagged(const unsigned long long* p, V& v, unsigned tag, int stride = 1) {
    unsigned u[sizeof(V) / 4];
    bool ok = true;
#pragma unroll
    for (int h = 0; h < (int)(sizeof(V) / 4); ++h) {
        const unsigned long long x = __hip_atomic_load(&p[h * stride], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok &= (unsigned)(x >> 32) == tag;
        u[h] = (unsigned)x;
    }
    __builtin_memcpy(&v, u, sizeof(V));
    return ok;
}

// The 32-bit head of a pivot key's integer image: orders keys as key_code does except that keys
// differing only below their top 32 bits compare equal (such ties are settled from the full
// keys).  The sentinels -1 and -2 map to 1 and 0.
__device__ __forceinline__ unsigned key_head(double k) {
    return k >= 0.0 ? (unsigned)((unsigned long long)__double_as_longlong(k) >> 32) + 2u : (k == -1.0 ? 1u : 0u);
}
__device__ __forceinline__ unsigned key_head(float k) {  // the whole key: no ties beyond equal keys
    return k >= 0.0f ? __float_as_uint(k) + 2u : (k == -1.0f ? 1u : 0u);
}
// A publication word: the column's 16-bit generation, the candidate's panel row (16 bits; rows
// stay below 32768) and its key head.
__device__ __forceinline__ unsigned long long flag_word(unsigned gen, int idx, unsigned head) {
    return ((unsigned long long)(gen & 0xFFFFu) << 48) | ((unsigned long long)(idx & 0xFFFF) << 32) | head;
}

// Maximum over the wave of unsigned values, in every lane: prefix maxima inside each row of 16
// lanes (DPP row_shr 1, 2, 4, 8), then row 0 into row 1 and row 2 into row 3 (row_bcast:15),
// then rows 0-1 into rows 2-3 (row_bcast:31), so that lane 63 holds the maximum.  A lane without
// a source reads 0, the identity of the unsigned maximum.
__device__ __forceinline__ unsigned wave_max_u32(unsigned v) {
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false));
    return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}
__device__ __forceinline__ double lane_val(double v, int l) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ float lane_val(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ zd lane_val(zd v, int l) { return zd(lane_val(v.re, l), lane_val(v.im, l)); }

// The wave's first maximum in every lane: the largest key (compared as key_code images: the
// high words first, then among the lanes that share the largest high word the low words), ties
// to the lowest row.  Usually the first DPP maximum already names a single lane.
template <typename K>
__device__ __forceinline__ void wave_argmax(K& key, int& idx) {
    const unsigned long long code = key_code(key);
    const unsigned hi = (unsigned)(code >> 32), lo = (unsigned)code;
    bool in = hi == wave_max_u32(hi);
    unsigned long long m = __ballot(in);
    if (__popcll(m) > 1) {
        const unsigned lm = wave_max_u32(in ? lo : 0u);
        in = in && lo == lm;
        m = __ballot(in);
        if (__popcll(m) > 1) {  // equal keys: the lowest row (rows are >= 0, so ~row orders them)
            const unsigned im = wave_max_u32(in ? ~(unsigned)idx : 0u);
            in = in && ~(unsigned)idx == im;
            m = __ballot(in);
        }
    }
    const int wl = __ffsll((long long)m) - 1;
    key = lane_val(key, wl);
    idx = __builtin_amdgcn_readlane(idx, wl);
}

// Workgroup-wide first maximum in every thread: the largest key, ties to the lowest row.  The
// wave's winner by DPP, then the waves' winners in LDS behind one barrier; `par`
// alternates between calls (callers pass a workgroup barrier between two calls, so a call's
// writes never meet the previous call's reads).
template <typename K>
__device__ __forceinline__ void wg_argmax(K& key, int& idx, K (*wk)[PTW / 64], int (*wi)[PTW / 64], int& par) {
    wave_argmax(key, idx);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        wk[par][wave] = key;
        wi[par][wave] = idx;
    }
    __syncthreads();
    key = wk[par][0];
    idx = wi[par][0];
#pragma unroll
    for (int q = 1; q < PTW / 64; ++q)
        if (better(wk[par][q], wi[par][q], key, idx)) {
            key = wk[par][q];
            idx = wi[par][q];
        }
    par ^= 1;
}

#pragma clang fp contract(off)
// Factors the m x ncols panel at P (leading dimension lda), whose top-left element is the
// global diagonal element (row0, row0).  Workgroup w owns panel rows [w*chunk, (w+1)*chunk)
// (chunk <= 256), one per thread of waves 1-4, and keeps them in LDS for the whole panel.
// Wave 0 holds no row: it exchanges the columns' candidates.
//
// Column exchange (COOP, several workgroups): wave 0 writes the workgroup's slot -- its best
// candidate (key, row index) with that row, and row j if it owns it -- as generation-tagged
// words, and publishes one 64-bit word for the column (its generation, the candidate's row and
// the top 32 bits of its key) without waiting for the slot's stores.  Words and slots alternate
// by column parity: a workgroup overwrites parity p only after every workgroup has published the
// next column, which each does only after it has finished reading parity p.  Wave 0 of every
// workgroup polls the words of the column's parity (L2-missing loads) until all carry the
// column's generation; the largest key head names the winner, and only when several workgroups
// share it are their full keys read from the slots.  Then the winner's row and row j come from
// the slots in one round trip, read again until every word carries the column's tag.  No L2
// writeback or invalidate is involved.  Every spin is bounded by `timeout` ticks of the 100 MHz
// wall clock; on expiry the abort flag stops every workgroup and info is set to -1.  A single
// workgroup takes its own candidate without an exchange.
//
// Pipelining: once column j's multipliers are formed, column j+1 is updated and the next
// candidate chosen (one barrier), wave 0 brings the candidate's row and row j+1 up to date,
// publishes them, waits for column j+1's exchange and swaps the two rows involved, while waves
// 1-4 finish column j's elimination on every other row.  The rows wave 0 swaps are the two it
// brought up to date (the global winner is some workgroup's candidate), so the two never touch
// the same row; the pivot row is double-buffered by column parity.
template <typename T, bool COOP>
__global__ __launch_bounds__(PTW) void panel_getf2(T* __restrict__ P, size_t lda, int row0, int m, int ncols,
                                                   int chunk, int* __restrict__ ipiv, int* __restrict__ info,
                                                   unsigned long long* __restrict__ flags,
                                                   unsigned* __restrict__ abort_flag,
                                                   unsigned long long* __restrict__ slots, unsigned gen0,
                                                   long long timeout, int nwg_pin, int pin,
                                                   const unsigned* __restrict__ slot_tab) {
    typedef typename PanelElem<T>::key_t K;
    constexpr int NB = PanelElem<T>::NB;
    constexpr int SW = slot_words<T>(), SR = slot_row<T>(), H = (int)(sizeof(T) / 4);
    extern __shared__ __align__(16) unsigned char smem[];
    T* S = (T*)smem;  // chunk x NB, column-major with leading dimension chunk | 1
    __shared__ T prow_c[2][NB];  // the pivot row of column j in prow_c[j & 1]
    __shared__ K wk[2][PTW / 64];
    __shared__ int wi[2][PTW / 64];
    __shared__ int win_lds;
    __shared__ int ok_lds;
    __shared__ unsigned tab[2 * MAX_WG];  // pinned: word offset of slot (par, q) at par * MAX_WG + q

    // pin >= 0: only the workgroups dispatched to that half of the chip (XCDs 4*pin .. 4*pin+3;
    // workgroups go to XCD blockIdx % 8) take part, nwg_pin of them; the others leave at once
    int nwg = COOP ? (int)gridDim.x : 1, w = COOP ? (int)blockIdx.x : 0;
    if (COOP && pin >= 0) {
        const int xb = (int)(blockIdx.x & 7);
        if ((xb >> 2) != pin) return;
        w = (int)(blockIdx.x >> 3) * 4 + (xb & 3);
        if (w >= nwg_pin) return;
        nwg = nwg_pin;
    }
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (*info != 0) return;  // an earlier panel hit a zero pivot or failed
    const int rb = w * chunk;
    const int ls = chunk | 1;  // LDS column stride: odd, so a row's entries (stride ls) fall in different banks
    const int nr = (m - rb) < chunk ? ((m - rb) > 0 ? (m - rb) : 0) : chunk;
    const int tr = t - 64;  // this thread's row (waves 1-4): panel row rb + tr
    const bool has = tr >= 0 && tr < nr;
    const int g = rb + tr;
    int argpar = 0;

    for (int c = 0; c < ncols; ++c)
        if (has) S[c * ls + tr] = P[(size_t)g + (size_t)c * lda];
    if (COOP && pin >= 0)
        for (int i = t; i < 2 * MAX_WG; i += PTW) tab[i] = slot_tab[i];
    __syncthreads();

    auto slot_ptr = [&](int par, int q) -> unsigned long long* {
        if (pin >= 0) return slots + tab[par * MAX_WG + q];
        return slots + ((size_t)par * nwg + q) * SW;
    };
    // wave 0, several workgroups: this workgroup's slot for column jj and its word; lane c holds
    // column c of the candidate row (vc) and of row jj (vn, when this workgroup owns row jj)
    auto publish = [&](int jj, K key, int idx, T vc, T vn) {
        if (!COOP) return;
        const bool cand = key > (K)-2;
        const bool ownj = jj >= rb && jj < rb + nr;
        const bool col = lane < ncols;  // ncols <= NB <= 64: one column per lane
        // tagged words need no acknowledgement before the publication word: a reader that sees
        // the word before the slot's words have landed sees old tags and reads again
        unsigned long long* my = slot_ptr(jj & 1, w);
        const unsigned tag = gen0 + (unsigned)jj + 1u;
        if (cand && col) put_tagged(&my[SR + lane], vc, tag, NB);
        if (ownj && col) put_tagged(&my[SR + H * NB + lane], vn, tag, NB);
        if (lane == 0) {
            put_tagged(&my[0], key, tag);
            put_tagged(&my[SR - 1], (unsigned)idx, tag);
            __hip_atomic_store(&flags[w * 8 + (jj & 1)], flag_word(tag, idx, key_head(key)), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
    };
    // wave 0: column jj's winner (a single workgroup's winner is its candidate idx with row vc),
    // the two rows swapped where this workgroup holds them, the pivot row in prow_c[jj & 1] and
    // its index in win_lds; ok_lds = 0 on abort
    auto exchange = [&](int jj, int idx, T vc, T vn) {
        int ok = 1, bidx = idx;
        T pv = vc, rv = vn;
        if (COOP) {
            const int par = jj & 1;
            const unsigned tag = gen0 + (unsigned)jj + 1u;
            const unsigned long long want = (unsigned long long)(tag & 0xFFFFu);
            const long long t0 = wall_clock64();
            // lane 0 checks the abort flag and the time limit; the wave leaves together
            auto give_up = [&]() -> bool {
                // a short wait only polls: the abort flag is read once the wait passes 20 us
                const long long el = wall_clock64() - t0;
                if (el < 2000 && el <= timeout) return false;
                int stop = 0;
                if (lane == 0) {
                    if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
                        stop = 1;
                    } else if (el > timeout) {
                        __hip_atomic_store(abort_flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        stop = 1;
                    }
                }
                return __shfl(stop, 0, 64) != 0;
            };
            unsigned long long f0 = 0, f1 = 0;
            for (;;) {
                f0 = lane < nwg ? ld_flag(&flags[lane * 8 + par]) : 0ull;
                f1 = lane + 64 < nwg ? ld_flag(&flags[(lane + 64) * 8 + par]) : 0ull;
                const bool mine = (lane >= nwg || (f0 >> 48) == want) && (lane + 64 >= nwg || (f1 >> 48) == want);
                if (__all(mine)) break;
                if (give_up()) {
                    ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            int bq = 0;  // the winning workgroup
            if (ok) {
                const unsigned h0 = lane < nwg ? (unsigned)f0 : 0u;
                const unsigned h1 = lane + 64 < nwg ? (unsigned)f1 : 0u;
                const unsigned hmax = wave_max_u32(h0 > h1 ? h0 : h1);
                const bool t0m = lane < nwg && h0 == hmax, t1m = lane + 64 < nwg && h1 == hmax;
                const unsigned long long b0 = __ballot(t0m), b1 = __ballot(t1m);
                if (__popcll(b0) + __popcll(b1) == 1) {
                    bq = b0 ? __ffsll((long long)b0) - 1 : 64 + __ffsll((long long)b1) - 1;
                    const unsigned long long fw = __shfl(bq < 64 ? f0 : f1, bq & 63, 64);
                    bidx = (int)((fw >> 32) & 0xFFFFu);
                } else {
                    // several workgroups share the largest key head: compare their full keys
                    for (;;) {
                        K bk = (K)-3;
                        int bi = 0x7fffffff;
                        bool good = true;
                        for (int q = lane; q < nwg; q += 64) {
                            if (!(q < 64 ? t0m : t1m)) continue;
                            const unsigned long long* sq = slot_ptr(par, q);
                            K kq;
                            unsigned iq;
                            good &= get_tagged(&sq[0], kq, tag);
                            good &= get_tagged(&sq[SR - 1], iq, tag);
                            if (better(kq, (int)iq, bk, bi)) {
                                bk = kq;
                                bi = (int)iq;
                            }
                        }
#pragma unroll
                        for (int off = 32; off >= 1; off >>= 1) {
                            const K ok2 = __shfl_xor(bk, off, 64);
                            const int oi = __shfl_xor(bi, off, 64);
                            if (better(ok2, oi, bk, bi)) {
                                bk = ok2;
                                bi = oi;
                            }
                        }
                        if (__all(good)) {
                            bq = bi / chunk;
                            bidx = bi;
                            break;
                        }
                        if (give_up()) {
                            ok = 0;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
            }
            if (ok) {
                // the winner's row and row jj (one round trip, repeated while the slots' words
                // have not landed); the winner's index came with its word, and its key is the
                // pivot key of the row's entry in column jj
                const unsigned long long* win = slot_ptr(par, bq);
                const unsigned long long* own = slot_ptr(par, jj / chunk);
                for (;;) {
                    bool good = true;
                    if (lane < ncols) {
                        good &= get_tagged(&win[SR + lane], pv, tag, NB);
                        good &= get_tagged(&own[SR + H * NB + lane], rv, tag, NB);
                    }
                    if (__all(good)) break;
                    if (give_up()) {
                        ok = 0;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
        }
        if (ok) {
            // swap rows jj and bidx where this workgroup holds them (not before a zero pivot: the
            // crate stops before its swap), and keep the pivot row for every thread
            if (pivot_key(lane_val(pv, jj), bidx, jj) > (K)0 && bidx != jj && lane < ncols) {
                if (bidx >= rb && bidx < rb + nr) S[lane * ls + (bidx - rb)] = rv;
                if (jj >= rb && jj < rb + nr) S[lane * ls + (jj - rb)] = pv;
            }
            if (lane < ncols) prow_c[jj & 1][lane] = pv;
            if (lane == 0) win_lds = bidx;
        }
        if (lane == 0) ok_lds = ok;
    };

    {
        // column 0: the workgroup's candidate, published and exchanged
        K key = has ? pivot_key(S[tr], g, 0) : (K)-2;
        int idx = has ? g : 0x7fffffff;
        wg_argmax(key, idx, wk, wi, argpar);
        if (wave == 0) {
            T vc = T(0), vn = T(0);
            if (lane < ncols) {
                if (key > (K)-2) vc = S[lane * ls + (idx - rb)];
                if (rb == 0 && nr > 0) vn = S[lane * ls];
            }
            publish(0, key, idx, vc, vn);
            exchange(0, idx, vc, vn);
        }
        __syncthreads();
    }

    for (int j = 0; j < ncols; ++j) {
        if (!ok_lds) {
            if (w == 0 && t == 0) *info = -1;
            break;
        }
        const T* prow = prow_c[j & 1];
        const int p = win_lds;
        if (!(pivot_key(prow[j], p, j) > (K)0)) {
            // |pivot| <= 0: singular (src/lapack/mod.rs:136-138); a NaN diagonal carries +inf
            if (w == 0 && t == 0) *info = row0 + j + 1;
            break;
        }
        if (w == 0 && t == 0) ipiv[row0 + j] = row0 + p;

        const T d = prow[j];
        const bool mine = has && g > j;
        if (j + 1 < ncols) {
            // multipliers of column j and column j+1's update, then the next candidate
            K key = (K)-2;
            int idx = 0x7fffffff;
            if (mine) {
                const T f = S[j * ls + tr] / d;
                S[j * ls + tr] = f;
                const T v = S[(j + 1) * ls + tr] - f * prow[j + 1];
                S[(j + 1) * ls + tr] = v;
                key = pivot_key(v, g, j + 1);
                idx = g;
            }
            wg_argmax(key, idx, wk, wi, argpar);
            const int lc = (key > (K)-2) ? idx - rb : -1;
            const int ln = (j + 1 >= rb && j + 1 < rb + nr) ? j + 1 - rb : -1;
            if (wave == 0) {
                // the candidate's row and row j+1 brought up to date (a lane per column),
                // published, then column j+1's exchange
                T vc = T(0), vn = T(0);
                if (lane < ncols) {
                    if (lc >= 0) vc = S[lane * ls + lc];
                    if (ln >= 0) vn = S[lane * ls + ln];
                    if (lane >= j + 2) {
                        if (lc >= 0) {
                            vc = vc - S[j * ls + lc] * prow[lane];
                            S[lane * ls + lc] = vc;
                        }
                        if (ln >= 0 && ln != lc) {
                            vn = vn - S[j * ls + ln] * prow[lane];
                            S[lane * ls + ln] = vn;
                        }
                    }
                    if (ln >= 0 && ln == lc) vn = vc;
                }
                publish(j + 1, key, idx, vc, vn);
                exchange(j + 1, idx, vc, vn);
            } else if (mine && tr != lc && tr != ln) {
                // column j's elimination on every other row
                const T f = S[j * ls + tr];
                int c = j + 2;
                for (; c + 8 <= ncols; c += 8) {
                    T v[8];
#pragma unroll
                    for (int q = 0; q < 8; ++q) v[q] = S[(c + q) * ls + tr];
#pragma unroll
                    for (int q = 0; q < 8; ++q) S[(c + q) * ls + tr] = v[q] - f * prow[c + q];
                }
                for (; c < ncols; ++c) S[c * ls + tr] -= f * prow[c];
            }
            __syncthreads();
        } else {
            if (mine) S[j * ls + tr] = S[j * ls + tr] / d;
            __syncthreads();
        }
    }
    for (int c = 0; c < ncols; ++c)
        if (has) P[(size_t)g + (size_t)c * lda] = S[c * ls + tr];
}
#pragma clang fp contract(on)

template <typename T>
int launch_panel(hipStream_t s, Workspace* ws, T* Ppanel, size_t lda, size_t row0, size_t m, size_t ncols,
                 int* ipiv, int* info) {
    constexpr int NB = PanelElem<T>::NB;
    static bool attr_set = false;
    if (!attr_set) {
        AL_HIP(hipFuncSetAttribute((const void*)panel_getf2<T, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)((CHUNK_MAX + 1) * NB * sizeof(T))));
        AL_HIP(hipFuncSetAttribute((const void*)panel_getf2<T, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)((CHUNK_MAX + 1) * NB * sizeof(T))));
        attr_set = true;
    }
    if (ncols > (size_t)NB) {
        set_error("getrf: panel wider than the panel kernel holds");
        return AL_ERR_ARG;
    }
    size_t nwg = (m + CHUNK_MAX - 1) / CHUNK_MAX;
    if (nwg > (size_t)MAX_WG) nwg = MAX_WG;
    const size_t chunk = (m + nwg - 1) / nwg;
    if (chunk > (size_t)CHUNK_MAX) {
        set_error("getrf: panel taller than the cooperative kernel holds");
        return AL_ERR_ARG;
    }
    // publication words carry 16-bit generations: start a new epoch (words cleared) before one
    // would wrap, so a word left from an earlier panel never matches a current generation
    if (ws->gen + (unsigned)ncols >= 0xFFFFu) {
        AL_CHECK(fill_u32(s, ws->flags, (size_t)MAX_WG * 16, 0u));
        AL_CHECK(fill_u32(s, (unsigned*)ws->slots, ws->slots_bytes / sizeof(unsigned), 0u));  // tags 0: never current
        ws->gen = 0;
    }
    const long long timeout = wait_limit_ticks();
    const unsigned gen0 = ws->gen;
    ws->gen += (unsigned)ncols;
    const int ticket = prof_begin(s);
    if (nwg > 1 && ws->pin_half >= 0 && g_pinning.load(std::memory_order_relaxed)) {
        // every workgroup on one half of the chip, the exchange area's words in memory near it;
        // eight workgroups per four taking part (the other half's leave at once)
        const unsigned grid = 8u * (unsigned)((nwg + 3) / 4);
        unsigned long long* area = (unsigned long long*)ws->slots;
        hipLaunchKernelGGL((panel_getf2<T, true>), dim3(grid), dim3(PTW), (chunk | 1) * NB * sizeof(T), s, Ppanel, lda,
                           (int)row0, (int)m, (int)ncols, (int)chunk, ipiv, info, area + ws->pin_flags,
                           ws->abort_flag, area, gen0, timeout, (int)nwg, ws->pin_half, ws->slot_tab);
    } else if (nwg > 1) {
        hipLaunchKernelGGL((panel_getf2<T, true>), dim3((unsigned)nwg), dim3(PTW), (chunk | 1) * NB * sizeof(T), s, Ppanel,
                           lda, (int)row0, (int)m, (int)ncols, (int)chunk, ipiv, info, (unsigned long long*)ws->flags, ws->abort_flag,
                           (unsigned long long*)ws->slots, gen0, timeout, (int)nwg, -1, nullptr);
    } else {
        hipLaunchKernelGGL((panel_getf2<T, false>), dim3(1), dim3(PTW), (chunk | 1) * NB * sizeof(T), s, Ppanel, lda,
                           (int)row0, (int)m, (int)ncols, (int)chunk, ipiv, info, (unsigned long long*)ws->flags, ws->abort_flag,
                           (unsigned long long*)ws->slots, gen0, timeout, 1, -1, nullptr);
    }
    AL_HIP(hipGetLastError());
    if (ticket >= 0) {
        // unblocked LU of an m x ncols panel: m*ncols^2 - ncols^3/3 multiply-adds' worth of flops
        // (four times that for complex entries); the panel read and written once
        const double dm = (double)m, db = (double)ncols;
        const bool cplx = sizeof(T) == 2 * sizeof(double);
        prof_end(ticket, s, cplx ? "zpanel_getf2" : "panel_getf2",
                 (cplx ? 4.0 : 1.0) * (dm * db * db - db * db * db / 3.0), 2.0 * sizeof(T) * dm * db);
    }
    return AL_OK;
}

// Factors the m x n block at (k0, k0) of A with 64-wide cooperative panels; swaps are applied
// inside columns [k0, k0 + n) only -- the caller applies them outside.
int panel_blocked(hipStream_t s, Workspace* ws, double* A, size_t lda, size_t M, size_t k0, size_t n, int* ipiv,
                  int* info) {
    for (size_t kk = k0; kk < k0 + n; kk += PNB) {
        const size_t ib = (k0 + n - kk) < (size_t)PNB ? (k0 + n - kk) : (size_t)PNB;
        AL_CHECK(launch_panel<double>(s, ws, A + kk + kk * lda, lda, kk, M - kk, ib, ipiv, info));
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

// LU without pivoting of the nb x nb (nb <= 64) diagonal block at A, in place, in one
// workgroup with the block in LDS: per column, the multipliers below the diagonal, then the
// rank-1 update of the block right of and below it (fused multiply-adds).  A zero or
// non-finite pivot sets info to its 1-based global column (col0 + j + 1) and stops; the
// caller then factors with pivoting instead.  Used by the factorization in a given row order.
__global__ __launch_bounds__(256) void getf2_nopiv_k(double* __restrict__ A, size_t lda, int nb, int col0,
                                                     int* __restrict__ info) {
    if (*info != 0) return;
    __shared__ double S[PNB][PNB + 1];
    const int t = threadIdx.x;
    const int i = t & (PNB - 1), q = t >> 6;  // this thread: row i, columns q, q + 4, q + 8, ..
    for (int c = q; c < nb; c += 4) S[i][c] = i < nb ? A[(size_t)i + (size_t)c * lda] : 0.0;
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
        const double piv = S[j][j];
        if (!(fabs(piv) > 0.0) || !(fabs(piv) < INFINITY)) {  // uniform: every thread reads the same value
            if (t == 0) *info = col0 + j + 1;
            break;
        }
        const double l = S[i][j] / piv;  // row i's multiplier (used only below the diagonal)
        __syncthreads();                  // every thread has read column j
        if (i > j && i < nb) {
            if (q == 0) S[i][j] = l;
            for (int c = j + 1 + ((q - j - 1) & 3); c < nb; c += 4) S[i][c] = fma(-l, S[j][c], S[i][c]);
        }
        __syncthreads();
    }
    for (int c = q; c < nb; c += 4)
        if (i < nb) A[(size_t)i + (size_t)c * lda] = S[i][c];
}

// The m x n outer panel at (k0, k0) of A factored without pivoting, 64 columns at a time: the
// diagonal block's LU, L21 := A21 U11^{-1}, then U12 and the in-panel trailing product on the
// outer panel's later columns.
int panel_nopiv(hipStream_t s, double* A, size_t lda, size_t M, size_t k0, size_t n, int* info) {
    for (size_t kk = k0; kk < k0 + n; kk += PNB) {
        const size_t ib = (k0 + n - kk) < (size_t)PNB ? (k0 + n - kk) : (size_t)PNB;
        double* Akk = A + kk + kk * lda;
        hipLaunchKernelGGL(getf2_nopiv_k, dim3(1), dim3(256), 0, s, Akk, lda, (int)ib, (int)kk, info);
        AL_HIP(hipGetLastError());
        if (M > kk + ib) AL_CHECK(trsm_f64(s, false, false, false, false, M - kk - ib, ib, 1.0, Akk, lda, Akk + ib, lda, info));
        const size_t rest = k0 + n - kk - ib;
        if (rest > 0) {
            double* A12 = Akk + ib * lda;
            AL_CHECK(trsm_f64(s, true, true, false, true, ib, rest, 1.0, Akk, lda, A12, lda, info));
            if (M > kk + ib)
                AL_CHECK(gemm_f64(s, false, false, M - kk - ib, rest, ib, -1.0, Akk + ib, lda, A12, lda, 1.0, A12 + ib,
                                  lda));
        }
    }
    return AL_OK;
}

}  // namespace

int panel_f32(hipStream_t s, Workspace* ws, float* P, size_t lda, size_t row0, size_t m, size_t ncols, int* ipiv,
              int* info) {
    return launch_panel<float>(s, ws, P, lda, row0, m, ncols, ipiv, info);
}

int panel_z(hipStream_t s, Workspace* ws, zd* P, size_t lda, size_t row0, size_t m, size_t ncols, int* ipiv,
            int* info) {
    return launch_panel<zd>(s, ws, P, lda, row0, m, ncols, ipiv, info);
}
int panel_z_width() { return PanelElem<zd>::NB; }

namespace {

std::atomic<long long> g_wait_limit{0};

// Cooperative panel workgroups in flight, per device (see coop_ledger.hpp), with a pool of
// timing-free events that mark the end of each reserving factorization.
struct PanelLedger {
    CoopLedger<hipEvent_t> ledger;
    std::mutex pool_mu;
    std::vector<hipEvent_t> pool;
};
constexpr int MAX_DEVICES = 64;
PanelLedger g_ledgers[MAX_DEVICES];

PanelLedger* ledger_of_current() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) return nullptr;
    PanelLedger* L = &g_ledgers[dev];
    // half the CUs: a panel pinned to one half of the chip must find all its workgroups' CUs there
    if (L->ledger.capacity() == 0) L->ledger.set_capacity(std::max(1, cu_count() / 2));
    return L;
}

bool event_done(hipEvent_t e) {
    const hipError_t r = hipEventQuery(e);
    if (r == hipErrorNotReady) return false;
    if (r != hipSuccess) (void)hipGetLastError();  // a failed event frees its reservation too
    return true;
}

// Workgroups of the largest (first) cooperative panel of an m-row factorization; 0 when the
// panels run in one workgroup, which waits on nobody.
int panel_units(size_t m) {
    if (m <= (size_t)CHUNK_MAX) return 0;
    size_t nwg = (m + CHUNK_MAX - 1) / CHUNK_MAX;
    return (int)(nwg > (size_t)MAX_WG ? (size_t)MAX_WG : nwg);
}

}  // namespace

long long wait_limit_ticks() {
    const long long v = g_wait_limit.load(std::memory_order_relaxed);
    return v > 0 ? v : 200000000LL;  // default: 2 s of the 100 MHz wall clock
}

void set_wait_limit_ticks(long long ticks) { g_wait_limit.store(ticks, std::memory_order_relaxed); }

int coop_reserve(hipStream_t s, size_t m, int* units_out) {
    *units_out = 0;
    const int units = panel_units(m);
    if (units == 0) return AL_OK;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) return AL_OK;
    PanelLedger* L = ledger_of_current();
    if (!L) return AL_OK;
    auto recycle = [L](hipEvent_t e) {
        std::lock_guard<std::mutex> lk(L->pool_mu);
        L->pool.push_back(e);
    };
    *units_out = L->ledger.acquire(units, event_done, recycle);
    return AL_OK;
}

int coop_hold(hipStream_t s, int units) {
    if (units <= 0) return AL_OK;
    PanelLedger* L = ledger_of_current();
    if (!L) return AL_OK;
    hipEvent_t e = nullptr;
    {
        std::lock_guard<std::mutex> lk(L->pool_mu);
        if (!L->pool.empty()) {
            e = L->pool.back();
            L->pool.pop_back();
        }
    }
    if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) e = nullptr;
    if (!e || hipEventRecord(e, s) != hipSuccess) {
        // no event to tie the reservation to: wait for the work, then return it
        if (e) {
            std::lock_guard<std::mutex> lk(L->pool_mu);
            L->pool.push_back(e);
        }
        (void)hipGetLastError();
        const hipError_t r = hipStreamSynchronize(s);
        L->ledger.release(units);
        AL_HIP(r);
        return AL_OK;
    }
    L->ledger.hold_until(e, units);
    return AL_OK;
}

void coop_release(int units) {
    if (units <= 0) return;
    PanelLedger* L = ledger_of_current();
    if (L) L->ledger.release(units);
}

int coop_in_use() {
    PanelLedger* L = ledger_of_current();
    if (!L) return 0;
    auto recycle = [L](hipEvent_t e) {
        std::lock_guard<std::mutex> lk(L->pool_mu);
        L->pool.push_back(e);
    };
    return L->ledger.in_use(event_done, recycle);
}

int getrf_f64(hipStream_t s, Workspace* ws, size_t M, size_t N, double* A, size_t lda, int* ipiv, int* info,
              double* perm) {
    if (M > (size_t)CHUNK_MAX * MAX_WG || N > INT32_MAX) {
        set_error("getrf: more rows than the cooperative panel holds (32768)");
        return AL_ERR_ARG;
    }
    // the panels' workgroups wait on each other: reserve room for them beside the cooperative
    // panels of other factorizations in flight, and hold it until this one's last kernel ends
    int units = 0;
    AL_CHECK(coop_reserve(s, M, &units));
    const int st = getrf_enqueue(s, ws, M, N, A, lda, ipiv, info, perm);
    if (st != AL_OK) {
        // kernels already enqueued may still run: let them finish before the room is returned
        (void)hipStreamSynchronize(s);
        if (ws->side) (void)hipStreamSynchronize(ws->side);
        coop_release(units);
        return st;
    }
    return coop_hold(s, units);
}

int ensure_side_stream(Workspace* ws) {
    if (!ws->side) AL_HIP(hipStreamCreateWithFlags(&ws->side, hipStreamNonBlocking));
    if (!ws->ev_side) AL_HIP(hipEventCreateWithFlags(&ws->ev_side, hipEventDisableTiming));
    if (!ws->ev_main) AL_HIP(hipEventCreateWithFlags(&ws->ev_main, hipEventDisableTiming));
    return AL_OK;
}

// Blocked right-looking LU with look-ahead on two streams (when there is more than one outer
// panel).  The side stream carries the chain that bounds the run: it factors an outer panel
// (256 columns as four cooperative 64-column panels with their in-panel updates), then applies
// that panel's interchanges, U12 solve and trailing product to the next panel's columns only,
// and factors the next panel.  The caller's stream applies the same interchanges to every other
// column (and to perm), solves U12 there and runs the large trailing product, while the side
// stream factors the next panel.  The two streams touch disjoint columns; two events order
// them: the side stream's preparation of a panel waits for the main stream's previous update
// (that update covers the panel's columns), and the main stream's update waits for the panel.
// Each stream composes the interchanges in its own pair buffer.  The pivots, and every
// operation applied to any column, are those of the single-stream order.
namespace {

// The blocked factorization with look-ahead; pivot false: no row interchanges at all (ipiv and
// perm untouched), the panels factored by panel_nopiv.
int getrf_la(hipStream_t s, Workspace* ws, size_t M, size_t N, double* A, size_t lda, int* ipiv, int* info,
             double* perm, bool pivot) {
    const size_t kmax = M < N ? M : N;
    AL_CHECK(fill_i32(s, info, 1, 0));
    AL_CHECK(fill_u32(s, ws->abort_flag, 1, 0u));
    if (perm && pivot) AL_CHECK(perm_init(s, perm, M));
    auto panel = [&](hipStream_t q, size_t k, size_t jb) -> int {
        return pivot ? panel_blocked(q, ws, A, lda, M, k, jb, ipiv, info) : panel_nopiv(q, A, lda, M, k, jb, info);
    };
    auto swaps = [&](hipStream_t q, size_t k, size_t jb, double* A0, size_t n0, double* A1, size_t n1, double* pv,
                     int* pairs) -> int {
        if (!pivot) return AL_OK;
        return swap_block_f64(q, ws, ipiv, k, jb, A0, n0, A1, n1, lda, pv, info, pairs);
    };
    auto update = [&](hipStream_t q, size_t k, size_t jb, size_t c0, size_t c1) -> int {
        // U12 and the trailing product of the outer panel [k, k + jb) on columns [c0, c1)
        if (c1 <= c0) return AL_OK;
        double* A12 = A + k + c0 * lda;
        AL_CHECK(trsm_f64(q, true, true, false, true, jb, c1 - c0, 1.0, A + k + k * lda, lda, A12, lda, info));
        if (M > k + jb)
            AL_CHECK(gemm_f64(q, false, false, M - k - jb, c1 - c0, jb, -1.0, A + (k + jb) + k * lda, lda, A12, lda,
                              1.0, A + (k + jb) + c0 * lda, lda));
        return AL_OK;
    };
    if (kmax <= (size_t)NBO) {
        for (size_t k = 0; k < kmax; k += NBO) {
            const size_t jb = (kmax - k) < (size_t)NBO ? (kmax - k) : (size_t)NBO;
            AL_CHECK(panel(s, k, jb));
            const size_t right = (k + jb < N) ? N - k - jb : 0;
            AL_CHECK(swaps(s, k, jb, A, k, A + (k + jb) * lda, right, perm, nullptr));
            AL_CHECK(update(s, k, jb, k + jb, N));
        }
        return AL_OK;
    }
    AL_CHECK(ensure_side_stream(ws));
    hipStream_t side = ws->side;
    AL_HIP(hipEventRecord(ws->ev_main, s));
    AL_HIP(hipStreamWaitEvent(side, ws->ev_main, 0));
    AL_CHECK(panel(side, 0, NBO));
    for (size_t k = 0; k < kmax; k += NBO) {
        const size_t jb = (kmax - k) < (size_t)NBO ? (kmax - k) : (size_t)NBO;
        const size_t ke = k + jb;
        if (ke >= kmax) {  // the last panel: the columns left of it (and right of kmax when N > M)
            AL_HIP(hipEventRecord(ws->ev_side, side));
            AL_HIP(hipStreamWaitEvent(s, ws->ev_side, 0));
            AL_CHECK(swaps(s, k, jb, A, k, A + ke * lda, N - ke, perm, ws->pairs2));
            AL_CHECK(update(s, k, jb, ke, N));
            break;
        }
        const size_t ke2 = (kmax - ke) < (size_t)NBO ? kmax : ke + NBO;
        // side: the next panel's columns, once the main stream's last update of them is done
        AL_HIP(hipStreamWaitEvent(side, ws->ev_main, 0));
        AL_CHECK(swaps(side, k, jb, A + ke * lda, ke2 - ke, nullptr, 0, nullptr, ws->pairs));
        AL_CHECK(update(side, k, jb, ke, ke2));
        AL_HIP(hipEventRecord(ws->ev_side, side));
        AL_CHECK(panel(side, ke, ke2 - ke));
        // main: every other column
        AL_HIP(hipStreamWaitEvent(s, ws->ev_side, 0));
        AL_CHECK(swaps(s, k, jb, A, k, A + ke2 * lda, N - ke2, perm, ws->pairs2));
        AL_CHECK(update(s, k, jb, ke2, N));
        AL_HIP(hipEventRecord(ws->ev_main, s));
    }
    return AL_OK;
}

}  // namespace

int getrf_enqueue(hipStream_t s, Workspace* ws, size_t M, size_t N, double* A, size_t lda, int* ipiv, int* info,
                  double* perm) {
    return getrf_la(s, ws, M, N, A, lda, ipiv, info, perm, true);
}

int getrf_nopiv_f64(hipStream_t s, Workspace* ws, size_t n, double* A, size_t lda, int* info) {
    return getrf_la(s, ws, n, n, A, lda, nullptr, info, nullptr, false);
}

int getrs_f64(hipStream_t s, Workspace* ws, size_t n, size_t nrhs, const double* LU, size_t lda, const int* ipiv,
              const double* perm, double* B, size_t ldb, const int* info) {
    if (n == 0 || nrhs == 0) return AL_OK;
    AL_CHECK(ws_reserve(ws, n, n * nrhs));
    // the solves' waits raise this flag when they give up; it must not carry over from an
    // earlier call on this workspace
    AL_CHECK(fill_u32(s, ws->abort_flag, 1, 0u));
    if (!perm) {
        AL_CHECK(perm_from_ipiv(s, ipiv, n, ws->perm));
        perm = ws->perm;
    }
    // apply the interchanges as one gather into the staging block, solve there, copy back
    AL_CHECK(gather_rows(s, B, ldb, n, nrhs, perm, ws->wbuf));
    if (trsv_fits(nrhs)) {
        AL_CHECK(trsv_solve<double>(s, ws, true, true, n, nrhs, LU, lda, ws->wbuf, n));
        AL_CHECK(trsv_solve<double>(s, ws, false, false, n, nrhs, LU, lda, ws->wbuf, n));
    } else {
        AL_CHECK(trsm_f64(s, true, true, false, true, n, nrhs, 1.0, LU, lda, ws->wbuf, n, info));
        AL_CHECK(trsm_f64(s, true, false, false, false, n, nrhs, 1.0, LU, lda, ws->wbuf, n, info));
    }
    AL_CHECK(copy_block(s, ws->wbuf, n, nrhs, B, ldb));
    return AL_OK;
}

size_t panel_slots_bytes() {
    const int words = std::max(slot_words<double>(), std::max(slot_words<float>(), slot_words<zd>()));
    static_assert(slot_words<double>() <= SLOT_PAD && slot_words<zd>() <= SLOT_PAD && 3 * SLOT_PAD <= XBLK_WORDS,
                  "three slots per 8 KiB block");
    static_assert(MAX_WG * 8 <= XBLK_WORDS, "the publication words fit one block");
    const size_t plain = (size_t)2 * MAX_WG * (size_t)words * sizeof(unsigned long long);
    const size_t pinned = (size_t)XBLOCKS * XBLK_WORDS * sizeof(unsigned long long);
    return std::max(plain, pinned);
}

namespace {

// Round trips between two workgroups on XCD 0 (workgroups 0 and 8) and, at the same time, two on
// XCD 4 (workgroups 4 and 12), through relaxed agent-scope words in each 8 KiB block of the
// exchange area; every workgroup records the XCD it runs on.  Each spin is bounded; a block whose
// round trips did not complete records -1 and raises a stop word that ends every pair's loop.
__global__ __launch_bounds__(64) void xcd_calib_k(unsigned long long* __restrict__ area, int nblk, int rts,
                                                  long long* __restrict__ out, unsigned* __restrict__ xcc) {
    const int g = (int)blockIdx.x;
    if (threadIdx.x != 0) return;
    xcc[g] = (unsigned)__builtin_amdgcn_s_getreg((3 << 11) | 20) & 15u;  // HW_REG_XCC_ID
    if ((g & 3) != 0 || g >= 16) return;
    const int pair = (g >> 2) & 1;  // 0: XCD 0, 1: XCD 4
    const bool lead = g < 8;
    unsigned long long* stop = area + 512;
    for (int b = 0; b < nblk; ++b) {
        if (__hip_atomic_load(stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0ull) break;
        unsigned long long* fa = area + (size_t)b * XBLK_WORDS + pair * 64;
        unsigned long long* fb = fa + 16;
        const long long t0 = wall_clock64();
        bool bad = false;
        for (int i = 0; i < rts && !bad; ++i) {
            const unsigned long long v = ((unsigned long long)(b + 1) << 32) | (unsigned long long)(2 * i + 1);
            long long spins = 0;
            if (lead) {
                __hip_atomic_store(fa, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                while (__hip_atomic_load(fb, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != v)
                    if (++spins > 100000) {
                        bad = true;
                        break;
                    }
            } else {
                while (__hip_atomic_load(fa, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != v)
                    if (++spins > 100000) {
                        bad = true;
                        break;
                    }
                __hip_atomic_store(fb, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (lead) out[pair * nblk + b] = bad ? -1 : wall_clock64() - t0;
        if (bad) {
            __hip_atomic_store(stop, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
    }
}

}  // namespace

void set_panel_pinning(bool on) { g_pinning.store(on, std::memory_order_relaxed); }

// Decides whether the cooperative panels of this workspace run on one half of the chip.  The
// agent-scope words of the exchange travel to the memory that holds them, and an 8 KiB block of
// the area is nearer to some XCDs than to others (round trips of about 0.6, 0.8 or 1.05 us between
// two workgroups of one XCD, by block); a column's exchange is a few such trips.  A calibration
// launch measures every block from XCD 0 and from XCD 4.  If workgroup i runs on XCD i % 8, the
// half (XCDs 0-3 or 4-7) that more blocks are clearly nearer to is chosen, and its XUSED nearest
// blocks hold the publication words and the slots (ws->pin_flags, ws->slot_tab); with too few
// such blocks the panels run anywhere.  Then the area is cleared.
int panel_exchange_init(hipStream_t s, Workspace* ws) {
    ws->pin_half = -1;
    unsigned long long* area = (unsigned long long*)ws->slots;
    AL_CHECK(fill_u32(s, (unsigned*)area, ws->slots_bytes / sizeof(unsigned), 0u));
    if (ws->slots_bytes >= (size_t)XBLOCKS * XBLK_WORDS * sizeof(unsigned long long)) {
        const int nblk = XBLOCKS;
        long long* d_out = nullptr;
        AL_HIP(hipMalloc((void**)&d_out, 2 * (size_t)nblk * sizeof(long long) + 64 * sizeof(unsigned)));
        unsigned* d_xcc = (unsigned*)(d_out + 2 * nblk);
        std::vector<long long> t(2 * (size_t)nblk, 0);
        std::vector<unsigned> xcc(16, 99u);
        hipLaunchKernelGGL(xcd_calib_k, dim3(16), dim3(64), 0, s, area, nblk, 12, d_out, d_xcc);
        hipError_t r = hipGetLastError();
        if (r == hipSuccess) r = hipMemcpyAsync(t.data(), d_out, t.size() * sizeof(long long), hipMemcpyDeviceToHost, s);
        if (r == hipSuccess) r = hipMemcpyAsync(xcc.data(), d_xcc, 16 * sizeof(unsigned), hipMemcpyDeviceToHost, s);
        if (r == hipSuccess) r = hipStreamSynchronize(s);
        (void)hipFree(d_out);
        AL_HIP(r);
        bool ok = true;
        for (int g = 0; g < 16; ++g) ok = ok && xcc[g] == (unsigned)(g % 8);
        for (int b = 0; b < nblk && ok; ++b) ok = t[b] > 0 && t[nblk + b] > 0;
        if (ok) {
            // blocks at least 10% nearer to one half, by half, nearest first
            std::vector<std::pair<long long, int>> near[2];
            for (int b = 0; b < nblk; ++b) {
                const long long a = t[b], c = t[nblk + b];
                if (a * 10 < c * 9) near[0].push_back({a, b});
                if (c * 10 < a * 9) near[1].push_back({c, b});
            }
            const int h = near[1].size() > near[0].size() ? 1 : 0;
            if ((int)near[h].size() >= XUSED) {
                std::sort(near[h].begin(), near[h].end());
                std::vector<unsigned> tab(2 * MAX_WG);
                for (int k = 0; k < 2 * MAX_WG; ++k)
                    tab[k] = (unsigned)near[h][1 + k / 3].second * XBLK_WORDS + (unsigned)(k % 3) * SLOT_PAD;
                if (!ws->slot_tab) AL_HIP(hipMalloc((void**)&ws->slot_tab, tab.size() * sizeof(unsigned)));
                AL_HIP(hipMemcpy(ws->slot_tab, tab.data(), tab.size() * sizeof(unsigned), hipMemcpyHostToDevice));
                ws->pin_flags = (size_t)near[h][0].second * XBLK_WORDS;
                ws->pin_half = h;
            }
        }
        AL_CHECK(fill_u32(s, (unsigned*)area, ws->slots_bytes / sizeof(unsigned), 0u));
    }
    AL_HIP(hipStreamSynchronize(s));
    return AL_OK;
}
size_t panel_flags_count() { return (size_t)MAX_WG * 16; }

}  // namespace al
