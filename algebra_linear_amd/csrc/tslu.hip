// Tournament-pivoted 64-column panel (TSLU, the panel of communication-avoiding LU) for the
// FP32 factorization of the mixed-precision solve.
//
// Partial pivoting needs one grid-wide argmax per column: 64 exchanges per panel, each a few
// microseconds on this chip.  A tournament needs none per column.  The panel's rows are split
// into leaves of 256; every leaf runs Gaussian elimination with partial pivoting on its rows
// in registers (thread = row) and nominates the 64 rows it pivoted on.  Leaves are merged four
// at a time: the last workgroup of a group to finish (an atomic arrival counter, no waiting)
// runs the same elimination on the original values of the 4 x 64 nominees and nominates 64 of
// them, up to a single root.  The root's elimination of its nominees is the LU of the panel's
// pivot rows (L11 \ U11); it also writes the pivots as the usual sequence of row interchanges.
// The caller then applies the interchanges and solves L21 U11 = A21 for the other rows.
//
// The pivots differ from partial pivoting's (the factorization is still P A = L U with |L| <=
// a small bound in practice); the FP64 drop-in path keeps the crate's partial pivoting.
#include <cfloat>
#include <utility>

#include "al_internal.hpp"

namespace al {
namespace {

constexpr int W = 64;        // panel width
constexpr int LEAF = 256;    // rows per leaf (one per thread)
constexpr int ARITY = 4;     // children per merge (4 x 64 nominees = 256 rows)
constexpr int MAX_LEVELS = 8;

typedef unsigned long long u64;

struct Tree {
    int levels;                   // number of levels, leaves included (the last has one node)
    int count[MAX_LEVELS];        // nodes per level
    int offset[MAX_LEVELS];       // first node id of each level
};

constexpr int CW = W / 4;    // columns per wave

// LDS of one panel workgroup.
struct TsluLds {
    float lmul[2][LEAF];  // multipliers of the current step, by local row
    float prow[2][W];     // the pivot row; each wave writes and reads only its own 16 columns
    int win[2];           // local index of the current pivot row
    float piv[W];         // pivots in elimination order
    int winners[W];       // global row of each pivot
    int last;
    int occ[W], where_[W];
};

// Maximum over the wave of values >= -1, on the VALU: prefix maxima inside each row of 16
// lanes (DPP row_shr 1, 2, 4, 8), then row 0 into row 1 and row 2 into row 3 (row_bcast:15),
// then rows 0-1 into rows 2-3 (row_bcast:31); lane 63 ends with the maximum.
__device__ __forceinline__ float wave_max_dpp(float v) {
    const int id = __float_as_int(-2.f);
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(id, __float_as_int(v), 0x111, 0xf, 0xf, false)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(id, __float_as_int(v), 0x112, 0xf, 0xf, false)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(id, __float_as_int(v), 0x114, 0xf, 0xf, false)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(id, __float_as_int(v), 0x118, 0xf, 0xf, false)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(id, __float_as_int(v), 0x142, 0xa, 0xf, false)));
    v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(id, __float_as_int(v), 0x143, 0xc, 0xf, false)));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Gaussian elimination with partial pivoting of up to 256 rows x 64 columns held by the
// workgroup: lane l of wave q holds rows l, l+64, l+128, l+192 (slots a = 0..3) in columns
// 16q..16q+15.  Step j: the wave owning column j finds the first maximum of |x| over the active
// rows (lane-local over the slots, then a wave max and the lowest lane holding it), scales the
// column by the pivot's reciprocal and publishes the multipliers; after one barrier every wave
// takes the pivot row's values in its own columns from the lane that holds them and updates
// its active rows.  A NaN counts as the largest magnitude, so it ends up on the diagonal.
// On return step[a] is the step at which slot a's row was chosen (-1 if never) and the chosen
// rows hold their L \ U entries.
// The owner wave's half of step j (column cj = j % 16 of this wave, a compile-time index so the
// row registers are never indexed dynamically): first maximum of |x| over the active rows
// (lane-local over the slots, then the wave), the column scaled by the pivot's reciprocal,
// and the multipliers, the pivot's position, value and row published through LDS.  A NaN
// counts as the largest magnitude, so it ends up on the diagonal.
template <int CJ>
__device__ __forceinline__ void owner_part(float (&x)[4][CW], const bool (&active)[4], const int (&row)[4],
                                           TsluLds& L, int lane, int b, int j) {
    float m = -1.f, mx = 0.f;
    int ma = 0, mrow = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const float v = !active[a] ? -1.f : (x[a][CJ] != x[a][CJ] ? INFINITY : fabsf(x[a][CJ]));
        const bool gt = v > m;
        m = gt ? v : m;
        ma = gt ? a : ma;
        mx = gt ? x[a][CJ] : mx;
        mrow = gt ? row[a] : mrow;
    }
    const float wm = wave_max_dpp(m);
    const unsigned long long bal = __ballot(m == wm);
    int wl = __ffsll((long long)bal) - 1;
    wl = wl < 0 ? 0 : wl;
    const int wa = __builtin_amdgcn_readlane(ma, wl);
    const float piv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mx), wl));
    const float rp = 1.f / piv;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const bool upd = active[a] && !(a == wa && lane == wl);
        const float l = upd ? x[a][CJ] * rp : 0.f;
        x[a][CJ] = upd ? l : x[a][CJ];
        L.lmul[b][lane + 64 * a] = l;
    }
    if (lane == wl) {
        L.win[b] = wl + 64 * wa;
        L.piv[j] = piv;
        L.winners[j] = mrow;
    }
}

template <int... Cs>
__device__ __forceinline__ void owner_dispatch(std::integer_sequence<int, Cs...>, int cj, float (&x)[4][CW],
                                               const bool (&active)[4], const int (&row)[4], TsluLds& L, int lane,
                                               int b, int j) {
    ((cj == Cs ? owner_part<Cs>(x, active, row, L, lane, b, j) : void()), ...);
}

// Gaussian elimination with partial pivoting of up to 256 rows x 64 columns held by the
// workgroup: lane l of wave q holds rows l, l+64, l+128, l+192 (slots a = 0..3) in columns
// 16q..16q+15.  Step j: the wave owning column j picks the pivot and publishes the
// multipliers; after one barrier every wave takes the pivot row's values in its own columns
// (written to LDS by the lane that holds the row) and updates its active rows.  The step
// loop stays rolled so the kernel's code fits the instruction cache.  On return step[a] is
// the step at which slot a's row was chosen (-1 if never); the chosen rows hold L \ U.
__device__ __forceinline__ void gepp2d(float (&x)[4][CW], bool (&active)[4], const int (&row)[4], int (&step)[4],
                                       TsluLds& L) {
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
#pragma unroll
    for (int a = 0; a < 4; ++a) step[a] = -1;
#pragma nounroll
    for (int j = 0; j < W; ++j) {
        const int b = j & 1, qj = j >> 4, cj = j & 15;
        if (q == qj) owner_dispatch(std::make_integer_sequence<int, CW>{}, cj, x, active, row, L, lane, b, j);
        __syncthreads();
        const int local = L.win[b];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const bool hit = lane + 64 * a == local;
            step[a] = hit ? j : step[a];
            active[a] = active[a] && !hit;
        }
        // columns of this wave right of j: all of them beyond the owner, c > cj in the owner
        const int c_from = q > qj ? 0 : (q == qj ? cj + 1 : CW);
        if (c_from < CW) {
#pragma unroll
            for (int a = 0; a < 4; ++a)
                if (lane + 64 * a == local)
#pragma unroll
                    for (int c = 0; c < CW; ++c) L.prow[b][q * CW + c] = x[a][c];
            __builtin_amdgcn_wave_barrier();
            float u[CW];
#pragma unroll
            for (int c = 0; c < CW; ++c) u[c] = L.prow[b][q * CW + c];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const float l = L.lmul[b][lane + 64 * a];
#pragma unroll
                for (int c = 0; c < CW; ++c) {
                    const float v = fmaf(-l, u[c], x[a][c]);
                    x[a][c] = (active[a] && c >= c_from) ? v : x[a][c];
                }
            }
        }
    }
}

// One launch per panel.  Leaves are workgroups 0..count[0]-1; node n of level l stores its 64
// nominees' original rows at cand[node][s][0..64) and their global row indices at cidx[node][s].
__global__ __launch_bounds__(256, 1) void tslu_k(const float* __restrict__ F, size_t ld, int J, int m, Tree tree,
                                              float* __restrict__ cand, int* __restrict__ cidx,
                                              unsigned* __restrict__ counters, float* __restrict__ D,
                                              int* __restrict__ ipiv, int* __restrict__ info) {
    __shared__ TsluLds L;
    const int t = threadIdx.x, lane = t & 63, q = t >> 6;

    // leaf: local rows i = lane + 64 a of rows J + 256 b ..
    int node = blockIdx.x;
    int level = 0;
    int row[4];
    bool active[4];
    float x[4][CW], x0[4][CW];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int r = blockIdx.x * LEAF + lane + 64 * a;
        active[a] = r < m;
        row[a] = J + r;
#pragma unroll
        for (int c = 0; c < CW; ++c)
            x[a][c] = active[a] ? F[(size_t)row[a] + (size_t)(J + q * CW + c) * ld] : 0.f;
    }
    for (;;) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int c = 0; c < CW; ++c) x0[a][c] = x[a][c];
        int step[4];
        gepp2d(x, active, row, step, L);
        if (level == tree.levels - 1) {
            // root: L11 \ U11 in pivot order (winners and pivots are in LDS already)
#pragma unroll
            for (int a = 0; a < 4; ++a)
                if (step[a] >= 0)
#pragma unroll
                    for (int c = 0; c < CW; ++c) D[step[a] + (q * CW + c) * W] = x[a][c];
            break;
        }
        // nominate: original values and row index of the rows chosen as pivots
#pragma unroll
        for (int a = 0; a < 4; ++a)
            if (step[a] >= 0) {
                float4* dst = (float4*)(cand + ((size_t)node * W + step[a]) * W + q * CW);
#pragma unroll
                for (int v = 0; v < CW / 4; ++v)
                    dst[v] = float4{x0[a][4 * v], x0[a][4 * v + 1], x0[a][4 * v + 2], x0[a][4 * v + 3]};
                if (q == 0) cidx[node * W + step[a]] = row[a];
            }
        __threadfence();
        __syncthreads();
        const int parent_level = level + 1;
        const int pidx = (node - tree.offset[level]) / ARITY;
        const int parent = tree.offset[parent_level] + pidx;
        const int first_child = pidx * ARITY;
        const int nchild = min(ARITY, tree.count[level] - first_child);
        if (t == 0) {
            const unsigned prev = __hip_atomic_fetch_add(&counters[parent], 1u, __ATOMIC_ACQ_REL,
                                                         __HIP_MEMORY_SCOPE_AGENT);
            const int last = prev == (unsigned)(nchild - 1);
            if (last) counters[parent] = 0u;  // every child has arrived: ready for the next panel
            L.last = last;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        __syncthreads();
        if (!L.last) return;
        // this workgroup merges the parent's children: slot a holds child a, lane = nominee
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int child = tree.offset[level] + first_child + a;
            active[a] = a < nchild;
            if (active[a]) {
                const float4* src = (const float4*)(cand + ((size_t)child * W + lane) * W + q * CW);
#pragma unroll
                for (int v = 0; v < CW / 4; ++v) {
                    const float4 f = src[v];
                    x[a][4 * v] = f.x;
                    x[a][4 * v + 1] = f.y;
                    x[a][4 * v + 2] = f.z;
                    x[a][4 * v + 3] = f.w;
                }
                row[a] = cidx[child * W + lane];
            } else {
                row[a] = 0x7FFFFFFF;
#pragma unroll
                for (int c = 0; c < CW; ++c) x[a][c] = 0.f;
            }
        }
        node = parent;
        level = parent_level;
    }
    // root only from here: the pivots as a sequence of interchanges of rows J + s, built by
    // wave 0 with every lane k tracking occ (the original row now at position J + k) and
    // where (the position of original row J + k).  A winner from below the block is still at
    // its own index when its turn comes: a position below the block changes hands only when
    // its occupant is the winner being placed.
    __syncthreads();
    if (q != 0) return;
    const int my_winner = L.winners[lane];
    const bool sane = __all(my_winner >= J && my_winner < J + m);
    if (!sane) {  // cannot happen with a correct elimination; never hand out a wild row index
        ipiv[J + lane] = J + lane;
        if (lane == 0) atomicCAS(info, 0, -1);
        return;
    }
    int occ = J + lane, where = J + lane, my_p = J + lane;
    for (int s = 0; s < W; ++s) {
        const int w = __builtin_amdgcn_readlane(my_winner, s);
        const int p = w < J + W ? __builtin_amdgcn_readlane(where, w - J) : w;
        const int d = __builtin_amdgcn_readlane(occ, s);  // an original row of the block
        my_p = lane == s ? p : my_p;
        if (p != J + s) {
            occ = (p < J + W && lane == p - J) ? d : occ;
            where = lane == d - J ? p : where;
            occ = lane == s ? w : occ;
            where = (w < J + W && lane == w - J) ? J + s : where;
        }
    }
    ipiv[J + lane] = my_p;
    const bool zero = !(L.piv[lane] != 0.f);
    const unsigned long long zb = __ballot(zero);
    if (lane == 0 && zb != 0ull) atomicCAS(info, 0, J + (__ffsll((long long)zb) - 1) + 1);
}

// The rest of the panel once its rows are interchanged: rows J..J+63 receive L11 \ U11 from
// D; every row below solves x U11 = a (one thread per row, U11 broadcast from LDS).
__global__ __launch_bounds__(256) void tslu_l21_k(float* __restrict__ F, size_t ld, int J, int m,
                                                  const float* __restrict__ D) {
    __shared__ float U[W][W + 1];
    const int t = threadIdx.x;
    if (blockIdx.x == 0) {
        for (int idx = t; idx < W * W; idx += 256) {
            const int i = idx & 63, c = idx >> 6;
            F[(size_t)(J + i) + (size_t)(J + c) * ld] = D[i + c * W];
        }
        return;
    }
    for (int idx = t; idx < W * W; idx += 256) {
        const int i = idx & 63, c = idx >> 6;
        U[i][c] = c >= i ? D[i + c * W] : 0.f;
    }
    __syncthreads();
    const int r = W + (blockIdx.x - 1) * 256 + t;
    if (r >= m) return;
    float* rowp = F + (size_t)(J + r);
    float x[W];
#pragma unroll
    for (int c = 0; c < W; ++c) x[c] = rowp[(size_t)(J + c) * ld];
#pragma unroll
    for (int j = 0; j < W; ++j) {
        x[j] = x[j] / U[j][j];
#pragma unroll
        for (int c = j + 1; c < W; ++c) x[c] = fmaf(-x[j], U[j][c], x[c]);
    }
#pragma unroll
    for (int c = 0; c < W; ++c) rowp[(size_t)(J + c) * ld] = x[c];
}

Tree make_tree(int m) {
    Tree tr{};
    int count = (m + LEAF - 1) / LEAF, off = 0, l = 0;
    for (;;) {
        tr.count[l] = count;
        tr.offset[l] = off;
        off += count;
        ++l;
        if (count == 1) break;
        count = (count + ARITY - 1) / ARITY;
    }
    tr.levels = l;
    return tr;
}

}  // namespace

size_t tslu_scratch_bytes(size_t m) {
    const Tree tr = make_tree((int)m);
    const int nodes = tr.offset[tr.levels - 1] + 1;
    return (size_t)nodes * W * W * sizeof(float) + (size_t)nodes * W * sizeof(int) + (size_t)nodes * sizeof(unsigned) +
           W * W * sizeof(float) + 1024;
}

int tslu_panel(hipStream_t s, float* F, size_t ld, size_t J, size_t m, int* ipiv, int* info, void* scratch,
               unsigned* counters) {
    if (m < (size_t)W || m % W != 0 || m > (size_t)LEAF * 4096) {
        set_error("tslu: panel height must be a positive multiple of 64");
        return AL_ERR_ARG;
    }
    const Tree tr = make_tree((int)m);
    if (tr.levels > MAX_LEVELS) {
        set_error("tslu: panel too tall");
        return AL_ERR_ARG;
    }
    const int nodes = tr.offset[tr.levels - 1] + 1;
    float* cand = (float*)scratch;
    int* cidx = (int*)(cand + (size_t)nodes * W * W);
    float* D = (float*)(cidx + (size_t)nodes * W);
    const int ticket = prof_begin(s);
    hipLaunchKernelGGL(tslu_k, dim3((unsigned)tr.count[0]), dim3(256), 0, s, F, ld, (int)J, (int)m, tr, cand, cidx,
                       counters, D, ipiv, info);
    AL_HIP(hipGetLastError());
    if (ticket >= 0) {
        const double dm = (double)m, b = W;
        prof_end(ticket, s, "tslu_tree", (dm + (double)(nodes - tr.count[0]) * 256.0) * b * b, 4.0 * dm * b);
    }
    return AL_OK;
}

int tslu_finish(hipStream_t s, float* F, size_t ld, size_t J, size_t m, const void* scratch) {
    const Tree tr = make_tree((int)m);
    const int nodes = tr.offset[tr.levels - 1] + 1;
    const float* D = (const float*)((const char*)scratch + (size_t)nodes * W * W * sizeof(float) +
                                    (size_t)nodes * W * sizeof(int));
    const unsigned grid = 1u + (unsigned)((m - W + 255) / 256);
    hipLaunchKernelGGL(tslu_l21_k, dim3(grid), dim3(256), 0, s, F, ld, (int)J, (int)m, D);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

}  // namespace al
