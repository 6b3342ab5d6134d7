// Tournament-pivoted 64-column panel (TSLU, the panel of communication-avoiding LU) for the
// FP32 factorization of the mixed-precision solve.
//
// Partial pivoting needs one grid-wide argmax per column: 64 exchanges per panel, each a few
// microseconds on this chip.  A tournament needs none per column.  The panel's rows are split
// into leaves of 256; every leaf runs Gaussian elimination with partial pivoting on its rows
// in registers (thread = row) and nominates the 64 rows it pivoted on.  Leaves are merged four
// at a time: the last workgroup of a group to finish (an atomic arrival counter, no waiting)
// runs the same elimination on the original values of the 4 x 64 nominees and nominates 64 of
// them, up to a single root.  Every level is a chain of 64 latency-bound steps; a 16384-row
// panel has four.  Wider workgroups (512 threads: leaves of 512, 8-way merges, three levels)
// make each step slower by more than the level they save.  The root's elimination of its nominees is the LU of the panel's
// pivot rows (L11 \ U11); it also writes the pivots as the usual sequence of row interchanges.
// The caller then applies the interchanges and solves L21 U11 = A21 for the other rows.
//
// The pivots differ from partial pivoting's (the factorization is still P A = L U with |L| <=
// a small bound in practice); the FP64 drop-in path keeps the crate's partial pivoting.
#include <cfloat>
#include <climits>
#include <utility>

#include "al_internal.hpp"
#include "solve64.hpp"

namespace al {
namespace {

constexpr int W = 64;        // panel width
constexpr int NT = 256;      // threads per workgroup
constexpr int NW = NT / 64;  // waves per workgroup
constexpr int LEAF = NT;     // rows per leaf (one per thread)
constexpr int ARITY = NT / W;  // children per merge (4 x 64 nominees = 256 rows)
constexpr int MAX_LEVELS = 8;

typedef unsigned long long u64;

struct Tree {
    int levels;                   // number of levels, leaves included (the last has one node)
    int count[MAX_LEVELS];        // nodes per level
    int offset[MAX_LEVELS];       // first node id of each level
};

constexpr int KB = 8;        // columns per elimination block

// One wave's candidate of a step: its row's values in the current block and its row index,
// 64 bytes so that the winner's record is found with a shift.
struct TsluCand {
    float v[KB];
    int row;
    int pad[7];
};

// LDS of one panel workgroup.
struct TsluLds {
    int mag[2][NW];           // per wave: its best candidate's key ..
    TsluCand cand[2][NW];     // .. and that candidate
    __attribute__((aligned(16))) float ur[KB][W];  // the block's pivot rows right of the block (U12 after the solve)
    float l11[KB][KB];        // the block's multipliers among its pivot rows
    float piv[W];             // pivots in elimination order
    int winners[W];           // global row of each pivot
    int last;
};

// Maximum over the wave of the int images of floats that are >= 0 or -1 (non-negative floats
// order like their bit patterns; -1.0f is negative as an int), on the VALU: prefix maxima inside
// each row of 16 lanes (DPP row_shr 1, 2, 4, 8), then row 0 into row 1 and row 2 into row 3
// (row_bcast:15), then rows 0-1 into rows 2-3 (row_bcast:31); lane 63 ends with the maximum.
// A lane without a source reads INT_MIN, the identity of max, which lets the compiler fold
// each stage into one v_max_i32 with a DPP operand.
__device__ __forceinline__ int wave_max_dpp(int v) {
    constexpr int lo = INT_MIN;
    v = max(v, __builtin_amdgcn_update_dpp(lo, v, 0x111, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(lo, v, 0x112, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(lo, v, 0x114, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(lo, v, 0x118, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(lo, v, 0x142, 0xa, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(lo, v, 0x143, 0xc, 0xf, false));
    return __builtin_amdgcn_readlane(v, 63);
}

__device__ __forceinline__ float lane_value(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Step J of the elimination, inside the block of columns [C0, C0 + KB).  Each candidate row's
// key is |x[J]|'s bit pattern with its low 8 bits replaced by the row's wave and lane
// (inverted, so that on equal magnitudes the lower wave and lane win): non-negative floats
// order like their bit patterns, so one integer maximum over the wave (DPP) names the wave's
// winner, and one over the four waves' winners names the step's pivot row.  The comparison
// ignores the low 8 of 23 mantissa bits: the pivot is within 2^-15 of the largest magnitude
// (the pivot value itself is exact).  A NaN counts as the largest magnitude; rows no longer
// candidates have negative keys.  The winning lane of each wave publishes its row's block
// values; after one barrier every thread reads the winning wave's record and the active rows
// eliminate column J inside the block.  The step is one dependency chain, so it is kept free
// of branches: the multiplier is x[J] times the pivot's reciprocal (v_rcp_f32, 1 ulp) and
// rows that do not update select their old values.
static_assert(NW <= 4, "two bits of the pivot key name the wave");
template <int C0, int J>
__device__ __forceinline__ void narrow_step(float (&x)[W], bool& active, int row, int& step, TsluLds& L, int lane,
                                            int wave) {
    constexpr int b = J & 1;
    constexpr int NV = C0 + KB - J;  // block columns from J on
    const int tag = ((NW - 1 - wave) << 6) | (63 - lane);
    const float ax = fminf(fabsf(x[J]), __builtin_inff());  // minNum: a NaN becomes +inf
    const int key = ((active ? __float_as_int(ax) : __float_as_int(-1.f)) & ~0xFF) | tag;
    const int wmi = wave_max_dpp(key);
    if (key == wmi) {  // exactly one lane: keys are distinct
        L.mag[b][wave] = wmi;
        L.cand[b][wave].row = row;
#pragma unroll
        for (int c = 0; c < NV; ++c) L.cand[b][wave].v[c] = x[J + c];
    }
    __syncthreads();
    int mx = L.mag[b][0];
#pragma unroll
    for (int w = 1; w < NW; ++w) mx = max(mx, L.mag[b][w]);
    const int best = NW - 1 - ((mx >> 6) & 3);
    const TsluCand& cb = L.cand[b][best];
    const int prow = cb.row;
    const float piv = cb.v[0];
    float u[NV];
#pragma unroll
    for (int c = 1; c < NV; ++c) u[c] = cb.v[c];
    const float r = __builtin_amdgcn_rcpf(piv);
    const bool is_piv = active && row == prow;
    const bool upd = active && !is_piv;
    const float l = x[J] * r;
    x[J] = upd ? l : x[J];
#pragma unroll
    for (int c = 1; c < NV; ++c) {
        const float v = fmaf(-l, u[c], x[J + c]);
        x[J + c] = upd ? v : x[J + c];
    }
    step = is_piv ? J : step;
    active = upd;
    if (threadIdx.x == 0) {
        L.piv[J] = piv;
        L.winners[J] = prow;
    }
}

// After the KB steps of block [C0, C0 + KB): the rows right of the block take the block's
// rank-8 update.  The pivot rows publish their entries right of the block (A12) and their
// multipliers among each other (L11); after one barrier every candidate row turns its 8
// multipliers l into w = l L11^{-1} (28 FMAs against broadcast LDS values) and subtracts w A12,
// which is l U12 with U12 = L11^{-1} A12 never formed.  A pivot row k takes w = e_k - e_k L11^{-1},
// which turns its own A12 row into row k of U12.
// No barrier is needed before the pivot rows write ur and l11: every thread last read them in
// the previous block's update, and each narrow step since then ends at a barrier.
template <int C0>
__device__ __forceinline__ void block_update(float (&x)[W], bool active, int step, TsluLds& L) {
    constexpr int REST = W - C0 - KB;
    if (REST == 0) return;
    const int k = step - C0;
    const bool piv_here = step >= C0 && step < C0 + KB;
    if (piv_here) {
#pragma unroll
        for (int c = 0; c < REST; ++c) L.ur[k][c] = x[C0 + KB + c];
#pragma unroll
        for (int i = 0; i < KB; ++i) L.l11[k][i] = i < k ? x[C0 + i] : 0.f;
    }
    __syncthreads();
    if (piv_here || active) {
        // v L11 = a by back substitution, a = e_k for a pivot row, its multipliers otherwise
        float v[KB];
#pragma unroll
        for (int i = KB - 1; i >= 0; --i) {
            float t = piv_here ? (i == k ? 1.f : 0.f) : x[C0 + i];
#pragma unroll
            for (int j = i + 1; j < KB; ++j) t = fmaf(-v[j], L.l11[j][i], t);
            v[i] = t;
        }
#pragma unroll
        for (int i = 0; i < KB; ++i) v[i] = piv_here ? (i == k ? 1.f : 0.f) - v[i] : v[i];
#pragma unroll
        for (int c = 0; c < REST; ++c) {
            float t = x[C0 + KB + c];
#pragma unroll
            for (int i = 0; i < KB; ++i) t = fmaf(-v[i], L.ur[i][c], t);
            x[C0 + KB + c] = t;
        }
    }
}

template <int C0, int... Ks>
__device__ __forceinline__ void elim_block(std::integer_sequence<int, Ks...>, float (&x)[W], bool& active, int row,
                                           int& step, TsluLds& L, int lane, int wave) {
    (narrow_step<C0, C0 + Ks>(x, active, row, step, L, lane, wave), ...);
    block_update<C0>(x, active, step, L);
}

template <int... Bs>
__device__ __forceinline__ void elim_all(std::integer_sequence<int, Bs...>, float (&x)[W], bool& active, int row,
                                         int& step, TsluLds& L, int lane, int wave) {
    (elim_block<Bs * KB>(std::make_integer_sequence<int, KB>{}, x, active, row, step, L, lane, wave), ...);
}

// Gaussian elimination with partial pivoting of up to NT rows (one per thread, 64 columns in
// registers), blocked by 8 columns: inside a block each step updates only the block's
// columns; the rest of the row is updated once per block with a rank-8 product.  On return
// `step` is the step at which this thread's row was chosen (-1 if never); chosen rows hold
// L \ U.
__device__ __forceinline__ void gepp_rows(float (&x)[W], bool active, int row, int& step, TsluLds& L) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    step = -1;
    elim_all(std::make_integer_sequence<int, W / KB>{}, x, active, row, step, L, lane, wave);
}

// One launch per panel.  Leaves are workgroups 0..count[0]-1; node n of level l stores its 64
// nominees' original rows at cand[node][s][0..64) and their global row indices at cidx[node][s].
__global__ __launch_bounds__(NT, 1) void tslu_k(const float* __restrict__ F, size_t ld, int J, int m, Tree tree,
                                                 float* __restrict__ cand, int* __restrict__ cidx,
                                                 unsigned* __restrict__ counters, float* __restrict__ D,
                                                 int* __restrict__ ipiv, int* __restrict__ info,
                                                 int* __restrict__ pairs) {
    __shared__ TsluLds L;
    const int t = threadIdx.x, lane = t & 63, q = t >> 6;

    // leaf: thread t holds row J + LEAF b + t
    int node = blockIdx.x;
    int level = 0;
    float x[W];
    const int r0 = blockIdx.x * LEAF + t;
    bool active = r0 < m;
    int row = J + r0;
#pragma unroll
    for (int c = 0; c < W; ++c) x[c] = active ? F[(size_t)row + (size_t)(J + c) * ld] : 0.f;
    for (;;) {
        // the original values travel with a nomination: keep them in LDS-free registers
        float x0[W];
#pragma unroll
        for (int c = 0; c < W; ++c) x0[c] = x[c];
        int step;
        gepp_rows(x, active, row, step, L);
        if (level == tree.levels - 1) {
            // root: L11 \ U11 in pivot order (winners and pivots are in LDS already)
            if (step >= 0) {  // row-major: row `step` of L11 \ U11 is 64 consecutive floats
                float4* dst = (float4*)(D + step * W);
#pragma unroll
                for (int v = 0; v < W / 4; ++v) dst[v] = float4{x[4 * v], x[4 * v + 1], x[4 * v + 2], x[4 * v + 3]};
            }
            break;
        }
        // nominate: original values and row index of the rows chosen as pivots
        if (step >= 0) {
            float4* dst = (float4*)(cand + ((size_t)node * W + step) * W);
#pragma unroll
            for (int v = 0; v < W / 4; ++v) dst[v] = float4{x0[4 * v], x0[4 * v + 1], x0[4 * v + 2], x0[4 * v + 3]};
            cidx[node * W + step] = row;
        }
        // each wave's nominee stores reach the L2 before the barrier; thread 0's acquire-release
        // arrival then writes the L2 back once for the whole workgroup (and, for the merging
        // workgroup, invalidates once), instead of every wave fencing on its own
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
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
        }
        __syncthreads();
        if (!L.last) return;
        // this workgroup merges the parent's children: thread t holds nominee t
        const int child = tree.offset[level] + first_child + t / W;
        active = t < nchild * W;
        if (active) {
            const float4* src = (const float4*)(cand + ((size_t)child * W + (t % W)) * W);
#pragma unroll
            for (int v = 0; v < W / 4; ++v) {
                const float4 f = src[v];
                x[4 * v] = f.x;
                x[4 * v + 1] = f.y;
                x[4 * v + 2] = f.z;
                x[4 * v + 3] = f.w;
            }
            row = cidx[child * W + (t % W)];
        } else {
            row = 0x7FFFFFFF;
#pragma unroll
            for (int c = 0; c < W; ++c) x[c] = 0.f;
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
        if (lane == 0) {
            atomicCAS(info, 0, -1);
            if (pairs) pairs[0] = 0;
        }
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
    if (pairs) {
        // the interchanges composed into (destination, source) row pairs, sources being the
        // original rows: position J + lane now holds original row occ; original row J + lane,
        // if it left the block, now sits at where (below it).  These are exactly the positions
        // the sequence changes, top block first.
        const bool top = occ != J + lane;
        const bool down = where >= J + W;
        const unsigned long long bt = __ballot(top), bd = __ballot(down);
        const unsigned long long below = (1ull << lane) - 1ull;
        const int nt = __popcll(bt);
        if (top) {
            const int r = __popcll(bt & below);
            pairs[1 + 2 * r] = J + lane;
            pairs[2 + 2 * r] = occ;
        }
        if (down) {
            const int r = nt + __popcll(bd & below);
            pairs[1 + 2 * r] = where;
            pairs[2 + 2 * r] = J + lane;
        }
        if (lane == 0) pairs[0] = nt + __popcll(bd);
    }
    const bool zero = !(L.piv[lane] != 0.f);
    const unsigned long long zb = __ballot(zero);
    if (lane == 0 && zb != 0ull) atomicCAS(info, 0, J + (__ffsll((long long)zb) - 1) + 1);
}

// The rest of the panel once its rows are interchanged: rows J..J+63 receive L11 \ U11 from
// D (row-major); every row below solves x U11 = a, one thread per row.  U11 is the same for
// every thread: each workgroup stages it in LDS once, and the 64-step chain reads it as
// broadcast LDS loads (all lanes one address), which the compiler issues well ahead of the
// FMAs that use them.  Each pivot is applied as a multiplication by its reciprocal
// (v_rcp_f32, 1 ulp).  The last nu workgroups instead solve U12 := L11^{-1} A12 on columns
// [u_c0, u_c1) of rows J..J+63, with L11 from D: nothing here depends on the other parts.
__global__ __launch_bounds__(256) void tslu_l21_k(float* __restrict__ F, size_t ld, int J, int m,
                                                  const float* __restrict__ D, int nl, int u_c0, int u_c1, int nu) {
    const int t = threadIdx.x;
    if (blockIdx.x == 0) {
        for (int idx = t; idx < W * W; idx += 256) {
            const int i = idx & 63, c = idx >> 6;
            F[(size_t)(J + i) + (size_t)(J + c) * ld] = D[i * W + c];
        }
        return;
    }
    if ((int)blockIdx.x > nl) {
        solve_l64_cols<true>(F, ld, J, u_c0, u_c1, D, W, ((int)blockIdx.x - nl - 1) * 4 + (t >> 6), nu * 4);
        return;
    }
    __shared__ __align__(16) float U[W * W];
    for (int idx = t; idx < W * W / 4; idx += 256) ((float4*)U)[idx] = ((const float4*)D)[idx];
    __syncthreads();
    const int r = W + (blockIdx.x - 1) * 256 + t;
    if (r >= m) return;
    float* const row0 = F + (size_t)(J + r) + (size_t)J * ld;  // the row's entry in column J
    float x[W];
    {
        const float* p = row0;
#pragma unroll
        for (int c = 0; c < W; ++c) {
            x[c] = *p;
            p += ld;
        }
    }
#pragma unroll
    for (int j = 0; j < W; ++j) {
        const float* uj = U + j * W;  // row j of U11
        x[j] *= __builtin_amdgcn_rcpf(uj[j]);
#pragma unroll
        for (int c = j + 1; c < W; ++c) x[c] = fmaf(-x[j], uj[c], x[c]);
    }
    float* p = row0;
#pragma unroll
    for (int c = 0; c < W; ++c) {
        *p = x[c];
        p += ld;
    }
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
               unsigned* counters, int* pairs) {
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
    hipLaunchKernelGGL(tslu_k, dim3((unsigned)tr.count[0]), dim3(NT), 0, s, F, ld, (int)J, (int)m, tr, cand, cidx,
                       counters, D, ipiv, info, pairs);
    AL_HIP(hipGetLastError());
    if (ticket >= 0) {
        const double dm = (double)m, b = W;
        prof_end(ticket, s, "tslu_tree", (dm + (double)(nodes - tr.count[0]) * (double)LEAF) * b * b, 4.0 * dm * b);
    }
    return AL_OK;
}

int tslu_finish(hipStream_t s, float* F, size_t ld, size_t J, size_t m, const void* scratch, int u_c0, int u_c1) {
    const Tree tr = make_tree((int)m);
    const int nodes = tr.offset[tr.levels - 1] + 1;
    const float* D = (const float*)((const char*)scratch + (size_t)nodes * W * W * sizeof(float) +
                                    (size_t)nodes * W * sizeof(int));
    const int nl = (int)((m - W + 255) / 256);
    const int groups = u_c1 > u_c0 ? (u_c1 - u_c0 + S64_G - 1) / S64_G : 0;
    const int nu = (groups + 3) / 4;
    hipLaunchKernelGGL(tslu_l21_k, dim3((unsigned)(1 + nl + nu)), dim3(256), 0, s, F, ld, (int)J, (int)m, D, nl, u_c0,
                       u_c1, nu);
    AL_HIP(hipGetLastError());
    return AL_OK;
}

}  // namespace al
