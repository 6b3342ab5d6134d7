// Shared internals of the MI355X (gfx950) dense solve library: status plumbing, the
// device-side building blocks the kernels share, and the host-side launch helpers.
// Matrices are column-major; element (i, j) of a matrix with leading dimension ld sits
// at [i + j*ld], as in the algebra_linear crate (src/blas.rs:1-9).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/algebra_linear_hip.h"
#include "zcomplex.hpp"

namespace al {

// ---------------------------------------------------------------------------------------
// host-side status handling
// ---------------------------------------------------------------------------------------

// Records the message a failed call leaves for al_last_error(); thread-local.
void set_error(const std::string& msg);
int fail_hip(hipError_t e, const char* what);

#define AL_HIP(expr)                                         \
    do {                                                     \
        hipError_t _e = (expr);                              \
        if (_e != hipSuccess) return ::al::fail_hip(_e, #expr); \
    } while (0)

#define AL_CHECK(expr)             \
    do {                           \
        int _s = (expr);           \
        if (_s != AL_OK) return _s; \
    } while (0)

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// ---------------------------------------------------------------------------------------
// device helpers shared by the tiled kernels
// ---------------------------------------------------------------------------------------

// Tile (tm, tn) of this workgroup in a tiles_m x tiles_n grid of output tiles.  Workgroups
// are dispatched round-robin over the 8 XCDs; each XCD is given a contiguous range of tiles,
// walked in groups of 8 tile-rows (column-major inside a group), so the operand panels a
// group shares stay in that XCD's L2.
__device__ inline void tile_coords(int tiles_m, int tiles_n, int& tm, int& tn) {
    const int nb = tiles_m * tiles_n;
    const int bid = blockIdx.x;
    // blocks are dispatched round-robin over the 8 XCDs; give XCD x a contiguous range
    const int x = bid & 7, slot = bid >> 3;
    const int base = nb >> 3, rem = nb & 7;
    const int lin = x * base + (x < rem ? x : rem) + slot;
    // grouped ordering: 8 tile-rows per group, column-major inside the group
    constexpr int G = 8;
    const int per_group = G * tiles_n;
    const int g = lin / per_group;
    const int first_m = g * G;
    const int gsize = (tiles_m - first_m) < G ? (tiles_m - first_m) : G;
    const int in = lin - g * per_group;
    tm = first_m + in % gsize;
    tn = in / gsize;
}

// ---------------------------------------------------------------------------------------
// per-call device workspace: the cooperative panel kernels' exchange slots, counters and
// the status word every kernel of a factorization checks before doing work.
// ---------------------------------------------------------------------------------------
struct Workspace {
    int device = -1;
    // status word (int32): 0 = fine, k > 0 = first zero pivot (1-based column), < 0 device
    // error (a bounded spin timed out).
    int* info = nullptr;
    unsigned* abort_flag = nullptr;  // raised by a timed-out spin; stops every cooperating workgroup
    unsigned* flags = nullptr;     // per-workgroup publication flags of the cooperative panels
    unsigned gen = 0;              // last generation used on `flags` (host side, monotonic)
    double* slots = nullptr;       // panel exchange area
    size_t slots_bytes = 0;
    int pin_half = -1;             // cooperative panels on XCDs 4*pin_half..+3 (-1: anywhere)
    size_t pin_flags = 0;          // pinned: word offset of the publication words in `slots`
    unsigned* slot_tab = nullptr;  // pinned: word offset of each exchange slot in `slots`
    int* ipiv32 = nullptr;         // int32 pivots on the device
    size_t ipiv_len = 0;
    int* pairs = nullptr;          // composed row permutation of one pivot block
    int* pairs2 = nullptr;         // the same for the main stream while the panel stream runs ahead
    double* perm = nullptr;        // row origins of a whole factorization (exact integers)
    size_t perm_len = 0;
    double* wbuf = nullptr;        // right-hand-side staging for the solves
    size_t wbuf_len = 0;
    unsigned* trsv_flags = nullptr;  // per-block completion flags of the one-launch solves
    size_t trsv_flags_len = 0;
    unsigned trsv_gen = 0;         // last generation used on trsv_flags (host side)
    void* mbuf = nullptr;          // mixed-precision solver arena (factor, vectors, partial sums)
    size_t mbuf_bytes = 0;
    void* fbuf = nullptr;          // its FP64 fallback: a copy of A and the pivots
    size_t fbuf_bytes = 0;
    void* hnorms = nullptr;        // pinned host words the refinement loop reads back
    unsigned mgen = 0;             // last generation of the no-pivot panel flag (host side)
    void* gbuf = nullptr;          // bf16 operand staging of the device-API bf16 GEMM
    size_t gbuf_bytes = 0;
    void* bbuf = nullptr;          // batched solver: FP32 work matrices
    size_t bbuf_bytes = 0;
    void* zbuf = nullptr;          // complex twins: real-form staging (R(L), R(A), E, W)
    size_t zbuf_bytes = 0;
    void* cbuf = nullptr;          // FP64 Cholesky: a diagonal block's inverse (128 x 128)
    size_t cbuf_bytes = 0;
    // a second stream (default priority) for the panel factorization that runs ahead of the
    // trailing update in the FP32 LU, and the two events that order the streams
    hipStream_t side = nullptr;
    hipEvent_t ev_side = nullptr, ev_main = nullptr;
};

// The workspace of `stream` for the calling thread (created on first use).
int workspace_on(hipStream_t stream, Workspace** out);


// Grows the one-launch triangular solve's flag array (entries, zero-filled when new).
int ws_reserve_flags(Workspace* ws, size_t entries);

// Clears the cooperative panels' exchange area and decides whether this workspace's panels run
// on one half of the chip (a short calibration launch on `s`, then a wait for it).
int panel_exchange_init(hipStream_t s, Workspace* ws);
void set_panel_pinning(bool on);  // false: cooperative panels run anywhere (test hook)

// One-launch blocked triangular solve of op = lower/upper, unit or not, with nrhs <= 16
// right-hand sides in X (n x nrhs, FP64); the matrix may be FP64 or FP32.
bool trsv_fits(size_t nrhs);
// Dinv (FP32 matrices, nrhs <= 4): the 128 x 128 diagonal blocks' inverses from
// trsv_diag_inverse, applied as products instead of substitutions; may be null.
template <typename TA>
int trsv_solve(hipStream_t s, Workspace* ws, bool lower, bool unit, size_t n, size_t nrhs, const TA* A, size_t lda,
               double* X, size_t ldx, const double* Dinv = nullptr);
// FP64 inverses of the 128-row diagonal blocks of the unit-lower and the upper triangle of an
// FP32 LU factor (n x n in A): trsv_diag_inverse_len(n) doubles each.
size_t trsv_diag_inverse_len(size_t n);
int trsv_diag_inverse(hipStream_t s, size_t n, const float* A, size_t lda, double* Linv, double* Uinv);

// Grows the workspace's perm (n entries) and staging buffer (nw doubles).
int ws_reserve(Workspace* ws, size_t n, size_t nw);

// The calling thread's workspace on the current device (created on first use).
int workspace(Workspace** out, size_t ipiv_len);

// Per-thread HIP stream used by the host-slice API (default priority).
int thread_stream(hipStream_t* out);

// ---------------------------------------------------------------------------------------
// device kernels' host launchers (all asynchronous on `stream`)
// ---------------------------------------------------------------------------------------

// C := alpha*op(A)*op(B) + beta*C; tri: 0 full, 1 lower triangle only, 2 upper only.
int gemm_f64(hipStream_t s, bool ta, bool tb, size_t m, size_t n, size_t k, double alpha, const double* A,
             size_t lda, const double* B, size_t ldb, double beta, double* C, size_t ldc, int tri = 0);
// The FP64 GEMM restricted to a complex triangle of C's real view (2m' x n' with row r in
// complex row r >> 1): upper = false -> op = (N, N), complex rows >= columns kept; upper = true
// -> op = (T, N), complex rows <= columns kept.  m is the real row count.
int gemm_f64_ctri(hipStream_t s, bool upper, size_t m, size_t n, size_t k, double alpha, const double* A, size_t lda,
                  const double* B, size_t ldb, double beta, double* C, size_t ldc);
// The same on FP32 matrix cores (v_mfma_f32_16x16x4_f32), FP32 accumulation.
int gemm_f32(hipStream_t s, bool ta, bool tb, size_t m, size_t n, size_t k, float alpha, const float* A, size_t lda,
             const float* B, size_t ldb, float beta, float* C, size_t ldc, int tri = 0);
// FP32 untransposed GEMM with 16-byte accesses (gemm_f32 routes to it when sgemm_nn_fits).
bool sgemm_nn_fits(size_t m, size_t k, const float* A, size_t lda, const float* B, size_t ldb, const float* C,
                   size_t ldc);
int sgemm_nn(hipStream_t s, int m, int n, int k, float alpha, const float* A, size_t lda, const float* B, size_t ldb,
             float beta, float* C, size_t ldc);

// Triangular solve op(T) X = alpha B (left) or X op(T) = alpha B (right), in place in B.
// trans selects op(T) = T^T (used by the Cholesky panel); `info` (nullable) turns the call
// into a no-op once a factorization has failed.
int trsm_f64(hipStream_t s, bool left, bool lower, bool trans, bool unit, size_t m, size_t n, double alpha,
             const double* T, size_t ldt, double* B, size_t ldb, const int* info = nullptr);

// Row interchanges ipiv[k0..k0+jb) (int32, global row indices, jb <= 256) applied to the n0
// columns at A0 and the n1 columns at A1 (leading dimension lda) and, if perm is not null, to
// the row-origin vector perm -- with the result of the sequential swaps, as one gather.
// pairs: the composed-permutation buffer (1 + 2 * 512 ints); ws->pairs when null.  Swaps
// that may run concurrently on two streams need buffers of their own.
int swap_block_f64(hipStream_t s, Workspace* ws, const int* ipiv, size_t k0, size_t jb, double* A0, size_t n0,
                   double* A1, size_t n1, size_t lda, double* perm, const int* info, int* pairs = nullptr);
// The same for an FP32 matrix (no perm vector).
int swap_block_f32(hipStream_t s, Workspace* ws, const int* ipiv, size_t k0, size_t jb, float* A0, size_t n0,
                   float* A1, size_t n1, size_t lda, const int* info, int* pairs = nullptr);
// The same for a complex matrix (lda in complex elements).
int swap_block_z(hipStream_t s, Workspace* ws, const int* ipiv, size_t k0, size_t jb, zd* A0, size_t n0, zd* A1,
                 size_t n1, size_t lda, const int* info);
// Applies (destination, source) row pairs already composed (as tslu_panel writes them) to the
// columns [0, n0) of A0 and [0, n1) of A1; nothing when *info != 0.
int swap_apply_f32(hipStream_t s, const int* pairs, float* A0, size_t n0, float* A1, size_t n1, size_t lda,
                   const int* info);
// Applies nl pivot blocks' pair lists (each from one 64-pivot tournament, at pairs + l * stride)
// in order to columns [0, n0) of A0 and [0, n1) of A1.
int swap_apply_multi_f32(hipStream_t s, const int* pairs, size_t stride, int nl, float* A0, size_t n0, float* A1,
                         size_t n1, size_t lda, const int* info);
int perm_init(hipStream_t s, double* perm, size_t n);                          // perm[i] = i
int perm_from_ipiv(hipStream_t s, const int* ipiv, size_t n, double* perm);    // compose all pivots
int gather_rows(hipStream_t s, const double* B, size_t ldb, size_t n, size_t nrhs, const double* perm, double* W);
int copy_block(hipStream_t s, const double* W, size_t n, size_t nrhs, double* B, size_t ldb);

// Creates ws->side (a second stream, default priority) and the two events that order it with
// the caller's stream, once per workspace.
int ensure_side_stream(Workspace* ws);

// Compute units of the current device (256 on MI355X).
int cu_count();

// Limit, in ticks of the 100 MHz wall clock, of every bounded device-side wait (the cooperative
// panel's column exchange, the one-launch solves' block flags).  Default 2 s; a test may lower
// it to force the give-up path from the host (al_debug_set_wait_limit).
long long wait_limit_ticks();
void set_wait_limit_ticks(long long ticks);  // <= 0 restores the default

// Cooperative-panel room (coop_ledger.hpp): coop_reserve waits until the panels of an m-row
// factorization fit beside those of unfinished ones and returns the units reserved (0 when its
// panels use one workgroup, or while `s` is being captured into a graph); coop_hold ties them to
// an event recorded on `s` now; coop_release returns them at once.  coop_in_use: units held.
int coop_reserve(hipStream_t s, size_t m, int* units);
int coop_hold(hipStream_t s, int units);
void coop_release(int units);
int coop_in_use();

// Blocked LU with partial pivoting of the m x n matrix A (m >= n for the panel form).  If perm
// is not null it receives the row origins (perm[i] = original row now at row i, m entries).
int getrf_f64(hipStream_t s, Workspace* ws, size_t m, size_t n, double* A, size_t lda, int* ipiv, int* info,
              double* perm = nullptr);
// The same without the cooperative-panel reservation (the caller holds one).
int getrf_enqueue(hipStream_t s, Workspace* ws, size_t m, size_t n, double* A, size_t lda, int* ipiv, int* info,
                  double* perm);
// LU without any row interchange of the n x n matrix A (already in a good row order), with the
// same blocking and look-ahead; info > 0: a zero or non-finite pivot at that 1-based column.
int getrf_nopiv_f64(hipStream_t s, Workspace* ws, size_t n, double* A, size_t lda, int* info);

// One cooperative 64-wide (ncols <= 64) partial-pivoting panel of an FP32 matrix: the m x
// ncols block at P (rows row0.. of the whole matrix), pivots written to ipiv[row0..] as
// global row indices, swaps applied inside the panel only.
int panel_f32(hipStream_t s, Workspace* ws, float* P, size_t lda, size_t row0, size_t m, size_t ncols, int* ipiv,
              int* info);
// The same panel for a complex matrix, ncols <= panel_z_width() (32), |Re| + |Im| pivots.
int panel_z(hipStream_t s, Workspace* ws, zd* P, size_t lda, size_t row0, size_t m, size_t ncols, int* ipiv,
            int* info);
int panel_z_width();

// Complex twins (complex.hip).  Complex matrices are interleaved (re, im) with leading
// dimensions in complex elements.
// C := alpha op(A) op(B) + beta C (never conjugated), beta*C formed for every entry.
int zgemm_dev(hipStream_t s, Workspace* ws, bool ta, bool tb, size_t m, size_t n, size_t k, zd alpha, const zd* A,
              size_t lda, const zd* B, size_t ldb, zd beta, zd* C, size_t ldc);
// Complex LU with partial pivoting (m >= n for the panel form), pivots int32 global rows.
int zgetrf_dev(hipStream_t s, Workspace* ws, size_t m, size_t n, zd* A, size_t lda, int* ipiv, int* info);
// B (n x nrhs) := A^-1 B from zgetrf_dev's factors.
int zgetrs_dev(hipStream_t s, Workspace* ws, size_t n, size_t nrhs, const zd* LU, size_t lda, const int* ipiv,
               zd* B, size_t ldb, const int* info = nullptr);
// Complex Cholesky, lower (A = L L^H) or upper (A = U^H U); only that triangle is read/written.
int zpotrf_dev(hipStream_t s, Workspace* ws, size_t n, zd* A, size_t lda, bool upper, int* info);

// Tournament-pivoted 64-column panel of an FP32 matrix F (rows J..J+m, columns J..J+64):
// tslu_panel chooses the pivots (ipiv[J..J+64), global indices) and the LU of the pivot rows;
// after the interchanges are applied to the panel's columns, tslu_finish writes L11 \ U11,
// solves the rows below and, in the same launch, U12 := L11^{-1} A12 on columns [u_c0, u_c1) of
// rows J..J+64 (none when u_c1 <= u_c0).  scratch: tslu_scratch_bytes(m) bytes; counters:
// zero-filled words (one per tree node, left at zero by every panel).
size_t tslu_scratch_bytes(size_t m);
int tslu_panel(hipStream_t s, float* F, size_t ld, size_t J, size_t m, int* ipiv, int* info, void* scratch,
               unsigned* counters, int* pairs = nullptr);
int tslu_finish(hipStream_t s, float* F, size_t ld, size_t J, size_t m, const void* scratch, int u_c0, int u_c1);

// Solve with a getrf factorization: B (n x nrhs) := A^{-1} B.  perm (nullable) is getrf's row
// origin vector; without it the pivots are composed first.
int getrs_f64(hipStream_t s, Workspace* ws, size_t n, size_t nrhs, const double* LU, size_t lda, const int* ipiv,
              const double* perm, double* B, size_t ldb, const int* info = nullptr);

// bf16 staging of a GEMM's operands, rows of kp >= k entries (zero beyond k), rounded to
// nearest even: At[r][p] = L[r + p*ld] (L m x k) and Bt[c][p] = U[p + c*ld] (U k x n).
int bf16_rows(hipStream_t s, size_t m, size_t k, size_t kp, const float* L, size_t ld, __bf16* At);
int bf16_cols(hipStream_t s, size_t k, size_t kp, size_t n, const float* U, size_t ld, __bf16* Bt);
// Split staging of an FP32 product for three bf16 products: each value x becomes hi = bf16(x)
// and lo = bf16(x - hi); At rows are [Lhi | Lhi | Llo] and Bt rows [Uhi | Ulo | Uhi] (3k
// entries each), so gemm_bf16 with depth 3k gives Lhi Uhi + Lhi Ulo + Llo Uhi, the product to
// about 2^-16 relative instead of bf16's 2^-8.  k % 64 == 0 keeps the depth a multiple of 64.
int bf16x3_rows(hipStream_t s, size_t m, size_t k, const float* L, size_t ld, __bf16* At);
int bf16x3_cols(hipStream_t s, size_t k, size_t n, const float* U, size_t ld, __bf16* Bt);
// C (m x n, FP32) := alpha sum_p At[i][p] Bt[j][p] + beta C, bf16 products accumulated in
// FP32 on the matrix cores; k % 64 == 0.
int gemm_bf16(hipStream_t s, size_t m, size_t n, size_t k, float alpha, const __bf16* At, const __bf16* Bt, float beta,
              float* C, size_t ldc);

// A low-precision factorization kept for reuse: the FP32 LU (np x np, np = n rounded up to a
// multiple of 64) and its row origins.
struct MixedFactor {
    float* F = nullptr;
    double* perm = nullptr;
    size_t n = 0, np = 0;
    int breakdown = 0;  // 1-based column of an exactly zero pivot, or 0
};
inline size_t mixed_padded(size_t n) { return (n + 63) / 64 * 64; }

// Mixed-precision solve X := A^{-1} B (A, B untouched; synchronizes s).  With `reuse` the
// factorization is taken from it instead of computed; with `keep` (buffers allocated by the
// caller for mixed_padded(n)) a computed factorization is copied there.
int gesv_mixed(hipStream_t s, Workspace* ws, size_t n, size_t nrhs, const double* A, size_t lda, const double* B,
               size_t ldb, double* X, size_t ldx, int factor, const al_ir_options* opt, al_ir_stats* stats,
               const MixedFactor* reuse = nullptr, MixedFactor* keep = nullptr);

// Batched solver (orders <= 512, size-bucketed persistent kernels): see batch.hip.  work:
// batch_scratch_bytes(count) device bytes; W_pool (mixed mode): an FP32 pool shaped like A_pool.
// Uses the workspace's side stream and events.
size_t batch_scratch_bytes(size_t count);
int batch_solve(hipStream_t s, Workspace* ws, size_t count, const al_batch_entry* d_entries, double* A_pool,
                double* B_pool, float* W_pool, int* d_info, int* d_steps, void* work, bool mixed, int max_iter);
// The same FP64 solve with its scratch in a static device array (al_dev_batch_solve_noalloc).
int batch_solve_noalloc(hipStream_t s, size_t count, const al_batch_entry* d_entries, double* A_pool, double* B_pool,
                        int* d_info, int* d_steps);
// The per-bucket launch configuration (bucket 0: n <= 64, 1: <= 128, 2: <= 256, 3: <= 512).
int batch_tune_set(int bucket, int kb, int wgs_per_cu);
int batch_tune_get(int bucket, int* kb, int* wgs_per_cu);

// Level-2 and support routines (blas2.hip).
// y := alpha op(A) x + beta y with the crate's summation order (bit for bit), A m x n.
int gemv_f64(hipStream_t s, bool trans, size_t m, size_t n, double alpha, const double* A, size_t lda,
             const double* x, double beta, double* y);
// *out (device) := the norm of the m x n matrix A ('M', '1'/'O', 'I', 'F'/'E'); scratch holds
// lange_scratch_doubles(m, n) doubles.
size_t lange_scratch_doubles(size_t m, size_t n);
int lange_f64(hipStream_t s, char norm, size_t m, size_t n, const double* A, size_t lda, double* out,
              double* scratch);
// In-place inverse from getrf's factors and pivots (device int32); *info = 1-based index of a
// zero on U's diagonal (the factor is then left untouched), else 0.  Uses ws->perm and ws->wbuf.
int getri_f64(hipStream_t s, Workspace* ws, size_t n, double* LU, size_t lda, const int* ipiv, int* info);
// One triangle of C := alpha V W^T + alpha W V^T + beta C (V, W m x k).
int syr2k_f64(hipStream_t s, bool upper, size_t m, size_t k, double alpha, const double* V, size_t lv,
              const double* W, size_t lw, double beta, double* C, size_t lc);

// Lower Cholesky in place; only the lower triangle is read or written.
int potrf_f64(hipStream_t s, Workspace* ws, size_t n, double* A, size_t lda, int* info);

// Launch instrumentation (al_prof_*): when enabled, prof_begin records a start event on the
// stream and returns a ticket; prof_end records the stop event with the launch's algorithmic
// flops and bytes.  Both are no-ops (ticket -1) when disabled.
int prof_begin(hipStream_t s);
void prof_end(int ticket, hipStream_t s, const char* kernel_class, double flops, double bytes);

// Fill helpers (kernels, never hipMemset, so the sequence stays graph-capturable).
int fill_i32(hipStream_t s, int* p, size_t n, int v);
int fill_u32(hipStream_t s, unsigned* p, size_t n, unsigned v);
int iota_i32(hipStream_t s, int* p, size_t n);  // p[i] = i

}  // namespace al
