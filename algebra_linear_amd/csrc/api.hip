// The C ABI of include/algebra_linear_hip.h: shape checks with the crate's own messages,
// per-thread streams and device buffers, host<->HBM staging, and status mapping.
#include <hip/hip_runtime.h>

#include <atomic>
#include <climits>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "al_internal.hpp"

namespace al {

size_t panel_slots_bytes();
size_t panel_flags_count();

namespace {

thread_local std::string g_last_error;

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

struct ThreadCtx {
    int device = -1;
    hipStream_t stream = nullptr;
    DevBuf bufs[4];
    std::unordered_map<hipStream_t, Workspace> ws;  // keyed by the stream the work runs on
    ~ThreadCtx() {
        for (auto& b : bufs)
            if (b.p) (void)hipFree(b.p);
        for (auto& kv : ws) {
            Workspace& w = kv.second;
            if (w.info) (void)hipFree(w.info);
            if (w.slots) (void)hipFree(w.slots);
            if (w.slot_tab) (void)hipFree(w.slot_tab);
            if (w.flags) (void)hipFree(w.flags);
            if (w.ipiv32) (void)hipFree(w.ipiv32);
            if (w.pairs) (void)hipFree(w.pairs);
            if (w.pairs2) (void)hipFree(w.pairs2);
            if (w.perm) (void)hipFree(w.perm);
            if (w.wbuf) (void)hipFree(w.wbuf);
            if (w.trsv_flags) (void)hipFree(w.trsv_flags);
            if (w.mbuf) (void)hipFree(w.mbuf);
            if (w.fbuf) (void)hipFree(w.fbuf);
            if (w.hnorms) (void)hipHostFree(w.hnorms);
            if (w.gbuf) (void)hipFree(w.gbuf);
            if (w.bbuf) (void)hipFree(w.bbuf);
            if (w.zbuf) (void)hipFree(w.zbuf);
            if (w.cbuf) (void)hipFree(w.cbuf);
            if (w.ev_side) (void)hipEventDestroy(w.ev_side);
            if (w.ev_main) (void)hipEventDestroy(w.ev_main);
            if (w.side) (void)hipStreamDestroy(w.side);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }
};

thread_local ThreadCtx g_ctx;

int ensure_device() {
    if (g_ctx.device >= 0) return AL_OK;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
        set_error("no HIP device visible (algebra_linear_hip needs an MI355X / gfx950)");
        return AL_ERR_NO_DEVICE;
    }
    int dev = 0;
    AL_HIP(hipGetDevice(&dev));
    g_ctx.device = dev;
    return AL_OK;
}

int buffer(int slot, size_t bytes, void** out) {
    DevBuf& b = g_ctx.bufs[slot];
    if (b.cap < bytes) {
        if (b.p) AL_HIP(hipFree(b.p));
        b.p = nullptr;
        b.cap = 0;
        size_t want = bytes < 256 ? 256 : bytes;
        hipError_t e = hipMalloc(&b.p, want);
        if (e != hipSuccess) {
            b.p = nullptr;
            set_error("device allocation failed");
            return AL_ERR_OOM;
        }
        b.cap = want;
    }
    *out = b.p;
    return AL_OK;
}

bool mul_ok(size_t a, size_t b, size_t* out) {
    if (a != 0 && b > SIZE_MAX / a) return false;
    *out = a * b;
    return true;
}

bool len_is(size_t len, size_t a, size_t b) {
    size_t p;
    return mul_ok(a, b, &p) && p == len;
}

int bad_shape(const char* msg) {
    set_error(msg);
    return AL_BAD_SHAPE;
}

int upload(int slot, const double* h, size_t n, double** d, hipStream_t s) {
    void* p;
    AL_CHECK(buffer(slot, n * sizeof(double), &p));
    *d = (double*)p;
    if (n) AL_HIP(hipMemcpyAsync(p, h, n * sizeof(double), hipMemcpyHostToDevice, s));
    return AL_OK;
}

int download(double* h, const double* d, size_t n, hipStream_t s) {
    if (n) AL_HIP(hipMemcpyAsync(h, d, n * sizeof(double), hipMemcpyDeviceToHost, s));
    return AL_OK;
}

int finish(hipStream_t s) {
    AL_HIP(hipStreamSynchronize(s));
    return AL_OK;
}

int status_from_info(int info) {
    if (info == 0) return AL_OK;
    if (info > 0) {
        set_error("matrix is singular (or not positive definite) to working precision");
        return AL_SINGULAR;
    }
    set_error("device-side timeout in a cooperative panel barrier");
    return AL_ERR_DEVICE;
}

int gemm_host(const char* name, bool ta, bool tb, double alpha, const double* a, size_t a_len, const double* b,
              size_t b_len, double beta, double* c, size_t c_len, size_t m, size_t n, size_t k) {
    // messages of src/blas.rs:202-204, :221-223, :238-240, :253-255
    std::string nm(name);
    if (!len_is(a_len, ta ? k : m, ta ? m : k)) return bad_shape((nm + (ta ? ": a must be k*m" : ": a must be m*k")).c_str());
    if (!len_is(b_len, tb ? n : k, tb ? k : n)) return bad_shape((nm + (tb ? ": b must be n*k" : ": b must be k*n")).c_str());
    if (!len_is(c_len, m, n)) return bad_shape((nm + ": c must be m*n").c_str());
    if (m == 0 || n == 0) return AL_OK;
    if ((!a && a_len) || (!b && b_len) || !c) {
        set_error("null pointer");
        return AL_ERR_ARG;
    }
    AL_CHECK(ensure_device());
    hipStream_t s;
    AL_CHECK(thread_stream(&s));
    double *dA = nullptr, *dB = nullptr, *dC = nullptr;
    const bool need_ab = !(k == 0 || alpha == 0.0);
    if (need_ab) {
        AL_CHECK(upload(0, a, a_len, &dA, s));
        AL_CHECK(upload(1, b, b_len, &dB, s));
    }
    if (beta != 0.0) {
        AL_CHECK(upload(2, c, c_len, &dC, s));
    } else {
        void* p;
        AL_CHECK(buffer(2, c_len * sizeof(double), &p));
        dC = (double*)p;
    }
    AL_CHECK(gemm_f64(s, ta, tb, m, n, k, alpha, dA, ta ? k : m, dB, tb ? n : k, beta, dC, m));
    AL_CHECK(download(c, dC, c_len, s));
    return finish(s);
}

int workspace_for(hipStream_t s, Workspace** out, size_t ipiv_len) {
    Workspace& w = g_ctx.ws[s];
    if (!w.info) {
        void* p;
        // info (int) and the abort flag, each on its own 64-byte line
        AL_HIP(hipMalloc(&p, 256));
        AL_HIP(hipMemset(p, 0, 256));
        w.info = (int*)p;
        w.abort_flag = (unsigned*)((char*)p + 64);
        AL_HIP(hipMalloc((void**)&w.flags, panel_flags_count() * sizeof(unsigned)));
        AL_HIP(hipMemset(w.flags, 0, panel_flags_count() * sizeof(unsigned)));
        // hipMemset runs on the legacy null stream, which does not order work on non-blocking
        // streams (the library's own and, typically, the caller's): finish it before any kernel
        // can read these words
        AL_HIP(hipDeviceSynchronize());
        w.gen = 0;
        w.slots_bytes = panel_slots_bytes();
        AL_HIP(hipMalloc((void**)&w.slots, w.slots_bytes));
        AL_CHECK(panel_exchange_init(s, &w));
        AL_HIP(hipMalloc((void**)&w.pairs, (1 + 2 * 512) * sizeof(int)));
        AL_HIP(hipMalloc((void**)&w.pairs2, (1 + 2 * 512) * sizeof(int)));
        w.device = g_ctx.device;
    }
    if (w.ipiv_len < ipiv_len) {
        if (w.ipiv32) AL_HIP(hipFree(w.ipiv32));
        w.ipiv32 = nullptr;
        AL_HIP(hipMalloc((void**)&w.ipiv32, (ipiv_len < 64 ? 64 : ipiv_len) * sizeof(int)));
        w.ipiv_len = ipiv_len < 64 ? 64 : ipiv_len;
    }
    *out = &w;
    return AL_OK;
}

// Runs getrf on host data already validated; downloads the factors and pivots.
int getrf_host(double* a, size_t M, size_t N, size_t* ipiv, hipStream_t s, Workspace* ws, double** dA_out) {
    double* dA;
    AL_CHECK(upload(0, a, M * N, &dA, s));
    const size_t kmax = M < N ? M : N;
    // identity pivots first, so that even a partial factorization never holds a wild index
    AL_CHECK(iota_i32(s, ws->ipiv32, kmax));
    AL_CHECK(ws_reserve(ws, M, 0));
    AL_CHECK(getrf_f64(s, ws, M, N, dA, M, ws->ipiv32, ws->info, ws->perm));
    int info = 0;
    std::vector<int> piv(kmax);
    AL_CHECK(download(a, dA, M * N, s));
    if (kmax) AL_HIP(hipMemcpyAsync(piv.data(), ws->ipiv32, kmax * sizeof(int), hipMemcpyDeviceToHost, s));
    AL_HIP(hipMemcpyAsync(&info, ws->info, sizeof(int), hipMemcpyDeviceToHost, s));
    AL_CHECK(finish(s));
    for (size_t i = 0; i < kmax; ++i) ipiv[i] = (size_t)piv[i];
    if (dA_out) *dA_out = dA;
    return status_from_info(info);
}

// Copies a solve's result back and reports a device-side wait that gave up during it (its
// rows of the result are NaN then) as AL_ERR_DEVICE.
int download_solution(double* h, const double* d, size_t n, hipStream_t s, Workspace* ws) {
    unsigned aborted = 0;
    AL_CHECK(download(h, d, n, s));
    AL_HIP(hipMemcpyAsync(&aborted, ws->abort_flag, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    AL_CHECK(finish(s));
    if (aborted) {
        set_error("device-side timeout in a triangular solve's block wait");
        return AL_ERR_DEVICE;
    }
    return AL_OK;
}

int check_pivots(const size_t* ipiv, size_t n) {
    for (size_t k = 0; k < n; ++k)
        if (ipiv[k] >= n) {
            set_error("dgetrs: pivot index out of range");
            return AL_ERR_ARG;
        }
    return AL_OK;
}

}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }

int fail_hip(hipError_t e, const char* what) {
    g_last_error = std::string("HIP error: ") + hipGetErrorString(e) + " in " + what;
    return AL_ERR_DEVICE;
}

int thread_stream(hipStream_t* out) {
    AL_CHECK(ensure_device());
    if (!g_ctx.stream) AL_HIP(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
    *out = g_ctx.stream;
    return AL_OK;
}

int ws_reserve(Workspace* ws, size_t n, size_t nw) {
    if (ws->perm_len < n) {
        if (ws->perm) AL_HIP(hipFree(ws->perm));
        ws->perm = nullptr;
        ws->perm_len = 0;
        AL_HIP(hipMalloc((void**)&ws->perm, n * sizeof(double)));
        ws->perm_len = n;
    }
    if (ws->wbuf_len < nw) {
        if (ws->wbuf) AL_HIP(hipFree(ws->wbuf));
        ws->wbuf = nullptr;
        ws->wbuf_len = 0;
        AL_HIP(hipMalloc((void**)&ws->wbuf, nw * sizeof(double)));
        ws->wbuf_len = nw;
    }
    return AL_OK;
}

int workspace_on(hipStream_t s, Workspace** out) { return workspace_for(s, out, 1); }

int cu_count() {
    static std::atomic<int> cus{0};
    int v = cus.load(std::memory_order_relaxed);
    if (v == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
            v = 256;
        cus.store(v, std::memory_order_relaxed);
    }
    return v;
}

int ws_reserve_flags(Workspace* ws, size_t entries) {
    if (ws->trsv_flags_len < entries) {
        if (ws->trsv_flags) AL_HIP(hipFree(ws->trsv_flags));
        ws->trsv_flags = nullptr;
        ws->trsv_flags_len = 0;
        AL_HIP(hipMalloc((void**)&ws->trsv_flags, entries * sizeof(unsigned)));
        AL_HIP(hipMemset(ws->trsv_flags, 0, entries * sizeof(unsigned)));
        AL_HIP(hipDeviceSynchronize());  // see workspace_for: the null stream orders nothing here
        ws->trsv_flags_len = entries;
    }
    return AL_OK;
}

int workspace(Workspace** out, size_t ipiv_len) {
    hipStream_t s;
    AL_CHECK(thread_stream(&s));
    return workspace_for(s, out, ipiv_len);
}

}  // namespace al

using namespace al;

extern "C" {

const char* al_last_error(void) { return g_last_error.c_str(); }

const char* al_version(void) { return "algebra_linear_hip 0.1.0 (gfx950)"; }

int al_device_count(void) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) return 0;
    return count;
}

int al_dgemm(double alpha, const double* a, size_t a_len, const double* b, size_t b_len, double beta, double* c,
             size_t c_len, size_t m, size_t n, size_t k) {
    return gemm_host("dgemm", false, false, alpha, a, a_len, b, b_len, beta, c, c_len, m, n, k);
}
int al_dgemm_tn(double alpha, const double* a, size_t a_len, const double* b, size_t b_len, double beta, double* c,
                size_t c_len, size_t m, size_t n, size_t k) {
    return gemm_host("dgemm_tn", true, false, alpha, a, a_len, b, b_len, beta, c, c_len, m, n, k);
}
int al_dgemm_nt(double alpha, const double* a, size_t a_len, const double* b, size_t b_len, double beta, double* c,
                size_t c_len, size_t m, size_t n, size_t k) {
    return gemm_host("dgemm_nt", false, true, alpha, a, a_len, b, b_len, beta, c, c_len, m, n, k);
}
int al_dgemm_tt(double alpha, const double* a, size_t a_len, const double* b, size_t b_len, double beta, double* c,
                size_t c_len, size_t m, size_t n, size_t k) {
    return gemm_host("dgemm_tt", true, true, alpha, a, a_len, b, b_len, beta, c, c_len, m, n, k);
}

int al_dtrsm(double alpha, const double* a, size_t a_len, double* b, size_t b_len, size_t m, size_t n, int left,
             int lower, int unit) {
    // src/blas.rs:1385-1397
    if (!len_is(b_len, m, n)) return bad_shape("dtrsm: b must be m*n");
    if (left && !len_is(a_len, m, m)) return bad_shape("dtrsm: a must be m*m for the left side");
    if (!left && !len_is(a_len, n, n)) return bad_shape("dtrsm: a must be n*n for the right side");
    if (m == 0 || n == 0) return AL_OK;
    hipStream_t s;
    AL_CHECK(thread_stream(&s));
    double *dA, *dB;
    AL_CHECK(upload(0, a, a_len, &dA, s));
    AL_CHECK(upload(1, b, b_len, &dB, s));
    AL_CHECK(trsm_f64(s, left != 0, lower != 0, false, unit != 0, m, n, alpha, dA, left ? m : n, dB, m));
    AL_CHECK(download(b, dB, b_len, s));
    return finish(s);
}

int al_dsyrk(double alpha, const double* a, size_t a_len, double beta, double* c, size_t c_len, size_t n, size_t k,
             int upper) {
    // src/blas.rs:610-611
    if (!len_is(a_len, n, k)) return bad_shape("dsyrk: a must be n*k");
    if (!len_is(c_len, n, n)) return bad_shape("dsyrk: c must be n*n");
    if (n == 0) return AL_OK;
    hipStream_t s;
    AL_CHECK(thread_stream(&s));
    double *dA = nullptr, *dC;
    if (k) AL_CHECK(upload(0, a, a_len, &dA, s));
    AL_CHECK(upload(2, c, c_len, &dC, s));  // the other triangle must come back unchanged
    AL_CHECK(gemm_f64(s, false, true, n, n, k, alpha, dA, n, dA, n, beta, dC, n, upper ? 2 : 1));
    AL_CHECK(download(c, dC, c_len, s));
    return finish(s);
}

int al_dgetrf(double* a, size_t a_len, size_t n, size_t* ipiv, size_t ipiv_len, double* det) {
    // src/lapack/mod.rs:112-118
    if (!len_is(a_len, n, n) || ipiv_len != n) return bad_shape("dgetrf: a must be n*n and ipiv n");
    if (n == 0) {
        if (det) *det = 1.0;
        return AL_OK;
    }
    hipStream_t s;
    AL_CHECK(thread_stream(&s));
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, n));
    int st = getrf_host(a, n, n, ipiv, s, ws, nullptr);
    if (st != AL_OK) return st;
    if (det) {
        // the crate's determinant: sign flip per interchange, then the running product, in
        // pivot order (src/lapack/mod.rs:141-150)
        double d = 1.0;
        for (size_t k = 0; k < n; ++k) {
            if (ipiv[k] != k) d = -d;
            d *= a[k + k * n];
        }
        *det = d;
    }
    return AL_OK;
}

int al_dgetrf_rect(double* a, size_t a_len, size_t nrows, size_t ncols, size_t* ipiv, size_t ipiv_len) {
    // src/lapack/mod.rs:203-205
    if (nrows < ncols || !len_is(a_len, nrows, ncols) || ipiv_len != ncols)
        return bad_shape("dgetrf_rect: need nrows >= ncols, a nrows*ncols and ipiv ncols");
    if (ncols == 0) return AL_OK;
    hipStream_t s;
    AL_CHECK(thread_stream(&s));
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, ncols));
    return getrf_host(a, nrows, ncols, ipiv, s, ws, nullptr);
}

int al_dgetrs_nrhs(const double* lu, size_t lu_len, const size_t* ipiv, size_t ipiv_len, double* b, size_t b_len,
                   size_t n, size_t nrhs) {
    if (!len_is(lu_len, n, n) || ipiv_len != n || !len_is(b_len, n, nrhs))
        return bad_shape("dgetrs: lu must be n*n, ipiv n and b n*nrhs");
    if (n == 0 || nrhs == 0) return AL_OK;
    AL_CHECK(check_pivots(ipiv, n));
    hipStream_t s;
    AL_CHECK(thread_stream(&s));
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, n));
    // compose the pivots into row origins on the host (O(n)), as the crate applies them
    std::vector<double> perm(n);
    {
        std::vector<size_t> p(n);
        for (size_t i = 0; i < n; ++i) p[i] = i;
        for (size_t k = 0; k < n; ++k) std::swap(p[k], p[ipiv[k]]);
        for (size_t i = 0; i < n; ++i) perm[i] = (double)p[i];
    }
    AL_CHECK(ws_reserve(ws, n, n * nrhs));
    double *dLU, *dB;
    AL_CHECK(upload(0, lu, lu_len, &dLU, s));
    AL_CHECK(upload(1, b, b_len, &dB, s));
    AL_HIP(hipMemcpyAsync(ws->perm, perm.data(), n * sizeof(double), hipMemcpyHostToDevice, s));
    AL_CHECK(getrs_f64(s, ws, n, nrhs, dLU, n, nullptr, ws->perm, dB, n));
    return download_solution(b, dB, b_len, s, ws);
}

int al_dgetrs(const double* lu, size_t lu_len, const size_t* ipiv, size_t ipiv_len, double* b, size_t b_len) {
    return al_dgetrs_nrhs(lu, lu_len, ipiv, ipiv_len, b, b_len, b_len, 1);
}

int al_dgesv_nrhs(double* a, size_t a_len, double* b, size_t b_len, size_t n, size_t nrhs, size_t* ipiv) {
    if (!len_is(a_len, n, n) || !len_is(b_len, n, nrhs)) return bad_shape("dgesv: a must be n*n and b n*nrhs");
    if (n == 0) return AL_OK;
    hipStream_t s;
    AL_CHECK(thread_stream(&s));
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, n));
    std::vector<size_t> piv(n);
    double* dA;
    int st = getrf_host(a, n, n, piv.data(), s, ws, &dA);
    if (ipiv) memcpy(ipiv, piv.data(), n * sizeof(size_t));
    if (st != AL_OK) return st;
    if (nrhs == 0) return AL_OK;
    double* dB;
    AL_CHECK(upload(1, b, b_len, &dB, s));
    AL_CHECK(ws_reserve(ws, n, n * nrhs));
    AL_CHECK(getrs_f64(s, ws, n, nrhs, dA, n, ws->ipiv32, ws->perm, dB, n));
    return download_solution(b, dB, b_len, s, ws);
}

int al_dgesv(double* a, size_t a_len, double* b, size_t b_len) {
    // src/lapack/mod.rs:309-323: n = b.len()
    const size_t n = b_len;
    if (!len_is(a_len, n, n)) return bad_shape("dgesv: a must be n*n with n = b.len()");
    return al_dgesv_nrhs(a, a_len, b, b_len, n, 1, nullptr);
}

int al_dpotrf(double* a, size_t a_len, size_t n) {
    if (!len_is(a_len, n, n)) return bad_shape("dpotrf: a must be n*n");
    if (n == 0) return AL_OK;
    hipStream_t s;
    AL_CHECK(thread_stream(&s));
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, 1));
    double* dA;
    AL_CHECK(upload(0, a, a_len, &dA, s));
    AL_CHECK(potrf_f64(s, ws, n, dA, n, ws->info));
    int info = 0;
    AL_CHECK(download(a, dA, a_len, s));
    AL_HIP(hipMemcpyAsync(&info, ws->info, sizeof(int), hipMemcpyDeviceToHost, s));
    AL_CHECK(finish(s));
    return status_from_info(info);
}

int al_dgesv_mixed(const double* a, size_t a_len, double* b, size_t b_len, size_t n, size_t nrhs, int factor,
                   const al_ir_options* opt, al_ir_stats* stats) {
    if (!len_is(a_len, n, n)) return bad_shape("dgesv_mixed: a must be n*n");
    if (!len_is(b_len, n, nrhs)) return bad_shape("dgesv_mixed: b must be n*nrhs");
    if (n == 0 || nrhs == 0) {
        if (stats) *stats = al_ir_stats{0, 0, 0, 1, 0.0, 0.0};
        return AL_OK;
    }
    hipStream_t s;
    AL_CHECK(thread_stream(&s));
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, 1));
    double *dA, *dB;
    void* dX;
    AL_CHECK(upload(0, a, a_len, &dA, s));
    AL_CHECK(upload(1, b, b_len, &dB, s));
    AL_CHECK(buffer(2, b_len * sizeof(double), &dX));
    const int st = gesv_mixed(s, ws, n, nrhs, dA, n, dB, n, (double*)dX, n, factor, opt, stats);
    if (st == AL_OK || st == AL_CONVERGENCE) {
        AL_CHECK(download(b, (double*)dX, b_len, s));
        AL_CHECK(finish(s));
    }
    return st;
}

// ------------------------------------------------------------------------------------------
// device-pointer API
// ------------------------------------------------------------------------------------------

// The device API runs on the caller's stream; NULL is HIP's null (default) stream.
static int dev_stream(void* stream, hipStream_t* s) {
    AL_CHECK(ensure_device());
    *s = (hipStream_t)stream;
    return AL_OK;
}

int al_dev_dgemm(int trans_a, int trans_b, size_t m, size_t n, size_t k, double alpha, const double* A, size_t lda,
                 const double* B, size_t ldb, double beta, double* C, size_t ldc, void* stream) {
    if (ldc < m || lda < (trans_a ? k : m) || ldb < (trans_b ? n : k)) return bad_shape("dgemm: leading dimension too small");
    hipStream_t s;
    AL_CHECK(dev_stream(stream, &s));
    return gemm_f64(s, trans_a != 0, trans_b != 0, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

int al_dev_dsyrk(int upper, size_t n, size_t k, double alpha, const double* A, size_t lda, double beta, double* C,
                 size_t ldc, void* stream) {
    if (ldc < n || lda < n) return bad_shape("dsyrk: leading dimension too small");
    hipStream_t s;
    AL_CHECK(dev_stream(stream, &s));
    return gemm_f64(s, false, true, n, n, k, alpha, A, lda, A, lda, beta, C, ldc, upper ? 2 : 1);
}

int al_dev_dtrsm(int left, int lower, int unit, size_t m, size_t n, double alpha, const double* A, size_t lda,
                 double* B, size_t ldb, void* stream) {
    if (ldb < m || lda < (left ? m : n)) return bad_shape("dtrsm: leading dimension too small");
    hipStream_t s;
    AL_CHECK(dev_stream(stream, &s));
    return trsm_f64(s, left != 0, lower != 0, false, unit != 0, m, n, alpha, A, lda, B, ldb);
}

int al_dev_dgetrf(size_t m, size_t n, double* A, size_t lda, int32_t* d_ipiv, int32_t* d_info, void* stream) {
    if (lda < m || m < n) return bad_shape("dgetrf: need lda >= m >= n");
    hipStream_t s;
    AL_CHECK(dev_stream(stream, &s));
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, 1));
    const size_t kmax = m < n ? m : n;
    if (kmax == 0) return fill_i32(s, d_info, 1, 0);
    // identity pivots first (a kernel, not a memset, so the sequence can be graph-captured)
    AL_CHECK(iota_i32(s, d_ipiv, kmax));
    return getrf_f64(s, ws, m, n, A, lda, d_ipiv, d_info);
}

int al_dev_dgetrs(size_t n, size_t nrhs, const double* LU, size_t lda, const int32_t* d_ipiv, double* B, size_t ldb,
                  void* stream) {
    if (lda < n || ldb < n) return bad_shape("dgetrs: leading dimension too small");
    hipStream_t s;
    AL_CHECK(dev_stream(stream, &s));
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, 1));
    return getrs_f64(s, ws, n, nrhs, LU, lda, d_ipiv, nullptr, B, ldb);
}

int al_dev_dpotrf(size_t n, double* A, size_t lda, int32_t* d_info, void* stream) {
    if (lda < n) return bad_shape("dpotrf: leading dimension too small");
    hipStream_t s;
    AL_CHECK(dev_stream(stream, &s));
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, 1));
    return potrf_f64(s, ws, n, A, lda, d_info);
}

int al_dev_sgemm(int trans_a, int trans_b, size_t m, size_t n, size_t k, float alpha, const float* A, size_t lda,
                 const float* B, size_t ldb, float beta, float* C, size_t ldc, void* stream) {
    if (ldc < m || lda < (trans_a ? k : m) || ldb < (trans_b ? n : k)) return bad_shape("sgemm: leading dimension too small");
    hipStream_t s;
    AL_CHECK(dev_stream(stream, &s));
    return gemm_f32(s, trans_a != 0, trans_b != 0, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

int al_dev_gemm_bf16(size_t m, size_t n, size_t k, float alpha, const float* A, size_t lda, const float* B, size_t ldb,
                     float beta, float* C, size_t ldc, void* stream) {
    if (ldc < m || lda < m || ldb < k) return bad_shape("gemm_bf16: leading dimension too small");
    hipStream_t s;
    AL_CHECK(dev_stream(stream, &s));
    if (m == 0 || n == 0) return AL_OK;
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, 1));
    const size_t kp = (k + 63) / 64 * 64;
    if (kp == 0) return gemm_f32(s, false, false, m, n, 0, alpha, A, lda, B, ldb, beta, C, ldc);  // C := beta C
    const size_t abytes = (m * kp * 2 + 255) / 256 * 256, bytes = abytes + n * kp * 2;
    if (ws->gbuf_bytes < bytes) {
        if (ws->gbuf) AL_HIP(hipFree(ws->gbuf));
        ws->gbuf = nullptr;
        ws->gbuf_bytes = 0;
        AL_HIP(hipMalloc(&ws->gbuf, bytes));
        ws->gbuf_bytes = bytes;
    }
    __bf16* At = (__bf16*)ws->gbuf;
    __bf16* Bt = (__bf16*)((char*)ws->gbuf + abytes);
    AL_CHECK(bf16_rows(s, m, k, kp, A, lda, At));
    AL_CHECK(bf16_cols(s, k, kp, n, B, ldb, Bt));
    return gemm_bf16(s, m, n, kp, alpha, At, Bt, beta, C, ldc);
}

int al_dev_dgesv_mixed(size_t n, size_t nrhs, const double* A, size_t lda, const double* B, size_t ldb, double* X,
                       size_t ldx, int factor, const al_ir_options* opt, al_ir_stats* stats, void* stream) {
    if (lda < n || ldb < n || ldx < n) return bad_shape("dgesv_mixed: leading dimension too small");
    hipStream_t s;
    AL_CHECK(dev_stream(stream, &s));
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, 1));
    return gesv_mixed(s, ws, n, nrhs, A, lda, B, ldb, X, ldx, factor, opt, stats);
}

int al_dev_batch_solve(size_t count, const al_batch_entry* d_entries, double* A_pool, size_t a_pool_len,
                       double* B_pool, int32_t* d_info, int32_t* d_steps, int factor, void* stream) {
    if (factor != AL_FACTOR_FP64 && factor != AL_FACTOR_FP32) {
        set_error("batch_solve: factor must be AL_FACTOR_FP64 or AL_FACTOR_FP32");
        return AL_ERR_ARG;
    }
    if (count > (size_t)INT32_MAX) {
        set_error("batch_solve: too many systems");
        return AL_ERR_ARG;
    }
    hipStream_t s;
    AL_CHECK(dev_stream(stream, &s));
    if (count == 0) return AL_OK;
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, 1));
    const bool mixed = factor == AL_FACTOR_FP32;
    const size_t sbytes = (batch_scratch_bytes(count) + 255) / 256 * 256;
    const size_t need = sbytes + (mixed ? a_pool_len * sizeof(float) : 0);
    if (ws->bbuf_bytes < need) {
        if (ws->bbuf) AL_HIP(hipFree(ws->bbuf));
        ws->bbuf = nullptr;
        ws->bbuf_bytes = 0;
        AL_HIP(hipMalloc(&ws->bbuf, need));
        ws->bbuf_bytes = need;
    }
    float* W = (float*)((char*)ws->bbuf + sbytes);
    return batch_solve(s, ws, count, d_entries, A_pool, B_pool, W, d_info, d_steps, ws->bbuf, mixed, 10);
}

int al_dev_batch_solve_noalloc(size_t count, const al_batch_entry* d_entries, double* A_pool, double* B_pool,
                               int32_t* d_info, int32_t* d_steps, void* stream) {
    if (count > (size_t)INT32_MAX) {
        set_error("batch_solve_noalloc: too many systems");
        return AL_ERR_ARG;
    }
    int devs = 0;
    if (hipGetDeviceCount(&devs) != hipSuccess || devs == 0) {
        set_error("no HIP device visible (algebra_linear_hip needs an MI355X / gfx950)");
        return AL_ERR_NO_DEVICE;
    }
    return batch_solve_noalloc((hipStream_t)stream, count, d_entries, A_pool, B_pool, d_info, d_steps);
}

int al_batch_tune_set(int bucket, int kb, int wgs_per_cu) { return batch_tune_set(bucket, kb, wgs_per_cu); }

int al_batch_tune_get(int bucket, int* kb, int* wgs_per_cu) { return batch_tune_get(bucket, kb, wgs_per_cu); }

// ------------------------------------------------------------------------------------------
// support routines: dgemv, dgemv_trans, dlange, dtrsv, dtrsv_transpose, dgetri, dsyr2k
// ------------------------------------------------------------------------------------------

static int gemv_host(const char* nm, bool trans, double alpha, const double* a, size_t a_len, const double* x,
                     size_t x_len, double beta, double* y, size_t y_len, size_t m, size_t n) {
    // messages of src/blas.rs:103-105 and :145-147
    const std::string name(nm);
    if (!len_is(a_len, m, n)) return bad_shape((name + ": a must be m*n").c_str());
    if (x_len != (trans ? m : n)) return bad_shape((name + (trans ? ": x must have m entries" : ": x must have n entries")).c_str());
    if (y_len != (trans ? n : m)) return bad_shape((name + (trans ? ": y must have n entries" : ": y must have m entries")).c_str());
    if (y_len == 0) return AL_OK;
    hipStream_t s;
    AL_CHECK(thread_stream(&s));
    double *dA = nullptr, *dX = nullptr, *dY;
    if (a_len) AL_CHECK(upload(0, a, a_len, &dA, s));
    if (x_len) AL_CHECK(upload(1, x, x_len, &dX, s));
    AL_CHECK(upload(2, y, y_len, &dY, s));
    AL_CHECK(gemv_f64(s, trans, m, n, alpha, dA, m ? m : 1, dX, beta, dY));
    AL_CHECK(download(y, dY, y_len, s));
    return finish(s);
}

int al_dgemv(double alpha, const double* a, size_t a_len, const double* x, size_t x_len, double beta, double* y,
             size_t y_len, size_t m, size_t n) {
    return gemv_host("dgemv", false, alpha, a, a_len, x, x_len, beta, y, y_len, m, n);
}

int al_dgemv_trans(double alpha, const double* a, size_t a_len, const double* x, size_t x_len, double beta, double* y,
                   size_t y_len, size_t m, size_t n) {
    return gemv_host("dgemv_trans", true, alpha, a, a_len, x, x_len, beta, y, y_len, m, n);
}

int al_dlange(char norm, const double* a, size_t a_len, size_t m, size_t n, double* out) {
    if (!len_is(a_len, m, n)) return bad_shape("dlange: a must be m*n");  // src/blas.rs:639
    if (!out) {
        set_error("dlange: null output");
        return AL_ERR_ARG;
    }
    hipStream_t s;
    AL_CHECK(thread_stream(&s));
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, 1));
    AL_CHECK(ws_reserve(ws, 0, lange_scratch_doubles(m, n) + 1));
    double* dA = nullptr;
    if (a_len) AL_CHECK(upload(0, a, a_len, &dA, s));
    double* dOut = ws->wbuf + lange_scratch_doubles(m, n);
    AL_CHECK(lange_f64(s, norm, m, n, dA, m ? m : 1, dOut, ws->wbuf));
    AL_CHECK(download(out, dOut, 1, s));
    return finish(s);
}

static int trsv_host(const char* nm, bool trans, const double* a, size_t a_len, double* x, size_t x_len, size_t n,
                     int lower, int unit) {
    // messages of src/blas.rs:1356-1357 and :1367-1368
    const std::string name(nm);
    if (!len_is(a_len, n, n)) return bad_shape((name + ": a must be n*n").c_str());
    if (x_len != n) return bad_shape((name + ": x must be n").c_str());
    if (n == 0) return AL_OK;
    hipStream_t s;
    AL_CHECK(thread_stream(&s));
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, 1));
    double *dA, *dX;
    AL_CHECK(upload(0, a, a_len, &dA, s));
    AL_CHECK(upload(1, x, x_len, &dX, s));
    AL_CHECK(fill_u32(s, ws->abort_flag, 1, 0u));
    if (trans)
        AL_CHECK(trsm_f64(s, true, lower != 0, true, unit != 0, n, 1, 1.0, dA, n, dX, n));
    else
        AL_CHECK(trsv_solve<double>(s, ws, lower != 0, unit != 0, n, 1, dA, n, dX, n));
    return download_solution(x, dX, x_len, s, ws);
}

int al_dtrsv(const double* a, size_t a_len, double* x, size_t x_len, size_t n, int lower, int unit) {
    return trsv_host("dtrsv", false, a, a_len, x, x_len, n, lower, unit);
}

int al_dtrsv_transpose(const double* a, size_t a_len, double* x, size_t x_len, size_t n, int lower, int unit) {
    return trsv_host("dtrsv_transpose", true, a, a_len, x, x_len, n, lower, unit);
}

int al_dgetri(double* lu, size_t lu_len, size_t n, const size_t* ipiv, size_t ipiv_len) {
    // src/lapack/mod.rs:258-260
    if (!len_is(lu_len, n, n) || ipiv_len != n) return bad_shape("dgetri: lu must be n*n and ipiv n");
    if (n == 0) return AL_OK;
    AL_CHECK(check_pivots(ipiv, n));
    if (n > INT32_MAX) {
        set_error("dgetri: order beyond 2^31");
        return AL_ERR_ARG;
    }
    hipStream_t s;
    AL_CHECK(thread_stream(&s));
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, n));
    std::vector<int> piv(n);
    for (size_t i = 0; i < n; ++i) piv[i] = (int)ipiv[i];
    double* dLU;
    AL_CHECK(upload(0, lu, lu_len, &dLU, s));
    AL_HIP(hipMemcpyAsync(ws->ipiv32, piv.data(), n * sizeof(int), hipMemcpyHostToDevice, s));
    AL_CHECK(getri_f64(s, ws, n, dLU, n, ws->ipiv32, ws->info));
    int info = 0;
    AL_HIP(hipMemcpyAsync(&info, ws->info, sizeof(int), hipMemcpyDeviceToHost, s));
    AL_CHECK(finish(s));
    if (info != 0) {
        set_error("dgetri: U has a zero on its diagonal");
        return AL_SINGULAR;  // the factor is left as it was, as the crate leaves it
    }
    AL_CHECK(download(lu, dLU, lu_len, s));
    return finish(s);
}

int al_dsyr2k(int upper, size_t m, size_t k, double alpha, const double* v, size_t v_len, const double* w,
              size_t w_len, double beta, double* c, size_t c_len) {
    if (!len_is(v_len, m, k)) return bad_shape("dsyr2k: v must be m*k");
    if (!len_is(w_len, m, k)) return bad_shape("dsyr2k: w must be m*k");
    if (!len_is(c_len, m, m)) return bad_shape("dsyr2k: c must be m*m");
    if (m == 0) return AL_OK;
    hipStream_t s;
    AL_CHECK(thread_stream(&s));
    double *dV = nullptr, *dW = nullptr, *dC;
    if (v_len) AL_CHECK(upload(0, v, v_len, &dV, s));
    if (w_len) AL_CHECK(upload(1, w, w_len, &dW, s));
    AL_CHECK(upload(2, c, c_len, &dC, s));  // the other triangle must come back unchanged
    AL_CHECK(syr2k_f64(s, upper != 0, m, k, alpha, dV, m, dW, m, beta, dC, m));
    AL_CHECK(download(c, dC, c_len, s));
    return finish(s);
}

int al_dev_dgemv(int trans, size_t m, size_t n, double alpha, const double* A, size_t lda, const double* x, double beta,
                 double* y, void* stream) {
    if (lda < (m ? m : 1)) return bad_shape("dgemv: leading dimension too small");
    hipStream_t s;
    AL_CHECK(dev_stream(stream, &s));
    return gemv_f64(s, trans != 0, m, n, alpha, A, lda, x, beta, y);
}

int al_dev_dlange(char norm, size_t m, size_t n, const double* A, size_t lda, double* d_out, void* stream) {
    if (lda < (m ? m : 1)) return bad_shape("dlange: leading dimension too small");
    hipStream_t s;
    AL_CHECK(dev_stream(stream, &s));
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, 1));
    AL_CHECK(ws_reserve(ws, 0, lange_scratch_doubles(m, n)));
    return lange_f64(s, norm, m, n, A, lda, d_out, ws->wbuf);
}

int al_dev_dtrsv(int lower, int trans, int unit, size_t n, const double* A, size_t lda, double* x, void* stream) {
    if (lda < (n ? n : 1)) return bad_shape("dtrsv: leading dimension too small");
    hipStream_t s;
    AL_CHECK(dev_stream(stream, &s));
    if (n == 0) return AL_OK;
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, 1));
    if (trans) return trsm_f64(s, true, lower != 0, true, unit != 0, n, 1, 1.0, A, lda, x, n);
    AL_CHECK(fill_u32(s, ws->abort_flag, 1, 0u));
    return trsv_solve<double>(s, ws, lower != 0, unit != 0, n, 1, A, lda, x, n);
}

int al_dev_dgetri(size_t n, double* LU, size_t lda, const int32_t* d_ipiv, int32_t* d_info, void* stream) {
    if (lda < (n ? n : 1)) return bad_shape("dgetri: leading dimension too small");
    hipStream_t s;
    AL_CHECK(dev_stream(stream, &s));
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, 1));
    return getri_f64(s, ws, n, LU, lda, d_ipiv, d_info);
}

int al_dev_dsyr2k(int upper, size_t m, size_t k, double alpha, const double* V, size_t lv, const double* W, size_t lw,
                  double beta, double* C, size_t ldc, void* stream) {
    if (ldc < (m ? m : 1) || lv < (m ? m : 1) || lw < (m ? m : 1)) return bad_shape("dsyr2k: leading dimension too small");
    hipStream_t s;
    AL_CHECK(dev_stream(stream, &s));
    return syr2k_f64(s, upper != 0, m, k, alpha, V, lv, W, lw, beta, C, ldc);
}

void al_debug_set_wait_limit(long long ticks) { set_wait_limit_ticks(ticks); }

int al_debug_coop_units_in_use(void) {
    if (ensure_device() != AL_OK) return 0;
    return coop_in_use();
}

int al_debug_panel_pin_half(void* stream) {
    hipStream_t s;
    if (dev_stream(stream, &s) != AL_OK) return -1;
    Workspace* ws;
    if (workspace_for(s, &ws, 1) != AL_OK) return -1;
    return ws->pin_half;
}

void al_debug_set_panel_pinning(int on) { set_panel_pinning(on != 0); }

int al_dev_dlaswp(size_t ncols, double* A, size_t lda, size_t k0, size_t k1, const int32_t* d_ipiv, void* stream) {
    if (k1 < k0) return bad_shape("dlaswp: k1 < k0");
    if (k1 > k0 && lda < k1) return bad_shape("dlaswp: leading dimension too small");
    hipStream_t s;
    AL_CHECK(dev_stream(stream, &s));
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, 1));
    for (size_t k = k0; k < k1; k += 256) {
        const size_t jb = (k1 - k) < 256 ? (k1 - k) : 256;
        AL_CHECK(swap_block_f64(s, ws, d_ipiv, k, jb, A, ncols, nullptr, 0, lda, nullptr, nullptr));
    }
    return AL_OK;
}

// ---- complex twins (complex.hip) --------------------------------------------------------

static zd zc(al_complex v) { return zd(v.re, v.im); }

static int zgemm_host(const char* name, bool ta, bool tb, al_complex alpha, const al_complex* a, size_t a_len,
                      const al_complex* b, size_t b_len, al_complex beta, al_complex* c, size_t c_len, size_t m,
                      size_t n, size_t k) {
    // messages of src/blas.rs:1024-1026, :1043-1045, :1060-1062, :1072-1074
    std::string nm(name);
    if (!len_is(a_len, ta ? k : m, ta ? m : k)) return bad_shape((nm + (ta ? ": a must be k*m" : ": a must be m*k")).c_str());
    if (!len_is(b_len, tb ? n : k, tb ? k : n)) return bad_shape((nm + (tb ? ": b must be n*k" : ": b must be k*n")).c_str());
    if (!len_is(c_len, m, n)) return bad_shape((nm + ": c must be m*n").c_str());
    if (m == 0 || n == 0) return AL_OK;
    if ((!a && a_len) || (!b && b_len) || !c) {
        set_error("null pointer");
        return AL_ERR_ARG;
    }
    hipStream_t s;
    AL_CHECK(thread_stream(&s));
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, 1));
    double *dA = nullptr, *dB = nullptr, *dC;
    if (k) {
        AL_CHECK(upload(0, (const double*)a, 2 * a_len, &dA, s));
        AL_CHECK(upload(1, (const double*)b, 2 * b_len, &dB, s));
    }
    AL_CHECK(upload(2, (const double*)c, 2 * c_len, &dC, s));
    AL_CHECK(zgemm_dev(s, ws, ta, tb, m, n, k, zc(alpha), (const zd*)dA, ta ? k : m, (const zd*)dB, tb ? n : k,
                       zc(beta), (zd*)dC, m));
    AL_CHECK(download((double*)c, dC, 2 * c_len, s));
    return finish(s);
}

int al_zgemm(al_complex alpha, const al_complex* a, size_t a_len, const al_complex* b, size_t b_len, al_complex beta,
             al_complex* c, size_t c_len, size_t m, size_t n, size_t k) {
    return zgemm_host("zgemm", false, false, alpha, a, a_len, b, b_len, beta, c, c_len, m, n, k);
}
int al_zgemm_tn(al_complex alpha, const al_complex* a, size_t a_len, const al_complex* b, size_t b_len,
                al_complex beta, al_complex* c, size_t c_len, size_t m, size_t n, size_t k) {
    return zgemm_host("zgemm_tn", true, false, alpha, a, a_len, b, b_len, beta, c, c_len, m, n, k);
}
int al_zgemm_nt(al_complex alpha, const al_complex* a, size_t a_len, const al_complex* b, size_t b_len,
                al_complex beta, al_complex* c, size_t c_len, size_t m, size_t n, size_t k) {
    return zgemm_host("zgemm_nt", false, true, alpha, a, a_len, b, b_len, beta, c, c_len, m, n, k);
}
int al_zgemm_tt(al_complex alpha, const al_complex* a, size_t a_len, const al_complex* b, size_t b_len,
                al_complex beta, al_complex* c, size_t c_len, size_t m, size_t n, size_t k) {
    return zgemm_host("zgemm_tt", true, true, alpha, a, a_len, b, b_len, beta, c, c_len, m, n, k);
}

// Factors M x N complex host data in place on the device; pivots downloaded 0-based.
static int zgetrf_host(al_complex* a, size_t M, size_t N, size_t* ipiv, hipStream_t s, Workspace* ws,
                       double** dA_out) {
    double* dA;
    AL_CHECK(upload(0, (const double*)a, 2 * M * N, &dA, s));
    const size_t kmax = M < N ? M : N;
    AL_CHECK(iota_i32(s, ws->ipiv32, kmax));
    AL_CHECK(zgetrf_dev(s, ws, M, N, (zd*)dA, M, ws->ipiv32, ws->info));
    int info = 0;
    std::vector<int> piv(kmax);
    AL_CHECK(download((double*)a, dA, 2 * M * N, s));
    if (kmax) AL_HIP(hipMemcpyAsync(piv.data(), ws->ipiv32, kmax * sizeof(int), hipMemcpyDeviceToHost, s));
    AL_HIP(hipMemcpyAsync(&info, ws->info, sizeof(int), hipMemcpyDeviceToHost, s));
    AL_CHECK(finish(s));
    for (size_t i = 0; i < kmax; ++i) ipiv[i] = (size_t)piv[i];
    if (dA_out) *dA_out = dA;
    return status_from_info(info);
}

int al_zgetrf(al_complex* a, size_t a_len, size_t n, size_t* ipiv, size_t ipiv_len) {
    // src/lapack/complex.rs:67-73
    if (!len_is(a_len, n, n) || ipiv_len != n) return bad_shape("zgetrf: a must be n*n and ipiv n");
    if (n == 0) return AL_OK;
    hipStream_t s;
    AL_CHECK(thread_stream(&s));
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, n));
    return zgetrf_host(a, n, n, ipiv, s, ws, nullptr);
}

int al_zgetrf_rect(al_complex* a, size_t a_len, size_t nrows, size_t ncols, size_t* ipiv, size_t ipiv_len) {
    // src/lapack/complex.rs:169-175
    if (nrows < ncols || !len_is(a_len, nrows, ncols) || ipiv_len != ncols)
        return bad_shape("zgetrf_rect: need nrows >= ncols, a nrows*ncols and ipiv ncols");
    if (ncols == 0) return AL_OK;
    hipStream_t s;
    AL_CHECK(thread_stream(&s));
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, ncols));
    return zgetrf_host(a, nrows, ncols, ipiv, s, ws, nullptr);
}

int al_zgesv(al_complex* a, size_t a_len, al_complex* b, size_t b_len) {
    // src/lapack/complex.rs:279-290: n is b's length
    const size_t n = b_len;
    if (!len_is(a_len, n, n)) return bad_shape("zgesv: a must be n*n for n = b.len()");
    if (n == 0) return AL_OK;
    hipStream_t s;
    AL_CHECK(thread_stream(&s));
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, n));
    std::vector<size_t> piv(n);
    double* dA;
    const int st = zgetrf_host(a, n, n, piv.data(), s, ws, &dA);
    if (st != AL_OK) return st;  // b untouched, as the crate's early return leaves it
    double* dB;
    AL_CHECK(upload(1, (const double*)b, 2 * n, &dB, s));
    AL_CHECK(zgetrs_dev(s, ws, n, 1, (const zd*)dA, n, ws->ipiv32, (zd*)dB, n, ws->info));
    return download_solution((double*)b, dB, 2 * n, s, ws);
}

int al_zpotrf(al_complex* a, size_t a_len, size_t n, int upper) {
    // src/lapack/complex.rs:344-347
    if (!len_is(a_len, n, n)) return bad_shape("zpotrf: a must be n*n");
    if (n == 0) return AL_OK;
    hipStream_t s;
    AL_CHECK(thread_stream(&s));
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, 1));
    double* dA;
    AL_CHECK(upload(0, (const double*)a, 2 * a_len, &dA, s));
    AL_CHECK(zpotrf_dev(s, ws, n, (zd*)dA, n, upper != 0, ws->info));
    int info = 0;
    AL_CHECK(download((double*)a, dA, 2 * a_len, s));
    AL_HIP(hipMemcpyAsync(&info, ws->info, sizeof(int), hipMemcpyDeviceToHost, s));
    AL_CHECK(finish(s));
    return status_from_info(info);
}

int al_dev_zgemm(int trans_a, int trans_b, size_t m, size_t n, size_t k, al_complex alpha, const al_complex* A,
                 size_t lda, const al_complex* B, size_t ldb, al_complex beta, al_complex* C, size_t ldc,
                 void* stream) {
    if (ldc < m || lda < (trans_a ? k : m) || ldb < (trans_b ? n : k)) return bad_shape("zgemm: leading dimension too small");
    hipStream_t s;
    AL_CHECK(dev_stream(stream, &s));
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, 1));
    return zgemm_dev(s, ws, trans_a != 0, trans_b != 0, m, n, k, zc(alpha), (const zd*)A, lda, (const zd*)B, ldb,
                     zc(beta), (zd*)C, ldc);
}

int al_dev_zgetrf(size_t m, size_t n, al_complex* A, size_t lda, int32_t* d_ipiv, int32_t* d_info, void* stream) {
    if (lda < m || m < n) return bad_shape("zgetrf: need lda >= m >= n");
    hipStream_t s;
    AL_CHECK(dev_stream(stream, &s));
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, 1));
    const size_t kmax = m < n ? m : n;
    if (kmax == 0) return fill_i32(s, d_info, 1, 0);
    AL_CHECK(iota_i32(s, d_ipiv, kmax));
    return zgetrf_dev(s, ws, m, n, (zd*)A, lda, d_ipiv, d_info);
}

int al_dev_zgetrs(size_t n, size_t nrhs, const al_complex* LU, size_t lda, const int32_t* d_ipiv, al_complex* B,
                  size_t ldb, void* stream) {
    if (lda < n || ldb < n) return bad_shape("zgetrs: leading dimension too small");
    hipStream_t s;
    AL_CHECK(dev_stream(stream, &s));
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, 1));
    return zgetrs_dev(s, ws, n, nrhs, (const zd*)LU, lda, d_ipiv, (zd*)B, ldb);
}

int al_dev_zpotrf(int upper, size_t n, al_complex* A, size_t lda, int32_t* d_info, void* stream) {
    if (lda < n) return bad_shape("zpotrf: leading dimension too small");
    hipStream_t s;
    AL_CHECK(dev_stream(stream, &s));
    Workspace* ws;
    AL_CHECK(workspace_for(s, &ws, 1));
    return zpotrf_dev(s, ws, n, (zd*)A, lda, upper != 0, d_info);
}

}  // extern "C"
