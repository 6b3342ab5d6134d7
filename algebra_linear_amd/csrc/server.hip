// Continuous-batching solve server: requests of mixed order and right-hand-side count are
// admitted while HBM allows, run concurrently on several streams, and small systems that are
// waiting together are solved by one batched launch.
//
// Admission: requests are taken in arrival order (small ones may join an earlier small one's
// launch, see below); the head of the queue is admitted when its
// HBM footprint fits what the running requests and the resident factorizations leave of the
// budget (a request larger than the whole budget runs alone).  Each of `slots` worker threads
// owns one HIP stream (default priority), grow-only device buffers and its share of the
// library's per-thread workspaces; a worker that frees up takes the next admissible request, so
// requests that arrive while others run start as soon as a slot and memory are free and their
// kernels overlap on the device.  A worker that takes a small request (order <= 512, no
// resident factorization involved) also takes every other waiting small request of the same
// precision that fits the budget, and solves them all with one persistent batched launch.
// A request may keep its factorization resident (its footprint stays reserved) and later
// requests may solve new right-hand sides with it.  A released factor's device block goes to a
// pool that later keep requests draw from (held blocks are capped at a quarter of the budget),
// and the workers' buffers only grow, so a steady request stream allocates no device memory.
//
// Every path is deterministic: a request's result is bit-for-bit the same alone or in a mix.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "al_internal.hpp"

namespace al {
namespace {

constexpr size_t SMALL_MAX = 512;
constexpr size_t BATCH_MAX = 1024;

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Pending {
    al_request r;
    uint64_t ticket;
    double t_submit;
    size_t footprint;
};

struct Resident {
    int kind = AL_BATCH_GESV;
    int factor = AL_FACTOR_FP64;
    size_t n = 0;
    double* A = nullptr;     // FP64: the factors; mixed: a copy of A for the refinement
    int* ipiv = nullptr;     // FP64 LU pivots
    MixedFactor mf;          // mixed: FP32 LU and row origins
    void* mem = nullptr;     // one allocation holding all of the above
    size_t mem_bytes = 0;    // its size
    size_t footprint = 0;
};

// Device blocks of released resident factors, kept for later keep requests so that steady
// keep/release traffic allocates nothing: a request takes the smallest kept block of at least
// its size (and at most twice it); blocks beyond a quarter of the HBM budget are freed.
struct FactorPool {
    std::mutex mu;
    std::multimap<size_t, void*> blocks;
    size_t held = 0, cap = 0;
    hipError_t take(size_t bytes, void** p, size_t* got) {
        {
            std::lock_guard<std::mutex> lk(mu);
            auto it = blocks.lower_bound(bytes);
            if (it != blocks.end() && it->first <= 2 * bytes) {
                *p = it->second;
                *got = it->first;
                held -= it->first;
                blocks.erase(it);
                return hipSuccess;
            }
        }
        *got = bytes;
        return hipMalloc(p, bytes);
    }
    void give(void* p, size_t bytes) {
        if (!p) return;
        {
            std::lock_guard<std::mutex> lk(mu);
            if (held + bytes <= cap) {
                blocks.emplace(bytes, p);
                held += bytes;
                return;
            }
        }
        (void)hipFree(p);
    }
    void clear() {
        std::lock_guard<std::mutex> lk(mu);
        for (auto& kv : blocks) (void)hipFree(kv.second);
        blocks.clear();
        held = 0;
    }
};

// Grow-only device buffer.
struct DBuf {
    void* p = nullptr;
    size_t cap = 0;
    int need(size_t bytes) {
        if (cap >= bytes) return AL_OK;
        if (p) AL_HIP(hipFree(p));
        p = nullptr;
        cap = 0;
        AL_HIP(hipMalloc(&p, bytes < 256 ? 256 : bytes));
        cap = bytes < 256 ? 256 : bytes;
        return AL_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

bool is_small(const al_request& r) { return r.n <= SMALL_MAX && r.factor_id < 0 && !r.keep; }

size_t footprint_of(const al_request& r) {
    const size_t n = r.n, nr = r.nrhs;
    size_t f = 8 * n * n + 16 * n * (nr ? nr : 1) + 4096;
    if (r.factor_id >= 0) f = 16 * n * (nr ? nr : 1) + 4096;  // the factor is already resident
    if (r.factor == AL_FACTOR_FP32) {
        const size_t np = mixed_padded(n);
        f += 4 * np * np + 8 * n * 16 * (n / 256 + 2);
    }
    return f;
}

}  // namespace
}  // namespace al

struct al_server {
    int device = 0;
    size_t budget = 0, in_use = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<al::Pending> queue;
    std::unordered_map<uint64_t, al_result> results;
    std::unordered_map<int64_t, al::Resident> factors;
    al::FactorPool pool;
    uint64_t next_ticket = 1;
    int64_t next_factor = 1;
    int running = 0;
    bool stop = false;
    std::vector<std::thread> workers;
};

namespace al {
namespace {

struct Worker {
    al_server* S;
    hipStream_t s = nullptr;
    DBuf dA, dB, dX, dE, dInfo, dSteps;

    // Solve one request on this worker's stream; fills res (status, steps, factor_id).
    int solve_one(const al_request& r, al_result& res, Resident* reuse, Resident* keep) {
        const size_t n = r.n, nr = r.nrhs;
        const size_t nb = n * nr * sizeof(double);
        AL_CHECK(dB.need(nb));
        AL_CHECK(dX.need(nb));
        double* B = (double*)dB.p;
        double* X = (double*)dX.p;
        if (nb) AL_HIP(hipMemcpyAsync(B, r.b, nb, hipMemcpyHostToDevice, s));
        double* A = nullptr;
        if (reuse) {
            A = reuse->A;
        } else if (keep) {
            A = keep->A;
            AL_HIP(hipMemcpyAsync(A, r.a, n * n * sizeof(double), hipMemcpyHostToDevice, s));
        } else {
            AL_CHECK(dA.need(n * n * sizeof(double) + 8 * n + 256));
            A = (double*)dA.p;
            AL_HIP(hipMemcpyAsync(A, r.a, n * n * sizeof(double), hipMemcpyHostToDevice, s));
        }
        res.steps = 0;
        if (r.factor == AL_FACTOR_FP32) {
            Workspace* ws;
            AL_CHECK(workspace_on(s, &ws));
            al_ir_stats st{};
            const int rc = gesv_mixed(s, ws, n, nr, A, n, B, n, X, n, AL_FACTOR_FP32, nullptr, &st,
                                      reuse ? &reuse->mf : nullptr, keep ? &keep->mf : nullptr);
            if (rc != AL_OK && rc != AL_SINGULAR && rc != AL_CONVERGENCE) {
                res.status = rc;  // e.g. a device-side timeout: no result, nothing kept
                return AL_OK;
            }
            res.status = rc;
            res.steps = st.fell_back ? -1 : st.iterations;
        } else {
            int* ipiv = reuse ? reuse->ipiv : keep ? keep->ipiv : (int*)((char*)dA.p + n * n * sizeof(double));
            AL_CHECK(dInfo.need(256));
            int* info = (int*)dInfo.p;
            if (nb) AL_HIP(hipMemcpyAsync(X, B, nb, hipMemcpyDeviceToDevice, s));
            int hinfo = 0;
            const bool chol = r.kind == AL_BATCH_POSV;
            if (!reuse) {
                AL_CHECK(chol ? al_dev_dpotrf(n, A, n, info, s) : al_dev_dgetrf(n, n, A, n, ipiv, info, s));
                AL_HIP(hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s));
                AL_HIP(hipStreamSynchronize(s));
            }
            unsigned aborted = 0;
            if (hinfo == 0 && nr) {
                if (chol) {
                    AL_CHECK(trsm_f64(s, true, true, false, false, n, nr, 1.0, A, n, X, n));  // L y = b
                    AL_CHECK(trsm_f64(s, true, true, true, false, n, nr, 1.0, A, n, X, n));   // L^T x = y
                } else {
                    AL_CHECK(al_dev_dgetrs(n, nr, A, n, ipiv, X, n, s));
                    // a solve whose block waits gave up leaves NaN rows and raises the flag
                    Workspace* ws;
                    AL_CHECK(workspace_on(s, &ws));
                    AL_HIP(hipMemcpyAsync(&aborted, ws->abort_flag, sizeof(unsigned), hipMemcpyDeviceToHost, s));
                    AL_HIP(hipStreamSynchronize(s));
                }
            }
            // info < 0: a bounded wait of the factorization gave up, so neither the factors nor
            // a solution exist; the request fails and nothing is kept
            if (hinfo < 0 || aborted) {
                set_error("server: device-side timeout in a cooperative wait");
                res.status = AL_ERR_DEVICE;
            } else {
                res.status = hinfo > 0 ? AL_SINGULAR : AL_OK;
            }
        }
        if (nb && (res.status == AL_OK || res.status == AL_CONVERGENCE)) AL_HIP(hipMemcpyAsync(r.x, X, nb, hipMemcpyDeviceToHost, s));
        AL_HIP(hipStreamSynchronize(s));
        return AL_OK;
    }

    // Solve a group of small requests with one batched launch.
    int solve_batch(const std::vector<Pending>& g, std::vector<al_result>& out) {
        size_t a_len = 0, b_len = 0;
        std::vector<al_batch_entry> ents(g.size());
        for (size_t i = 0; i < g.size(); ++i) {
            const al_request& r = g[i].r;
            ents[i] = al_batch_entry{a_len, b_len, (int32_t)r.n, (int32_t)r.nrhs, r.kind, 0};
            a_len += r.n * r.n;
            b_len += r.n * r.nrhs;
        }
        AL_CHECK(dA.need(a_len * sizeof(double)));
        AL_CHECK(dB.need(b_len * sizeof(double)));
        AL_CHECK(dE.need(g.size() * sizeof(al_batch_entry)));
        AL_CHECK(dInfo.need(g.size() * sizeof(int)));
        AL_CHECK(dSteps.need(g.size() * sizeof(int)));
        double* A = (double*)dA.p;
        double* B = (double*)dB.p;
        for (size_t i = 0; i < g.size(); ++i) {
            const al_request& r = g[i].r;
            if (r.n) AL_HIP(hipMemcpyAsync(A + ents[i].a_off, r.a, r.n * r.n * sizeof(double), hipMemcpyHostToDevice, s));
            if (r.n * r.nrhs)
                AL_HIP(hipMemcpyAsync(B + ents[i].b_off, r.b, r.n * r.nrhs * sizeof(double), hipMemcpyHostToDevice, s));
        }
        AL_HIP(hipMemcpyAsync(dE.p, ents.data(), g.size() * sizeof(al_batch_entry), hipMemcpyHostToDevice, s));
        AL_CHECK(al_dev_batch_solve(g.size(), (const al_batch_entry*)dE.p, A, a_len, B, (int*)dInfo.p, (int*)dSteps.p,
                                    g[0].r.factor == AL_FACTOR_FP32 ? AL_FACTOR_FP32 : AL_FACTOR_FP64, s));
        std::vector<int> info(g.size()), steps(g.size());
        AL_HIP(hipMemcpyAsync(info.data(), dInfo.p, g.size() * sizeof(int), hipMemcpyDeviceToHost, s));
        AL_HIP(hipMemcpyAsync(steps.data(), dSteps.p, g.size() * sizeof(int), hipMemcpyDeviceToHost, s));
        AL_HIP(hipStreamSynchronize(s));
        for (size_t i = 0; i < g.size(); ++i) {
            const al_request& r = g[i].r;
            out[i].status = info[i] > 0 ? AL_SINGULAR : info[i] < 0 ? AL_ERR_ARG : AL_OK;
            out[i].steps = steps[i];
            if (info[i] == 0 && r.n * r.nrhs)
                AL_HIP(hipMemcpyAsync(r.x, B + ents[i].b_off, r.n * r.nrhs * sizeof(double), hipMemcpyDeviceToHost, s));
        }
        AL_HIP(hipStreamSynchronize(s));
        return AL_OK;
    }

    void run() {
        if (hipSetDevice(S->device) != hipSuccess) return;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return;
        for (;;) {
            std::vector<Pending> group;
            Resident* reuse = nullptr;
            {
                std::unique_lock<std::mutex> lk(S->mu);
                S->cv_work.wait(lk, [&] {
                    if (S->stop) return true;
                    if (S->queue.empty()) return false;
                    const size_t f = S->queue.front().footprint;
                    return S->in_use + f <= S->budget || S->running == 0;
                });
                if (S->stop && S->queue.empty()) break;
                if (S->stop) {
                    // shutting down: fail what is left
                    while (!S->queue.empty()) {
                        al_result res{};
                        res.status = AL_ERR_DEVICE;
                        res.factor_id = -1;
                        S->results[S->queue.front().ticket] = res;
                        S->queue.pop_front();
                    }
                    S->cv_done.notify_all();
                    break;
                }
                group.push_back(S->queue.front());
                S->queue.pop_front();
                S->in_use += group[0].footprint;
                if (is_small(group[0].r)) {
                    // every waiting small request of the same precision joins the launch (they
                    // may pass larger requests: they take little of the device and of its time)
                    for (auto it = S->queue.begin(); it != S->queue.end() && group.size() < BATCH_MAX;) {
                        if (is_small(it->r) && it->r.factor == group[0].r.factor &&
                            S->in_use + it->footprint <= S->budget) {
                            S->in_use += it->footprint;
                            group.push_back(*it);
                            it = S->queue.erase(it);
                        } else {
                            ++it;
                        }
                    }
                } else if (group[0].r.factor_id >= 0) {
                    auto it = S->factors.find(group[0].r.factor_id);
                    if (it != S->factors.end()) reuse = &it->second;
                }
                ++S->running;
            }
            const double t_start = now_ms();
            std::vector<al_result> out(group.size());
            for (auto& o : out) o.factor_id = -1;
            int rc = AL_OK;
            Resident kept;
            bool kept_ok = false;
            if (group.size() > 1 || is_small(group[0].r)) {
                rc = solve_batch(group, out);
            } else {
                const al_request& r = group[0].r;
                if (r.factor_id >= 0 && !reuse) {
                    set_error("server: unknown resident factor");
                    rc = AL_ERR_ARG;
                } else if (reuse && (reuse->n != r.n || reuse->factor != r.factor || reuse->kind != r.kind)) {
                    set_error("server: the resident factor does not match the request");
                    rc = AL_ERR_ARG;
                } else {
                    Resident* keep = nullptr;
                    if (r.keep) {
                        // one allocation: A (FP64), pivots, and for the mixed mode the FP32 LU + origins
                        const size_t n = r.n, np = mixed_padded(n);
                        size_t bytes = n * n * 8 + ((n * 4 + 255) / 256) * 256;
                        if (r.factor == AL_FACTOR_FP32) bytes += np * np * 4 + np * 8 + 256;
                        hipError_t e = S->pool.take(bytes, &kept.mem, &kept.mem_bytes);
                        if (e != hipSuccess) {
                            (void)hipGetLastError();
                            set_error("server: device allocation failed for a resident factor");
                            rc = AL_ERR_OOM;
                        } else {
                            char* p = (char*)kept.mem;
                            kept.A = (double*)p;
                            p += n * n * 8;
                            kept.ipiv = (int*)p;
                            p += ((n * 4 + 255) / 256) * 256;
                            if (r.factor == AL_FACTOR_FP32) {
                                kept.mf.F = (float*)p;
                                p += np * np * 4;
                                kept.mf.perm = (double*)p;
                            }
                            kept.kind = r.kind;
                            kept.factor = r.factor;
                            kept.n = n;
                            kept.footprint = group[0].footprint;
                            keep = &kept;
                        }
                    }
                    if (rc == AL_OK) rc = solve_one(r, out[0], reuse, keep);
                    // only a factorization that completed (status AL_OK) may be reused
                    kept_ok = keep && rc == AL_OK && out[0].status == AL_OK;
                    if (keep && !kept_ok) S->pool.give(kept.mem, kept.mem_bytes);
                }
            }
            const double t_end = now_ms();
            {
                std::lock_guard<std::mutex> lk(S->mu);
                --S->running;
                for (size_t i = 0; i < group.size(); ++i) {
                    al_result& res = out[i];
                    if (rc != AL_OK) res.status = rc;
                    res.queue_ms = t_start - group[i].t_submit;
                    res.service_ms = t_end - t_start;
                    res.batch_size = (int)group.size();
                    if (i == 0 && kept_ok) {
                        const int64_t id = S->next_factor++;
                        S->factors[id] = kept;
                        res.factor_id = id;  // its footprint stays reserved until released
                    } else {
                        S->in_use -= group[i].footprint;
                    }
                    S->results[group[i].ticket] = res;
                }
            }
            S->cv_done.notify_all();
            S->cv_work.notify_all();
        }
        dA.release();
        dB.release();
        dX.release();
        dE.release();
        dInfo.release();
        dSteps.release();
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
};

}  // namespace
}  // namespace al

using namespace al;

extern "C" al_server* al_server_create(size_t hbm_budget_bytes, int slots) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        set_error("no HIP device visible (algebra_linear_hip needs an MI355X / gfx950)");
        return nullptr;
    }
    al_server* S = new al_server;
    (void)hipGetDevice(&S->device);
    S->budget = hbm_budget_bytes;
    S->pool.cap = hbm_budget_bytes / 4;
    const int k = slots < 1 ? 1 : (slots > 16 ? 16 : slots);
    for (int i = 0; i < k; ++i)
        S->workers.emplace_back([S] {
            Worker w{S};
            w.run();
        });
    return S;
}

extern "C" int al_server_submit(al_server* S, const al_request* r, uint64_t* ticket) {
    if (!S || !r || !ticket) {
        set_error("server_submit: null argument");
        return AL_ERR_ARG;
    }
    if (r->kind != AL_BATCH_GESV && r->kind != AL_BATCH_POSV) {
        set_error("server_submit: kind must be AL_BATCH_GESV or AL_BATCH_POSV");
        return AL_ERR_ARG;
    }
    if (r->factor != AL_FACTOR_FP64 && r->factor != AL_FACTOR_FP32) {
        set_error("server_submit: factor must be AL_FACTOR_FP64 or AL_FACTOR_FP32");
        return AL_ERR_ARG;
    }
    if (r->n > 32704 || (r->factor_id < 0 && r->n && !r->a) || (r->n * r->nrhs && (!r->b || !r->x))) {
        set_error("server_submit: missing operand or order beyond 32704");
        return AL_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) {
        set_error("server_submit: the server is shutting down");
        return AL_ERR_ARG;
    }
    Pending p{*r, S->next_ticket++, now_ms(), footprint_of(*r)};
    *ticket = p.ticket;
    S->queue.push_back(p);
    S->cv_work.notify_one();
    return AL_OK;
}

extern "C" int al_server_wait(al_server* S, uint64_t ticket, al_result* out) {
    if (!S || !out) {
        set_error("server_wait: null argument");
        return AL_ERR_ARG;
    }
    std::unique_lock<std::mutex> lk(S->mu);
    if (ticket == 0 || ticket >= S->next_ticket) {
        set_error("server_wait: unknown ticket");
        return AL_ERR_ARG;
    }
    S->cv_done.wait(lk, [&] { return S->results.count(ticket) != 0; });
    *out = S->results[ticket];
    S->results.erase(ticket);
    return AL_OK;
}

extern "C" int al_server_release_factor(al_server* S, int64_t factor_id) {
    if (!S) {
        set_error("server_release_factor: null server");
        return AL_ERR_ARG;
    }
    void* mem = nullptr;
    size_t bytes = 0;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->factors.find(factor_id);
        if (it == S->factors.end()) {
            set_error("server_release_factor: unknown factor");
            return AL_ERR_ARG;
        }
        mem = it->second.mem;
        bytes = it->second.mem_bytes;
        S->in_use -= it->second.footprint;
        S->factors.erase(it);
    }
    S->pool.give(mem, bytes);
    S->cv_work.notify_all();
    return AL_OK;
}

extern "C" void al_server_destroy(al_server* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (auto& t : S->workers) t.join();
    for (auto& kv : S->factors) (void)hipFree(kv.second.mem);
    S->pool.clear();
    delete S;
}
