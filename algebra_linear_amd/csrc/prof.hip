// Launch instrumentation behind al_prof_*: HIP events bracket each instrumented launch on
// the stream it runs on, tagged with the launch's algorithmic flops and bytes, so a caller
// can read achieved FLOP/s and GB/s per kernel class over exactly the launches it made.
#include <atomic>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "al_internal.hpp"

namespace al {
namespace {

struct Rec {
    hipEvent_t start = nullptr, stop = nullptr;
    std::string cls;
    double flops = 0.0, bytes = 0.0;
    bool done = false;
};

std::atomic<bool> g_on{false};
std::mutex g_mu;
std::vector<Rec> g_recs;

}  // namespace

int prof_begin(hipStream_t s) {
    if (!g_on.load(std::memory_order_relaxed)) return -1;
    Rec r;
    if (hipEventCreate(&r.start) != hipSuccess) return -1;
    if (hipEventCreate(&r.stop) != hipSuccess) {
        (void)hipEventDestroy(r.start);
        return -1;
    }
    if (hipEventRecord(r.start, s) != hipSuccess) {
        (void)hipEventDestroy(r.start);
        (void)hipEventDestroy(r.stop);
        return -1;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    g_recs.push_back(r);
    return (int)g_recs.size() - 1;
}

void prof_end(int ticket, hipStream_t s, const char* kernel_class, double flops, double bytes) {
    if (ticket < 0) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if ((size_t)ticket >= g_recs.size()) return;
    Rec& r = g_recs[ticket];
    (void)hipEventRecord(r.stop, s);
    r.cls = kernel_class;
    r.flops = flops;
    r.bytes = bytes;
    r.done = true;
}

}  // namespace al

extern "C" {

void al_prof_enable(int on) { al::g_on.store(on != 0); }

void al_prof_reset(void) {
    std::lock_guard<std::mutex> lk(al::g_mu);
    for (auto& r : al::g_recs) {
        if (r.start) (void)hipEventSynchronize(r.stop);
        if (r.start) (void)hipEventDestroy(r.start);
        if (r.stop) (void)hipEventDestroy(r.stop);
    }
    al::g_recs.clear();
}

int al_prof_query(const char* kernel_class, int* launches, double* total_ms, double* flops, double* bytes) {
    std::lock_guard<std::mutex> lk(al::g_mu);
    int n = 0;
    double ms = 0.0, fl = 0.0, by = 0.0;
    for (auto& r : al::g_recs) {
        if (!r.done || r.cls != kernel_class) continue;
        hipError_t e = hipEventSynchronize(r.stop);
        if (e != hipSuccess) return al::fail_hip(e, "hipEventSynchronize");
        float t = 0.f;
        e = hipEventElapsedTime(&t, r.start, r.stop);
        if (e != hipSuccess) return al::fail_hip(e, "hipEventElapsedTime");
        ++n;
        ms += t;
        fl += r.flops;
        by += r.bytes;
    }
    if (launches) *launches = n;
    if (total_ms) *total_ms = ms;
    if (flops) *flops = fl;
    if (bytes) *bytes = by;
    return AL_OK;
}

}  // extern "C"
