// CPU unit test of the cooperative-panel ledger (algebra_linear_amd/csrc/coop_ledger.hpp) with
// stand-in events: an event is an index into a table of completion flags that the test (or a
// helper thread) raises.  Prints "ok" and exits 0 when every check holds.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "../../algebra_linear_amd/csrc/coop_ledger.hpp"

static std::atomic<bool> g_done[16];
static std::vector<int> g_recycled;

#define CHECK(c)                                                     \
    do {                                                             \
        if (!(c)) {                                                  \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            std::exit(1);                                            \
        }                                                            \
    } while (0)

int main() {
    auto done = [](int e) { return g_done[e].load(); };
    auto recycle = [](int e) { g_recycled.push_back(e); };
    const auto poll = std::chrono::microseconds(50);
    al::CoopLedger<int> L(256);

    // nothing reserved: a request of 0 units takes nothing
    CHECK(L.acquire(0, done, recycle, poll) == 0);
    CHECK(L.in_use(done, recycle) == 0);

    // two panels of 128 workgroups fit together on 256 CUs
    CHECK(L.acquire(128, done, recycle, poll) == 128);
    L.hold_until(0, 128);
    CHECK(L.acquire(94, done, recycle, poll) == 94);
    L.hold_until(1, 94);
    CHECK(L.in_use(done, recycle) == 222);

    // a third of 94 does not fit (222 + 94 > 256): it must wait until event 0 completes
    std::atomic<bool> got{false};
    std::thread t([&] {
        const int u = L.acquire(94, done, recycle, poll);
        CHECK(u == 94);
        got = true;
    });
    std::this_thread::sleep_for(std::chrono::milliseconds(30));
    CHECK(!got.load());  // still waiting: the reservations of events 0 and 1 hold
    g_done[0] = true;    // the first factorization's last kernel ends
    t.join();
    CHECK(got.load());
    CHECK(g_recycled.size() == 1 && g_recycled[0] == 0);  // its event went back to the pool
    L.hold_until(2, 94);
    CHECK(L.in_use(done, recycle) == 188);  // 94 (event 1) + 94 (event 2)

    // a request above the capacity is clamped and runs once everything else has finished
    g_done[1] = true;
    g_done[2] = true;
    CHECK(L.in_use(done, recycle) == 0);
    CHECK(L.acquire(1000, done, recycle, poll) == 256);
    L.release(256);  // work not enqueued: returned at once
    CHECK(L.in_use(done, recycle) == 0);

    // many threads reserving and completing at random never exceed the capacity
    std::atomic<int> peak{0}, live{0};
    std::atomic<int> next_ev{3};
    std::vector<std::thread> ts;
    for (int k = 0; k < 8; ++k)
        ts.emplace_back([&, k] {
            for (int r = 0; r < 20; ++r) {
                const int want = 60 + (k * 7 + r * 13) % 69;  // 60..128
                const int u = L.acquire(want, done, recycle, poll);
                const int now = (live += u);
                int p = peak.load();
                while (now > p && !peak.compare_exchange_weak(p, now)) {
                }
                CHECK(now <= 256);
                std::this_thread::sleep_for(std::chrono::microseconds(200));
                live -= u;
                L.release(u);
            }
        });
    for (auto& th : ts) th.join();
    CHECK(peak.load() <= 256 && peak.load() > 128);  // concurrency happened, within capacity
    CHECK(L.in_use(done, recycle) == 0);
    (void)next_ev;
    std::printf("ok\n");
    return 0;
}
