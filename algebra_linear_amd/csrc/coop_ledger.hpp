// Process-wide account of the workgroups of cooperative kernels that spin on each other.
//
// The FP64 panel factorization (panel_getf2) spreads a panel over up to 128 workgroups that
// exchange a pivot candidate per column and wait for one another; each holds 128 KiB of LDS,
// so at most one fits on a CU.  Its waits can only end if every workgroup of the panel becomes
// resident.  Several factorizations running at once on different streams (the server's
// workers, concurrent host threads) could otherwise fill every CU with spinning workgroups of
// unfinished panels and stall each other until their bounded waits give up.
//
// The ledger keeps the sum of the panel workgroups that may be in flight at or below the CU
// count: a factorization reserves its largest panel's workgroup count before it enqueues any
// kernel, waiting on the host while the reservations of unfinished factorizations leave too
// little; the reservation is tied to an event recorded after its last kernel and returned once
// that event has completed.  Any set of factorizations holding reservations then fits on the
// device at once, so every panel can complete (other kernels, which do not spin, only delay it).
//
// Header-only and free of HIP types so that the accounting can be unit-tested on the CPU with
// stand-in events: Ev is the event handle, Done(ev) tells whether the work before it finished.
#pragma once

#include <chrono>
#include <deque>
#include <mutex>
#include <thread>
#include <utility>

namespace al {

template <typename Ev>
class CoopLedger {
public:
    explicit CoopLedger(int capacity = 0) : capacity_(capacity) {}

    void set_capacity(int c) {
        std::lock_guard<std::mutex> lk(mu_);
        capacity_ = c;
    }
    int capacity() {
        std::lock_guard<std::mutex> lk(mu_);
        return capacity_;
    }
    // Units reserved now (in flight or about to be enqueued), after retiring finished entries.
    template <typename Done, typename Recycle>
    int in_use(Done done, Recycle recycle) {
        std::lock_guard<std::mutex> lk(mu_);
        retire(done, recycle);
        return used_;
    }

    // Reserves `units` (clamped to the capacity) once they fit beside the unfinished
    // reservations; polls the in-flight events every `poll` while they do not.  Returns the
    // units reserved, to be passed to hold_until() or release().
    template <typename Done, typename Recycle>
    int acquire(int units, Done done, Recycle recycle,
                std::chrono::microseconds poll = std::chrono::microseconds(20)) {
        if (units <= 0) return 0;
        for (;;) {
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (units > capacity_) units = capacity_ > 0 ? capacity_ : units;
                retire(done, recycle);
                if (used_ + units <= capacity_ || used_ == 0) {
                    used_ += units;
                    return units;
                }
            }
            std::this_thread::sleep_for(poll);
        }
    }
    // The reservation stays until `ev` (recorded after the reserving work) has completed.
    void hold_until(Ev ev, int units) {
        if (units <= 0) return;
        std::lock_guard<std::mutex> lk(mu_);
        inflight_.push_back(std::make_pair(ev, units));
    }
    // Returns a reservation at once (the work was not enqueued, or has been waited for).
    void release(int units) {
        if (units <= 0) return;
        std::lock_guard<std::mutex> lk(mu_);
        used_ -= units;
    }

private:
    template <typename Done, typename Recycle>
    void retire(Done done, Recycle recycle) {
        for (auto it = inflight_.begin(); it != inflight_.end();) {
            if (done(it->first)) {
                used_ -= it->second;
                recycle(it->first);
                it = inflight_.erase(it);
            } else {
                ++it;
            }
        }
    }

    std::mutex mu_;
    int capacity_ = 0;
    int used_ = 0;
    std::deque<std::pair<Ev, int>> inflight_;
};

}  // namespace al
