// Per-kernel timing with HIP events on the stream the kernels run on.
#pragma once

#include <map>
#include <string>
#include <vector>

#include "common.hpp"

namespace demsph {

class KernelTimer {
public:
    bool enabled = false;

    // Records an event pair around one launch.  Call begin(), launch, end().
    void begin(const char* name, hipStream_t s)
    {
        if (!enabled) return;
        Rec r;
        r.name = name;
        r.a = take();
        r.b = take();
        DEMSPH_HIP(hipEventRecord(r.a, s));
        pending_.push_back(r);
    }
    void end(hipStream_t s)
    {
        if (!enabled || pending_.empty()) return;
        DEMSPH_HIP(hipEventRecord(pending_.back().b, s));
    }
    // Converts recorded event pairs into totals.  Each event is waited on
    // before it is read: a stream synchronisation alone has been seen to
    // leave an event's timestamp not yet readable ("device not ready") while
    // other host threads share the hardware queues.  A pair that still cannot
    // be read is counted in dropped() and its events are destroyed, not
    // pooled; timing never fails the work it measures, and the pending list
    // is taken over before the loop so that no event is both pending and
    // pooled.
    void harvest()
    {
        std::vector<Rec> recs;
        recs.swap(pending_);
        for (auto& r : recs) {
            float ms = 0.f;
            hipError_t e = hipEventSynchronize(r.b);
            if (e == hipSuccess) e = hipEventSynchronize(r.a);
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, r.a, r.b);
            if (e == hipSuccess) {
                auto& t = totals_[r.name];
                t.first += 1;
                t.second += ms;
                pool_.push_back(r.a);
                pool_.push_back(r.b);
            } else {
                (void)hipGetLastError();
                dropped_++;
                (void)hipEventDestroy(r.a);
                (void)hipEventDestroy(r.b);
            }
        }
    }
    long dropped() const { return dropped_; }
    void reset()
    {
        totals_.clear();
    }
    const std::map<std::string, std::pair<long, double>>& totals() const { return totals_; }
    // Exchanges the accumulated totals (lets a caller time its own launches
    // and then put back what was recorded before).
    void swap_totals(std::map<std::string, std::pair<long, double>>& other) { totals_.swap(other); }
    ~KernelTimer()
    {
        for (auto& r : pending_) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        for (auto e : pool_) (void)hipEventDestroy(e);
    }

private:
    struct Rec {
        std::string name;
        hipEvent_t a, b;
    };
    hipEvent_t take()
    {
        if (!pool_.empty()) {
            hipEvent_t e = pool_.back();
            pool_.pop_back();
            return e;
        }
        hipEvent_t e;
        DEMSPH_HIP(hipEventCreate(&e));
        return e;
    }
    std::vector<Rec> pending_;
    std::vector<hipEvent_t> pool_;
    long dropped_ = 0;
    std::map<std::string, std::pair<long, double>> totals_;
};

// Launches a kernel with optional event timing.
#define DEMSPH_LAUNCH(timer, name, kernel, grid, block, shmem, stream, ...)          \
    do {                                                                             \
        (timer).begin(name, stream);                                                 \
        hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);         \
        (timer).end(stream);                                                         \
        DEMSPH_HIP(hipGetLastError());                                               \
    } while (0)

}  // namespace demsph
