// timing.h -- optional per-kernel HIP-event timers of a context (tosca_ctx_enable_timing).
// When enabled, every launch made through TOSCA_LAUNCH is bracketed by two events recorded on
// the launch stream; tosca_ctx_timing_* synchronise on them and return per-kernel totals.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

namespace tosca {

struct KernelTimer {
    bool on = false;
    struct Pending { const char *name; hipEvent_t a, b; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    std::map<std::string, std::pair<double, long>> acc;

    hipEvent_t get() {
        if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }
    void collect() {
        for (auto &p : pending) {
            float ms = 0.f;
            (void)hipEventSynchronize(p.b);
            (void)hipEventElapsedTime(&ms, p.a, p.b);
            auto &v = acc[p.name];
            v.first += ms;
            v.second += 1;
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
    }
    void reset() { collect(); acc.clear(); }
    void release() {
        collect();
        for (auto e : pool) (void)hipEventDestroy(e);
        pool.clear();
    }
};

// the timer of the context whose entry point is running on this host thread (or null)
extern thread_local KernelTimer *g_timer;

}  // namespace tosca

#define TOSCA_LAUNCH(NAME, KERNEL, GRID, BLOCK, SHM, STREAM, ...)                                   \
    do {                                                                                         \
        tosca::KernelTimer *T_ = tosca::g_timer;                                                 \
        hipEvent_t a_ = nullptr;                                                                 \
        if (T_ && T_->on) { a_ = T_->get(); (void)hipEventRecord(a_, STREAM); }                  \
        hipLaunchKernelGGL(KERNEL, GRID, BLOCK, SHM, STREAM, __VA_ARGS__);                       \
        if (T_ && T_->on) {                                                                      \
            hipEvent_t b_ = T_->get();                                                           \
            (void)hipEventRecord(b_, STREAM);                                                    \
            T_->pending.push_back({NAME, a_, b_});                                               \
            if (T_->pending.size() > 8192) T_->collect();                                       \
        }                                                                                        \
    } while (0)
