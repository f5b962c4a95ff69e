// Optional per-pass timing with HIP events on the solver stream. When
// enabled, every kernel group of a step is bracketed by two events; collect()
// waits for them and accumulates milliseconds and launch counts per pass.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <vector>

#include "engine.hpp"

namespace phx {

class PassProfiler {
  public:
    static constexpr int kMaxPasses = 16;

    ~PassProfiler() {
        for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
        for (const Pending& p : pending_) {
            (void)hipEventDestroy(p.start);
            (void)hipEventDestroy(p.stop);
        }
    }

    void set_enabled(bool on) { enabled_ = on; }
    bool enabled() const { return enabled_; }

    // Marks the start of pass `kind`; returns a token for end().
    int begin(hipStream_t stream, int kind) {
        if (!enabled_) return -1;
        Pending p{acquire(), acquire(), kind};
        PHX_HIP(hipEventRecord(p.start, stream));
        pending_.push_back(p);
        return static_cast<int>(pending_.size()) - 1;
    }
    void end(hipStream_t stream, int token) noexcept {
        if (token < 0) return;
        if (hipEventRecord(pending_[static_cast<size_t>(token)].stop, stream) != hipSuccess) failed_ = true;
    }

    // Waits for all recorded passes and folds them into the totals.
    void collect() {
        require(!failed_, PHX_ERROR_DEVICE, "hipEventRecord failed while profiling");
        for (const Pending& p : pending_) {
            PHX_HIP(hipEventSynchronize(p.stop));
            float ms = 0.0F;
            PHX_HIP(hipEventElapsedTime(&ms, p.start, p.stop));
            total_ms_[static_cast<size_t>(p.kind)] += ms;
            launches_[static_cast<size_t>(p.kind)] += 1u;
            pool_.push_back(p.start);
            pool_.push_back(p.stop);
        }
        pending_.clear();
    }
    void reset() {
        collect();
        total_ms_.fill(0.0);
        launches_.fill(0u);
    }
    double total_ms(int kind) const { return total_ms_[static_cast<size_t>(kind)]; }
    uint64_t launches(int kind) const { return launches_[static_cast<size_t>(kind)]; }

  private:
    struct Pending {
        hipEvent_t start;
        hipEvent_t stop;
        int kind;
    };
    hipEvent_t acquire() {
        if (!pool_.empty()) {
            hipEvent_t e = pool_.back();
            pool_.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        PHX_HIP(hipEventCreate(&e));
        return e;
    }
    bool enabled_ = false;
    bool failed_ = false;
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> pool_;
    std::array<double, kMaxPasses> total_ms_{};
    std::array<uint64_t, kMaxPasses> launches_{};
};

// RAII bracket for one pass.
class PassScope {
  public:
    PassScope(PassProfiler& p, hipStream_t s, int kind) : p_(p), s_(s), token_(p.begin(s, kind)) {}
    ~PassScope() { p_.end(s_, token_); }
    PassScope(const PassScope&) = delete;
    PassScope& operator=(const PassScope&) = delete;

  private:
    PassProfiler& p_;
    hipStream_t s_;
    int token_;
};

}  // namespace phx
