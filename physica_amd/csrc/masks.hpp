// Host side of a solver's neighbour masks: the exact and skin mask buffers,
// the device flags of a step, and the skin radius, adapted from the predicted
// displacements of a previous step read back without blocking.
#pragma once

#include "engine.hpp"

namespace phx {

// Smallest float t with sqrtf(t) >= h: a candidate at squared distance d2
// lies inside the support (sqrtf(d2) < h) exactly when d2 < t.
float sqrt_support_threshold(float h);

// Squared-distance bound of an exact mask that keeps every pair on which any
// kernel of the support h can be non-zero: inside sqrtf(d2) < h (spiky,
// cubic), d2 < h * h (poly6), and the FP64 radius of the Hessian products,
// with a relative margin for the FP32 rounding of d2.
float exact_mask_threshold(float h);

// Pinned host buffer for asynchronous read-back.
class PinnedBuffer {
  public:
    PinnedBuffer() = default;
    PinnedBuffer(const PinnedBuffer&) = delete;
    PinnedBuffer& operator=(const PinnedBuffer&) = delete;
    ~PinnedBuffer() {
        if (ptr_ != nullptr) (void)hipHostFree(ptr_);
    }
    void resize(size_t bytes) {
        if (bytes <= bytes_) return;
        if (ptr_ != nullptr) (void)hipHostFree(ptr_);
        ptr_ = nullptr;
        PHX_HIP(hipHostMalloc(&ptr_, bytes, hipHostMallocDefault));
        PHX_HIP(hipHostGetDevicePointer(&device_, ptr_, 0u));
        bytes_ = bytes;
    }
    void* data() { return ptr_; }
    // the same memory as the device addresses it
    void* device() { return device_; }

  private:
    void* ptr_ = nullptr;
    void* device_ = nullptr;
    size_t bytes_ = 0;
};

class MaskSet {
  public:
    MaskSet() = default;
    MaskSet(const MaskSet&) = delete;
    MaskSet& operator=(const MaskSet&) = delete;
    ~MaskSet();

    // Masks of n particles and flags for `predicting_passes` displacement
    // slots (one per pass that writes predicted positions).
    void allocate(uint32_t n, uint32_t predicting_passes, float support_radius);
    // Adopts the skin of the newest completed read-back, if any, and clears the flags;
    // with full_stencil the step's masks are marked unusable (overflow), so
    // every pass walks the whole 27-cell stencil.
    void begin_step(hipStream_t stream, bool full_stencil = false);
    // Enqueues the read-back of this step's displacement flags.
    void end_step(hipStream_t stream);

    // The skin mask keeps d2 < t_skin; a pass may use it while every particle
    // moved at most sqrt(limit2) (negative: never).
    float t_skin() const;
    float limit2() const;
    float skin() const { return skin_; }
    uint32_t flag_count() const { return flag_count_; }

    uint16_t* exact() { return exact_.data(); }
    uint16_t* skin_mask() { return skin_mask_.data(); }
    uint32_t* flags() { return flags_.data(); }
    // per particle: walks the full stencil (written by the mask pass)
    uint8_t* crowded() { return crowded_.data(); }

  private:
    float h_ = 0.0F;
    float skin_ = 0.3F;  // in units of h
    uint32_t flag_count_ = 0;
    DeviceBuffer<uint16_t> exact_, skin_mask_;
    DeviceBuffer<uint8_t> crowded_;
    DeviceBuffer<uint32_t> flags_;
    // A ring of read-backs: a step's copy may still be queued behind its
    // last pass when the next step starts (always so when the host waits
    // inside a step), so begin_step adopts the newest one that completed.
    static constexpr uint32_t kReadbacks = 4;
    PinnedBuffer flags_host_;  // kReadbacks copies of the flags
    hipEvent_t readback_[kReadbacks] = {};
    bool readback_pending_[kReadbacks] = {};
    uint32_t readback_next_ = 0;  // slot of the next end_step
};

}  // namespace phx
