// Skin adaptation and flag handling of the neighbour masks.
#include "masks.hpp"

#include <cmath>
#include <cstring>

#include "neighbor_mask.hpp"

namespace phx {

float sqrt_support_threshold(float h) {
    float t = h * h;
    while (std::sqrt(t) < h) t = std::nextafter(t, INFINITY);
    for (;;) {
        const float below = std::nextafter(t, 0.0F);
        if (below <= 0.0F || std::sqrt(below) < h) break;
        t = below;
    }
    return t;
}

float exact_mask_threshold(float h) { return std::fmax(sqrt_support_threshold(h), h * h) * (1.0F + 4e-6F); }

MaskSet::~MaskSet() {
    for (hipEvent_t e : readback_) {
        if (e == nullptr) continue;
        (void)hipEventSynchronize(e);
        (void)hipEventDestroy(e);
    }
}

void MaskSet::allocate(uint32_t n, uint32_t predicting_passes, float support_radius) {
    h_ = support_radius;
    exact_.resize(static_cast<size_t>(kStencil) * n);
    skin_mask_.resize(static_cast<size_t>(kStencil) * n);
    crowded_.resize(n);
    flag_count_ = kFlagDisplacement + predicting_passes;
    flags_.resize(flag_count_);
    flags_host_.resize(sizeof(uint32_t) * flag_count_ * kReadbacks);
    for (hipEvent_t& e : readback_)
        if (e == nullptr) PHX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

// Picks the skin from the largest predicted displacement of the newest step
// whose read-back completed: twice the displacement (both particles of a
// pair move) with a 30 % margin, in [0.02, 0.6] h. Older read-backs are
// dropped.
void MaskSet::begin_step(hipStream_t stream, bool full_stencil) {
    for (uint32_t back = 1; back <= kReadbacks; ++back) {
        const uint32_t k = (readback_next_ + kReadbacks - back) % kReadbacks;
        if (!readback_pending_[k] || hipEventQuery(readback_[k]) != hipSuccess) continue;
        const auto* flags = static_cast<const uint32_t*>(flags_host_.data()) + static_cast<size_t>(k) * flag_count_;
        float moved2 = 0.0F;
        for (uint32_t f = kFlagDisplacement; f < flag_count_; ++f) {
            float v;
            std::memcpy(&v, &flags[f], sizeof(float));
            moved2 = std::fmax(moved2, v);
        }
        const float moved = std::sqrt(moved2) / h_;
        skin_ = std::fmin(0.6F, std::fmax(0.02F, 2.6F * moved + 0.01F));
        // this read-back and every older one are consumed
        for (uint32_t older = back; older <= kReadbacks; ++older)
            readback_pending_[(readback_next_ + kReadbacks - older) % kReadbacks] = false;
        break;
    }
    fill_u32(stream, flags_.data(), flag_count_, 0u);
    if (full_stencil) fill_u32(stream, flags_.data() + kFlagOverflow, 1u, 1u);
}

void MaskSet::end_step(hipStream_t stream) {
    const uint32_t k = readback_next_;
    // a slot still in flight (the host is kReadbacks steps ahead) is reused
    // only after its copy ends
    if (readback_pending_[k]) PHX_HIP(hipEventSynchronize(readback_[k]));
    auto* host = static_cast<uint32_t*>(flags_host_.data()) + static_cast<size_t>(k) * flag_count_;
    PHX_HIP(hipMemcpyAsync(host, flags_.data(), sizeof(uint32_t) * flag_count_, hipMemcpyDeviceToHost, stream));
    PHX_HIP(hipEventRecord(readback_[k], stream));
    readback_pending_[k] = true;
    readback_next_ = (k + 1u) % kReadbacks;
}

// R = h (1 + skin); the mask keeps d2 < R^2 (with a relative margin), and a
// pass may use it while every particle moved at most (R - h) / 2.
float MaskSet::t_skin() const {
    const float radius = h_ * (1.0F + skin_);
    return radius * radius * (1.0F + 1e-5F);
}

float MaskSet::limit2() const {
    const float radius = h_ * (1.0F + skin_);
    const float limit = 0.5F * (radius - h_ * (1.0F + 2e-5F));
    return limit > 0.0F ? limit * limit : -1.0F;
}

}  // namespace phx
