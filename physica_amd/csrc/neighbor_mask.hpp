// Per-particle neighbour masks over the 27-cell stencil.
//
// For particle i and stencil slot k (offset_x outer, offset_z inner), bit b
// of mask[k * N + i] marks the b-th particle (in sorted order) of that cell
// as a neighbour. Walking slots in order and bits in ascending order visits
// the marked candidates in exactly the reference's summation order (27 cells,
// ascending original index), so a sum over the marked candidates equals the
// full-stencil sum whenever every unmarked candidate contributes an exact
// zero. Cells with more than kMaskBits particles cannot be encoded: a particle
// with such a cell among its 27 is marked crowded by the pass that builds the
// step's masks, and every pass walks the full stencil for that particle alone.
#pragma once

#include <cstdint>

#include "device_math.hpp"

namespace phx {

constexpr int kStencil = 27;
constexpr uint32_t kMaskBits = 16;

// Device flags of a step: [0] masks unusable for every particle (boundary
// samples); [1 + k] largest squared displacement |x* - x|^2 (float bits)
// written by the k-th predicting pass.
constexpr uint32_t kFlagOverflow = 0;
constexpr uint32_t kFlagDisplacement = 1;

// Linear index of cell (x, y, z), x fastest.
__device__ __forceinline__ uint32_t cell_of(const SortedGrid& g, int x, int y, int z) {
    return (static_cast<uint32_t>(z) * g.cy + static_cast<uint32_t>(y)) * g.cx + static_cast<uint32_t>(x);
}

// Cell-index offset of each stencil slot k = (ox + 1) * 9 + (oy + 1) * 3 +
// (oz + 1): ox + oy Cx + oz Cx Cy (two's complement). Filled by the first 27
// threads of the block; callers synchronise before use.
__device__ __forceinline__ void fill_slot_deltas(const SortedGrid& g, int* delta) {
    const uint32_t k = threadIdx.x;
    if (k < static_cast<uint32_t>(kStencil)) {
        const int ox = static_cast<int>(k) / 9 - 1;
        const int oy = (static_cast<int>(k) / 3) % 3 - 1;
        const int oz = static_cast<int>(k) % 3 - 1;
        delta[k] = ox + oy * static_cast<int>(g.cx) + oz * static_cast<int>(g.cx) * static_cast<int>(g.cy);
    }
}

// Copies the 27 masks of particle i into its LDS column and returns the
// bitmap of slots with a non-empty mask.
__device__ __forceinline__ uint32_t load_lane_column(const uint16_t* __restrict__ masks, uint32_t n, uint32_t i,
                                                     uint16_t* column, uint32_t stride) {
    uint32_t nonzero = 0u;
#pragma unroll
    for (int k = 0; k < kStencil; ++k) {
        const uint16_t m = masks[static_cast<size_t>(k) * n + i];
        column[static_cast<uint32_t>(k) * stride] = m;
        nonzero |= m != 0u ? 1u << k : 0u;
    }
    return nonzero;
}

// Marked candidates of one lane in the reference order, one candidate per
// loop trip; `nonzero` lists the slots to visit, `center` is the lane's cell.
// visit() should not branch on its candidate: a wave then runs one trip per
// candidate of its busiest lane with no other divergence.
template <class Visit>
__device__ __forceinline__ void walk_masked(const uint32_t* __restrict__ offsets, uint32_t center, const int* delta,
                                            const uint16_t* column, uint32_t stride, uint32_t nonzero, Visit&& visit) {
    uint32_t m = 0u;
    uint32_t base = 0u;
    while ((m | nonzero) != 0u) {
        if (m == 0u) {
            const uint32_t k = static_cast<uint32_t>(__builtin_ctz(nonzero));
            nonzero &= nonzero - 1u;
            m = column[k * stride];
            base = offsets[center + static_cast<uint32_t>(delta[k])];
        }
        const uint32_t b = static_cast<uint32_t>(__builtin_ctz(m));
        m &= m - 1u;
        visit(base + b);
    }
}

// walk_masked with the memory latency of each step taken off the critical
// path: the base index of the next non-empty slot (slot_base(k): the sorted
// index of the first particle of stencil cell k, or its place in a staged
// copy) is fetched when the current slot starts, and the data of the next
// kDepth candidates (load(s), which returns a value) is fetched before the
// current one is processed (process(value)), in the same order as
// walk_masked visits them. load() must be safe to call with any index the
// bases give (lanes without a next candidate load `fallback`, a valid index,
// and discard it).
template <int kDepth, class SlotBase, class Load, class Process>
__device__ __forceinline__ void walk_slots_prefetch(SlotBase&& slot_base, const uint16_t* column, uint32_t stride,
                                                    uint32_t nonzero, uint32_t fallback, Load&& load,
                                                    Process&& process) {
    auto fetch = [&](uint32_t& mask, uint32_t& base) {
        mask = 0u;
        if (nonzero != 0u) {
            const uint32_t k = static_cast<uint32_t>(__builtin_ctz(nonzero));
            nonzero &= nonzero - 1u;
            mask = column[k * stride];
            base = slot_base(k);
        }
    };
    uint32_t m = 0u, base = 0u, m_next = 0u, base_next = 0u;
    fetch(m, base);
    fetch(m_next, base_next);
    // the next candidate in the reference order, or fallback when none is left
    auto next = [&](bool& have) -> uint32_t {
        if (m == 0u) {
            m = m_next;
            base = base_next;
            fetch(m_next, base_next);
        }
        have = m != 0u;
        const uint32_t s = base + static_cast<uint32_t>(__builtin_ctz(m));
        m &= m - 1u;
        return have ? s : fallback;
    };
    using Data = decltype(load(0u));
    Data data[kDepth];
    bool have[kDepth];
#pragma unroll
    for (int d = 0; d < kDepth; ++d) data[d] = load(next(have[d]));
    // candidates arrive in order, so a lane is done at its first empty slot;
    // the ring is unrolled, so every data[d] stays in registers
    for (;;) {
#pragma unroll
        for (int d = 0; d < kDepth; ++d) {
            if (!have[d]) return;
            const Data current = data[d];
            data[d] = load(next(have[d]));
            // keeps the loads issued ahead of this candidate's arithmetic
            // (the compiler would otherwise sink them to their use)
            asm volatile("" ::: "memory");
            process(current);
        }
    }
}

// walk_slots_prefetch over the grid's cell offsets.
template <int kDepth = 1, class Load, class Process>
__device__ __forceinline__ void walk_masked_prefetch(const uint32_t* __restrict__ offsets, uint32_t center,
                                                     const int* delta, const uint16_t* column, uint32_t stride,
                                                     uint32_t nonzero, uint32_t fallback, Load&& load,
                                                     Process&& process) {
    walk_slots_prefetch<kDepth>([&](uint32_t k) { return offsets[center + static_cast<uint32_t>(delta[k])]; },
                                column, stride, nonzero, fallback, load, process);
}

// Clears particle i's own bit from its LDS column (slot 13, its own cell,
// whose first sorted particle is cell_first) for the sums that skip j == i;
// returns the bitmap of non-empty slots without slot 13 if that empties it.
__device__ __forceinline__ uint32_t drop_self(uint32_t i, uint32_t cell_first, uint16_t* column, uint32_t stride,
                                              uint32_t nonzero) {
    constexpr uint32_t kCenter = 13u;
    const uint32_t bit = i - cell_first;
    const uint32_t m = column[kCenter * stride] & ~(bit < kMaskBits ? 1u << bit : 0u);
    column[kCenter * stride] = static_cast<uint16_t>(m);
    return m == 0u ? nonzero & ~(1u << kCenter) : nonzero;
}

// Workgroup index with consecutive ranges of indices on one XCD (MI355X
// dispatches workgroup b to XCD b % 8). gridDim.x must be a multiple of 8.
__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t per_xcd = gridDim.x / 8u;
    return (blockIdx.x % 8u) * per_xcd + blockIdx.x / 8u;
}

// Builds both masks of particle i at position p: `exact` marks d2 < t_exact,
// `skin` marks d2 < t_skin, with d2 = dot(p - x_s, p - x_s) evaluated as the
// kernels do. The exact mask is also copied to `column` and its non-empty
// slots are added to the bitmap `nonzero`, if given. Returns false if a
// stencil cell holds more than kMaskBits particles (the masks of that cell are
// then truncated and must not be used).
__device__ __forceinline__ bool build_masks(const SortedGrid& g, const float4* __restrict__ xm, uint32_t n, uint32_t i,
                                            Float3 p, int cx, int cy, int cz, float t_exact, float t_skin,
                                            uint16_t* __restrict__ exact, uint16_t* __restrict__ skin,
                                            uint16_t* column = nullptr, uint32_t stride = 0u,
                                            uint32_t* nonzero = nullptr, const float* cell_bound = nullptr,
                                            float h = 0.0F) {
    bool complete = true;
    int k = 0;
    const float bound_i = cell_bound != nullptr ? cell_bound[cell_of(g, cx, cy, cz)] : 0.0F;
    // not unrolled: the unrolled 27 cells take 95 VGPRs (5 waves per SIMD)
    // and the pass 20 % longer than the loops at 39 VGPRs
#pragma unroll 1
    for (int ox = -1; ox <= 1; ++ox) {
        const int X = cx + ox;
#pragma unroll 1
        for (int oy = -1; oy <= 1; ++oy) {
            const int Y = cy + oy;
#pragma unroll 1
            for (int oz = -1; oz <= 1; ++oz, ++k) {
                const int Z = cz + oz;
                uint32_t me = 0u, ms = 0u;
                if (X >= 0 && X < static_cast<int>(g.cx) && Y >= 0 && Y < static_cast<int>(g.cy) && Z >= 0 &&
                    Z < static_cast<int>(g.cz)) {
                    const uint32_t c = static_cast<uint32_t>(Z) * g.cy * g.cx + static_cast<uint32_t>(Y) * g.cx +
                                       static_cast<uint32_t>(X);
                    const uint32_t first = g.offsets[c];
                    uint32_t last = g.offsets[c + 1u];
                    if (last - first > kMaskBits) {
                        complete = false;
                        last = first + kMaskBits;
                    }
                    // with per-cell bounds: a pair within h after both moved
                    // by at most their cells' bounds was within their sum
                    // plus h before (the relative margin covers rounding)
                    float t_cell = t_skin;
                    if (cell_bound != nullptr) {
                        const float radius = h + bound_i + cell_bound[c];
                        t_cell = radius * radius * (1.0F + 2e-5F);
                    }
                    for (uint32_t s = first; s < last; ++s) {
                        const float4 q = xm[s];
                        const Float3 d = sub(p, Float3{q.x, q.y, q.z});
                        const float d2 = dot(d, d);
                        const uint32_t bit = 1u << (s - first);
                        me |= d2 < t_exact ? bit : 0u;
                        ms |= d2 < t_cell ? bit : 0u;
                    }
                }
                exact[static_cast<size_t>(k) * n + i] = static_cast<uint16_t>(me);
                skin[static_cast<size_t>(k) * n + i] = static_cast<uint16_t>(ms);
                if (column != nullptr) column[static_cast<uint32_t>(k) * stride] = static_cast<uint16_t>(me);
                if (nonzero != nullptr && me != 0u) *nonzero |= 1u << k;
            }
        }
    }
    return complete;
}

// A step's skin masks and the flags that decide, pass by pass, whether they
// apply: a pass at positions that moved at most sqrt(flags[kFlagDisplacement
// + slot]) from the positions the masks were built at may use them while
// that is within limit2 and no cell overflowed.
struct MaskView {
    const uint16_t* masks;  // [kStencil][n]; null: always scan the full stencil
    const uint32_t* flags;
    uint32_t n;
    float limit2;
    // exact masks of one set of positions: they apply unless flags[kFlagOverflow]
    // is raised; no displacement slot is read
    uint32_t exact = 0u;
    // per particle (sorted order): non-zero where a stencil cell holds more
    // than kMaskBits particles; null: no particle is crowded
    const uint8_t* crowded = nullptr;
};

// Whether particle i walks its masks in a pass reading displacement slot `slot`.
__device__ __forceinline__ bool masks_apply(const MaskView& m, uint32_t slot, uint32_t i) {
    if (m.masks == nullptr || m.flags[kFlagOverflow] != 0u) return false;
    if (m.crowded != nullptr && m.crowded[i] != 0u) return false;
    return m.exact != 0u || __uint_as_float(m.flags[kFlagDisplacement + slot]) <= m.limit2;
}

// Visits the candidates of particle i in the reference order: the marked
// ones of its masks (copied to its LDS column) or every particle of the 27
// cells around (cx, cy, cz).
template <class Visit>
__device__ __forceinline__ void visit_candidates(const SortedGrid& g, const MaskView& m, bool masked, uint32_t i, int cx,
                                                 int cy, int cz, const int* delta, uint16_t* column, uint32_t stride,
                                                 Visit&& visit) {
    if (masked) {
        const uint32_t nonzero = load_lane_column(m.masks, m.n, i, column, stride);
        walk_masked(g.offsets, cell_of(g, cx, cy, cz), delta, column, stride, nonzero, visit);
    } else {
        for_each_candidate(g, cx, cy, cz, visit);
    }
}

// Visits the candidates of particle i like visit_candidates, where visit(s)
// returns whether candidate s is kept, and writes the kept candidates of each
// stencil slot k as the mask out[k * n + i] (every slot is written). Returns
// false if a kept candidate lies beyond the first kMaskBits particles of its
// cell: the masks are then incomplete and must not be used.
template <class Visit>
__device__ __forceinline__ bool visit_and_mask(const SortedGrid& g, const MaskView& m, bool masked, uint32_t i,
                                               int cx, int cy, int cz, const int* delta, uint16_t* column,
                                               uint32_t stride, uint16_t* __restrict__ out, uint32_t n,
                                               Visit&& visit) {
    uint32_t written = 0u;
    auto emit = [&](uint32_t k, uint32_t bits) {
        for (; written < k; ++written) out[static_cast<size_t>(written) * n + i] = 0u;
        out[static_cast<size_t>(k) * n + i] = static_cast<uint16_t>(bits);
        written = k + 1u;
    };
    bool complete = true;
    if (masked) {
        uint32_t nonzero = load_lane_column(m.masks, m.n, i, column, stride);
        const uint32_t center = cell_of(g, cx, cy, cz);
        while (nonzero != 0u) {
            const uint32_t k = static_cast<uint32_t>(__builtin_ctz(nonzero));
            nonzero &= nonzero - 1u;
            uint32_t candidates = column[k * stride];
            const uint32_t base = g.offsets[center + static_cast<uint32_t>(delta[k])];
            uint32_t bits = 0u;
            while (candidates != 0u) {
                const uint32_t b = static_cast<uint32_t>(__builtin_ctz(candidates));
                candidates &= candidates - 1u;
                if (visit(base + b)) bits |= 1u << b;
            }
            emit(k, bits);
        }
    } else {
        uint32_t k = 0u;
#pragma unroll 1
        for (int ox = -1; ox <= 1; ++ox) {
            const int X = cx + ox;
#pragma unroll 1
            for (int oy = -1; oy <= 1; ++oy) {
                const int Y = cy + oy;
#pragma unroll 1
                for (int oz = -1; oz <= 1; ++oz, ++k) {
                    const int Z = cz + oz;
                    if (X < 0 || X >= static_cast<int>(g.cx) || Y < 0 || Y >= static_cast<int>(g.cy) || Z < 0 ||
                        Z >= static_cast<int>(g.cz))
                        continue;
                    const uint32_t c = cell_of(g, X, Y, Z);
                    const uint32_t first = g.offsets[c];
                    const uint32_t last = g.offsets[c + 1u];
                    uint32_t bits = 0u;
                    for (uint32_t s = first; s < last; ++s) {
                        if (visit(s)) {
                            if (s - first < kMaskBits) {
                                bits |= 1u << (s - first);
                            } else {
                                complete = false;
                            }
                        }
                    }
                    emit(k, bits);
                }
            }
        }
    }
    for (; written < static_cast<uint32_t>(kStencil); ++written) out[static_cast<size_t>(written) * n + i] = 0u;
    return complete;
}

// Skin mask of particle i at p from SoA positions: bit b of slot k marks the
// b-th particle of that stencil cell with squared distance below t. Returns
// false if a stencil cell holds more than kMaskBits particles.
__device__ __forceinline__ bool build_skin_mask(const SortedGrid& g, CVec pos, uint32_t n, uint32_t i, Float3 p,
                                                int cx, int cy, int cz, float t, uint16_t* __restrict__ masks) {
    bool complete = true;
    int k = 0;
#pragma unroll 1
    for (int ox = -1; ox <= 1; ++ox) {
        const int X = cx + ox;
#pragma unroll 1
        for (int oy = -1; oy <= 1; ++oy) {
            const int Y = cy + oy;
#pragma unroll 1
            for (int oz = -1; oz <= 1; ++oz, ++k) {
                const int Z = cz + oz;
                uint32_t bits = 0u;
                if (X >= 0 && X < static_cast<int>(g.cx) && Y >= 0 && Y < static_cast<int>(g.cy) && Z >= 0 &&
                    Z < static_cast<int>(g.cz)) {
                    const uint32_t c = cell_of(g, X, Y, Z);
                    const uint32_t first = g.offsets[c];
                    uint32_t last = g.offsets[c + 1u];
                    if (last - first > kMaskBits) {
                        complete = false;
                        last = first + kMaskBits;
                    }
                    for (uint32_t s = first; s < last; ++s) {
                        const Float3 d = sub(p, load(pos, s));
                        bits |= dot(d, d) < t ? 1u << (s - first) : 0u;
                    }
                }
                masks[static_cast<size_t>(k) * n + i] = static_cast<uint16_t>(bits);
            }
        }
    }
    return complete;
}

// Block-wide maximum of a non-negative float, folded into *target with one
// atomic per block (non-negative IEEE floats order like their bit patterns).
template <uint32_t kBlockSize>
__device__ __forceinline__ void block_max_to(float value, uint32_t* target) {
    __shared__ float wave_max[kBlockSize / 64];
    float v = value;
    for (int offset = 32; offset > 0; offset >>= 1) v = fmaxf(v, __shfl_xor(v, offset, 64));
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = threadIdx.x >> 6;
    if (lane == 0u) wave_max[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0u) {
        float m = wave_max[0];
        for (uint32_t w = 1; w < kBlockSize / 64; ++w) m = fmaxf(m, wave_max[w]);
        atomicMax(target, __float_as_uint(m));
    }
}

}  // namespace phx
