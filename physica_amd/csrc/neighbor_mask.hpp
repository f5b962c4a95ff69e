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

// walk_slots_prefetch that hands each candidate's place to process: its
// sorted index s, stencil slot k and bit b within that cell's mask, as
// process(value, s, k, b).
template <int kDepth, class SlotBase, class Load, class Process>
__device__ __forceinline__ void walk_slots_prefetch_tagged(SlotBase&& slot_base, const uint16_t* column,
                                                           uint32_t stride, uint32_t nonzero, uint32_t fallback,
                                                           Load&& load, Process&& process) {
    auto fetch = [&](uint32_t& mask, uint32_t& base, uint32_t& slot) {
        mask = 0u;
        if (nonzero != 0u) {
            slot = static_cast<uint32_t>(__builtin_ctz(nonzero));
            nonzero &= nonzero - 1u;
            mask = column[slot * stride];
            base = slot_base(slot);
        }
    };
    uint32_t m = 0u, base = 0u, slot = 0u, m_next = 0u, base_next = 0u, slot_next = 0u;
    fetch(m, base, slot);
    fetch(m_next, base_next, slot_next);
    auto next = [&](bool& have, uint32_t& s, uint32_t& k, uint32_t& b) -> uint32_t {
        if (m == 0u) {
            m = m_next;
            base = base_next;
            slot = slot_next;
            fetch(m_next, base_next, slot_next);
        }
        have = m != 0u;
        b = static_cast<uint32_t>(__builtin_ctz(m));
        s = base + b;
        k = slot;
        m &= m - 1u;
        return have ? s : fallback;
    };
    using Data = decltype(load(0u));
    Data data[kDepth];
    bool have[kDepth];
    uint32_t ss[kDepth], ks[kDepth], bs[kDepth];
#pragma unroll
    for (int d = 0; d < kDepth; ++d) data[d] = load(next(have[d], ss[d], ks[d], bs[d]));
    for (;;) {
#pragma unroll
        for (int d = 0; d < kDepth; ++d) {
            if (!have[d]) return;
            const Data current = data[d];
            const uint32_t s = ss[d], k = ks[d], b = bs[d];
            data[d] = load(next(have[d], ss[d], ks[d], bs[d]));
            asm volatile("" ::: "memory");
            process(current, s, k, b);
        }
    }
}

// Every candidate of the 27 cells around (cx, cy, cz) in the reference order
// (for_each_candidate) with the memory latency taken off the critical path as
// in walk_slots_prefetch: the range of the next in-grid stencil cell is
// fetched when a cell starts, and the records of the next kDepth candidates
// before the current one is processed. process(value, s, k, first) gets the
// candidate's sorted index s, stencil slot k and its cell's first index.
// load() must accept `fallback`. It pays where a lane walks long lists with
// few waves to hide its loads (crowded cells in a scene that does not fill
// the GPU); with the GPU full, its bookkeeping costs more than it hides.
template <int kDepth, class Load, class Process>
__device__ __forceinline__ void walk_stencil_prefetch(const SortedGrid& g, int cx, int cy, int cz, uint32_t fallback,
                                                      Load&& load, Process&& process) {
    struct Range {
        uint32_t s, last, first, k;
    };
    uint32_t slot = 0u;  // next stencil slot to fetch
    // the next in-grid stencil cell (empty or not); slot 27 once none is left
    auto fetch = [&](Range& r) {
        r = Range{0u, 0u, 0u, static_cast<uint32_t>(kStencil)};
        while (slot < static_cast<uint32_t>(kStencil)) {
            const uint32_t k = slot++;
            const int X = cx + static_cast<int>(k / 9u) - 1;
            const int Y = cy + static_cast<int>((k / 3u) % 3u) - 1;
            const int Z = cz + static_cast<int>(k % 3u) - 1;
            if (X < 0 || X >= static_cast<int>(g.cx) || Y < 0 || Y >= static_cast<int>(g.cy) || Z < 0 ||
                Z >= static_cast<int>(g.cz))
                continue;
            const uint32_t c = cell_of(g, X, Y, Z);
            r.first = g.offsets[c];
            r.s = r.first;
            r.last = g.offsets[c + 1u];
            r.k = k;
            break;
        }
    };
    Range cur, ahead;
    fetch(cur);
    fetch(ahead);
    auto next = [&](bool& have, uint32_t& s, uint32_t& k, uint32_t& first) -> uint32_t {
        while (cur.s == cur.last && cur.k < static_cast<uint32_t>(kStencil)) {
            cur = ahead;
            fetch(ahead);
        }
        have = cur.s != cur.last;
        s = cur.s;
        k = cur.k;
        first = cur.first;
        cur.s += have ? 1u : 0u;
        return have ? s : fallback;
    };
    using Data = decltype(load(0u));
    Data data[kDepth];
    bool have[kDepth];
    uint32_t ss[kDepth], ks[kDepth], fs[kDepth];
#pragma unroll
    for (int d = 0; d < kDepth; ++d) data[d] = load(next(have[d], ss[d], ks[d], fs[d]));
    for (;;) {
#pragma unroll
        for (int d = 0; d < kDepth; ++d) {
            if (!have[d]) return;
            const Data current = data[d];
            const uint32_t s = ss[d], k = ks[d], first = fs[d];
            data[d] = load(next(have[d], ss[d], ks[d], fs[d]));
            asm volatile("" ::: "memory");
            process(current, s, k, first);
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

// ------------------------------------------------ crowded cells, wave-shared
// A particle next to a crowded cell walks every candidate of its 27 cells.
// Inside a crowded cell every particle has the same 27 cells, so a wave whose
// crowded lanes all sit in one cell fetches each run of 64 candidates with
// one coalesced load per lane and hands candidate k to every lane from scalar
// registers (v_readlane): the walk has no memory access per candidate. Each
// lane still adds its own terms in the reference order, so its sums are
// unchanged bit for bit.

// The value `value` holds on lane `lane`, on every lane (whole dwords).
template <class T>
__device__ __forceinline__ T broadcast_from(const T& value, uint32_t lane) {
    static_assert(sizeof(T) % 4 == 0, "records are whole dwords");
    constexpr int kWords = static_cast<int>(sizeof(T) / 4);
    uint32_t words[kWords];
    __builtin_memcpy(words, &value, sizeof(T));
#pragma unroll
    for (int w = 0; w < kWords; ++w)
        words[w] = static_cast<uint32_t>(
            __builtin_amdgcn_readlane(static_cast<int>(words[w]), static_cast<int>(lane)));
    T out;
    __builtin_memcpy(&out, words, sizeof(T));
    return out;
}

// Whether the wave walks its crowded particles' stencil together: every lane
// of the wave active, at least one crowded lane, and every crowded lane in
// one cell. `leader` is then the first crowded lane. The result is uniform
// over the wave.
__device__ __forceinline__ bool shared_crowded_cell(bool crowded, uint32_t cell, uint32_t& leader) {
    const uint64_t active = __ballot(1);
    const uint64_t lanes = __ballot(crowded);
    if (active != ~0ull || lanes == 0ull) return false;
    leader = static_cast<uint32_t>(__builtin_ctzll(lanes));
    const uint32_t shared = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(cell), leader));
    return __ballot(!crowded || cell == shared) == ~0ull;
}

// The candidates of the 27 cells around the leader lane's cell (cx, cy, cz)
// in the reference order, 64 at a time: lane l loads candidate base + l
// (load(s) returns its record), then process(record, s) runs for each
// candidate in turn on the lanes with `participate`. Every lane of the wave
// must call it together, in uniform control flow.
template <class Load, class Process>
__device__ __forceinline__ void walk_stencil_shared(const SortedGrid& g, uint32_t leader, int cx, int cy, int cz,
                                                    bool participate, Load&& load, Process&& process) {
    const int x = __builtin_amdgcn_readlane(cx, static_cast<int>(leader));
    const int y = __builtin_amdgcn_readlane(cy, static_cast<int>(leader));
    const int z = __builtin_amdgcn_readlane(cz, static_cast<int>(leader));
    const uint32_t lane = __lane_id();
    for_each_cell(g, x, y, z, [&](uint32_t c) {
        const uint32_t first = g.offsets[c];
        const uint32_t last = g.offsets[c + 1u];
        for (uint32_t base = first; base < last; base += 64u) {
            const uint32_t count = min(64u, last - base);
            const auto mine = load(base + min(lane, count - 1u));
            for (uint32_t k = 0; k < count; ++k) {
                const auto record = broadcast_from(mine, k);
                if (participate) process(record, base + k);
            }
        }
    });
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
    // per cell: the largest displacement (float bits, rounded up) of its
    // particles from the positions the masks were built at, over the position
    // sets written so far in the step; null: only the step-wide limit applies.
    // Where that limit fails, a particle whose own displacement plus the
    // largest of its stencil cells' stays within pair_limit keeps its masks.
    const uint32_t* cell_moved = nullptr;
    float pair_limit = -1.0F;
};

// The displacement |moved| rounded up for the per-cell bounds: the relative
// margin covers the rounding of the difference, the dot product and the
// square root.
__device__ __forceinline__ float displacement_bound(Float3 moved) {
    return sqrtf(dot(moved, moved)) * (1.0F + 4e-6F);
}

// Raises cell_moved[cell] to at least `moved` for each lane's cell. The lanes
// of one cell are consecutive in sorted order, so a segmented maximum over the
// wave leaves one atomic per run of equal cells. Every lane of the wave calls
// this; a lane without a particle passes cell = ~0u.
__device__ __forceinline__ void raise_cell_bound(uint32_t* cell_moved, uint32_t cell, float moved) {
    const int lane = static_cast<int>(__lane_id());
    float v = moved;
#pragma unroll
    for (int offset = 1; offset < 64; offset <<= 1) {
        const float up = __shfl_up(v, offset);
        const uint32_t up_cell = __shfl_up(cell, offset);
        if (lane >= offset && up_cell == cell) v = fmaxf(v, up);
    }
    const uint32_t next_cell = __shfl_down(cell, 1);
    if (cell != ~0u && (lane == 63 || next_cell != cell)) atomicMax(cell_moved + cell, __float_as_uint(v));
}

// Whether particle i is crowded in the step of these masks (it then walks the
// full stencil in every pass).
__device__ __forceinline__ bool is_crowded(const MaskView& m, uint32_t i) {
    return m.crowded != nullptr && m.crowded[i] != 0u;
}

// Candidate records of the wave-shared walks (whole dwords).
struct PointMass {
    Float3 x;
    float m;
};
struct PointMassLambda {
    Float3 x;
    float m;
    float lambda;
};
struct PointVector {
    Float3 x;
    Float3 v;
};

// Whether particle i walks its masks in a pass reading displacement slot `slot`.
__device__ __forceinline__ bool masks_apply(const MaskView& m, uint32_t slot, uint32_t i) {
    if (m.masks == nullptr || m.flags[kFlagOverflow] != 0u) return false;
    if (m.crowded != nullptr && m.crowded[i] != 0u) return false;
    return m.exact != 0u || __uint_as_float(m.flags[kFlagDisplacement + slot]) <= m.limit2;
}

// masks_apply, and where the step-wide displacement limit fails, the
// per-cell rule: particle i may keep its masks if own() (its displacement
// bound) plus the largest cell_moved of its stencil stays within pair_limit.
// A pair within h at the current positions was then within h + pair_limit at
// the positions the masks were built at. A wave whose lanes decided both ways
// would run the masked and the full walk one after the other, so a wave walks
// its masks only if every lane that is not crowded may. Every lane of the
// wave with a particle calls this.
template <class Own>
__device__ __forceinline__ bool masks_apply(const MaskView& m, uint32_t slot, uint32_t i, const SortedGrid& g, int cx,
                                            int cy, int cz, Own&& own) {
    if (m.masks == nullptr || m.flags[kFlagOverflow] != 0u) return false;
    const bool crowded = m.crowded != nullptr && m.crowded[i] != 0u;
    if (m.exact != 0u || __uint_as_float(m.flags[kFlagDisplacement + slot]) <= m.limit2) return !crowded;
    if (m.cell_moved == nullptr) return false;
    bool may = crowded;  // crowded lanes walk the full stencil, whatever the wave does
    if (!crowded) {
        float largest = 0.0F;
#pragma unroll 1
        for (int oz = -1; oz <= 1; ++oz) {
            const int Z = cz + oz;
            if (Z < 0 || Z >= static_cast<int>(g.cz)) continue;
#pragma unroll 1
            for (int oy = -1; oy <= 1; ++oy) {
                const int Y = cy + oy;
                if (Y < 0 || Y >= static_cast<int>(g.cy)) continue;
                for (int ox = -1; ox <= 1; ++ox) {
                    const int X = cx + ox;
                    if (X < 0 || X >= static_cast<int>(g.cx)) continue;
                    largest = fmaxf(largest, __uint_as_float(m.cell_moved[cell_of(g, X, Y, Z)]));
                }
            }
        }
        may = own() + largest <= m.pair_limit;
    }
    return !crowded && __all(may) != 0;
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
        // one candidate per loop trip, as in walk_masked (a loop per slot
        // would run, slot after slot, as many trips as the wave's fullest
        // lane in that slot); a lane writes a slot's mask as it moves on
        uint32_t nonzero = load_lane_column(m.masks, m.n, i, column, stride);
        const uint32_t center = cell_of(g, cx, cy, cz);
        constexpr uint32_t kNone = static_cast<uint32_t>(kStencil);
        uint32_t candidates = 0u, base = 0u, bits = 0u, k = kNone;
        while ((candidates | nonzero) != 0u) {
            if (candidates == 0u) {
                if (k != kNone) emit(k, bits);
                k = static_cast<uint32_t>(__builtin_ctz(nonzero));
                nonzero &= nonzero - 1u;
                candidates = column[k * stride];
                base = g.offsets[center + static_cast<uint32_t>(delta[k])];
                bits = 0u;
            }
            const uint32_t b = static_cast<uint32_t>(__builtin_ctz(candidates));
            candidates &= candidates - 1u;
            if (visit(base + b)) bits |= 1u << b;
        }
        if (k != kNone) emit(k, bits);
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

// A candidate's record and sorted index.
template <class Record>
struct Tagged {
    Record r;
    uint32_t s;
};

// Visits the candidates of particle i like visit_candidates, as
// pair(load(s), s): load(s) fetches candidate s's record. With kDepth > 0 the
// walks fetch kDepth records ahead of the one processed
// (walk_masked_prefetch, walk_stencil_prefetch); with 0 they are the plain
// walks, for launches that fill the GPU. A wave whose crowded particles all
// sit in one cell walks their stencil together (walk_stencil_shared); its
// other lanes walk their own candidates first. Every lane of the wave must
// call it together, in uniform control flow.
template <int kDepth, class Load, class Pair>
__device__ __forceinline__ void visit_records(const SortedGrid& g, const MaskView& m, bool masked, uint32_t i, int cx,
                                              int cy, int cz, const int* delta, uint16_t* column, uint32_t stride,
                                              Load&& load, Pair&& pair) {
    uint32_t leader = 0u;
    const bool crowded = is_crowded(m, i);
    const bool shared = shared_crowded_cell(crowded, cell_of(g, cx, cy, cz), leader);
    if (!shared || !crowded) {
        if constexpr (kDepth == 0) {
            visit_candidates(g, m, masked, i, cx, cy, cz, delta, column, stride, [&](uint32_t s) { pair(load(s), s); });
        } else if (masked) {
            using Record = decltype(load(0u));
            const uint32_t nonzero = load_lane_column(m.masks, m.n, i, column, stride);
            walk_masked_prefetch<kDepth>(
                g.offsets, cell_of(g, cx, cy, cz), delta, column, stride, nonzero, i,
                [&](uint32_t s) { return Tagged<Record>{load(s), s}; },
                [&](const Tagged<Record>& t) { pair(t.r, t.s); });
        } else {
            walk_stencil_prefetch<kDepth>(g, cx, cy, cz, i, load,
                                          [&](const auto& r, uint32_t s, uint32_t, uint32_t) { pair(r, s); });
        }
    }
    if (shared) walk_stencil_shared(g, leader, cx, cy, cz, crowded, load, pair);
}

// visit_records where keep(record, s) says whether candidate s is kept, and
// the kept candidates of each stencil slot k are written as the mask
// out[k * n + i] (every slot is written) like visit_and_mask. Crowded lanes
// of a wave-shared cell write no masks (a crowded particle's masks are never
// read). Returns false if a kept candidate lies beyond the first kMaskBits
// particles of its cell.
template <int kDepth, class Load, class Keep>
__device__ __forceinline__ bool visit_records_and_mask(const SortedGrid& g, const MaskView& m, bool masked,
                                                       uint32_t i, int cx, int cy, int cz, const int* delta,
                                                       uint16_t* column, uint32_t stride,
                                                       uint16_t* __restrict__ out, uint32_t n, Load&& load,
                                                       Keep&& keep) {
    uint32_t leader = 0u;
    const bool crowded = is_crowded(m, i);
    const bool shared = shared_crowded_cell(crowded, cell_of(g, cx, cy, cz), leader);
    bool complete = true;
    if (!shared || !crowded) {
        if constexpr (kDepth == 0) {
            complete = visit_and_mask(g, m, masked, i, cx, cy, cz, delta, column, stride, out, n,
                                      [&](uint32_t s) { return keep(load(s), s); });
        } else {
            constexpr uint32_t kNone = static_cast<uint32_t>(kStencil);
            uint32_t written = 0u, open = kNone, bits = 0u;
            // closes the open slot (zero masks before it) and opens slot k
            auto open_slot = [&](uint32_t k) {
                if (open != kNone) {
                    for (; written < open; ++written) out[static_cast<size_t>(written) * n + i] = 0u;
                    out[static_cast<size_t>(open) * n + i] = static_cast<uint16_t>(bits);
                    written = open + 1u;
                }
                open = k;
                bits = 0u;
            };
            auto take = [&](const auto& r, uint32_t s, uint32_t k, uint32_t b) {
                if (k != open) open_slot(k);
                if (keep(r, s)) {
                    if (b < kMaskBits) {
                        bits |= 1u << b;
                    } else {
                        complete = false;
                    }
                }
            };
            if (masked) {
                const uint32_t nonzero = load_lane_column(m.masks, m.n, i, column, stride);
                const uint32_t center = cell_of(g, cx, cy, cz);
                walk_slots_prefetch_tagged<kDepth>(
                    [&](uint32_t k) { return g.offsets[center + static_cast<uint32_t>(delta[k])]; }, column, stride,
                    nonzero, i, load, take);
            } else {
                walk_stencil_prefetch<kDepth>(
                    g, cx, cy, cz, i, load,
                    [&](const auto& r, uint32_t s, uint32_t k, uint32_t first) { take(r, s, k, s - first); });
            }
            open_slot(kNone);
            for (; written < static_cast<uint32_t>(kStencil); ++written)
                out[static_cast<size_t>(written) * n + i] = 0u;
        }
    }
    if (shared) walk_stencil_shared(g, leader, cx, cy, cz, crowded, load, [&](const auto& r, uint32_t s) {
        (void)keep(r, s);
    });
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
