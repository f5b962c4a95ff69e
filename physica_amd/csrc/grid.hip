// Uniform-grid neighbour search for gfx950: counting sort by cell key.
//
// Semantics follow NeighborhoodSearch::build (neighborhood/neighborhood.cpp:
// 44-89, neighborhood/neighborhood.cu:20-70): cell key per particle, stable
// ascending sort of (key, index), cell_offsets[c] = lower_bound(keys, c).
// The implementation differs: keys are < C, so a histogram + exclusive scan
// gives the offsets directly, and a stable order within a cell comes from
// ranking each particle by original index among its cell's members: a scan
// of the members for cells of at most kRankScan particles, and for larger
// cells (any distribution, up to every particle in one cell) a workgroup per
// cell that radix-sorts the cell's indices, so the build stays O(N).
#include <algorithm>
#include <cmath>
#include <vector>

#include "engine.hpp"

namespace phx {
namespace {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kWaves = kBlock / 64;
constexpr uint32_t kScanItems = 8;
// cells with more members than this are ordered by big_cell_sort_kernel
constexpr uint32_t kRankScan = 128;
constexpr uint32_t kBigCellGroups = 512;  // workgroups of big_cell_sort_kernel

// Most cells a build can queue for big_cell_sort_kernel (each holds more
// than kRankScan particles).
size_t big_cell_capacity(uint32_t particles) { return static_cast<size_t>(particles) / (kRankScan + 1u) + 1u; }
constexpr uint32_t kScanTile = kBlock * kScanItems;

__global__ void __launch_bounds__(kBlock) fill_u32_kernel(uint32_t* data, size_t count, uint32_t value) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) data[i] = value;
}

__global__ void __launch_bounds__(kBlock) fill_f32_kernel(float* data, size_t count, float value) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) data[i] = value;
}

struct KeyBox {
    float minimum[3];
    float velocity[3];
    float h;
    float time;
    uint32_t res[3];
};

// key = (cz Cy + cy) Cx + cx of the cell of p - (min + t v_box), each axis
// clamped into the grid (key_kernel, neighborhood.cu:20-39).
__device__ __forceinline__ uint32_t cell_key(const KeyBox& box, Float3 p) {
    const int cx = cell_clamp((p.x - box.minimum[0] - box.time * box.velocity[0]) / box.h, box.res[0]);
    const int cy = cell_clamp((p.y - box.minimum[1] - box.time * box.velocity[1]) / box.h, box.res[1]);
    const int cz = cell_clamp((p.z - box.minimum[2] - box.time * box.velocity[2]) / box.h, box.res[2]);
    return (static_cast<uint32_t>(cz) * box.res[1] + static_cast<uint32_t>(cy)) * box.res[0] +
           static_cast<uint32_t>(cx);
}

// boundary_position (neighborhood/device.cuh:16-22): x0 + t v per axis, the
// product rounded before the sum.
__device__ __forceinline__ Float3 moved_sample(const BoundarySource& src, float time, uint32_t i) {
    return Float3{src.x[0][i] + time * src.v[0][i], src.x[1][i] + time * src.v[1][i],
                  src.x[2][i] + time * src.v[2][i]};
}

// Adds one to counts[key] for each lane with a particle and returns the
// lane's slot: the cell's count before its own addition (unique within the
// cell, in no particular order; the in-cell rank restores the reference's).
// Consecutive lanes with equal keys, common where the particle order follows
// space, take one atomic per run. Every lane of the wave calls this; a lane
// without a particle passes key = ~0u, which no cell has.
__device__ __forceinline__ uint32_t count_in_cell(uint32_t* counts, uint32_t key) {
    const uint32_t lane = __lane_id();
    const uint32_t previous = __shfl_up(key, 1);
    const bool head = lane == 0u || previous != key;
    const uint64_t heads = __ballot(head);
    const uint64_t upto = lane == 63u ? ~0ull : (2ull << lane) - 1ull;
    const uint32_t first = 63u - static_cast<uint32_t>(__builtin_clzll(heads & upto));  // lane 0 is a head
    uint32_t base = 0u;
    if (head && key != ~0u) {
        const uint64_t above = lane == 63u ? 0ull : heads >> (lane + 1u);
        const uint32_t run = above != 0ull ? static_cast<uint32_t>(__builtin_ctzll(above)) + 1u : 64u - lane;
        base = atomicAdd(&counts[key], run);
    }
    return __shfl(base, static_cast<int>(first)) + (lane - first);
}

// key_kernel, neighborhood.cu:20-39, fused with the cell histogram.
template <bool kPredict>
__global__ void __launch_bounds__(kBlock)
    key_count_kernel(uint32_t n, KeyBox box, CVec positions, PredictInput predict, uint32_t* keys, uint32_t* slots,
                     uint32_t* counts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t key = ~0u;
    if (i < n) {
        Float3 p = load(positions, i);
        if constexpr (kPredict) {
            const Float3 pv =
                add(load(predict.velocities, i), scale(add(predict.gravity, load(predict.controls, i)), predict.dt));
            p = add(p, scale(pv, predict.dt));
            store(predict.predicted, i, p);
        }
        key = cell_key(box, p);
        keys[i] = key;
    }
    const uint32_t slot = count_in_cell(counts, key);
    if (i < n) slots[i] = slot;
}

// key_count_kernel<true> of a batch: workgroup b serves particles
// [first, first + count) of one scene, whose gravity predicts x* and whose
// grid (box, resolution, cell base) keys it.
__global__ void __launch_bounds__(kBlock)
    batch_key_count_kernel(BatchView batch, CVec positions, PredictInput predict, uint32_t* keys, uint32_t* slots,
                           uint32_t* counts) {
    const BatchBlock blk = batch.blocks[blockIdx.x];
    const bool valid = threadIdx.x < blk.count;
    const uint32_t i = blk.first + threadIdx.x;
    uint32_t key = ~0u;
    if (valid) {
        const BatchScene sc = batch.scenes[blk.scene];
        const Float3 gravity{sc.gravity[0], sc.gravity[1], sc.gravity[2]};
        const Float3 pv = add(load(predict.velocities, i), scale(add(gravity, load(predict.controls, i)), predict.dt));
        const Float3 p = add(load(positions, i), scale(pv, predict.dt));
        store(predict.predicted, i, p);
        KeyBox box{};
        for (int a = 0; a < 3; ++a) {
            box.minimum[a] = sc.minimum[a];
            box.velocity[a] = sc.velocity[a];
            box.res[a] = sc.res[a];
        }
        box.h = batch.h;
        box.time = batch.time;
        key = sc.cell_base + cell_key(box, p);
        keys[i] = key;
    }
    const uint32_t slot = count_in_cell(counts, key);
    if (valid) slots[i] = slot;
}

// key_kernel with moving_positions (neighborhood.cu:20-39 as called at
// :63-69): the sample is first moved to x0 + t v.
__global__ void __launch_bounds__(kBlock)
    boundary_key_count_kernel(uint32_t count, KeyBox box, BoundarySource src, uint32_t* keys, uint32_t* slots,
                              uint32_t* counts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t key = ~0u;
    if (i < count) {
        key = cell_key(box, moved_sample(src, box.time, i));
        keys[i] = key;
    }
    const uint32_t slot = count_in_cell(counts, key);
    if (i < count) slots[i] = slot;
}

// Stable rank of each sample within its cell (as rank_gather_kernel), then
// its sorted record: position at t with volume, and velocity.
__global__ void __launch_bounds__(kBlock)
    boundary_rank_kernel(uint32_t count, float time, const uint32_t* keys, const uint32_t* offsets,
                         const uint32_t* unstable, BoundarySource src, uint32_t* perm, float4* x, float4* v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t key = keys[i];
    const uint32_t first = offsets[key];
    const uint32_t last = offsets[key + 1u];
    uint32_t rank = 0;
    for (uint32_t s = first; s < last; ++s) rank += unstable[s] < i ? 1u : 0u;
    const uint32_t pos = first + rank;
    perm[pos] = i;
    if (x != nullptr) {
        const Float3 p = moved_sample(src, time, i);
        x[pos] = make_float4(p.x, p.y, p.z, src.volume[i]);
        v[pos] = make_float4(src.v[0][i], src.v[1][i], src.v[2][i], 0.0F);
    }
}

__global__ void __launch_bounds__(kBlock)
    scatter_unstable_kernel(uint32_t n, const uint32_t* keys, const uint32_t* slots, const uint32_t* offsets,
                            uint32_t* unstable) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unstable[offsets[keys[i]] + slots[i]] = i;
}

// Final position of particle i = cell start + number of cell members with a
// smaller original index; equal keys therefore keep input order, which is the
// stable-sort order of cub::DeviceRadixSort::SortPairs (neighborhood.cu:57).
// A cell with more than kRankScan members is queued once (by its member that
// arrived first) for big_cell_sort_kernel instead.
__global__ void __launch_bounds__(kBlock)
    rank_gather_kernel(uint32_t n, const uint32_t* keys, const uint32_t* slots, const uint32_t* offsets,
                       const uint32_t* unstable, uint32_t* perm, Channels channels, uint32_t* big_cells,
                       uint32_t* big_count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t key = keys[i];
    const uint32_t first = offsets[key];
    const uint32_t last = offsets[key + 1u];
    if (last - first > kRankScan) {
        if (slots[i] == 0u) big_cells[atomicAdd(big_count, 1u)] = key;
        return;
    }
    uint32_t rank = 0;
    for (uint32_t s = first; s < last; ++s) rank += unstable[s] < i ? 1u : 0u;
    const uint32_t pos = first + rank;
    perm[pos] = i;
    for (uint32_t c = 0; c < channels.count; ++c) channels.dst[c][pos * channels.stride[c]] = channels.src[c][i];
}

__global__ void __launch_bounds__(kBlock)
    export_kernel(uint32_t n, const uint32_t* perm, const uint32_t* keys, uint64_t* sorted_keys,
                  uint32_t* sorted_indices) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t i = perm[s];
    if (sorted_keys != nullptr) sorted_keys[s] = static_cast<uint64_t>(keys[i]);
    if (sorted_indices != nullptr) sorted_indices[s] = i;
}

__global__ void __launch_bounds__(kBlock) copy_u32_kernel(const uint32_t* src, uint32_t* dst, size_t count) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = src[i];
}

// Every channel is loaded before the first is stored: the stores could alias
// the sources as far as the compiler knows, so a loop that stores each
// channel after its load waits out each scattered load in turn.
__global__ void __launch_bounds__(kBlock)
    scatter_channels_kernel(uint32_t n, const uint32_t* perm, Channels channels) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t i = perm[s];
    float v[kMaxChannels];
#pragma unroll
    for (int c = 0; c < kMaxChannels; ++c)
        if (static_cast<uint32_t>(c) < channels.count) v[c] = channels.src[c][s * channels.stride[c]];
#pragma unroll
    for (int c = 0; c < kMaxChannels; ++c)
        if (static_cast<uint32_t>(c) < channels.count) channels.dst[c][i] = v[c];
}

__global__ void __launch_bounds__(kBlock)
    gather_channels_kernel(uint32_t n, const uint32_t* perm, Channels channels) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t i = perm[s];
    float v[kMaxChannels];
#pragma unroll
    for (int c = 0; c < kMaxChannels; ++c)
        if (static_cast<uint32_t>(c) < channels.count) v[c] = channels.src[c][i];
#pragma unroll
    for (int c = 0; c < kMaxChannels; ++c)
        if (static_cast<uint32_t>(c) < channels.count) channels.dst[c][s * channels.stride[c]] = v[c];
}

// ------------------------------------------------------------------- scan
// Inclusive scan of one value per lane across the block (wave64 shuffles,
// then one LDS pass over the wave totals). Returns the exclusive prefix and
// the block total.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t value, uint32_t& total) {
    __shared__ uint32_t wave_totals[kBlock / 64];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = threadIdx.x >> 6;
    uint32_t inclusive = value;
    for (uint32_t offset = 1; offset < 64; offset <<= 1) {
        const uint32_t up = __shfl_up(inclusive, offset, 64);
        if (lane >= offset) inclusive += up;
    }
    if (lane == 63u) wave_totals[wave] = inclusive;
    __syncthreads();
    uint32_t wave_prefix = 0;
    total = 0;
    for (uint32_t w = 0; w < kBlock / 64; ++w) {
        const uint32_t t = wave_totals[w];
        if (w < wave) wave_prefix += t;
        total += t;
    }
    __syncthreads();
    return wave_prefix + inclusive - value;
}

__global__ void __launch_bounds__(kBlock) scan_reduce_kernel(const uint32_t* in, size_t count, uint32_t* partials) {
    const size_t base = static_cast<size_t>(blockIdx.x) * kScanTile;
    uint32_t sum = 0;
    for (uint32_t k = 0; k < kScanItems; ++k) {
        const size_t idx = base + static_cast<size_t>(k) * kBlock + threadIdx.x;
        if (idx < count) sum += in[idx];
    }
    uint32_t total;
    (void)block_exclusive_scan(sum, total);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

// Exclusive scan of the per-tile sums in one block: thread t owns a
// contiguous chunk.
__global__ void __launch_bounds__(kBlock) scan_partials_kernel(uint32_t* partials, uint32_t count) {
    const uint32_t chunk = (count + kBlock - 1) / kBlock;
    const uint32_t first = threadIdx.x * chunk;
    uint32_t sum = 0;
    for (uint32_t k = 0; k < chunk; ++k)
        if (first + k < count) sum += partials[first + k];
    uint32_t total;
    uint32_t prefix = block_exclusive_scan(sum, total);
    for (uint32_t k = 0; k < chunk; ++k)
        if (first + k < count) {
            const uint32_t v = partials[first + k];
            partials[first + k] = prefix;
            prefix += v;
        }
}

__global__ void __launch_bounds__(kBlock)
    scan_downsweep_kernel(const uint32_t* in, uint32_t* out, size_t count, const uint32_t* partials) {
    const size_t base = static_cast<size_t>(blockIdx.x) * kScanTile + static_cast<size_t>(threadIdx.x) * kScanItems;
    uint32_t items[kScanItems];
    uint32_t sum = 0;
    for (uint32_t k = 0; k < kScanItems; ++k) {
        items[k] = base + k < count ? in[base + k] : 0u;
        sum += items[k];
    }
    uint32_t total;
    uint32_t prefix = block_exclusive_scan(sum, total) + partials[blockIdx.x];
    for (uint32_t k = 0; k < kScanItems; ++k) {
        if (base + k < count) out[base + k] = prefix;
        prefix += items[k];
    }
}

// Lanes of this wave whose `digit` equals the lane's own, among the lanes in
// `active` (eight ballots, one per digit bit).
__device__ __forceinline__ uint64_t digit_peers(uint32_t digit, uint64_t active) {
    uint64_t peers = active;
#pragma unroll
    for (uint32_t b = 0; b < 8u; ++b) {
        const bool bit = ((digit >> b) & 1u) != 0u;
        const uint64_t set = __ballot(bit);
        peers &= bit ? set : ~set;
    }
    return peers;
}

// The cells queued by rank_gather_kernel, one workgroup per cell at a time:
// the cell's member indices (unstable[first .. last), distinct) are sorted
// ascending by an LSD radix sort of 8-bit digits over the index bits, each
// pass a digit histogram then a stable scatter in rounds of one member per
// lane (a member's place among equal digits is its wave rank plus the counts
// of earlier waves and rounds). The passes alternate between perm and slots
// (free after the scatter into `unstable`) and end in perm; then the
// members' channels are gathered. Every workgroup loops until the queue is
// empty, so the grid drains whatever the queue holds.
__global__ void __launch_bounds__(kBlock)
    big_cell_sort_kernel(const uint32_t* offsets, const uint32_t* unstable, uint32_t* slots, uint32_t* perm,
                         const uint32_t* big_cells, const uint32_t* big_count, uint32_t index_bits,
                         Channels channels) {
    __shared__ uint32_t base[256];
    __shared__ uint32_t wave_counts[kWaves][256];
    const uint32_t t = threadIdx.x;
    const uint32_t lane = t & 63u;
    const uint32_t wave = t >> 6;
    for (uint32_t w = 0; w < kWaves; ++w) wave_counts[w][t] = 0u;
    const uint32_t cells = *big_count;
    const uint32_t passes = (index_bits + 7u) / 8u;
    for (uint32_t q = blockIdx.x; q < cells; q += gridDim.x) {
        const uint32_t c = big_cells[q];
        const uint32_t first = offsets[c];
        const uint32_t k = offsets[c + 1u] - first;
        const uint32_t* src = unstable + first;
        uint32_t* dst = passes % 2u == 1u ? perm + first : slots + first;
        for (uint32_t p = 0; p < passes; ++p) {
            const uint32_t shift = 8u * p;
            __syncthreads();
            base[t] = 0u;
            __syncthreads();
            for (uint32_t e = t; e < k; e += kBlock) atomicAdd(&base[(src[e] >> shift) & 255u], 1u);
            __syncthreads();
            uint32_t total;
            const uint32_t start = block_exclusive_scan(base[t], total);
            base[t] = start;
            __syncthreads();
            for (uint32_t r = 0; r < k; r += kBlock) {
                const uint32_t e = r + t;
                const bool valid = e < k;
                const uint32_t v = valid ? src[e] : 0u;
                const uint32_t d = (v >> shift) & 255u;
                const uint64_t peers = digit_peers(d, __ballot(valid));
                const uint32_t rank = static_cast<uint32_t>(__popcll(peers & ((1ull << lane) - 1ull)));
                if (valid && rank == 0u) wave_counts[wave][d] = static_cast<uint32_t>(__popcll(peers));
                __syncthreads();
                if (valid) {
                    uint32_t place = base[d] + rank;
                    for (uint32_t w = 0; w < wave; ++w) place += wave_counts[w][d];
                    dst[place] = v;
                }
                __syncthreads();
                uint32_t round = 0u;
                for (uint32_t w = 0; w < kWaves; ++w) {
                    round += wave_counts[w][t];
                    wave_counts[w][t] = 0u;
                }
                base[t] += round;
                __syncthreads();
            }
            // the next pass reads what this one wrote
            src = dst;
            dst = dst == perm + first ? slots + first : perm + first;
        }
        __syncthreads();
        for (uint32_t e = t; e < k; e += kBlock) {
            const uint32_t pos = first + e;
            const uint32_t i = perm[pos];
            for (uint32_t ch = 0; ch < channels.count; ++ch)
                channels.dst[ch][pos * channels.stride[ch]] = channels.src[ch][i];
        }
    }
}

}  // namespace

void fill_u32(hipStream_t stream, uint32_t* data, size_t count, uint32_t value) {
    if (count == 0) return;
    fill_u32_kernel<<<blocks_for(count, kBlock), kBlock, 0, stream>>>(data, count, value);
    PHX_LAUNCH_CHECK();
}

void fill_f32(hipStream_t stream, float* data, size_t count, float value) {
    if (count == 0) return;
    fill_f32_kernel<<<blocks_for(count, kBlock), kBlock, 0, stream>>>(data, count, value);
    PHX_LAUNCH_CHECK();
}

size_t exclusive_scan_partials(size_t count) { return (count + kScanTile - 1) / kScanTile; }

namespace {
// The stable in-cell order of a fluid build: rank_gather_kernel for small
// cells, then big_cell_sort_kernel for the cells it queued (counts[C + 1]),
// both writing the permutation only; then every input is gathered into cell
// order, one sorted position per thread, so the stores are coalesced and
// only the loads scatter.
void order_cells(hipStream_t stream, uint32_t n, uint32_t cells, const GridScratch& scratch, const uint32_t* offsets,
                 uint32_t* perm, const Channels& gather) {
    const Channels none{};
    uint32_t* big_count = scratch.counts + static_cast<size_t>(cells) + 1u;
    rank_gather_kernel<<<blocks_for(n, kBlock), kBlock, 0, stream>>>(n, scratch.keys, scratch.slots, offsets,
                                                                     scratch.unstable, perm, none,
                                                                     scratch.big_cells, big_count);
    PHX_LAUNCH_CHECK();
    const uint32_t index_bits = n > 1u ? 32u - static_cast<uint32_t>(__builtin_clz(n - 1u)) : 1u;
    const uint32_t groups = static_cast<uint32_t>(std::min<size_t>(kBigCellGroups, big_cell_capacity(n)));
    big_cell_sort_kernel<<<groups, kBlock, 0, stream>>>(offsets, scratch.unstable, scratch.slots, perm,
                                                        scratch.big_cells, big_count, index_bits, none);
    PHX_LAUNCH_CHECK();
    if (gather.count > 0u) {
        gather_channels_kernel<<<blocks_for(n, kBlock), kBlock, 0, stream>>>(n, perm, gather);
        PHX_LAUNCH_CHECK();
    }
}
}  // namespace

void exclusive_scan(hipStream_t stream, const uint32_t* in, uint32_t* out, size_t count, uint32_t* partials) {
    if (count == 0) return;
    const size_t tiles = exclusive_scan_partials(count);
    require(tiles <= static_cast<size_t>(kBlock) * 4096u, PHX_ERROR_CAPACITY, "grid has too many cells for the scan");
    scan_reduce_kernel<<<static_cast<uint32_t>(tiles), kBlock, 0, stream>>>(in, count, partials);
    PHX_LAUNCH_CHECK();
    scan_partials_kernel<<<1, kBlock, 0, stream>>>(partials, static_cast<uint32_t>(tiles));
    PHX_LAUNCH_CHECK();
    scan_downsweep_kernel<<<static_cast<uint32_t>(tiles), kBlock, 0, stream>>>(in, out, count, partials);
    PHX_LAUNCH_CHECK();
}

// ------------------------------------------------------------------- host
void grid_resolution(const phx_domain_configuration& c, uint32_t res[3]) {
    // neighborhood.cpp:25-30 (FP32 ceil of the extent over h)
    res[0] = static_cast<uint32_t>(std::ceil((c.boundary.maximum.x - c.boundary.minimum.x) / c.support_radius));
    res[1] = static_cast<uint32_t>(std::ceil((c.boundary.maximum.y - c.boundary.minimum.y) / c.support_radius));
    res[2] = static_cast<uint32_t>(std::ceil((c.boundary.maximum.z - c.boundary.minimum.z) / c.support_radius));
}

void require_vec(const phx_vector_view& v, const char* what) {
    require(v.x != nullptr && v.y != nullptr && v.z != nullptr, PHX_ERROR_INVALID_ARGUMENT,
            std::string("null device pointer in ") + what);
}
void require_vec(const phx_const_vector_view& v, const char* what) {
    require(v.x != nullptr && v.y != nullptr && v.z != nullptr, PHX_ERROR_INVALID_ARGUMENT,
            std::string("null device pointer in ") + what);
}

Domain make_domain(const phx_domain_configuration* config) {
    require(config != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null domain configuration");
    Domain d;
    d.config = *config;
    const auto& c = d.config;
    require(std::isfinite(c.support_radius) && c.support_radius > 0.0F, PHX_ERROR_INVALID_ARGUMENT,
            "support_radius must be positive and finite");
    require(std::isfinite(c.time_step) && c.time_step > 0.0F, PHX_ERROR_INVALID_ARGUMENT,
            "time_step must be positive and finite");
    require(c.boundary.maximum.x > c.boundary.minimum.x && c.boundary.maximum.y > c.boundary.minimum.y &&
                c.boundary.maximum.z > c.boundary.minimum.z,
            PHX_ERROR_INVALID_ARGUMENT, "boundary box must have positive extent");
    require(c.boundary_particle_count == 0u || c.boundary_particles != nullptr, PHX_ERROR_INVALID_ARGUMENT,
            "null boundary_particles with a non-zero boundary_particle_count");
    grid_resolution(c, d.resolution);
    const uint64_t cells = static_cast<uint64_t>(d.resolution[0]) * d.resolution[1] * d.resolution[2];
    require(d.resolution[0] > 0 && d.resolution[1] > 0 && d.resolution[2] > 0, PHX_ERROR_INVALID_ARGUMENT,
            "empty grid");
    require(cells < (1ull << 31), PHX_ERROR_CAPACITY, "grid has 2^31 cells or more");
    require(d.resolution[0] < (1u << 24) && d.resolution[1] < (1u << 24) && d.resolution[2] < (1u << 24),
            PHX_ERROR_CAPACITY, "grid axis resolution must be below 2^24");
    d.cells = static_cast<uint32_t>(cells);
    return d;
}

Box collision_box(const phx_domain_configuration& c, uint64_t step) {
    // sph.cpp:21-35, pbf.cpp:17-31
    const float time = static_cast<float>(step) * c.time_step;
    const float rp = c.particle_radius;
    Box b{};
    b.lo[0] = c.boundary.minimum.x + time * c.boundary.velocity.x + rp;
    b.lo[1] = c.boundary.minimum.y + time * c.boundary.velocity.y + rp;
    b.lo[2] = c.boundary.minimum.z + time * c.boundary.velocity.z + rp;
    b.hi[0] = c.boundary.maximum.x + time * c.boundary.velocity.x - rp;
    b.hi[1] = c.boundary.maximum.y + time * c.boundary.velocity.y - rp;
    b.hi[2] = c.boundary.maximum.z + time * c.boundary.velocity.z - rp;
    b.vel[0] = c.boundary.velocity.x;
    b.vel[1] = c.boundary.velocity.y;
    b.vel[2] = c.boundary.velocity.z;
    b.no_slip = c.boundary.no_slip != 0u ? 1u : 0u;
    return b;
}

float reference_gradient_norm(const phx_domain_configuration& c) {
    // compute_reference_gradient_norm, sph.cpp:37-58: FP32 sum of |grad W|^2
    // plus |sum grad W|^2 over a cubic lattice of spacing 2 r_p.
    const float h = c.support_radius;
    const float diameter = 2.0F * c.particle_radius;
    const float coefficient = 8.0F / (kPi * h * h * h);
    float gx = 0.0F, gy = 0.0F, gz = 0.0F, squared = 0.0F;
    require(diameter > 0.0F, PHX_ERROR_INVALID_ARGUMENT, "particle_radius must be positive");
    for (float x = -h; x <= h; x += diameter)
        for (float y = -h; y <= h; y += diameter)
            for (float z = -h; z <= h; z += diameter) {
                const float distance = std::sqrt(x * x + y * y + z * z);
                if (distance == 0.0F || distance >= h) continue;
                const float q = distance / h;
                const float derivative = q <= 0.5F ? 18.0F * q * q - 12.0F * q : -6.0F * (1.0F - q) * (1.0F - q);
                const float s = coefficient * derivative / (h * distance);
                gx += s * x;
                gy += s * y;
                gz += s * z;
                squared += s * s * (x * x + y * y + z * z);
            }
    return gx * gx + gy * gy + gz * gz + squared;
}

float spiky_coefficient(float h) { return -45.0F / (kPi * std::pow(h, 6.0F)); }

size_t grid_scratch_bytes(uint32_t particles, uint32_t cells) {
    auto round = [](size_t b) { return (b + 255u) & ~static_cast<size_t>(255u); };
    const size_t n = particles;
    const size_t c = static_cast<size_t>(cells) + 1u;
    return round(n * 4u) * 3u + round((c + 1u) * 4u) + round(exclusive_scan_partials(c) * 4u) +
           round(big_cell_capacity(particles) * 4u) + 256u;
}

GridScratch carve_grid_scratch(void* base, size_t bytes, uint32_t particles, uint32_t cells) {
    require(bytes >= grid_scratch_bytes(particles, cells), PHX_ERROR_WORKSPACE_TOO_SMALL, "grid workspace too small");
    auto round = [](size_t b) { return (b + 255u) & ~static_cast<size_t>(255u); };
    uintptr_t p = (reinterpret_cast<uintptr_t>(base) + 255u) & ~static_cast<uintptr_t>(255u);
    const size_t n = particles;
    const size_t c = static_cast<size_t>(cells) + 1u;
    GridScratch s{};
    s.keys = reinterpret_cast<uint32_t*>(p);
    p += round(n * 4u);
    s.slots = reinterpret_cast<uint32_t*>(p);
    p += round(n * 4u);
    s.unstable = reinterpret_cast<uint32_t*>(p);
    p += round(n * 4u);
    s.counts = reinterpret_cast<uint32_t*>(p);
    p += round((c + 1u) * 4u);
    s.partials = reinterpret_cast<uint32_t*>(p);
    p += round(exclusive_scan_partials(c) * 4u);
    s.big_cells = reinterpret_cast<uint32_t*>(p);
    return s;
}

namespace {
KeyBox key_box(const Domain& domain, float time) {
    const auto& c = domain.config;
    KeyBox box{};
    box.minimum[0] = c.boundary.minimum.x;
    box.minimum[1] = c.boundary.minimum.y;
    box.minimum[2] = c.boundary.minimum.z;
    box.velocity[0] = c.boundary.velocity.x;
    box.velocity[1] = c.boundary.velocity.y;
    box.velocity[2] = c.boundary.velocity.z;
    box.h = c.support_radius;
    box.time = time;
    for (int a = 0; a < 3; ++a) box.res[a] = domain.resolution[a];
    return box;
}
}  // namespace

std::vector<float> pack_boundary_samples(const phx_domain_configuration& config) {
    const size_t b = config.boundary_particle_count;
    std::vector<float> host(7u * b);
    if (b == 0u) return host;
    require(config.boundary_particles != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null boundary_particles");
    for (size_t i = 0; i < b; ++i) {
        const phx_boundary_particle& q = config.boundary_particles[i];
        const float values[7] = {q.position.x, q.position.y, q.position.z, q.velocity.x, q.velocity.y, q.velocity.z,
                                 q.volume};
        for (size_t k = 0; k < 7u; ++k) {
            require(std::isfinite(values[k]), PHX_ERROR_INVALID_ARGUMENT, "non-finite boundary particle");
            host[k * b + i] = values[k];
        }
    }
    return host;
}

BoundarySource boundary_source(const float* base, uint32_t count) {
    const size_t b = count;
    return BoundarySource{{base, base + b, base + 2u * b}, {base + 3u * b, base + 4u * b, base + 5u * b},
                          base + 6u * b};
}

void BoundarySamples::upload(const phx_domain_configuration& config) {
    const std::vector<float> host = pack_boundary_samples(config);
    count_ = config.boundary_particle_count;
    if (count_ == 0u) return;
    data_.resize(host.size());
    PHX_HIP(hipMemcpy(data_.data(), host.data(), sizeof(float) * host.size(), hipMemcpyHostToDevice));
}

BoundarySource BoundarySamples::source() const { return boundary_source(data_.data(), count_); }

size_t boundary_scratch_bytes(uint32_t samples) {
    const size_t round = (static_cast<size_t>(samples) * 4u + 255u) & ~static_cast<size_t>(255u);
    return round * 3u + 256u;
}

BoundaryScratch carve_boundary_scratch(void* base, size_t bytes, uint32_t samples) {
    require(bytes >= boundary_scratch_bytes(samples), PHX_ERROR_WORKSPACE_TOO_SMALL,
            "boundary grid workspace too small");
    const size_t round = (static_cast<size_t>(samples) * 4u + 255u) & ~static_cast<size_t>(255u);
    uintptr_t p = (reinterpret_cast<uintptr_t>(base) + 255u) & ~static_cast<uintptr_t>(255u);
    BoundaryScratch s{};
    s.keys = reinterpret_cast<uint32_t*>(p);
    s.slots = reinterpret_cast<uint32_t*>(p + round);
    s.unstable = reinterpret_cast<uint32_t*>(p + 2u * round);
    return s;
}

void build_boundary_grid(hipStream_t stream, const Domain& domain, uint64_t step, const BoundarySource& samples,
                         uint32_t count, const BoundaryScratch& scratch, const GridScratch& grid_scratch,
                         uint32_t* perm, uint32_t* offsets, float4* x, float4* v) {
    const uint32_t cells = domain.cells;
    if (count == 0u) {
        fill_u32(stream, offsets, static_cast<size_t>(cells) + 1u, 0u);
        return;
    }
    const float time = static_cast<float>(step) * domain.config.time_step;
    const KeyBox box = key_box(domain, time);
    fill_u32(stream, grid_scratch.counts, static_cast<size_t>(cells) + 1u, 0u);
    const uint32_t grid = blocks_for(count, kBlock);
    boundary_key_count_kernel<<<grid, kBlock, 0, stream>>>(count, box, samples, scratch.keys, scratch.slots,
                                                           grid_scratch.counts);
    PHX_LAUNCH_CHECK();
    exclusive_scan(stream, grid_scratch.counts, offsets, static_cast<size_t>(cells) + 1u, grid_scratch.partials);
    scatter_unstable_kernel<<<grid, kBlock, 0, stream>>>(count, scratch.keys, scratch.slots, offsets,
                                                         scratch.unstable);
    PHX_LAUNCH_CHECK();
    boundary_rank_kernel<<<grid, kBlock, 0, stream>>>(count, time, scratch.keys, offsets, scratch.unstable, samples,
                                                      perm, x, v);
    PHX_LAUNCH_CHECK();
}

GridGeometry build_grid(hipStream_t stream, const Domain& domain, uint64_t step, CVec positions,
                        const PredictInput* predict, const Channels& gather, const GridScratch& scratch, uint32_t* perm,
                        uint32_t* offsets) {
    const auto& c = domain.config;
    const uint32_t n = c.particle_count;
    const uint32_t cells = domain.cells;
    const float time = static_cast<float>(step) * c.time_step;
    GridGeometry geo;
    for (int a = 0; a < 3; ++a) geo.resolution[a] = domain.resolution[a];
    // neighborhood.cpp:83-88: lookup origin min + t v_box, cell size h
    geo.origin[0] = c.boundary.minimum.x + time * c.boundary.velocity.x;
    geo.origin[1] = c.boundary.minimum.y + time * c.boundary.velocity.y;
    geo.origin[2] = c.boundary.minimum.z + time * c.boundary.velocity.z;
    geo.cell_size = c.support_radius;
    geo.time = time;

    fill_u32(stream, scratch.counts, static_cast<size_t>(cells) + 2u, 0u);  // histogram and big-cell count
    if (n == 0) {
        fill_u32(stream, offsets, static_cast<size_t>(cells) + 1u, 0u);
        return geo;
    }
    const KeyBox box = key_box(domain, time);
    const uint32_t grid = blocks_for(n, kBlock);
    if (predict != nullptr) {
        key_count_kernel<true><<<grid, kBlock, 0, stream>>>(n, box, positions, *predict, scratch.keys, scratch.slots,
                                                             scratch.counts);
    } else {
        key_count_kernel<false><<<grid, kBlock, 0, stream>>>(n, box, positions, PredictInput{}, scratch.keys,
                                                              scratch.slots, scratch.counts);
    }
    PHX_LAUNCH_CHECK();
    exclusive_scan(stream, scratch.counts, offsets, static_cast<size_t>(cells) + 1u, scratch.partials);
    scatter_unstable_kernel<<<grid, kBlock, 0, stream>>>(n, scratch.keys, scratch.slots, offsets, scratch.unstable);
    PHX_LAUNCH_CHECK();
    order_cells(stream, n, cells, scratch, offsets, perm, gather);
    return geo;
}

void build_batch_grid(hipStream_t stream, uint32_t particles, uint32_t cells, uint32_t block_count,
                      const BatchView& batch, CVec positions, const PredictInput& predict, const Channels& gather,
                      const GridScratch& scratch, uint32_t* perm, uint32_t* offsets) {
    fill_u32(stream, scratch.counts, static_cast<size_t>(cells) + 2u, 0u);  // histogram and big-cell count
    if (particles == 0) {
        fill_u32(stream, offsets, static_cast<size_t>(cells) + 1u, 0u);
        return;
    }
    batch_key_count_kernel<<<block_count, kBlock, 0, stream>>>(batch, positions, predict, scratch.keys, scratch.slots,
                                                               scratch.counts);
    PHX_LAUNCH_CHECK();
    exclusive_scan(stream, scratch.counts, offsets, static_cast<size_t>(cells) + 1u, scratch.partials);
    const uint32_t grid = blocks_for(particles, kBlock);
    scatter_unstable_kernel<<<grid, kBlock, 0, stream>>>(particles, scratch.keys, scratch.slots, offsets,
                                                         scratch.unstable);
    PHX_LAUNCH_CHECK();
    order_cells(stream, particles, cells, scratch, offsets, perm, gather);
}

void export_neighborhood(hipStream_t stream, const Neighborhood& nb, uint64_t* sorted_keys, uint32_t* sorted_indices,
                         uint32_t* offsets) {
    const uint32_t particles = nb.particles;
    if (particles > 0 && (sorted_keys != nullptr || sorted_indices != nullptr)) {
        export_kernel<<<blocks_for(particles, kBlock), kBlock, 0, stream>>>(particles, nb.perm.data(),
                                                                            nb.scratch.keys, sorted_keys, sorted_indices);
        PHX_LAUNCH_CHECK();
    }
    if (offsets != nullptr) {
        const size_t count = static_cast<size_t>(nb.cells) + 1u;
        copy_u32_kernel<<<blocks_for(count, kBlock), kBlock, 0, stream>>>(nb.offsets.data(), offsets, count);
        PHX_LAUNCH_CHECK();
    }
}

void export_sorted_keys(hipStream_t stream, const uint32_t* perm, const uint32_t* keys, uint32_t particles,
                        uint64_t* sorted_keys) {
    if (particles == 0) return;
    export_kernel<<<blocks_for(particles, kBlock), kBlock, 0, stream>>>(particles, perm, keys, sorted_keys, nullptr);
    PHX_LAUNCH_CHECK();
}

void scatter_to_original(hipStream_t stream, const uint32_t* perm, uint32_t particles, const Channels& channels) {
    if (particles == 0 || channels.count == 0) return;
    scatter_channels_kernel<<<blocks_for(particles, kBlock), kBlock, 0, stream>>>(particles, perm, channels);
    PHX_LAUNCH_CHECK();
}

void gather_to_sorted(hipStream_t stream, const uint32_t* perm, uint32_t particles, const Channels& channels) {
    if (particles == 0 || channels.count == 0) return;
    gather_channels_kernel<<<blocks_for(particles, kBlock), kBlock, 0, stream>>>(particles, perm, channels);
    PHX_LAUNCH_CHECK();
}

}  // namespace phx
