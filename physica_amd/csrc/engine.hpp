// Host-side engine internals shared by the solvers: error handling, device
// buffers, domain validation and the uniform-grid builder.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "device_math.hpp"
#include "physica_amd/physica.h"

namespace phx {

struct Error : std::runtime_error {
    phx_status status;
    Error(phx_status s, const std::string& message) : std::runtime_error(message), status(s) {}
};

#define PHX_HIP(call)                                                                                               \
    do {                                                                                                            \
        const hipError_t phx_err_ = (call);                                                                         \
        if (phx_err_ != hipSuccess)                                                                                 \
            throw ::phx::Error(phx_err_ == hipErrorOutOfMemory ? PHX_ERROR_OUT_OF_MEMORY : PHX_ERROR_DEVICE,        \
                               std::string(#call) + ": " + hipGetErrorString(phx_err_));                            \
    } while (0)

#define PHX_LAUNCH_CHECK() PHX_HIP(hipGetLastError())

inline void require(bool condition, phx_status status, const std::string& message) {
    if (!condition) throw Error(status, message);
}

// Owning device allocation (uninitialised, like the reference's
// ::cuda::no_init buffers, domain/domain.cpp:35-57).
template <class T>
class DeviceBuffer {
  public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t count) { resize(count); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& other) noexcept : ptr_(other.ptr_), count_(other.count_) {
        other.ptr_ = nullptr;
        other.count_ = 0;
    }
    DeviceBuffer& operator=(DeviceBuffer&& other) noexcept {
        if (this != &other) {
            release();
            ptr_ = other.ptr_;
            count_ = other.count_;
            other.ptr_ = nullptr;
            other.count_ = 0;
        }
        return *this;
    }
    ~DeviceBuffer() { release(); }

    void resize(size_t count) {
        if (count == count_) return;
        release();
        if (count > 0) PHX_HIP(hipMalloc(reinterpret_cast<void**>(&ptr_), count * sizeof(T)));
        count_ = count;
    }
    T* data() { return ptr_; }
    const T* data() const { return ptr_; }
    size_t size() const { return count_; }

  private:
    void release() {
        if (ptr_ != nullptr) (void)hipFree(ptr_);
        ptr_ = nullptr;
        count_ = 0;
    }
    T* ptr_ = nullptr;
    size_t count_ = 0;
};

// Three device float arrays of one length (VectorField, domain/domain.ixx:58-62).
struct VecBuffer {
    DeviceBuffer<float> x, y, z;
    void resize(size_t n) {
        x.resize(n);
        y.resize(n);
        z.resize(n);
    }
    Vec view() { return {x.data(), y.data(), z.data()}; }
    CVec cview() const { return {x.data(), y.data(), z.data()}; }
};

// FP64 SoA vector field (the adjoint views of domain/device.h).
struct DVec {
    double* x;
    double* y;
    double* z;
};

struct DVecBuffer {
    DeviceBuffer<double> x, y, z;
    void resize(size_t n) {
        x.resize(n);
        y.resize(n);
        z.resize(n);
    }
    DVec view() { return {x.data(), y.data(), z.data()}; }
};

// Validated copy of a DomainConfiguration plus derived grid quantities.
struct Domain {
    phx_domain_configuration config{};
    uint32_t resolution[3]{};
    uint32_t cells = 0;
};

Domain make_domain(const phx_domain_configuration* config);
void grid_resolution(const phx_domain_configuration& config, uint32_t resolution[3]);
Box collision_box(const phx_domain_configuration& config, uint64_t step);
float reference_gradient_norm(const phx_domain_configuration& config);
float spiky_coefficient(float support_radius);

inline CVec cvec(const phx_const_vector_view& v) { return {v.x, v.y, v.z}; }
inline CVec cvec(const phx_vector_view& v) { return {v.x, v.y, v.z}; }
inline Vec vec(const phx_vector_view& v) { return {v.x, v.y, v.z}; }
void require_vec(const phx_vector_view& v, const char* what);
void require_vec(const phx_const_vector_view& v, const char* what);

inline uint32_t blocks_for(uint64_t count, uint32_t block) { return static_cast<uint32_t>((count + block - 1) / block); }

// Per-particle float channels copied between original order (contiguous
// arrays) and sorted order (element s at sorted[s * stride], so four channels
// can fill one float4 array).
constexpr int kMaxChannels = 20;
struct Channels {
    const float* src[kMaxChannels];
    float* dst[kMaxChannels];
    uint32_t stride[kMaxChannels];
    uint32_t count;
};

class ChannelList {
  public:
    void add(const float* src, float* dst, uint32_t stride = 1u) {
        require(list_.count < static_cast<uint32_t>(kMaxChannels), PHX_ERROR_INVALID_ARGUMENT, "too many channels");
        require(src != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null channel source");
        list_.src[list_.count] = src;
        list_.dst[list_.count] = dst;
        list_.stride[list_.count] = stride;
        ++list_.count;
    }
    void add(CVec src, Vec dst) {
        add(src.x, dst.x);
        add(src.y, dst.y);
        add(src.z, dst.z);
    }
    // x, y, z into components 0..2 of a float4 array
    void add(CVec src, float4* dst) {
        float* base = reinterpret_cast<float*>(dst);
        add(src.x, base + 0, 4u);
        add(src.y, base + 1, 4u);
        add(src.z, base + 2, 4u);
    }
    // one scalar into component `component` of a float4 array
    void add(const float* src, float4* dst, int component) {
        add(src, reinterpret_cast<float*>(dst) + component, 4u);
    }
    const Channels& get() const { return list_; }

  private:
    Channels list_{};
};

// Optional predict fused into the key kernel (PBF predict_forward_kernel,
// pbf.cu:73-78): x* = x + dt (v + dt (g + u)), written in original order.
struct PredictInput {
    CVec velocities{};
    CVec controls{};
    Float3 gravity{};
    float dt = 0.0F;
    Vec predicted{};
};

// A batch of independent scenes stepped together: every scene's particles
// occupy one contiguous range of the batch's arrays (original and sorted
// order alike, in scene order), and every scene has its own grid, whose
// cells occupy one contiguous range of the batch's cell offsets. Workgroups
// are listed in a table, each serving consecutive particles of one scene, so
// a workgroup's grid, collision box and stencil slot table are its scene's.
struct BatchBlock {
    uint32_t scene;
    uint32_t first;  // first particle of the workgroup (batch index)
    uint32_t count;  // particles of the workgroup (0: padding)
    uint32_t pad;
};
struct BatchScene {
    uint32_t cell_base;  // first cell of the scene's grid in the batch's offsets
    uint32_t res[3];
    float minimum[3];
    float maximum[3];
    float velocity[3];
    float particle_radius;
    float gravity[3];
    uint32_t no_slip;
};
struct BatchView {
    const BatchBlock* blocks;  // null: a single scene
    const BatchScene* scenes;
    float time;  // step time of the grid and collision box (index * dt)
    float h;
};

// Scratch of one grid build, carved from a single byte buffer.
struct GridScratch {
    uint32_t* keys;       // cell key per particle, original order
    uint32_t* slots;      // arrival slot within the cell (atomic)
    uint32_t* unstable;   // particles grouped by cell, arbitrary order inside
    uint32_t* counts;     // C + 1 histogram, then the count of big cells
    uint32_t* partials;   // scan tile sums
    uint32_t* big_cells;  // cells whose members are radix-sorted
};
size_t grid_scratch_bytes(uint32_t particles, uint32_t cells);
GridScratch carve_grid_scratch(void* base, size_t bytes, uint32_t particles, uint32_t cells);

// Geometry of a built grid (Neighborhood::cell_resolution/cell_origin/
// cell_size/boundary_time, neighborhood/neighborhood.ixx:12-23).
struct GridGeometry {
    uint32_t resolution[3]{};
    float origin[3]{};
    float cell_size = 0.0F;
    float time = 0.0F;
};

// Counting-sort grid build (NeighborhoodSearch::build, neighborhood.cpp:44-89):
// key + atomic cell histogram, exclusive scan into cell offsets, unstable
// scatter, then a per-cell rank by original index that makes the order stable
// and gathers the requested channels into sorted order. Writes perm[N]
// (sorted position -> original index) and offsets[C+1].
GridGeometry build_grid(hipStream_t stream, const Domain& domain, uint64_t step, CVec positions,
                        const PredictInput* predict, const Channels& gather, const GridScratch& scratch, uint32_t* perm,
                        uint32_t* offsets);

// Device arrays of the boundary samples in original order (BoundaryView,
// neighborhood/device.h:26-36, without the time).
struct BoundarySource {
    const float* x[3];
    const float* v[3];
    const float* volume;
};

// The boundary samples of a configuration as seven host arrays of B floats
// (x, y, z, vx, vy, vz, volume), checked finite.
std::vector<float> pack_boundary_samples(const phx_domain_configuration& config);
// The seven arrays of a packed copy at `base` on the device.
BoundarySource boundary_source(const float* base, uint32_t count);

// Device copy of DomainConfiguration::boundary_particles (DeviceBoundary,
// domain/domain.ixx:25-39 and domain/domain.cpp): SoA position, velocity,
// volume, uploaded once from the host array at solver creation.
class BoundarySamples {
  public:
    void upload(const phx_domain_configuration& config);
    uint32_t count() const { return count_; }
    BoundarySource source() const;

  private:
    DeviceBuffer<float> data_;  // x, y, z, vx, vy, vz, volume, each count_ long
    uint32_t count_ = 0;
};

// Scratch and results of the boundary half of a grid build.
struct BoundaryScratch {
    uint32_t* keys;      // cell key per sample, original order
    uint32_t* slots;     // arrival slot within the cell
    uint32_t* unstable;  // samples grouped by cell, arbitrary order inside
};
size_t boundary_scratch_bytes(uint32_t samples);
BoundaryScratch carve_boundary_scratch(void* base, size_t bytes, uint32_t samples);

// The boundary half of NeighborhoodSearch::build (neighborhood.cu:63-69):
// keys of the samples moved to x0 + t v, a stable counting sort by cell
// (the fluid build's histogram and scan buffers are reused), perm[B] (sorted
// -> original index), offsets[C+1], and, when x/v are non-null, the sorted
// records of BoundaryGrid.
void build_boundary_grid(hipStream_t stream, const Domain& domain, uint64_t step, const BoundarySource& samples,
                         uint32_t count, const BoundaryScratch& scratch, const GridScratch& grid_scratch,
                         uint32_t* perm, uint32_t* offsets, float4* x, float4* v);

// The grid build of a batch (build_grid per scene, one launch per kernel):
// each particle predicts x* with its scene's gravity, takes the key of its
// cell in its scene's grid (time batch.time) plus the scene's cell base, and
// one counting sort over the batch's cells orders every scene at once, each
// scene's particles stably by index exactly as its own build would.
void build_batch_grid(hipStream_t stream, uint32_t particles, uint32_t cells, uint32_t block_count,
                      const BatchView& batch, CVec positions, const PredictInput& predict, const Channels& gather,
                      const GridScratch& scratch, uint32_t* perm, uint32_t* offsets);

// Owning grid state used by the solvers: scratch plus the permutation and
// offsets of the last build (kept in the step cache for replay).
struct Neighborhood {
    DeviceBuffer<uint8_t> scratch_bytes;
    GridScratch scratch{};
    DeviceBuffer<uint32_t> perm;
    DeviceBuffer<uint32_t> offsets;
    GridGeometry geometry;
    uint32_t particles = 0;
    uint32_t cells = 0;

    void allocate(uint32_t n, uint32_t c) {
        if (perm.size() == n && offsets.size() == static_cast<size_t>(c) + 1u && scratch_bytes.size() != 0) return;
        particles = n;
        cells = c;
        perm.resize(n);
        offsets.resize(static_cast<size_t>(c) + 1u);
        const size_t bytes = grid_scratch_bytes(n, c);
        scratch_bytes.resize(bytes);
        scratch = carve_grid_scratch(scratch_bytes.data(), bytes, n, c);
    }
    // `boundary` (may be null or empty) adds the sorted boundary samples.
    void build(hipStream_t stream, const Domain& domain, uint64_t step, CVec positions, const PredictInput* predict,
               const Channels& gather, const BoundarySamples* boundary = nullptr) {
        allocate(domain.config.particle_count, domain.cells);
        geometry = build_grid(stream, domain, step, positions, predict, gather, scratch, perm.data(), offsets.data());
        boundary_count = boundary != nullptr ? boundary->count() : 0u;
        if (boundary_count == 0u) return;
        if (boundary_perm.size() != boundary_count || boundary_offsets.size() != static_cast<size_t>(cells) + 1u) {
            boundary_perm.resize(boundary_count);
            boundary_offsets.resize(static_cast<size_t>(cells) + 1u);
            boundary_x.resize(boundary_count);
            boundary_v.resize(boundary_count);
            const size_t bytes = boundary_scratch_bytes(boundary_count);
            boundary_scratch_bytes_.resize(bytes);
            boundary_scratch = carve_boundary_scratch(boundary_scratch_bytes_.data(), bytes, boundary_count);
        }
        build_boundary_grid(stream, domain, step, boundary->source(), boundary_count, boundary_scratch, scratch,
                            boundary_perm.data(), boundary_offsets.data(), boundary_x.data(), boundary_v.data());
    }
    SortedGrid grid() const {
        return SortedGrid{offsets.data(),      geometry.resolution[0], geometry.resolution[1],
                          geometry.resolution[2], geometry.origin[0],     geometry.origin[1],
                          geometry.origin[2],     geometry.cell_size};
    }
    BoundaryGrid boundary() const {
        if (boundary_count == 0u) return BoundaryGrid{nullptr, nullptr, nullptr};
        return BoundaryGrid{boundary_offsets.data(), boundary_x.data(), boundary_v.data()};
    }

    uint32_t boundary_count = 0;
    DeviceBuffer<uint32_t> boundary_perm;
    DeviceBuffer<uint32_t> boundary_offsets;
    DeviceBuffer<float4> boundary_x, boundary_v;
    BoundaryScratch boundary_scratch{};

  private:
    DeviceBuffer<uint8_t> boundary_scratch_bytes_;
};

// Device-wide exclusive scan of count u32 values (in and out may alias).
size_t exclusive_scan_partials(size_t count);
void exclusive_scan(hipStream_t stream, const uint32_t* in, uint32_t* out, size_t count, uint32_t* partials);

// Writes the reference-layout neighborhood (u64 sorted keys, sorted indices,
// offsets) from a built Neighborhood.
void export_neighborhood(hipStream_t stream, const Neighborhood& nb, uint64_t* sorted_keys, uint32_t* sorted_indices,
                         uint32_t* offsets);

// sorted_keys[s] = keys[perm[s]] as u64 (the reference's sorted key array).
void export_sorted_keys(hipStream_t stream, const uint32_t* perm, const uint32_t* keys, uint32_t particles,
                        uint64_t* sorted_keys);

// Scatter sorted-order channels back to original order: dst[perm[s]] = src[s].
void scatter_to_original(hipStream_t stream, const uint32_t* perm, uint32_t particles, const Channels& channels);
// Gather original-order channels into sorted order: dst[s] = src[perm[s]].
void gather_to_sorted(hipStream_t stream, const uint32_t* perm, uint32_t particles, const Channels& channels);

void fill_u32(hipStream_t stream, uint32_t* data, size_t count, uint32_t value);
void fill_f32(hipStream_t stream, float* data, size_t count, float value);

}  // namespace phx
