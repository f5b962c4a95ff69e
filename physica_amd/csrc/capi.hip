// extern "C" entry points of libphysica_amd (include/physica_amd/physica.h).
// Exceptions never cross the ABI: each call maps them to a phx_status and a
// per-thread message.
#include <cstring>
#include <mutex>
#include <new>
#include <string>

#include "dfsph.hpp"
#include "pbf.hpp"
#include "sph.hpp"

struct phx_dfsph {
    phx::DfsphSolver impl;
    phx_dfsph(const phx_domain_configuration* d, const phx_dfsph_configuration* c, phx_execution_mode m, hipStream_t s)
        : impl(d, c, m, s) {}
};
struct phx_dfsph_step_cache {
    const phx_dfsph* owner = nullptr;
    phx::DfsphCache impl;
    bool valid = false;
};
struct phx_sph {
    phx::SphSolver impl;
    phx_sph(const phx_domain_configuration* d, const phx_sph_configuration* c, phx_execution_mode m, hipStream_t s)
        : impl(d, c, m, s) {}
};
struct phx_sph_step_cache {
    const phx_sph* owner = nullptr;
    phx::SphCache impl;
    bool valid = false;
};
struct phx_pbf {
    phx::PbfSolver impl;
    phx_pbf(const phx_domain_configuration* d, const phx_pbf_configuration* c, phx_execution_mode m, hipStream_t s)
        : impl(d, c, m, s) {}
};
struct phx_pbf_batch {
    phx::PbfBatch impl;
    phx_pbf_batch(const phx_domain_configuration* d, uint32_t count, const phx_pbf_configuration* c,
                  const phx_pbf_batch_configuration* b, hipStream_t s)
        : impl(d, count, c, b, s) {}
};
struct phx_pbf_step_cache {
    const phx_pbf* owner = nullptr;
    phx::PbfCache impl;
    bool valid = false;
};

namespace {

thread_local std::string g_last_error;

template <class F>
phx_status guarded(F&& body) {
    try {
        body();
        g_last_error.clear();
        return PHX_SUCCESS;
    } catch (const phx::Error& e) {
        g_last_error = e.what();
        return e.status;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return PHX_ERROR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return PHX_ERROR_DEVICE;
    } catch (...) {
        g_last_error = "unknown error";
        return PHX_ERROR_DEVICE;
    }
}

inline hipStream_t stream_of(phx_stream s) { return reinterpret_cast<hipStream_t>(s); }

constexpr uint32_t kBlock = 256;

// Reference-layout view (neighborhood/device.h:8-23).
struct RefGrid {
    const uint32_t* indices;
    const uint32_t* offsets;
    const uint32_t* boundary_indices;  // null when there are no boundary samples
    const uint32_t* boundary_offsets;
    uint32_t cx, cy, cz;
    float ox, oy, oz, cell;
};

// The reference-layout walk of particle i (NeighborhoodView, the cell of
// its topology position): per stencil cell (offset_x outer), the cell's
// particles j in sorted order, then its boundary samples b at
// x0 + t v (boundary_position, neighborhood/device.cuh:16-22).
template <class Fluid, class Boundary>
__device__ __forceinline__ void reference_walk(const RefGrid& g, const phx_boundary_view& boundary, phx::Float3 topo,
                                               Fluid&& fluid, Boundary&& sample) {
    using namespace phx;
    const int cx = cell_clamp((topo.x - g.ox) / g.cell, g.cx);
    const int cy = cell_clamp((topo.y - g.oy) / g.cell, g.cy);
    const int cz = cell_clamp((topo.z - g.oz) / g.cell, g.cz);
    for (int ox = -1; ox <= 1; ++ox)
        for (int oy = -1; oy <= 1; ++oy)
            for (int oz = -1; oz <= 1; ++oz) {
                const int X = cx + ox, Y = cy + oy, Z = cz + oz;
                if (X < 0 || X >= static_cast<int>(g.cx) || Y < 0 || Y >= static_cast<int>(g.cy) || Z < 0 ||
                    Z >= static_cast<int>(g.cz))
                    continue;
                const uint32_t c = (static_cast<uint32_t>(Z) * g.cy + static_cast<uint32_t>(Y)) * g.cx +
                                   static_cast<uint32_t>(X);
                for (uint32_t s = g.offsets[c]; s < g.offsets[c + 1u]; ++s) fluid(g.indices[s]);
                if (g.boundary_indices == nullptr) continue;
                for (uint32_t s = g.boundary_offsets[c]; s < g.boundary_offsets[c + 1u]; ++s) {
                    const uint32_t b = g.boundary_indices[s];
                    sample(b, Float3{boundary.position_x[b] + boundary.time * boundary.velocity_x[b],
                                     boundary.position_y[b] + boundary.time * boundary.velocity_y[b],
                                     boundary.position_z[b] + boundary.time * boundary.velocity_z[b]});
                }
            }
}

// density_kernel / density_kernel_gradient (density.cu:9-16)
__device__ __forceinline__ float density_w(phx::Float3 d, float h, int pbf) {
    return pbf != 0 ? phx::poly6(d, h) : phx::cubic(d, h);
}
__device__ __forceinline__ phx::Float3 density_g(phx::Float3 d, float h, int pbf) {
    return pbf != 0 ? phx::poly6_gradient(d, h) : phx::cubic_gradient(d, h);
}

// density_forward_kernel (density/density.cu:18-40) in original particle order
// over a reference-layout neighbourhood, each cell's boundary samples after
// its particles.
__global__ void __launch_bounds__(kBlock)
    reference_density_kernel(uint32_t n, float h, int pbf_kernel, phx::CVec topo, phx::CVec pos, const float* masses,
                             const float* rest_densities, phx_boundary_view boundary, RefGrid g, float* out) {
    using namespace phx;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Float3 p = load(pos, i);
    float result = 0.0F;
    reference_walk(
        g, boundary, load(topo, i), [&](uint32_t j) { result += masses[j] * density_w(sub(p, load(pos, j)), h, pbf_kernel); },
        [&](uint32_t b, Float3 xb) {
            result += rest_densities[i] * boundary.volumes[b] * density_w(sub(p, xb), h, pbf_kernel);
        });
    out[i] = result;
}

// density_jvp_kernel (density/density.cu:42-63); null tangent arrays count
// as zero.
__global__ void __launch_bounds__(kBlock)
    reference_density_jvp_kernel(uint32_t n, float h, int pbf_kernel, phx::CVec topo, phx::CVec pos,
                                 phx::CVec pos_t, const float* masses, const float* mass_t,
                                 const float* rest_densities, const float* rest_t, phx_boundary_view boundary,
                                 RefGrid g, float* out) {
    using namespace phx;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Float3 p = load(pos, i);
    const Float3 t = load(pos_t, i);
    const float rest_i = rest_densities != nullptr ? rest_densities[i] : 0.0F;
    const float rest_t_i = rest_t != nullptr ? rest_t[i] : 0.0F;
    float result = 0.0F;
    reference_walk(
        g, boundary, load(topo, i),
        [&](uint32_t j) {
            const Float3 d = sub(p, load(pos, j));
            const float mt = mass_t != nullptr ? mass_t[j] : 0.0F;
            result += mt * density_w(d, h, pbf_kernel) + masses[j] * dot(density_g(d, h, pbf_kernel), sub(t, load(pos_t, j)));
        },
        [&](uint32_t b, Float3 xb) {
            const Float3 d = sub(p, xb);
            result += boundary.volumes[b] *
                      (rest_t_i * density_w(d, h, pbf_kernel) + rest_i * dot(density_g(d, h, pbf_kernel), t));
        });
    out[i] = result;
}

// density_vjp_kernel (density/density.cu:65-98) in FP64, gather form: each
// particle collects the pair terms of its own density and of its
// neighbours' and accumulates (+=) into its outputs; null outputs are
// skipped.
__global__ void __launch_bounds__(kBlock)
    reference_density_vjp_kernel(uint32_t n, float h, int pbf_kernel, phx::CVec topo, phx::CVec pos,
                                 const float* masses, const float* rest_densities, phx_boundary_view boundary,
                                 RefGrid g, const double* rho_a, phx_adjoint_vector_view pos_a, double* mass_a,
                                 double* rest_a) {
    using namespace phx;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Float3 p = load(pos, i);
    const double out_a = rho_a[i];
    const float m_i = masses[i];
    const float rest_i = rest_densities != nullptr ? rest_densities[i] : 0.0F;
    double px = 0.0, py = 0.0, pz = 0.0, mass_c = 0.0, rest_c = 0.0;
    reference_walk(
        g, boundary, load(topo, i),
        [&](uint32_t j) {
            const Float3 d = sub(p, load(pos, j));
            const Float3 gr = density_g(d, h, pbf_kernel);
            const double weight = out_a * static_cast<double>(masses[j]) + rho_a[j] * static_cast<double>(m_i);
            px += static_cast<double>(gr.x) * weight;
            py += static_cast<double>(gr.y) * weight;
            pz += static_cast<double>(gr.z) * weight;
            mass_c += rho_a[j] * static_cast<double>(density_w(d, h, pbf_kernel));
        },
        [&](uint32_t b, Float3 xb) {
            const Float3 d = sub(p, xb);
            const double volume = static_cast<double>(boundary.volumes[b]);
            const Float3 gr = density_g(d, h, pbf_kernel);
            const double weight = out_a * static_cast<double>(rest_i) * volume;
            px += static_cast<double>(gr.x) * weight;
            py += static_cast<double>(gr.y) * weight;
            pz += static_cast<double>(gr.z) * weight;
            rest_c += out_a * volume * static_cast<double>(density_w(d, h, pbf_kernel));
        });
    if (pos_a.x != nullptr) {
        pos_a.x[i] += px;
        pos_a.y[i] += py;
        pos_a.z[i] += pz;
    }
    if (mass_a != nullptr) mass_a[i] += mass_c;
    if (rest_a != nullptr) rest_a[i] += rest_c;
}

// accumulate_scalar_kernel (domain/domain.cu:16-20): one element per lane.
__global__ void __launch_bounds__(kBlock) accumulate_kernel(const double* source, double* destination, size_t count) {
    const size_t k = static_cast<size_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (k < count) destination[k] += source[k];
}

// write_vectors_kernel (spectra/kernels.cu:8-14) for a block of kBlock
// particles: the three SoA components are read coalesced into LDS and the
// interleaved x, y, z written back as consecutive floats, so every global
// access is a coalesced dword.
__global__ void __launch_bounds__(kBlock) interleave_kernel(uint32_t n, phx_const_vector_view in, float* out) {
    __shared__ float staged[3u * kBlock];
    const uint32_t base = blockIdx.x * kBlock;
    const uint32_t i = base + threadIdx.x;
    if (i < n) {
        staged[3u * threadIdx.x + 0u] = in.x[i];
        staged[3u * threadIdx.x + 1u] = in.y[i];
        staged[3u * threadIdx.x + 2u] = in.z[i];
    }
    __syncthreads();
    const uint32_t valid = 3u * (n - base < kBlock ? n - base : kBlock);
    for (uint32_t k = threadIdx.x; k < valid; k += kBlock) out[3u * static_cast<size_t>(base) + k] = staged[k];
}

// Workspace of phx_build_neighborhood: the fluid grid scratch, then for B
// boundary samples their scratch, the packed samples (7 B floats) and the
// boundary permutation.
size_t round256(size_t b) { return (b + 255u) & ~static_cast<size_t>(255u); }
size_t neighborhood_workspace_bytes(const phx::Domain& d) {
    const size_t b = d.config.boundary_particle_count;
    size_t bytes = round256(phx::grid_scratch_bytes(d.config.particle_count, d.cells));
    if (b > 0u)
        bytes += round256(phx::boundary_scratch_bytes(static_cast<uint32_t>(b))) + round256(7u * b * sizeof(float)) +
                 round256(b * sizeof(uint32_t));
    return bytes;
}


// Shared by phx_build_neighborhood and its no-allocation variant.
void check_neighborhood_outputs(const phx::Domain& d, const phx_const_vector_view& positions,
                                const uint32_t* sorted_particle_indices, const uint32_t* cell_offsets) {
    phx::require(cell_offsets != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null cell_offsets");
    if (d.config.particle_count > 0) {
        phx::require_vec(positions, "positions");
        phx::require(sorted_particle_indices != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null sorted_particle_indices");
    }
}

phx::GridGeometry build_fluid_neighborhood(hipStream_t s, const phx::Domain& d, uint64_t step_index,
                                           const phx_const_vector_view& positions, const phx::GridScratch& scratch,
                                           uint64_t* sorted_keys, uint32_t* sorted_particle_indices,
                                           uint32_t* cell_offsets) {
    phx::ChannelList none;
    const phx::GridGeometry geo = phx::build_grid(s, d, step_index, phx::cvec(positions), nullptr, none.get(),
                                                  scratch, sorted_particle_indices, cell_offsets);
    if (sorted_keys != nullptr && d.config.particle_count > 0)
        phx::export_sorted_keys(s, sorted_particle_indices, scratch.keys, d.config.particle_count, sorted_keys);
    return geo;
}

void fill_neighborhood_view(const phx::Domain& d, const phx::GridGeometry& geo, uint64_t* sorted_keys,
                            uint32_t* sorted_particle_indices, uint32_t* cell_offsets, uint64_t* sorted_boundary_keys,
                            uint32_t* sorted_boundary_indices, uint32_t* boundary_cell_offsets,
                            phx_neighborhood_view* view) {
    view->sorted_keys = sorted_keys;
    view->sorted_particle_indices = sorted_particle_indices;
    view->cell_offsets = cell_offsets;
    view->sorted_boundary_keys = sorted_boundary_keys;
    view->sorted_boundary_indices = sorted_boundary_indices;
    view->boundary_cell_offsets = boundary_cell_offsets;
    view->particle_count = d.config.particle_count;
    view->boundary_count = d.config.boundary_particle_count;
    view->cells_x = geo.resolution[0];
    view->cells_y = geo.resolution[1];
    view->cells_z = geo.resolution[2];
    view->origin_x = geo.origin[0];
    view->origin_y = geo.origin[1];
    view->origin_z = geo.origin[2];
    view->cell_size = geo.cell_size;
}

// Fixed device scratch of phx_build_neighborhood_noalloc. Every GPU the
// library runs on holds its own copy until the process ends, and every call
// on one GPU shares it: calls that overlap on one GPU race. In host code the
// name `scratch` is only a host-side stand-in; kernels receive the per-device
// address that hipGetSymbolAddress returns.
constexpr size_t kNoAllocScratchBytes = PHX_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_address[kMaxDevices] = {};

// The scratch address on the device of `stream`, looked up once per device
// (that device current) on the first call there.
void* noalloc_scratch(hipStream_t stream) {
    int current = -1;
    PHX_HIP(hipGetDevice(&current));
    int device = current;
    if (stream != nullptr) PHX_HIP(hipStreamGetDevice(stream, &device));
    phx::require(device == current, PHX_ERROR_INVALID_ARGUMENT,
                 "phx_build_neighborhood_noalloc: the stream's device must be the current device");
    phx::require(device >= 0 && device < kMaxDevices, PHX_ERROR_UNSUPPORTED, "device ordinal above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_address[device] == nullptr) {
        void* ptr = nullptr;
        PHX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        phx::require(ptr != nullptr, PHX_ERROR_DEVICE, "hipGetSymbolAddress returned no scratch address");
        g_scratch_address[device] = ptr;
    }
    return g_scratch_address[device];
}

}  // namespace

extern "C" {

const char* phx_last_error(void) { return g_last_error.c_str(); }

const char* phx_version(void) { return "physica_amd 0.1.0 (gfx950)"; }

phx_status phx_accumulate(phx_stream stream, const double* source, double* destination, size_t count) {
    return guarded([&] {
        if (count == 0) return;
        phx::require(source != nullptr && destination != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null array");
        const size_t blocks = (count + kBlock - 1) / kBlock;
        phx::require(blocks < (1ull << 31), PHX_ERROR_CAPACITY, "count too large");
        accumulate_kernel<<<static_cast<uint32_t>(blocks), kBlock, 0, stream_of(stream)>>>(source, destination, count);
        PHX_LAUNCH_CHECK();
    });
}

phx_status phx_write_vectors(phx_stream stream, uint32_t count, phx_const_vector_view positions,
                             phx_const_vector_view velocities, float* positions_xyz, float* velocities_xyz) {
    return guarded([&] {
        if (count == 0u) return;
        const uint32_t blocks = phx::blocks_for(count, kBlock);
        if (positions_xyz != nullptr) {
            phx::require_vec(positions, "positions");
            interleave_kernel<<<blocks, kBlock, 0, stream_of(stream)>>>(count, positions, positions_xyz);
            PHX_LAUNCH_CHECK();
        }
        if (velocities_xyz != nullptr) {
            phx::require_vec(velocities, "velocities");
            interleave_kernel<<<blocks, kBlock, 0, stream_of(stream)>>>(count, velocities, velocities_xyz);
            PHX_LAUNCH_CHECK();
        }
    });
}

phx_status phx_grid_resolution(const phx_domain_configuration* domain, uint32_t resolution[3]) {
    return guarded([&] {
        phx::require(domain != nullptr && resolution != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null argument");
        phx::grid_resolution(*domain, resolution);
    });
}

phx_status phx_neighborhood_workspace_size(const phx_domain_configuration* domain, size_t* bytes) {
    return guarded([&] {
        phx::require(bytes != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null bytes");
        const phx::Domain d = phx::make_domain(domain);
        *bytes = neighborhood_workspace_bytes(d);
    });
}

phx_status phx_build_neighborhood(phx_stream stream, const phx_domain_configuration* domain, uint64_t step_index,
                                  phx_const_vector_view positions, void* workspace, size_t workspace_bytes,
                                  uint64_t* sorted_keys, uint32_t* sorted_particle_indices, uint32_t* cell_offsets,
                                  uint64_t* sorted_boundary_keys, uint32_t* sorted_boundary_indices,
                                  uint32_t* boundary_cell_offsets, phx_neighborhood_view* view) {
    return guarded([&] {
        const phx::Domain d = phx::make_domain(domain);
        const uint32_t n = d.config.particle_count;
        const uint32_t b = d.config.boundary_particle_count;
        const hipStream_t s = stream_of(stream);
        check_neighborhood_outputs(d, positions, sorted_particle_indices, cell_offsets);
        phx::require(workspace != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null workspace");
        phx::require(workspace_bytes >= neighborhood_workspace_bytes(d), PHX_ERROR_WORKSPACE_TOO_SMALL,
                     "neighbourhood workspace too small");
        const std::vector<float> packed = phx::pack_boundary_samples(d.config);
        auto* base = static_cast<uint8_t*>(workspace);
        const size_t grid_bytes = round256(phx::grid_scratch_bytes(n, d.cells));
        const phx::GridScratch scratch = phx::carve_grid_scratch(base, grid_bytes, n, d.cells);
        const phx::GridGeometry geo =
            build_fluid_neighborhood(s, d, step_index, positions, scratch, sorted_keys, sorted_particle_indices,
                                     cell_offsets);
        if (b > 0u) {
            uint8_t* p = base + grid_bytes;
            const size_t scratch_bytes = round256(phx::boundary_scratch_bytes(b));
            const phx::BoundaryScratch bs = phx::carve_boundary_scratch(p, scratch_bytes, b);
            p += scratch_bytes;
            auto* samples = reinterpret_cast<float*>(p);
            p += round256(7u * static_cast<size_t>(b) * sizeof(float));
            uint32_t* perm = sorted_boundary_indices != nullptr ? sorted_boundary_indices : reinterpret_cast<uint32_t*>(p);
            PHX_HIP(hipMemcpyAsync(samples, packed.data(), sizeof(float) * packed.size(), hipMemcpyHostToDevice, s));
            // scratch offsets when the caller does not want them
            uint32_t* offsets = boundary_cell_offsets;
            phx::DeviceBuffer<uint32_t> own;
            if (offsets == nullptr) {
                own.resize(static_cast<size_t>(d.cells) + 1u);
                offsets = own.data();
            }
            phx::build_boundary_grid(s, d, step_index, phx::boundary_source(samples, b), b, bs, scratch, perm, offsets,
                                     nullptr, nullptr);
            if (sorted_boundary_keys != nullptr) phx::export_sorted_keys(s, perm, bs.keys, b, sorted_boundary_keys);
            // the host copy of the samples and any owned offsets end with this call
            PHX_HIP(hipStreamSynchronize(s));
        } else if (boundary_cell_offsets != nullptr) {
            phx::fill_u32(s, boundary_cell_offsets, static_cast<size_t>(d.cells) + 1u, 0u);
        }
        if (view != nullptr)
            fill_neighborhood_view(d, geo, sorted_keys, sorted_particle_indices, cell_offsets, sorted_boundary_keys,
                                   sorted_boundary_indices, boundary_cell_offsets, view);
    });
}

phx_status phx_build_neighborhood_noalloc(phx_stream stream, const phx_domain_configuration* domain,
                                          uint64_t step_index, phx_const_vector_view positions,
                                          uint64_t* sorted_keys, uint32_t* sorted_particle_indices,
                                          uint32_t* cell_offsets, phx_neighborhood_view* view) {
    return guarded([&] {
        const phx::Domain d = phx::make_domain(domain);
        const uint32_t n = d.config.particle_count;
        const hipStream_t s = stream_of(stream);
        check_neighborhood_outputs(d, positions, sorted_particle_indices, cell_offsets);
        phx::require(d.config.boundary_particle_count == 0u, PHX_ERROR_UNSUPPORTED,
                     "phx_build_neighborhood_noalloc takes no boundary samples (they are uploaded from host memory "
                     "on every call); use phx_build_neighborhood");
        const size_t need = phx::grid_scratch_bytes(n, d.cells);
        phx::require(need <= kNoAllocScratchBytes, PHX_ERROR_CAPACITY,
                     "phx_build_neighborhood_noalloc: " + std::to_string(n) + " particles and " +
                         std::to_string(d.cells) + " cells need " + std::to_string(need) +
                         " bytes of scratch, above its fixed " + std::to_string(kNoAllocScratchBytes));
        void* base = noalloc_scratch(s);
        const phx::GridScratch scratch = phx::carve_grid_scratch(base, kNoAllocScratchBytes, n, d.cells);
        const phx::GridGeometry geo =
            build_fluid_neighborhood(s, d, step_index, positions, scratch, sorted_keys, sorted_particle_indices,
                                     cell_offsets);
        if (view != nullptr)
            fill_neighborhood_view(d, geo, sorted_keys, sorted_particle_indices, cell_offsets, nullptr, nullptr,
                                   nullptr, view);
    });
}

namespace {
// Checks of a reference-layout neighbourhood and boundary view shared by the
// density entry points; returns the device-side grid view.
RefGrid checked_grid(uint32_t particle_count, const phx_neighborhood_view& neighborhood,
                     const phx_boundary_view& boundary, const float* rest_densities) {
    phx::require(neighborhood.sorted_particle_indices != nullptr && neighborhood.cell_offsets != nullptr,
                 PHX_ERROR_INVALID_ARGUMENT, "null neighbourhood arrays");
    phx::require(neighborhood.boundary_count == boundary.count, PHX_ERROR_INVALID_ARGUMENT,
                 "boundary count does not match the neighbourhood's");
    if (boundary.count > 0u) {
        phx::require(neighborhood.sorted_boundary_indices != nullptr && neighborhood.boundary_cell_offsets != nullptr,
                     PHX_ERROR_INVALID_ARGUMENT, "null neighbourhood boundary arrays");
        phx::require(boundary.position_x != nullptr && boundary.position_y != nullptr &&
                         boundary.position_z != nullptr && boundary.velocity_x != nullptr &&
                         boundary.velocity_y != nullptr && boundary.velocity_z != nullptr &&
                         boundary.volumes != nullptr && rest_densities != nullptr,
                     PHX_ERROR_INVALID_ARGUMENT, "null boundary arrays or rest densities");
    }
    phx::require(neighborhood.particle_count == particle_count, PHX_ERROR_INVALID_ARGUMENT,
                 "neighbourhood particle_count does not match");
    phx::require(neighborhood.cells_x > 0 && neighborhood.cells_y > 0 && neighborhood.cells_z > 0 &&
                     neighborhood.cell_size > 0.0F,
                 PHX_ERROR_INVALID_ARGUMENT, "empty neighbourhood grid");
    const bool with_boundary = boundary.count > 0u;
    return RefGrid{neighborhood.sorted_particle_indices,
                   neighborhood.cell_offsets,
                   with_boundary ? neighborhood.sorted_boundary_indices : nullptr,
                   with_boundary ? neighborhood.boundary_cell_offsets : nullptr,
                   neighborhood.cells_x,
                   neighborhood.cells_y,
                   neighborhood.cells_z,
                   neighborhood.origin_x,
                   neighborhood.origin_y,
                   neighborhood.origin_z,
                   neighborhood.cell_size};
}
}  // namespace

phx_status phx_density_forward(phx_stream stream, uint32_t particle_count, float support_radius, int pbf_kernel,
                               phx_const_vector_view topology_positions, phx_const_vector_view positions,
                               phx_particle_parameters parameters, phx_neighborhood_view neighborhood,
                               phx_boundary_view boundary, float* densities) {
    return guarded([&] {
        if (particle_count == 0) return;
        phx::require_vec(topology_positions, "topology positions");
        phx::require_vec(positions, "positions");
        phx::require(parameters.masses != nullptr && densities != nullptr, PHX_ERROR_INVALID_ARGUMENT,
                     "null masses or densities");
        const RefGrid g = checked_grid(particle_count, neighborhood, boundary, parameters.rest_densities);
        reference_density_kernel<<<phx::blocks_for(particle_count, kBlock), kBlock, 0, stream_of(stream)>>>(
            particle_count, support_radius, pbf_kernel, phx::cvec(topology_positions), phx::cvec(positions),
            parameters.masses, parameters.rest_densities, boundary, g, densities);
        PHX_LAUNCH_CHECK();
    });
}

phx_status phx_density_jvp(phx_stream stream, uint32_t particle_count, float support_radius, int pbf_kernel,
                           phx_const_vector_view topology_positions, phx_const_vector_view positions,
                           phx_const_vector_view position_tangent, phx_particle_parameters parameters,
                           phx_particle_parameter_tangent parameter_tangent, phx_neighborhood_view neighborhood,
                           phx_boundary_view boundary, float* density_tangent) {
    return guarded([&] {
        if (particle_count == 0) return;
        phx::require_vec(topology_positions, "topology positions");
        phx::require_vec(positions, "positions");
        phx::require_vec(position_tangent, "position tangent");
        phx::require(parameters.masses != nullptr && density_tangent != nullptr, PHX_ERROR_INVALID_ARGUMENT,
                     "null masses or density tangent");
        const RefGrid g = checked_grid(particle_count, neighborhood, boundary, parameters.rest_densities);
        reference_density_jvp_kernel<<<phx::blocks_for(particle_count, kBlock), kBlock, 0, stream_of(stream)>>>(
            particle_count, support_radius, pbf_kernel, phx::cvec(topology_positions), phx::cvec(positions),
            phx::cvec(position_tangent), parameters.masses, parameter_tangent.masses, parameters.rest_densities,
            parameter_tangent.rest_densities, boundary, g, density_tangent);
        PHX_LAUNCH_CHECK();
    });
}

phx_status phx_density_vjp(phx_stream stream, uint32_t particle_count, float support_radius, int pbf_kernel,
                           phx_const_vector_view topology_positions, phx_const_vector_view positions,
                           phx_particle_parameters parameters, phx_neighborhood_view neighborhood,
                           phx_boundary_view boundary, const double* density_adjoint,
                           phx_adjoint_vector_view position_adjoint, phx_particle_parameter_adjoint parameter_adjoint) {
    return guarded([&] {
        if (particle_count == 0) return;
        phx::require_vec(topology_positions, "topology positions");
        phx::require_vec(positions, "positions");
        phx::require(parameters.masses != nullptr && density_adjoint != nullptr, PHX_ERROR_INVALID_ARGUMENT,
                     "null masses or density adjoint");
        phx::require((position_adjoint.x == nullptr) == (position_adjoint.y == nullptr) &&
                         (position_adjoint.x == nullptr) == (position_adjoint.z == nullptr),
                     PHX_ERROR_INVALID_ARGUMENT, "position adjoint must have all three or no components");
        const RefGrid g = checked_grid(particle_count, neighborhood, boundary, parameters.rest_densities);
        reference_density_vjp_kernel<<<phx::blocks_for(particle_count, kBlock), kBlock, 0, stream_of(stream)>>>(
            particle_count, support_radius, pbf_kernel, phx::cvec(topology_positions), phx::cvec(positions),
            parameters.masses, parameters.rest_densities, boundary, g, density_adjoint, position_adjoint,
            parameter_adjoint.masses, parameter_adjoint.rest_densities);
        PHX_LAUNCH_CHECK();
    });
}

// ------------------------------------------------------------------ DFSPH
phx_status phx_dfsph_create(const phx_domain_configuration* domain, const phx_dfsph_configuration* configuration,
                            phx_execution_mode mode, phx_stream stream, phx_dfsph** solver) {
    return guarded([&] {
        phx::require(solver != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null solver output");
        *solver = nullptr;
        *solver = new phx_dfsph(domain, configuration, mode, stream_of(stream));
    });
}

phx_status phx_dfsph_destroy(phx_dfsph* solver) {
    return guarded([&] { delete solver; });
}

phx_status phx_dfsph_allocate_step_cache(phx_dfsph* solver, phx_dfsph_step_cache** cache) {
    return guarded([&] {
        phx::require(solver != nullptr && cache != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null argument");
        *cache = nullptr;
        auto* c = new phx_dfsph_step_cache();
        c->owner = solver;
        c->impl.stream = solver->impl.stream();
        try {
            c->impl.allocate(solver->impl.domain().config.particle_count, solver->impl.domain().cells);
        } catch (...) {
            delete c;
            throw;
        }
        *cache = c;
    });
}

phx_status phx_dfsph_free_step_cache(phx_dfsph_step_cache* cache) {
    return guarded([&] { delete cache; });
}

phx_status phx_dfsph_forward_step(phx_dfsph* solver, const phx_sph_dfsph_state* state, const phx_control* control,
                                  const phx_sph_dfsph_parameters* parameters, phx_sph_dfsph_state* next_state,
                                  phx_dfsph_step_cache* cache) {
    return guarded([&] {
        phx::require(solver && state && control && parameters && next_state && cache, PHX_ERROR_INVALID_ARGUMENT,
                     "null argument");
        phx::require(cache->owner == solver, PHX_ERROR_INVALID_ARGUMENT, "step cache belongs to another solver");
        cache->valid = false;
        solver->impl.forward_step(*state, *control, *parameters, *next_state, cache->impl);
        cache->valid = true;
    });
}

phx_status phx_dfsph_reference_gradient_norm(const phx_dfsph* solver, float* value) {
    return guarded([&] {
        phx::require(solver != nullptr && value != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null argument");
        *value = solver->impl.gradient_norm();
    });
}

phx_status phx_dfsph_last_iterations(const phx_dfsph* solver, uint32_t iterations[2]) {
    return guarded([&] {
        phx::require(solver != nullptr && iterations != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null argument");
        iterations[0] = solver->impl.last_iterations()[0];
        iterations[1] = solver->impl.last_iterations()[1];
    });
}

phx_status phx_dfsph_cache_neighborhood(const phx_dfsph_step_cache* cache, uint64_t* sorted_keys,
                                        uint32_t* sorted_particle_indices, uint32_t* cell_offsets) {
    return guarded([&] {
        phx::require(cache != nullptr && cache->valid, PHX_ERROR_INVALID_ARGUMENT, "cache holds no step");
        phx::export_neighborhood(cache->impl.stream, cache->impl.neighborhood, sorted_keys, sorted_particle_indices,
                                 cell_offsets);
    });
}

phx_status phx_dfsph_cache_densities(const phx_dfsph_step_cache* cache, float* densities) {
    return guarded([&] {
        phx::require(cache != nullptr && cache->valid, PHX_ERROR_INVALID_ARGUMENT, "cache holds no step");
        phx::require(densities != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null output");
        phx::ChannelList ch;
        ch.add(cache->impl.densities.data(), densities);
        phx::scatter_to_original(cache->impl.stream, cache->impl.neighborhood.perm.data(),
                                 cache->impl.neighborhood.particles, ch.get());
    });
}

phx_status phx_dfsph_cache_non_pressure_accelerations(const phx_dfsph_step_cache* cache, phx_vector_view out) {
    return guarded([&] {
        phx::require(cache != nullptr && cache->valid, PHX_ERROR_INVALID_ARGUMENT, "cache holds no step");
        phx::require_vec(out, "output");
        phx::ChannelList ch;
        const float* np = reinterpret_cast<const float*>(cache->impl.non_pressure.data());
        ch.add(np + 0, out.x, 4u);
        ch.add(np + 1, out.y, 4u);
        ch.add(np + 2, out.z, 4u);
        phx::scatter_to_original(cache->impl.stream, cache->impl.neighborhood.perm.data(),
                                 cache->impl.neighborhood.particles, ch.get());
    });
}

phx_status phx_dfsph_jvp_step(phx_dfsph* solver, const phx_dfsph_step_cache* cache,
                              const phx_sph_dfsph_state_tangent* state_tangent,
                              const phx_control_tangent* control_tangent,
                              const phx_sph_dfsph_parameter_tangent* parameter_tangent,
                              phx_sph_dfsph_state_tangent* next_state_tangent) {
    return guarded([&] {
        phx::require(solver && cache && state_tangent && control_tangent && parameter_tangent && next_state_tangent,
                     PHX_ERROR_INVALID_ARGUMENT, "null argument");
        phx::require(cache->owner == solver && cache->valid, PHX_ERROR_INVALID_ARGUMENT,
                     "step cache holds no step of this solver");
        solver->impl.jvp_step(cache->impl, *state_tangent, *control_tangent, *parameter_tangent,
                              *next_state_tangent);
    });
}

phx_status phx_dfsph_vjp_step(phx_dfsph* solver, const phx_dfsph_step_cache* cache,
                              const phx_sph_dfsph_state_adjoint* next_state_adjoint,
                              phx_sph_dfsph_state_adjoint* previous_state_adjoint,
                              phx_control_adjoint* control_adjoint,
                              phx_sph_dfsph_parameter_adjoint* parameter_adjoint) {
    return guarded([&] {
        phx::require(solver && cache && next_state_adjoint && previous_state_adjoint && control_adjoint &&
                         parameter_adjoint,
                     PHX_ERROR_INVALID_ARGUMENT, "null argument");
        phx::require(cache->owner == solver && cache->valid, PHX_ERROR_INVALID_ARGUMENT,
                     "step cache holds no step of this solver");
        solver->impl.vjp_step(cache->impl, *next_state_adjoint, *previous_state_adjoint, *control_adjoint,
                              *parameter_adjoint);
    });
}

phx_status phx_dfsph_mask_skin(const phx_dfsph* solver, float* skin) {
    return guarded([&] {
        phx::require(solver != nullptr && skin != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null argument");
        *skin = solver->impl.skin();
    });
}

phx_status phx_dfsph_set_profiling(phx_dfsph* solver, int enabled) {
    return guarded([&] {
        phx::require(solver != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null solver");
        solver->impl.profiler().reset();
        solver->impl.profiler().set_enabled(enabled != 0);
    });
}

phx_status phx_dfsph_read_profile(phx_dfsph* solver, phx_pass_profile* profile, uint32_t capacity, uint32_t* count) {
    return guarded([&] {
        phx::require(solver != nullptr && count != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null argument");
        phx::PassProfiler& p = solver->impl.profiler();
        p.collect();
        const auto& d = solver->impl.domain();
        uint32_t k = 0;
        for (int pass = 0; pass < phx::kDfsphPassCount; ++pass) {
            if (p.launches(pass) == 0) continue;
            if (profile != nullptr && k < capacity) {
                profile[k].name = phx::dfsph_pass_name(pass);
                profile[k].total_ms = p.total_ms(pass);
                profile[k].launches = p.launches(pass);
                profile[k].bytes_per_particle =
                    phx::dfsph_pass_bytes(pass, d.config.particle_count, d.cells, solver->impl.warm_start());
            }
            ++k;
        }
        *count = k;
        p.reset();
    });
}

// ---------------------------------------------------- WCSPH, PCISPH, IISPH
phx_status phx_sph_create(const phx_domain_configuration* domain, const phx_sph_configuration* configuration,
                          phx_execution_mode mode, phx_stream stream, phx_sph** solver) {
    return guarded([&] {
        phx::require(solver != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null solver output");
        *solver = nullptr;
        *solver = new phx_sph(domain, configuration, mode, stream_of(stream));
    });
}

phx_status phx_sph_destroy(phx_sph* solver) {
    return guarded([&] { delete solver; });
}

phx_status phx_sph_allocate_step_cache(phx_sph* solver, phx_sph_step_cache** cache) {
    return guarded([&] {
        phx::require(solver != nullptr && cache != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null argument");
        *cache = nullptr;
        auto* c = new phx_sph_step_cache();
        c->owner = solver;
        c->impl.stream = solver->impl.stream();
        c->impl.method = solver->impl.method();
        try {
            c->impl.allocate(solver->impl.domain().config.particle_count, solver->impl.domain().cells);
        } catch (...) {
            delete c;
            throw;
        }
        *cache = c;
    });
}

phx_status phx_sph_free_step_cache(phx_sph_step_cache* cache) {
    return guarded([&] { delete cache; });
}

phx_status phx_sph_forward_step(phx_sph* solver, const phx_sph_state* state, const phx_control* control,
                                const phx_sph_parameters* parameters, phx_sph_state* next_state,
                                phx_sph_step_cache* cache) {
    return guarded([&] {
        phx::require(solver && state && control && parameters && next_state && cache, PHX_ERROR_INVALID_ARGUMENT,
                     "null argument");
        phx::require(cache->owner == solver, PHX_ERROR_INVALID_ARGUMENT, "step cache belongs to another solver");
        cache->valid = false;
        solver->impl.forward_step(*state, *control, *parameters, *next_state, cache->impl);
        cache->valid = true;
    });
}

phx_status phx_sph_reference_gradient_norm(const phx_sph* solver, float* value) {
    return guarded([&] {
        phx::require(solver != nullptr && value != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null argument");
        *value = solver->impl.gradient_norm();
    });
}

phx_status phx_sph_last_iterations(const phx_sph* solver, uint32_t* iterations) {
    return guarded([&] {
        phx::require(solver != nullptr && iterations != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null argument");
        *iterations = solver->impl.last_iterations();
    });
}

phx_status phx_sph_set_profiling(phx_sph* solver, int enabled) {
    return guarded([&] {
        phx::require(solver != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null solver");
        solver->impl.profiler().reset();
        solver->impl.profiler().set_enabled(enabled != 0);
    });
}

phx_status phx_sph_read_profile(phx_sph* solver, phx_pass_profile* profile, uint32_t capacity, uint32_t* count) {
    return guarded([&] {
        phx::require(solver != nullptr && count != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null argument");
        phx::PassProfiler& p = solver->impl.profiler();
        p.collect();
        const auto& d = solver->impl.domain();
        uint32_t k = 0;
        for (int pass = 0; pass < phx::kSphPassCount; ++pass) {
            if (p.launches(pass) == 0) continue;
            if (profile != nullptr && k < capacity) {
                profile[k].name = phx::sph_pass_name(pass);
                profile[k].total_ms = p.total_ms(pass);
                profile[k].launches = p.launches(pass);
                profile[k].bytes_per_particle =
                    phx::sph_pass_bytes(solver->impl.method(), pass, d.config.particle_count, d.cells);
            }
            ++k;
        }
        *count = k;
        p.reset();
    });
}

phx_status phx_sph_cache_neighborhood(const phx_sph_step_cache* cache, uint64_t* sorted_keys,
                                      uint32_t* sorted_particle_indices, uint32_t* cell_offsets) {
    return guarded([&] {
        phx::require(cache != nullptr && cache->valid, PHX_ERROR_INVALID_ARGUMENT, "cache holds no step");
        phx::export_neighborhood(cache->impl.stream, cache->impl.grid(), sorted_keys, sorted_particle_indices,
                                 cell_offsets);
    });
}

phx_status phx_sph_cache_densities(const phx_sph_step_cache* cache, float* densities) {
    return guarded([&] {
        phx::require(cache != nullptr && cache->valid, PHX_ERROR_INVALID_ARGUMENT, "cache holds no step");
        phx::require(densities != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null output");
        phx::ChannelList ch;
        ch.add(cache->impl.density_data(), densities);
        phx::scatter_to_original(cache->impl.stream, cache->impl.grid().perm.data(), cache->impl.grid().particles,
                                 ch.get());
    });
}

phx_status phx_sph_cache_pressures(const phx_sph_step_cache* cache, float* pressures) {
    return guarded([&] {
        phx::require(cache != nullptr && cache->valid, PHX_ERROR_INVALID_ARGUMENT, "cache holds no step");
        phx::require(pressures != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null output");
        phx::ChannelList ch;
        ch.add(cache->impl.pressure_data(), pressures);
        phx::scatter_to_original(cache->impl.stream, cache->impl.grid().perm.data(), cache->impl.grid().particles,
                                 ch.get());
    });
}

phx_status phx_sph_cache_non_pressure_accelerations(const phx_sph_step_cache* cache, phx_vector_view out) {
    return guarded([&] {
        phx::require(cache != nullptr && cache->valid, PHX_ERROR_INVALID_ARGUMENT, "cache holds no step");
        phx::require(cache->impl.method != PHX_SPH_WCSPH, PHX_ERROR_INVALID_ARGUMENT,
                     "WCSPH keeps no non-pressure accelerations");
        phx::require_vec(out, "output");
        phx::ChannelList ch;
        const float* np = reinterpret_cast<const float*>(cache->impl.non_pressure_data());
        ch.add(np + 0, out.x, 4u);
        ch.add(np + 1, out.y, 4u);
        ch.add(np + 2, out.z, 4u);
        phx::scatter_to_original(cache->impl.stream, cache->impl.grid().perm.data(), cache->impl.grid().particles,
                                 ch.get());
    });
}

phx_status phx_sph_jvp_step(phx_sph* solver, const phx_sph_step_cache* cache,
                            const phx_sph_state_tangent* state_tangent, const phx_control_tangent* control_tangent,
                            const phx_sph_parameter_tangent* parameter_tangent,
                            phx_sph_state_tangent* next_state_tangent) {
    return guarded([&] {
        phx::require(solver && cache && state_tangent && control_tangent && parameter_tangent && next_state_tangent,
                     PHX_ERROR_INVALID_ARGUMENT, "null argument");
        phx::require(cache->owner == solver && cache->valid, PHX_ERROR_INVALID_ARGUMENT,
                     "step cache holds no step of this solver");
        solver->impl.jvp_step(cache->impl, *state_tangent, *control_tangent, *parameter_tangent,
                              *next_state_tangent);
    });
}

phx_status phx_sph_vjp_step(phx_sph* solver, const phx_sph_step_cache* cache,
                            const phx_sph_state_adjoint* next_state_adjoint,
                            phx_sph_state_adjoint* previous_state_adjoint, phx_control_adjoint* control_adjoint,
                            phx_sph_parameter_adjoint* parameter_adjoint) {
    return guarded([&] {
        phx::require(solver && cache && next_state_adjoint && previous_state_adjoint && control_adjoint &&
                         parameter_adjoint,
                     PHX_ERROR_INVALID_ARGUMENT, "null argument");
        phx::require(cache->owner == solver && cache->valid, PHX_ERROR_INVALID_ARGUMENT,
                     "step cache holds no step of this solver");
        solver->impl.vjp_step(cache->impl, *next_state_adjoint, *previous_state_adjoint, *control_adjoint,
                              *parameter_adjoint);
    });
}

// -------------------------------------------------------------------- PBF
phx_status phx_pbf_create(const phx_domain_configuration* domain, const phx_pbf_configuration* configuration,
                          phx_execution_mode mode, phx_stream stream, phx_pbf** solver) {
    return guarded([&] {
        phx::require(solver != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null solver output");
        *solver = nullptr;
        *solver = new phx_pbf(domain, configuration, mode, stream_of(stream));
    });
}

phx_status phx_pbf_destroy(phx_pbf* solver) {
    return guarded([&] { delete solver; });
}

phx_status phx_pbf_allocate_step_cache(phx_pbf* solver, phx_pbf_step_cache** cache) {
    return guarded([&] {
        phx::require(solver != nullptr && cache != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null argument");
        *cache = nullptr;
        auto* c = new phx_pbf_step_cache();
        c->owner = solver;
        c->impl.record = solver->impl.mode() == PHX_MODE_DIFFERENTIABLE;
        try {
            c->impl.allocate(solver->impl.domain().config.particle_count, solver->impl.domain().cells,
                             solver->impl.iterations());
        } catch (...) {
            delete c;
            throw;
        }
        *cache = c;
    });
}

phx_status phx_pbf_free_step_cache(phx_pbf_step_cache* cache) {
    return guarded([&] { delete cache; });
}

phx_status phx_pbf_forward_step(phx_pbf* solver, const phx_pbf_state* state, const phx_control* control,
                                const phx_pbf_parameters* parameters, phx_pbf_state* next_state,
                                phx_pbf_step_cache* cache) {
    return guarded([&] {
        phx::require(solver && state && control && parameters && next_state && cache, PHX_ERROR_INVALID_ARGUMENT,
                     "null argument");
        phx::require(cache->owner == solver, PHX_ERROR_INVALID_ARGUMENT, "step cache belongs to another solver");
        cache->valid = false;
        solver->impl.forward_step(*state, *control, *parameters, *next_state, cache->impl);
        cache->valid = true;
    });
}

phx_status phx_pbf_jvp_step(phx_pbf* solver, const phx_pbf_parameters* parameters, const phx_pbf_step_cache* cache,
                            const phx_pbf_state_tangent* state_tangent, const phx_control_tangent* control_tangent,
                            const phx_pbf_parameter_tangent* parameter_tangent,
                            phx_pbf_state_tangent* next_state_tangent) {
    return guarded([&] {
        phx::require(solver && parameters && cache && state_tangent && control_tangent && parameter_tangent &&
                         next_state_tangent,
                     PHX_ERROR_INVALID_ARGUMENT, "null argument");
        phx::require(cache->owner == solver && cache->valid, PHX_ERROR_INVALID_ARGUMENT,
                     "step cache holds no step of this solver");
        solver->impl.jvp_step(*parameters, cache->impl, *state_tangent, *control_tangent, *parameter_tangent,
                              *next_state_tangent);
    });
}

phx_status phx_pbf_vjp_step(phx_pbf* solver, const phx_pbf_parameters* parameters, const phx_pbf_step_cache* cache,
                            const phx_pbf_state_adjoint* next_state_adjoint,
                            phx_pbf_state_adjoint* previous_state_adjoint, phx_control_adjoint* control_adjoint,
                            phx_pbf_parameter_adjoint* parameter_adjoint) {
    return guarded([&] {
        phx::require(solver && parameters && cache && next_state_adjoint && previous_state_adjoint && control_adjoint &&
                         parameter_adjoint,
                     PHX_ERROR_INVALID_ARGUMENT, "null argument");
        phx::require(cache->owner == solver && cache->valid, PHX_ERROR_INVALID_ARGUMENT,
                     "step cache holds no step of this solver");
        solver->impl.vjp_step(*parameters, cache->impl, *next_state_adjoint, *previous_state_adjoint,
                              *control_adjoint, *parameter_adjoint);
    });
}

phx_status phx_pbf_set_profiling(phx_pbf* solver, int enabled) {
    return guarded([&] {
        phx::require(solver != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null solver");
        solver->impl.profiler().reset();
        solver->impl.profiler().set_enabled(enabled != 0);
    });
}

phx_status phx_pbf_read_profile(phx_pbf* solver, phx_pass_profile* profile, uint32_t capacity, uint32_t* count) {
    return guarded([&] {
        phx::require(solver != nullptr && count != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null argument");
        phx::PassProfiler& p = solver->impl.profiler();
        p.collect();
        const auto& d = solver->impl.domain();
        uint32_t k = 0;
        for (int pass = 0; pass < phx::kPbfPassCount; ++pass) {
            if (p.launches(pass) == 0) continue;
            if (profile != nullptr && k < capacity) {
                profile[k].name = phx::pbf_pass_name(pass);
                profile[k].total_ms = p.total_ms(pass);
                profile[k].launches = p.launches(pass);
                profile[k].bytes_per_particle = phx::pbf_pass_bytes(pass, d.config.particle_count, d.cells);
            }
            ++k;
        }
        *count = k;
        p.reset();
    });
}

phx_status phx_pbf_cache_neighborhood(const phx_pbf_step_cache* cache, uint64_t* sorted_keys,
                                      uint32_t* sorted_particle_indices, uint32_t* cell_offsets) {
    return guarded([&] {
        phx::require(cache != nullptr && cache->valid, PHX_ERROR_INVALID_ARGUMENT, "cache holds no step");
        phx::export_neighborhood(cache->owner->impl.stream(), cache->impl.neighborhood, sorted_keys,
                                 sorted_particle_indices, cell_offsets);
    });
}

phx_status phx_pbf_cache_predicted_positions(const phx_pbf_step_cache* cache, phx_vector_view out) {
    return guarded([&] {
        phx::require(cache != nullptr && cache->valid, PHX_ERROR_INVALID_ARGUMENT, "cache holds no step");
        phx::require_vec(out, "output");
        const size_t bytes = sizeof(float) * cache->impl.neighborhood.particles;
        const hipStream_t s = cache->owner->impl.stream();
        PHX_HIP(hipMemcpyAsync(out.x, cache->impl.predicted.x.data(), bytes, hipMemcpyDeviceToDevice, s));
        PHX_HIP(hipMemcpyAsync(out.y, cache->impl.predicted.y.data(), bytes, hipMemcpyDeviceToDevice, s));
        PHX_HIP(hipMemcpyAsync(out.z, cache->impl.predicted.z.data(), bytes, hipMemcpyDeviceToDevice, s));
    });
}

phx_status phx_pbf_cache_last_iteration(const phx_pbf_step_cache* cache, float* densities, float* lambdas) {
    return guarded([&] {
        phx::require(cache != nullptr && cache->valid, PHX_ERROR_INVALID_ARGUMENT, "cache holds no step");
        phx::ChannelList ch;
        if (densities != nullptr) ch.add(cache->impl.densities.data(), densities);
        if (lambdas != nullptr) ch.add(cache->impl.lambdas.data(), lambdas);
        phx::scatter_to_original(cache->owner->impl.stream(), cache->impl.neighborhood.perm.data(),
                                 cache->impl.neighborhood.particles, ch.get());
    });
}

phx_status phx_pbf_cache_vorticity_magnitudes(const phx_pbf_step_cache* cache, float* magnitudes) {
    return guarded([&] {
        phx::require(cache != nullptr && cache->valid, PHX_ERROR_INVALID_ARGUMENT, "cache holds no step");
        phx::require(magnitudes != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null magnitudes");
        phx::ChannelList ch;
        ch.add(cache->impl.magnitudes.data(), magnitudes);
        phx::scatter_to_original(cache->owner->impl.stream(), cache->impl.neighborhood.perm.data(),
                                 cache->impl.neighborhood.particles, ch.get());
    });
}

phx_status phx_pbf_batch_create(const phx_domain_configuration* domains, uint32_t scene_count,
                                const phx_pbf_configuration* configuration,
                                const phx_pbf_batch_configuration* batch_configuration, phx_stream stream,
                                phx_pbf_batch** batch) {
    return guarded([&] {
        phx::require(batch != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null batch handle");
        *batch = nullptr;
        *batch = new phx_pbf_batch(domains, scene_count, configuration, batch_configuration, stream_of(stream));
    });
}

phx_status phx_pbf_batch_destroy(phx_pbf_batch* batch) {
    return guarded([&] { delete batch; });
}

phx_status phx_pbf_batch_particle_offsets(const phx_pbf_batch* batch, uint64_t* offsets) {
    return guarded([&] {
        phx::require(batch != nullptr && offsets != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null argument");
        const auto& o = batch->impl.offsets();
        std::memcpy(offsets, o.data(), sizeof(uint64_t) * o.size());
    });
}

phx_status phx_pbf_batch_run(phx_pbf_batch* batch, phx_pbf_state* state, const phx_control* control,
                             const phx_pbf_parameters* parameters, uint32_t steps) {
    return guarded([&] {
        phx::require(batch != nullptr && state != nullptr && control != nullptr && parameters != nullptr,
                     PHX_ERROR_INVALID_ARGUMENT, "null argument");
        batch->impl.run(*state, *control, *parameters, steps);
    });
}

phx_status phx_pbf_batch_kernels_per_step(const phx_pbf_batch* batch, uint32_t* kernels) {
    return guarded([&] {
        phx::require(batch != nullptr && kernels != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null argument");
        *kernels = batch->impl.kernels_per_step();
    });
}

}  // extern "C"
