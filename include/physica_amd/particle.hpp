// C++ host API over the physica_amd C ABI, mirroring the reference module
// physica.fluids.liquid.particle (fluids/liquid/particle/*.ixx): the same
// type and member names, the same "allocate_*, forward_step(state, control,
// parameters, next_state, cache)" flow, HIP device buffers in place of
// ::cuda::device_buffer and hipStream_t in place of ::cuda::stream_ref.
// Errors surface as physica_amd::Error exceptions (the reference's CCCL
// buffers throw on failure as well); the C ABI below never throws.
//
// Covered: Domain (with the tangent and adjoint fields), NeighborhoodSearch,
// density::{sph,pbf}_{forward,jvp,vjp}, SPH<WCSPH|PCISPH|IISPH|DFSPH> and PBF with
// forward_step, jvp_step and vjp_step. The solvers keep their own
// neighbourhood, densities and per-method cache in the C ABI's step cache;
// jvp_step/vjp_step read the step the cache recorded, so their state,
// control, parameter and next-state arguments are accepted for the
// reference's signature only (PBF::jvp/vjp_step read the parameters).
//
// Header-only; link with libphysica_amd.so and amdhip64.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstddef>
#include <cstdint>
#include <span>
#include <type_traits>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "physica_amd/physica.h"

namespace physica_amd {

class Error : public std::runtime_error {
  public:
    Error(phx_status status, const std::string& message) : std::runtime_error(message), status_(status) {}
    phx_status status() const { return status_; }

  private:
    phx_status status_;
};

inline void check(phx_status status, const char* what) {
    if (status != PHX_SUCCESS) throw Error(status, std::string(what) + ": " + phx_last_error());
}

inline void check_hip(hipError_t status, const char* what) {
    if (status != hipSuccess)
        throw Error(PHX_ERROR_DEVICE, std::string(what) + ": " + hipGetErrorString(status));
}

// Device allocation owned by the caller (stands in for ::cuda::device_buffer;
// uninitialised, as the reference's ::cuda::no_init buffers).
template <class T>
class device_buffer {
  public:
    device_buffer() = default;
    device_buffer(hipStream_t stream, std::size_t count) : stream_(stream), count_(count) {
        if (count_ > 0) check_hip(hipMalloc(reinterpret_cast<void**>(&data_), count_ * sizeof(T)), "hipMalloc");
    }
    device_buffer(const device_buffer&) = delete;
    device_buffer& operator=(const device_buffer&) = delete;
    device_buffer(device_buffer&& o) noexcept { swap(o); }
    device_buffer& operator=(device_buffer&& o) noexcept {
        swap(o);
        return *this;
    }
    ~device_buffer() {
        if (data_ != nullptr) (void)hipFree(data_);
    }
    T* data() { return data_; }
    const T* data() const { return data_; }
    std::size_t size() const { return count_; }
    void swap(device_buffer& o) noexcept {
        std::swap(data_, o.data_);
        std::swap(count_, o.count_);
        std::swap(stream_, o.stream_);
    }

  private:
    T* data_ = nullptr;
    std::size_t count_ = 0;
    hipStream_t stream_ = nullptr;
};

namespace fluids::liquid::particle {

// domain/domain.ixx:12-44
struct Vector3 {
    float x;
    float y;
    float z;
};

struct BoxBoundary {
    Vector3 minimum;
    Vector3 maximum;
    Vector3 velocity;
    bool no_slip{true};
};

struct BoundaryParticle {
    Vector3 position;
    Vector3 velocity;
    float volume;
};

struct DomainConfiguration {
    std::uint32_t particle_count;
    float time_step;
    float support_radius;
    float particle_radius;
    Vector3 gravity;
    BoxBoundary boundary;
    std::vector<BoundaryParticle> boundary_particles;

    phx_domain_configuration to_c() const {
        phx_domain_configuration c{};
        c.particle_count = particle_count;
        c.time_step = time_step;
        c.support_radius = support_radius;
        c.particle_radius = particle_radius;
        c.gravity = {gravity.x, gravity.y, gravity.z};
        c.boundary.minimum = {boundary.minimum.x, boundary.minimum.y, boundary.minimum.z};
        c.boundary.maximum = {boundary.maximum.x, boundary.maximum.y, boundary.maximum.z};
        c.boundary.velocity = {boundary.velocity.x, boundary.velocity.y, boundary.velocity.z};
        c.boundary.no_slip = boundary.no_slip ? 1u : 0u;
        c.boundary_particles = reinterpret_cast<const phx_boundary_particle*>(boundary_particles.data());
        c.boundary_particle_count = static_cast<std::uint32_t>(boundary_particles.size());
        return c;
    }
};
static_assert(sizeof(BoundaryParticle) == sizeof(phx_boundary_particle));

enum class ExecutionMode : std::uint32_t { forward, differentiable };

// domain/domain.ixx:46-119
struct ScalarField {
    device_buffer<float> values;
};

struct VectorField {
    device_buffer<float> x;
    device_buffer<float> y;
    device_buffer<float> z;

    phx_vector_view view() { return {x.data(), y.data(), z.data()}; }
    phx_const_vector_view const_view() const { return {x.data(), y.data(), z.data()}; }
};

struct ScalarAdjointField {
    device_buffer<double> values;
};

struct VectorAdjointField {
    device_buffer<double> x;
    device_buffer<double> y;
    device_buffer<double> z;

    phx_adjoint_vector_view view() { return {x.data(), y.data(), z.data()}; }
    phx_const_adjoint_vector_view const_view() const { return {x.data(), y.data(), z.data()}; }
};

struct ParticleState {
    VectorField positions;
    VectorField velocities;
    std::uint64_t step_index{};
};

struct Control {
    VectorField external_accelerations;
};

// domain/domain.ixx:76-90
struct ParticleStateTangent {
    VectorField positions;
    VectorField velocities;
};

struct ControlTangent {
    VectorField external_accelerations;
};

struct ParticleStateAdjoint {
    VectorAdjointField positions;
    VectorAdjointField velocities;
};

struct ControlAdjoint {
    VectorAdjointField external_accelerations;
};

struct ParticleParameters {
    device_buffer<float> masses;
    device_buffer<float> rest_densities;
    device_buffer<float> viscosities;
    device_buffer<float> surface_tensions;

    phx_particle_parameters view() const {
        return {masses.data(), rest_densities.data(), viscosities.data(), surface_tensions.data()};
    }
};

// domain/domain.ixx:99-113
struct ParticleParameterTangent {
    device_buffer<float> masses;
    device_buffer<float> rest_densities;
    device_buffer<float> viscosities;
    device_buffer<float> surface_tensions;

    phx_particle_parameter_tangent view() const {
        return {masses.data(), rest_densities.data(), viscosities.data(), surface_tensions.data()};
    }
};

struct ParticleParameterAdjoint {
    device_buffer<double> masses;
    device_buffer<double> rest_densities;
    device_buffer<double> viscosities;
    device_buffer<double> surface_tensions;

    phx_particle_parameter_adjoint view() {
        return {masses.data(), rest_densities.data(), viscosities.data(), surface_tensions.data()};
    }
};

// Allocation, clear, copy and upload (Domain, domain/domain.ixx:121-165).
class Domain {
  public:
    Domain(DomainConfiguration configuration, hipStream_t stream)
        : configuration(std::move(configuration)), stream(stream) {}

    const DomainConfiguration configuration;
    hipStream_t stream;

    ScalarField allocate_scalar_field(std::size_t count) const { return {device_buffer<float>(stream, count)}; }
    VectorField allocate_vector_field(std::size_t count) const {
        return {device_buffer<float>(stream, count), device_buffer<float>(stream, count),
                device_buffer<float>(stream, count)};
    }
    ParticleState allocate_particle_state() const {
        ParticleState s{allocate_vector_field(n()), allocate_vector_field(n()), 0};
        clear(s.positions);
        clear(s.velocities);
        return s;
    }
    Control allocate_control() const {
        Control c{allocate_vector_field(n())};
        clear(c.external_accelerations);
        return c;
    }
    ParticleParameters allocate_particle_parameters() const {
        return {device_buffer<float>(stream, n()), device_buffer<float>(stream, n()), device_buffer<float>(stream, n()),
                device_buffer<float>(stream, n())};
    }
    ScalarAdjointField allocate_scalar_adjoint_field(std::size_t count) const {
        return {device_buffer<double>(stream, count)};
    }
    VectorAdjointField allocate_vector_adjoint_field(std::size_t count) const {
        return {device_buffer<double>(stream, count), device_buffer<double>(stream, count),
                device_buffer<double>(stream, count)};
    }
    // tangents and adjoints start cleared
    ParticleStateTangent allocate_particle_state_tangent() const {
        ParticleStateTangent t{allocate_vector_field(n()), allocate_vector_field(n())};
        clear(t.positions);
        clear(t.velocities);
        return t;
    }
    ControlTangent allocate_control_tangent() const {
        ControlTangent t{allocate_vector_field(n())};
        clear(t.external_accelerations);
        return t;
    }
    ParticleStateAdjoint allocate_particle_state_adjoint() const {
        ParticleStateAdjoint a{allocate_vector_adjoint_field(n()), allocate_vector_adjoint_field(n())};
        clear(a.positions);
        clear(a.velocities);
        return a;
    }
    ControlAdjoint allocate_control_adjoint() const {
        ControlAdjoint a{allocate_vector_adjoint_field(n())};
        clear(a.external_accelerations);
        return a;
    }
    ParticleParameterTangent allocate_particle_parameter_tangent() const {
        ParticleParameterTangent t{device_buffer<float>(stream, n()), device_buffer<float>(stream, n()),
                                   device_buffer<float>(stream, n()), device_buffer<float>(stream, n())};
        clear(t);
        return t;
    }
    ParticleParameterAdjoint allocate_particle_parameter_adjoint() const {
        ParticleParameterAdjoint a{device_buffer<double>(stream, n()), device_buffer<double>(stream, n()),
                                   device_buffer<double>(stream, n()), device_buffer<double>(stream, n())};
        clear(a);
        return a;
    }
    void clear(ScalarField& f) const { clear(f.values); }
    void clear(ScalarAdjointField& f) const { clear(f.values); }
    void clear(VectorField& f) const {
        clear(f.x);
        clear(f.y);
        clear(f.z);
    }
    void clear(VectorAdjointField& f) const {
        clear(f.x);
        clear(f.y);
        clear(f.z);
    }
    void clear(ParticleParameterTangent& t) const {
        clear(t.masses);
        clear(t.rest_densities);
        clear(t.viscosities);
        clear(t.surface_tensions);
    }
    void clear(ParticleParameterAdjoint& a) const {
        clear(a.masses);
        clear(a.rest_densities);
        clear(a.viscosities);
        clear(a.surface_tensions);
    }
    template <class T>
    void clear(device_buffer<T>& b) const {
        check_hip(hipMemsetAsync(b.data(), 0, b.size() * sizeof(T), stream), "hipMemsetAsync");
    }
    void copy(const VectorAdjointField& source, VectorAdjointField& destination) const {
        copy(source.x, destination.x);
        copy(source.y, destination.y);
        copy(source.z, destination.z);
    }
    void copy(const ScalarAdjointField& source, ScalarAdjointField& destination) const {
        copy(source.values, destination.values);
    }
    void copy(const ParticleStateTangent& source, ParticleStateTangent& destination) const {
        copy(source.positions, destination.positions);
        copy(source.velocities, destination.velocities);
    }
    void copy(const ParticleStateAdjoint& source, ParticleStateAdjoint& destination) const {
        copy(source.positions, destination.positions);
        copy(source.velocities, destination.velocities);
    }
    void copy(const device_buffer<double>& source, device_buffer<double>& destination) const {
        check_hip(hipMemcpyAsync(destination.data(), source.data(), source.size() * sizeof(double),
                                 hipMemcpyDeviceToDevice, stream),
                  "hipMemcpyAsync");
    }
    // domain/domain.cpp:189-200 (destination += source)
    void accumulate(const ScalarAdjointField& source, ScalarAdjointField& destination) const {
        accumulate(source.values, destination.values);
    }
    void accumulate(const VectorAdjointField& source, VectorAdjointField& destination) const {
        accumulate(source.x, destination.x);
        accumulate(source.y, destination.y);
        accumulate(source.z, destination.z);
    }
    void accumulate(const ParticleStateAdjoint& source, ParticleStateAdjoint& destination) const {
        accumulate(source.positions, destination.positions);
        accumulate(source.velocities, destination.velocities);
    }
    void accumulate(const device_buffer<double>& source, device_buffer<double>& destination) const {
        check(phx_accumulate(stream, source.data(), destination.data(), source.size()), "accumulate");
    }
    void copy(const VectorField& source, VectorField& destination) const {
        copy(source.x, destination.x);
        copy(source.y, destination.y);
        copy(source.z, destination.z);
    }
    void copy(const ScalarField& source, ScalarField& destination) const { copy(source.values, destination.values); }
    void copy(const ParticleState& source, ParticleState& destination) const {
        copy(source.positions, destination.positions);
        copy(source.velocities, destination.velocities);
        destination.step_index = source.step_index;
    }
    void copy(const device_buffer<float>& source, device_buffer<float>& destination) const {
        check_hip(hipMemcpyAsync(destination.data(), source.data(), source.size() * sizeof(float),
                                 hipMemcpyDeviceToDevice, stream),
                  "hipMemcpyAsync");
    }
    void upload(std::span<const float> source, ScalarField& destination) const { upload(source, destination.values); }
    void upload(std::span<const float> source, device_buffer<float>& destination) const {
        check_hip(hipMemcpyAsync(destination.data(), source.data(), source.size() * sizeof(float), hipMemcpyHostToDevice,
                                 stream),
                  "hipMemcpyAsync");
    }
    // domain.cpp:206-218: AoS host vectors to SoA device fields
    void upload(std::span<const Vector3> source, VectorField& destination) const {
        std::vector<float> x(source.size()), y(source.size()), z(source.size());
        for (std::size_t k = 0; k < source.size(); ++k) {
            x[k] = source[k].x;
            y[k] = source[k].y;
            z[k] = source[k].z;
        }
        upload(x, destination.x);
        upload(y, destination.y);
        upload(z, destination.z);
        check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }

  private:
    std::size_t n() const { return configuration.particle_count; }
};

// DeviceBoundary (domain/domain.ixx:115-119): the configuration's boundary
// samples on the device, original order.
struct DeviceBoundary {
    VectorField positions;
    VectorField velocities;
    ScalarField volumes;

    DeviceBoundary(const Domain& domain)
        : positions(domain.allocate_vector_field(domain.configuration.boundary_particles.size())),
          velocities(domain.allocate_vector_field(domain.configuration.boundary_particles.size())),
          volumes(domain.allocate_scalar_field(domain.configuration.boundary_particles.size())) {
        const auto& samples = domain.configuration.boundary_particles;
        if (samples.empty()) return;
        std::vector<Vector3> x, v;
        std::vector<float> volume;
        for (const BoundaryParticle& b : samples) {
            x.push_back(b.position);
            v.push_back(b.velocity);
            volume.push_back(b.volume);
        }
        domain.upload(x, positions);
        domain.upload(v, velocities);
        domain.upload(volume, volumes);
        check_hip(hipStreamSynchronize(domain.stream), "hipStreamSynchronize");
    }
    // BoundaryView (neighborhood/interop.h:29-41) at the neighbourhood's time
    phx_boundary_view view(float time) const {
        return {positions.x.data(), positions.y.data(), positions.z.data(), velocities.x.data(), velocities.y.data(),
                velocities.z.data(), volumes.values.data(), static_cast<std::uint32_t>(volumes.values.size()), time};
    }
};

// neighborhood/neighborhood.ixx:12-23
struct Neighborhood {
    device_buffer<std::uint64_t> sorted_keys;
    device_buffer<std::uint32_t> sorted_particle_indices;
    device_buffer<std::uint32_t> cell_offsets;
    device_buffer<std::uint64_t> sorted_boundary_keys;
    device_buffer<std::uint32_t> sorted_boundary_indices;
    device_buffer<std::uint32_t> boundary_cell_offsets;
    std::array<std::uint32_t, 3u> cell_resolution{};
    Vector3 cell_origin{};
    float cell_size{};
    float boundary_time{};

    phx_neighborhood_view view() const {
        phx_neighborhood_view v{};
        v.sorted_keys = sorted_keys.data();
        v.sorted_particle_indices = sorted_particle_indices.data();
        v.cell_offsets = cell_offsets.data();
        v.sorted_boundary_keys = sorted_boundary_keys.data();
        v.sorted_boundary_indices = sorted_boundary_indices.data();
        v.boundary_cell_offsets = boundary_cell_offsets.data();
        v.particle_count = static_cast<std::uint32_t>(sorted_particle_indices.size());
        v.boundary_count = static_cast<std::uint32_t>(sorted_boundary_indices.size());
        v.cells_x = cell_resolution[0];
        v.cells_y = cell_resolution[1];
        v.cells_z = cell_resolution[2];
        v.origin_x = cell_origin.x;
        v.origin_y = cell_origin.y;
        v.origin_z = cell_origin.z;
        v.cell_size = cell_size;
        return v;
    }
};

// NeighborhoodSearch (neighborhood/neighborhood.ixx:25-44) over
// phx_build_neighborhood; the workspace replaces the reference's unsorted
// key buffers and sort scratch.
class NeighborhoodSearch {
  public:
    explicit NeighborhoodSearch(const Domain& domain) : domain_(domain), configuration_(domain.configuration.to_c()) {
        std::size_t bytes = 0;
        check(phx_neighborhood_workspace_size(&configuration_, &bytes), "phx_neighborhood_workspace_size");
        workspace_ = device_buffer<std::uint8_t>(domain.stream, bytes);
    }
    NeighborhoodSearch(const NeighborhoodSearch&) = delete;
    NeighborhoodSearch& operator=(const NeighborhoodSearch&) = delete;

    Neighborhood allocate() const {
        std::uint32_t resolution[3];
        check(phx_grid_resolution(&configuration_, resolution), "phx_grid_resolution");
        const std::size_t cells = static_cast<std::size_t>(resolution[0]) * resolution[1] * resolution[2];
        const std::size_t n = domain_.configuration.particle_count;
        const std::size_t b = domain_.configuration.boundary_particles.size();
        hipStream_t s = domain_.stream;
        Neighborhood nb{device_buffer<std::uint64_t>(s, n), device_buffer<std::uint32_t>(s, n),
                        device_buffer<std::uint32_t>(s, cells + 1u), device_buffer<std::uint64_t>(s, b),
                        device_buffer<std::uint32_t>(s, b), device_buffer<std::uint32_t>(s, cells + 1u)};
        nb.cell_resolution = {resolution[0], resolution[1], resolution[2]};
        return nb;
    }
    // NeighborhoodSearch::build (neighborhood/neighborhood.cpp:44-89)
    void build(std::uint64_t step_index, const VectorField& positions, Neighborhood& neighborhood) {
        phx_neighborhood_view v{};
        check(phx_build_neighborhood(domain_.stream, &configuration_, step_index, positions.const_view(),
                                     workspace_.data(), workspace_.size(), neighborhood.sorted_keys.data(),
                                     neighborhood.sorted_particle_indices.data(), neighborhood.cell_offsets.data(),
                                     neighborhood.sorted_boundary_keys.data(),
                                     neighborhood.sorted_boundary_indices.data(),
                                     neighborhood.boundary_cell_offsets.data(), &v),
              "NeighborhoodSearch::build");
        neighborhood.cell_resolution = {v.cells_x, v.cells_y, v.cells_z};
        neighborhood.cell_origin = {v.origin_x, v.origin_y, v.origin_z};
        neighborhood.cell_size = v.cell_size;
        neighborhood.boundary_time = static_cast<float>(step_index) * domain_.configuration.time_step;
    }

  private:
    const Domain& domain_;
    phx_domain_configuration configuration_;
    device_buffer<std::uint8_t> workspace_;
};

// density::sph_forward / pbf_forward (density/density.ixx:6-13); the
// boundary samples enter through `boundary` (null: the domain has none).
namespace density {
inline void forward(const Domain& domain, bool pbf_kernel, const VectorField& topology_positions,
                    const VectorField& positions, const ParticleParameters& parameters,
                    const Neighborhood& neighborhood, ScalarField& densities, const DeviceBoundary* boundary) {
    const phx_boundary_view b = boundary != nullptr ? boundary->view(neighborhood.boundary_time) : phx_boundary_view{};
    check(phx_density_forward(domain.stream, domain.configuration.particle_count, domain.configuration.support_radius,
                              pbf_kernel ? 1 : 0, topology_positions.const_view(), positions.const_view(),
                              parameters.view(), neighborhood.view(), b, densities.values.data()),
          "density::forward");
}
inline void sph_forward(const Domain& domain, const VectorField& topology_positions, const VectorField& positions,
                        const ParticleParameters& parameters, const Neighborhood& neighborhood, ScalarField& densities,
                        const DeviceBoundary* boundary = nullptr) {
    forward(domain, false, topology_positions, positions, parameters, neighborhood, densities, boundary);
}
inline void pbf_forward(const Domain& domain, const VectorField& topology_positions, const VectorField& positions,
                        const ParticleParameters& parameters, const Neighborhood& neighborhood, ScalarField& densities,
                        const DeviceBoundary* boundary = nullptr) {
    forward(domain, true, topology_positions, positions, parameters, neighborhood, densities, boundary);
}
// density::{sph,pbf}_jvp (density.cu:42-63): overwrites density_tangent
inline void jvp(const Domain& domain, bool pbf_kernel, const VectorField& topology_positions,
                const VectorField& positions, const VectorField& position_tangent, const ParticleParameters& parameters,
                const ParticleParameterTangent& parameter_tangent, const Neighborhood& neighborhood,
                ScalarField& density_tangent, const DeviceBoundary* boundary) {
    const phx_boundary_view b = boundary != nullptr ? boundary->view(neighborhood.boundary_time) : phx_boundary_view{};
    check(phx_density_jvp(domain.stream, domain.configuration.particle_count, domain.configuration.support_radius,
                          pbf_kernel ? 1 : 0, topology_positions.const_view(), positions.const_view(),
                          position_tangent.const_view(), parameters.view(), parameter_tangent.view(),
                          neighborhood.view(), b, density_tangent.values.data()),
          "density::jvp");
}
// density::{sph,pbf}_vjp (density.cu:65-98): accumulates (+=) the adjoints
inline void vjp(const Domain& domain, bool pbf_kernel, const VectorField& topology_positions,
                const VectorField& positions, const ParticleParameters& parameters, const Neighborhood& neighborhood,
                const ScalarAdjointField& density_adjoint, VectorAdjointField& position_adjoint,
                ParticleParameterAdjoint& parameter_adjoint, const DeviceBoundary* boundary) {
    const phx_boundary_view b = boundary != nullptr ? boundary->view(neighborhood.boundary_time) : phx_boundary_view{};
    check(phx_density_vjp(domain.stream, domain.configuration.particle_count, domain.configuration.support_radius,
                          pbf_kernel ? 1 : 0, topology_positions.const_view(), positions.const_view(), parameters.view(),
                          neighborhood.view(), b, density_adjoint.values.data(), position_adjoint.view(),
                          parameter_adjoint.view()),
          "density::vjp");
}
inline void sph_jvp(const Domain& domain, const VectorField& topology_positions, const VectorField& positions,
                    const VectorField& position_tangent, const ParticleParameters& parameters,
                    const ParticleParameterTangent& parameter_tangent, const Neighborhood& neighborhood,
                    ScalarField& density_tangent, const DeviceBoundary* boundary = nullptr) {
    jvp(domain, false, topology_positions, positions, position_tangent, parameters, parameter_tangent, neighborhood,
        density_tangent, boundary);
}
inline void pbf_jvp(const Domain& domain, const VectorField& topology_positions, const VectorField& positions,
                    const VectorField& position_tangent, const ParticleParameters& parameters,
                    const ParticleParameterTangent& parameter_tangent, const Neighborhood& neighborhood,
                    ScalarField& density_tangent, const DeviceBoundary* boundary = nullptr) {
    jvp(domain, true, topology_positions, positions, position_tangent, parameters, parameter_tangent, neighborhood,
        density_tangent, boundary);
}
inline void sph_vjp(const Domain& domain, const VectorField& topology_positions, const VectorField& positions,
                    const ParticleParameters& parameters, const Neighborhood& neighborhood,
                    const ScalarAdjointField& density_adjoint, VectorAdjointField& position_adjoint,
                    ParticleParameterAdjoint& parameter_adjoint, const DeviceBoundary* boundary = nullptr) {
    vjp(domain, false, topology_positions, positions, parameters, neighborhood, density_adjoint, position_adjoint,
        parameter_adjoint, boundary);
}
inline void pbf_vjp(const Domain& domain, const VectorField& topology_positions, const VectorField& positions,
                    const ParticleParameters& parameters, const Neighborhood& neighborhood,
                    const ScalarAdjointField& density_adjoint, VectorAdjointField& position_adjoint,
                    ParticleParameterAdjoint& parameter_adjoint, const DeviceBoundary* boundary = nullptr) {
    vjp(domain, true, topology_positions, positions, parameters, neighborhood, density_adjoint, position_adjoint,
        parameter_adjoint, boundary);
}
}  // namespace density

// sph/sph.ixx:39-193: the method descriptions. States of the three are empty.
struct WCSPH {
    struct Configuration {
        // VJP of a no-slip box collision: false (default) the reference's
        // (the whole velocity adjoint zeroed), true the exact transpose
        bool exact_no_slip_adjoint{false};
    };
    struct State {};
    struct StateTangent {};
    struct StateAdjoint {};
    struct Parameters {
        device_buffer<float> speed_of_sound;
        device_buffer<float> tait_exponent;
        device_buffer<float> boundary_surface_tension;
    };
    struct ParameterTangent {
        device_buffer<float> speed_of_sound;
        device_buffer<float> tait_exponent;
        device_buffer<float> boundary_surface_tension;
    };
    struct ParameterAdjoint {
        device_buffer<double> speed_of_sound;
        device_buffer<double> tait_exponent;
        device_buffer<double> boundary_surface_tension;
    };
};

struct PCISPH {
    // density_error_tolerance > 0: tolerance mode (an extension)
    struct Configuration {
        std::uint32_t pressure_iterations{6u};
        std::uint32_t checkpoint_interval{2u};
        float density_error_tolerance{0.0F};
        std::uint32_t max_iterations{100u};
        // VJP of a no-slip box collision: false (default) the reference's
        // (the whole velocity adjoint zeroed), true the exact transpose
        bool exact_no_slip_adjoint{false};
    };
    struct State {};
    struct StateTangent {};
    struct StateAdjoint {};
    struct Parameters {
        device_buffer<float> pressure_relaxation;
    };
    struct ParameterTangent {
        device_buffer<float> pressure_relaxation;
    };
    struct ParameterAdjoint {
        device_buffer<double> pressure_relaxation;
    };
};

struct IISPH {
    // density_error_tolerance > 0: tolerance mode (an extension)
    struct Configuration {
        std::uint32_t pressure_iterations{6u};
        std::uint32_t checkpoint_interval{2u};
        float density_error_tolerance{0.0F};
        std::uint32_t max_iterations{100u};
        // VJP of a no-slip box collision: false (default) the reference's
        // (the whole velocity adjoint zeroed), true the exact transpose
        bool exact_no_slip_adjoint{false};
    };
    struct State {};
    struct StateTangent {};
    struct StateAdjoint {};
    struct Parameters {
        device_buffer<float> jacobi_relaxation;
    };
    struct ParameterTangent {
        device_buffer<float> jacobi_relaxation;
    };
    struct ParameterAdjoint {
        device_buffer<double> jacobi_relaxation;
    };
};

// sph/sph.ixx:195-268. density_error_tolerance > 0 selects tolerance mode
// (an extension: each phase iterates until the FP64 mean compression error
// is within it, at most max_iterations times).
struct DFSPH {
    struct Configuration {
        std::uint32_t divergence_iterations{4u};
        std::uint32_t density_iterations{6u};
        std::uint32_t checkpoint_interval{2u};
        bool pressure_warm_start{true};
        float density_error_tolerance{0.0F};
        std::uint32_t max_iterations{100u};
        // VJP of a no-slip box collision: false (default) the reference's
        // (the whole velocity adjoint zeroed), true the exact transpose
        bool exact_no_slip_adjoint{false};
    };
    struct State {
        ScalarField warm_divergence_pressure;
        ScalarField warm_density_pressure;
    };
    struct StateTangent {
        ScalarField warm_divergence_pressure;
        ScalarField warm_density_pressure;
    };
    struct StateAdjoint {
        ScalarAdjointField warm_divergence_pressure;
        ScalarAdjointField warm_density_pressure;
    };
    struct Parameters {
        device_buffer<float> divergence_relaxation;
        device_buffer<float> density_relaxation;
    };
    struct ParameterTangent {
        device_buffer<float> divergence_relaxation;
        device_buffer<float> density_relaxation;
    };
    struct ParameterAdjoint {
        device_buffer<double> divergence_relaxation;
        device_buffer<double> density_relaxation;
    };
};

namespace detail {
// Owning handle of a C ABI step cache.
template <class Cache, phx_status (*Free)(Cache*)>
class CacheHandle {
  public:
    CacheHandle() = default;
    explicit CacheHandle(Cache* cache) : cache_(cache) {}
    CacheHandle(const CacheHandle&) = delete;
    CacheHandle& operator=(const CacheHandle&) = delete;
    CacheHandle(CacheHandle&& o) noexcept : cache_(std::exchange(o.cache_, nullptr)) {}
    CacheHandle& operator=(CacheHandle&& o) noexcept {
        std::swap(cache_, o.cache_);
        return *this;
    }
    ~CacheHandle() {
        if (cache_ != nullptr) (void)Free(cache_);
    }
    Cache* get() const { return cache_; }

  private:
    Cache* cache_ = nullptr;
};

inline phx_particle_state state_view(ParticleState& s) {
    return {s.positions.view(), s.velocities.view(), s.step_index};
}
inline phx_particle_state_tangent tangent_view(ParticleStateTangent& t) {
    return {t.positions.view(), t.velocities.view()};
}
inline phx_particle_state_adjoint adjoint_view(ParticleStateAdjoint& a) {
    return {a.positions.view(), a.velocities.view()};
}
}  // namespace detail

template <class Method>
class SPH;

// SPH<DFSPH>, sph/sph.ixx:273-372, over phx_dfsph_*.
template <>
class SPH<DFSPH> {
  public:
    struct State {
        ParticleState particles;
        DFSPH::State method;
    };
    struct StateTangent {
        ParticleStateTangent particles;
        DFSPH::StateTangent method;
    };
    struct StateAdjoint {
        ParticleStateAdjoint particles;
        DFSPH::StateAdjoint method;
    };
    struct Parameters {
        ParticleParameters particles;
        DFSPH::Parameters method;
    };
    struct ParameterTangent {
        ParticleParameterTangent particles;
        DFSPH::ParameterTangent method;
    };
    struct ParameterAdjoint {
        ParticleParameterAdjoint particles;
        DFSPH::ParameterAdjoint method;
    };
    using StepCache = detail::CacheHandle<phx_dfsph_step_cache, phx_dfsph_free_step_cache>;

    SPH(DomainConfiguration domain_configuration, DFSPH::Configuration method_configuration, ExecutionMode mode,
        hipStream_t stream)
        : domain(std::move(domain_configuration), stream) {
        const phx_domain_configuration d = domain.configuration.to_c();
        phx_dfsph_configuration c{};
        c.divergence_iterations = method_configuration.divergence_iterations;
        c.density_iterations = method_configuration.density_iterations;
        c.checkpoint_interval = method_configuration.checkpoint_interval;
        c.pressure_warm_start = method_configuration.pressure_warm_start ? 1u : 0u;
        c.density_error_tolerance = method_configuration.density_error_tolerance;
        c.max_iterations = method_configuration.max_iterations;
        c.exact_no_slip_adjoint = method_configuration.exact_no_slip_adjoint ? 1u : 0u;
        check(phx_dfsph_create(&d, &c, static_cast<phx_execution_mode>(mode), stream, &solver_), "SPH<DFSPH>");
    }
    SPH(const SPH&) = delete;
    SPH& operator=(const SPH&) = delete;
    SPH(SPH&&) = delete;
    SPH& operator=(SPH&&) = delete;
    ~SPH() {
        if (solver_ != nullptr) (void)phx_dfsph_destroy(solver_);
    }

    State allocate_state() const {
        State s{domain.allocate_particle_state(),
                {domain.allocate_scalar_field(count()), domain.allocate_scalar_field(count())}};
        domain.clear(s.method.warm_divergence_pressure);
        domain.clear(s.method.warm_density_pressure);
        return s;
    }
    Control allocate_control() const { return domain.allocate_control(); }
    Parameters allocate_parameters() const {
        return {domain.allocate_particle_parameters(),
                {device_buffer<float>(domain.stream, count()), device_buffer<float>(domain.stream, count())}};
    }
    StepCache allocate_step_cache() const {
        phx_dfsph_step_cache* cache = nullptr;
        check(phx_dfsph_allocate_step_cache(solver_, &cache), "allocate_step_cache");
        return StepCache(cache);
    }
    StateTangent allocate_state_tangent() const {
        StateTangent t{domain.allocate_particle_state_tangent(),
                       {domain.allocate_scalar_field(count()), domain.allocate_scalar_field(count())}};
        domain.clear(t.method.warm_divergence_pressure);
        domain.clear(t.method.warm_density_pressure);
        return t;
    }
    ControlTangent allocate_control_tangent() const { return domain.allocate_control_tangent(); }
    ParameterTangent allocate_parameter_tangent() const {
        ParameterTangent t{domain.allocate_particle_parameter_tangent(),
                           {device_buffer<float>(domain.stream, count()), device_buffer<float>(domain.stream, count())}};
        domain.clear(t.method.divergence_relaxation);
        domain.clear(t.method.density_relaxation);
        return t;
    }
    StateAdjoint allocate_state_adjoint() const {
        StateAdjoint a{domain.allocate_particle_state_adjoint(),
                       {domain.allocate_scalar_adjoint_field(count()), domain.allocate_scalar_adjoint_field(count())}};
        domain.clear(a.method.warm_divergence_pressure);
        domain.clear(a.method.warm_density_pressure);
        return a;
    }
    ControlAdjoint allocate_control_adjoint() const { return domain.allocate_control_adjoint(); }
    ParameterAdjoint allocate_parameter_adjoint() const {
        ParameterAdjoint a{domain.allocate_particle_parameter_adjoint(),
                           {device_buffer<double>(domain.stream, count()), device_buffer<double>(domain.stream, count())}};
        domain.clear(a.method.divergence_relaxation);
        domain.clear(a.method.density_relaxation);
        return a;
    }
    void copy_state(const State& source, State& destination) const {
        domain.copy(source.particles, destination.particles);
        domain.copy(source.method.warm_divergence_pressure, destination.method.warm_divergence_pressure);
        domain.copy(source.method.warm_density_pressure, destination.method.warm_density_pressure);
    }
    void copy_state_tangent(const StateTangent& source, StateTangent& destination) const {
        domain.copy(source.particles, destination.particles);
        domain.copy(source.method.warm_divergence_pressure, destination.method.warm_divergence_pressure);
        domain.copy(source.method.warm_density_pressure, destination.method.warm_density_pressure);
    }
    void copy_state_adjoint(const StateAdjoint& source, StateAdjoint& destination) const {
        domain.copy(source.particles, destination.particles);
        domain.copy(source.method.warm_divergence_pressure, destination.method.warm_divergence_pressure);
        domain.copy(source.method.warm_density_pressure, destination.method.warm_density_pressure);
    }
    void accumulate_state_adjoint(const StateAdjoint& source, StateAdjoint& destination) const {
        domain.accumulate(source.particles, destination.particles);
        domain.accumulate(source.method.warm_divergence_pressure, destination.method.warm_divergence_pressure);
        domain.accumulate(source.method.warm_density_pressure, destination.method.warm_density_pressure);
    }

    void forward_step(const State& state, const Control& control, const Parameters& parameters, State& next_state,
                      StepCache& cache) {
        phx_sph_dfsph_state s = to_c(const_cast<State&>(state));
        phx_sph_dfsph_state n = to_c(next_state);
        phx_control c{control.external_accelerations.const_view()};
        phx_sph_dfsph_parameters p{parameters.particles.view(),
                                   {parameters.method.divergence_relaxation.data(),
                                    parameters.method.density_relaxation.data()}};
        check(phx_dfsph_forward_step(solver_, &s, &c, &p, &n, cache.get()), "SPH<DFSPH>::forward_step");
        next_state.particles.step_index = n.particles.step_index;
    }
    // sph/sph.ixx:356-359; the step recorded in `cache` is differentiated
    void jvp_step(const State&, const Control&, const Parameters&, const State&, const StepCache& cache,
                  const StateTangent& state_tangent, const ControlTangent& control_tangent,
                  const ParameterTangent& parameter_tangent, StateTangent& next_state_tangent) {
        auto& st = const_cast<StateTangent&>(state_tangent);
        phx_sph_dfsph_state_tangent t{detail::tangent_view(st.particles),
                                      {st.method.warm_divergence_pressure.values.data(),
                                       st.method.warm_density_pressure.values.data()}};
        phx_control_tangent ct{control_tangent.external_accelerations.const_view()};
        phx_sph_dfsph_parameter_tangent pt{parameter_tangent.particles.view(),
                                           {parameter_tangent.method.divergence_relaxation.data(),
                                            parameter_tangent.method.density_relaxation.data()}};
        phx_sph_dfsph_state_tangent nt{detail::tangent_view(next_state_tangent.particles),
                                       {next_state_tangent.method.warm_divergence_pressure.values.data(),
                                        next_state_tangent.method.warm_density_pressure.values.data()}};
        check(phx_dfsph_jvp_step(solver_, cache.get(), &t, &ct, &pt, &nt), "SPH<DFSPH>::jvp_step");
    }
    // sph/sph.ixx:361-365: adjoints accumulate (+=)
    void vjp_step(const State&, const Control&, const Parameters&, const State&, const StepCache& cache,
                  const StateAdjoint& next_state_adjoint, StateAdjoint& previous_state_adjoint,
                  ControlAdjoint& control_adjoint, ParameterAdjoint& parameter_adjoint) {
        auto& na = const_cast<StateAdjoint&>(next_state_adjoint);
        phx_sph_dfsph_state_adjoint next{detail::adjoint_view(na.particles),
                                         {na.method.warm_divergence_pressure.values.data(),
                                          na.method.warm_density_pressure.values.data()}};
        phx_sph_dfsph_state_adjoint prev{detail::adjoint_view(previous_state_adjoint.particles),
                                         {previous_state_adjoint.method.warm_divergence_pressure.values.data(),
                                          previous_state_adjoint.method.warm_density_pressure.values.data()}};
        phx_control_adjoint ca{control_adjoint.external_accelerations.view()};
        phx_sph_dfsph_parameter_adjoint pa{parameter_adjoint.particles.view(),
                                           {parameter_adjoint.method.divergence_relaxation.data(),
                                            parameter_adjoint.method.density_relaxation.data()}};
        check(phx_dfsph_vjp_step(solver_, cache.get(), &next, &prev, &ca, &pa), "SPH<DFSPH>::vjp_step");
    }
    // iterations executed by the last forward_step (divergence, density)
    std::array<std::uint32_t, 2u> last_iterations() const {
        std::uint32_t it[2] = {0u, 0u};
        check(phx_dfsph_last_iterations(solver_, it), "last_iterations");
        return {it[0], it[1]};
    }

    Domain domain;

  private:
    static phx_sph_dfsph_state to_c(State& s) {
        return {detail::state_view(s.particles),
                {s.method.warm_divergence_pressure.values.data(), s.method.warm_density_pressure.values.data()}};
    }
    std::size_t count() const { return domain.configuration.particle_count; }
    phx_dfsph* solver_ = nullptr;
};

namespace detail {
// The per-method parts of SPH<WCSPH|PCISPH|IISPH> over phx_sph_*.
template <class Method>
struct SphMethod;

template <>
struct SphMethod<WCSPH> {
    static phx_sph_configuration configuration(const WCSPH::Configuration& c) {
        return {PHX_SPH_WCSPH, 0u, 2u, 0.0F, 0u, c.exact_no_slip_adjoint ? 1u : 0u};
    }
    template <class Buffers>
    static Buffers allocate(hipStream_t stream, std::size_t n) {
        using T = typename std::remove_reference_t<decltype(std::declval<Buffers>().speed_of_sound)>;
        return {T(stream, n), T(stream, n), T(stream, n)};
    }
    template <class Buffers, class F>
    static void each(Buffers& b, F&& f) {
        f(b.speed_of_sound);
        f(b.tait_exponent);
        f(b.boundary_surface_tension);
    }
    static phx_sph_method_parameters view(const WCSPH::Parameters& p) {
        return {p.speed_of_sound.data(), p.tait_exponent.data(), p.boundary_surface_tension.data(), nullptr, nullptr};
    }
    static phx_sph_method_parameter_tangent view(const WCSPH::ParameterTangent& p) {
        return {p.speed_of_sound.data(), p.tait_exponent.data(), p.boundary_surface_tension.data(), nullptr, nullptr};
    }
    static phx_sph_method_parameter_adjoint view(WCSPH::ParameterAdjoint& p) {
        return {p.speed_of_sound.data(), p.tait_exponent.data(), p.boundary_surface_tension.data(), nullptr, nullptr};
    }
};

template <>
struct SphMethod<PCISPH> {
    static phx_sph_configuration configuration(const PCISPH::Configuration& c) {
        return {PHX_SPH_PCISPH,          c.pressure_iterations, c.checkpoint_interval, c.density_error_tolerance,
                c.max_iterations,        c.exact_no_slip_adjoint ? 1u : 0u};
    }
    template <class Buffers>
    static Buffers allocate(hipStream_t stream, std::size_t n) {
        using T = typename std::remove_reference_t<decltype(std::declval<Buffers>().pressure_relaxation)>;
        return {T(stream, n)};
    }
    template <class Buffers, class F>
    static void each(Buffers& b, F&& f) {
        f(b.pressure_relaxation);
    }
    static phx_sph_method_parameters view(const PCISPH::Parameters& p) {
        return {nullptr, nullptr, nullptr, p.pressure_relaxation.data(), nullptr};
    }
    static phx_sph_method_parameter_tangent view(const PCISPH::ParameterTangent& p) {
        return {nullptr, nullptr, nullptr, p.pressure_relaxation.data(), nullptr};
    }
    static phx_sph_method_parameter_adjoint view(PCISPH::ParameterAdjoint& p) {
        return {nullptr, nullptr, nullptr, p.pressure_relaxation.data(), nullptr};
    }
};

template <>
struct SphMethod<IISPH> {
    static phx_sph_configuration configuration(const IISPH::Configuration& c) {
        return {PHX_SPH_IISPH,           c.pressure_iterations, c.checkpoint_interval, c.density_error_tolerance,
                c.max_iterations,        c.exact_no_slip_adjoint ? 1u : 0u};
    }
    template <class Buffers>
    static Buffers allocate(hipStream_t stream, std::size_t n) {
        using T = typename std::remove_reference_t<decltype(std::declval<Buffers>().jacobi_relaxation)>;
        return {T(stream, n)};
    }
    template <class Buffers, class F>
    static void each(Buffers& b, F&& f) {
        f(b.jacobi_relaxation);
    }
    static phx_sph_method_parameters view(const IISPH::Parameters& p) {
        return {nullptr, nullptr, nullptr, nullptr, p.jacobi_relaxation.data()};
    }
    static phx_sph_method_parameter_tangent view(const IISPH::ParameterTangent& p) {
        return {nullptr, nullptr, nullptr, nullptr, p.jacobi_relaxation.data()};
    }
    static phx_sph_method_parameter_adjoint view(IISPH::ParameterAdjoint& p) {
        return {nullptr, nullptr, nullptr, nullptr, p.jacobi_relaxation.data()};
    }
};
}  // namespace detail

// SPH<WCSPH>, SPH<PCISPH>, SPH<IISPH> (sph/sph.ixx:273-372) over phx_sph_*.
// The three method states are empty (sph/sph.ixx:41-43, 108-110, 154-156).
template <class Method>
class SPH {
    using M = detail::SphMethod<Method>;

  public:
    struct State {
        ParticleState particles;
        [[no_unique_address]] typename Method::State method;
    };
    struct StateTangent {
        ParticleStateTangent particles;
        [[no_unique_address]] typename Method::StateTangent method;
    };
    struct StateAdjoint {
        ParticleStateAdjoint particles;
        [[no_unique_address]] typename Method::StateAdjoint method;
    };
    struct Parameters {
        ParticleParameters particles;
        typename Method::Parameters method;
    };
    struct ParameterTangent {
        ParticleParameterTangent particles;
        typename Method::ParameterTangent method;
    };
    struct ParameterAdjoint {
        ParticleParameterAdjoint particles;
        typename Method::ParameterAdjoint method;
    };
    using StepCache = detail::CacheHandle<phx_sph_step_cache, phx_sph_free_step_cache>;

    SPH(DomainConfiguration domain_configuration, typename Method::Configuration method_configuration,
        ExecutionMode mode, hipStream_t stream)
        : domain(std::move(domain_configuration), stream) {
        const phx_domain_configuration d = domain.configuration.to_c();
        const phx_sph_configuration c = M::configuration(method_configuration);
        check(phx_sph_create(&d, &c, static_cast<phx_execution_mode>(mode), stream, &solver_), "SPH");
    }
    SPH(const SPH&) = delete;
    SPH& operator=(const SPH&) = delete;
    SPH(SPH&&) = delete;
    SPH& operator=(SPH&&) = delete;
    ~SPH() {
        if (solver_ != nullptr) (void)phx_sph_destroy(solver_);
    }

    State allocate_state() const { return {domain.allocate_particle_state(), {}}; }
    Control allocate_control() const { return domain.allocate_control(); }
    Parameters allocate_parameters() const {
        return {domain.allocate_particle_parameters(),
                M::template allocate<typename Method::Parameters>(domain.stream, count())};
    }
    StepCache allocate_step_cache() const {
        phx_sph_step_cache* cache = nullptr;
        check(phx_sph_allocate_step_cache(solver_, &cache), "allocate_step_cache");
        return StepCache(cache);
    }
    StateTangent allocate_state_tangent() const { return {domain.allocate_particle_state_tangent(), {}}; }
    ControlTangent allocate_control_tangent() const { return domain.allocate_control_tangent(); }
    ParameterTangent allocate_parameter_tangent() const {
        ParameterTangent t{domain.allocate_particle_parameter_tangent(),
                           M::template allocate<typename Method::ParameterTangent>(domain.stream, count())};
        M::each(t.method, [&](auto& b) { domain.clear(b); });
        return t;
    }
    StateAdjoint allocate_state_adjoint() const { return {domain.allocate_particle_state_adjoint(), {}}; }
    ControlAdjoint allocate_control_adjoint() const { return domain.allocate_control_adjoint(); }
    ParameterAdjoint allocate_parameter_adjoint() const {
        ParameterAdjoint a{domain.allocate_particle_parameter_adjoint(),
                           M::template allocate<typename Method::ParameterAdjoint>(domain.stream, count())};
        M::each(a.method, [&](auto& b) { domain.clear(b); });
        return a;
    }
    void copy_state(const State& source, State& destination) const { domain.copy(source.particles, destination.particles); }
    void copy_state_tangent(const StateTangent& source, StateTangent& destination) const {
        domain.copy(source.particles, destination.particles);
    }
    void copy_state_adjoint(const StateAdjoint& source, StateAdjoint& destination) const {
        domain.copy(source.particles, destination.particles);
    }
    void accumulate_state_adjoint(const StateAdjoint& source, StateAdjoint& destination) const {
        domain.accumulate(source.particles, destination.particles);
    }

    void forward_step(const State& state, const Control& control, const Parameters& parameters, State& next_state,
                      StepCache& cache) {
        phx_sph_state s{detail::state_view(const_cast<State&>(state).particles)};
        phx_sph_state n{detail::state_view(next_state.particles)};
        phx_control c{control.external_accelerations.const_view()};
        phx_sph_parameters p{parameters.particles.view(), M::view(parameters.method)};
        check(phx_sph_forward_step(solver_, &s, &c, &p, &n, cache.get()), "SPH::forward_step");
        next_state.particles.step_index = n.particles.step_index;
    }
    // sph/sph.ixx:356-359; the step recorded in `cache` is differentiated
    void jvp_step(const State&, const Control&, const Parameters&, const State&, const StepCache& cache,
                  const StateTangent& state_tangent, const ControlTangent& control_tangent,
                  const ParameterTangent& parameter_tangent, StateTangent& next_state_tangent) {
        phx_sph_state_tangent t{detail::tangent_view(const_cast<StateTangent&>(state_tangent).particles)};
        phx_control_tangent ct{control_tangent.external_accelerations.const_view()};
        phx_sph_parameter_tangent pt{parameter_tangent.particles.view(), M::view(parameter_tangent.method)};
        phx_sph_state_tangent nt{detail::tangent_view(next_state_tangent.particles)};
        check(phx_sph_jvp_step(solver_, cache.get(), &t, &ct, &pt, &nt), "SPH::jvp_step");
    }
    // sph/sph.ixx:361-365: adjoints accumulate (+=)
    void vjp_step(const State&, const Control&, const Parameters&, const State&, const StepCache& cache,
                  const StateAdjoint& next_state_adjoint, StateAdjoint& previous_state_adjoint,
                  ControlAdjoint& control_adjoint, ParameterAdjoint& parameter_adjoint) {
        phx_sph_state_adjoint next{detail::adjoint_view(const_cast<StateAdjoint&>(next_state_adjoint).particles)};
        phx_sph_state_adjoint prev{detail::adjoint_view(previous_state_adjoint.particles)};
        phx_control_adjoint ca{control_adjoint.external_accelerations.view()};
        phx_sph_parameter_adjoint pa{parameter_adjoint.particles.view(), M::view(parameter_adjoint.method)};
        check(phx_sph_vjp_step(solver_, cache.get(), &next, &prev, &ca, &pa), "SPH::vjp_step");
    }

    Domain domain;

  private:
    std::size_t count() const { return domain.configuration.particle_count; }
    phx_sph* solver_ = nullptr;
};

// PBF, pbf/pbf.ixx:14-95, over phx_pbf_*.
class PBF {
  public:
    struct Configuration {
        std::uint32_t pressure_iterations{5u};
        std::uint32_t checkpoint_interval{2u};
    };
    struct State {
        ParticleState particles;
    };
    struct StateTangent {
        ParticleStateTangent particles;
    };
    struct StateAdjoint {
        ParticleStateAdjoint particles;
    };
    struct Parameters {
        ParticleParameters particles;
        device_buffer<float> relaxation;
        device_buffer<float> artificial_pressure_strength;
        device_buffer<float> artificial_pressure_exponent;
        device_buffer<float> artificial_pressure_radius;
        device_buffer<float> xsph_viscosity;
        device_buffer<float> vorticity_confinement;
    };
    struct ParameterTangent {
        ParticleParameterTangent particles;
        device_buffer<float> relaxation;
        device_buffer<float> artificial_pressure_strength;
        device_buffer<float> artificial_pressure_exponent;
        device_buffer<float> artificial_pressure_radius;
        device_buffer<float> xsph_viscosity;
        device_buffer<float> vorticity_confinement;
    };
    struct ParameterAdjoint {
        ParticleParameterAdjoint particles;
        device_buffer<double> relaxation;
        device_buffer<double> artificial_pressure_strength;
        device_buffer<double> artificial_pressure_exponent;
        device_buffer<double> artificial_pressure_radius;
        device_buffer<double> xsph_viscosity;
        device_buffer<double> vorticity_confinement;
    };
    using StepCache = detail::CacheHandle<phx_pbf_step_cache, phx_pbf_free_step_cache>;

    PBF(DomainConfiguration domain_configuration, Configuration configuration, ExecutionMode mode, hipStream_t stream)
        : domain(std::move(domain_configuration), stream) {
        const phx_domain_configuration d = domain.configuration.to_c();
        const phx_pbf_configuration c{configuration.pressure_iterations, configuration.checkpoint_interval};
        check(phx_pbf_create(&d, &c, static_cast<phx_execution_mode>(mode), stream, &solver_), "PBF");
    }
    PBF(const PBF&) = delete;
    PBF& operator=(const PBF&) = delete;
    PBF(PBF&&) = delete;
    PBF& operator=(PBF&&) = delete;
    ~PBF() {
        if (solver_ != nullptr) (void)phx_pbf_destroy(solver_);
    }

    State allocate_state() const { return {domain.allocate_particle_state()}; }
    Control allocate_control() const { return domain.allocate_control(); }
    Parameters allocate_parameters() const {
        auto b = [&] { return device_buffer<float>(domain.stream, count()); };
        return {domain.allocate_particle_parameters(), b(), b(), b(), b(), b(), b()};
    }
    StepCache allocate_step_cache() const {
        phx_pbf_step_cache* cache = nullptr;
        check(phx_pbf_allocate_step_cache(solver_, &cache), "allocate_step_cache");
        return StepCache(cache);
    }
    StateTangent allocate_state_tangent() const { return {domain.allocate_particle_state_tangent()}; }
    ControlTangent allocate_control_tangent() const { return domain.allocate_control_tangent(); }
    ParameterTangent allocate_parameter_tangent() const {
        auto b = [&] { return device_buffer<float>(domain.stream, count()); };
        ParameterTangent t{domain.allocate_particle_parameter_tangent(), b(), b(), b(), b(), b(), b()};
        for (auto* f : {&t.relaxation, &t.artificial_pressure_strength, &t.artificial_pressure_exponent,
                        &t.artificial_pressure_radius, &t.xsph_viscosity, &t.vorticity_confinement})
            domain.clear(*f);
        return t;
    }
    StateAdjoint allocate_state_adjoint() const { return {domain.allocate_particle_state_adjoint()}; }
    ControlAdjoint allocate_control_adjoint() const { return domain.allocate_control_adjoint(); }
    ParameterAdjoint allocate_parameter_adjoint() const {
        auto b = [&] { return device_buffer<double>(domain.stream, count()); };
        ParameterAdjoint a{domain.allocate_particle_parameter_adjoint(), b(), b(), b(), b(), b(), b()};
        for (auto* f : {&a.relaxation, &a.artificial_pressure_strength, &a.artificial_pressure_exponent,
                        &a.artificial_pressure_radius, &a.xsph_viscosity, &a.vorticity_confinement})
            domain.clear(*f);
        return a;
    }
    void copy_state(const State& source, State& destination) const { domain.copy(source.particles, destination.particles); }
    void copy_state_tangent(const StateTangent& source, StateTangent& destination) const {
        domain.copy(source.particles, destination.particles);
    }
    void copy_state_adjoint(const StateAdjoint& source, StateAdjoint& destination) const {
        domain.copy(source.particles, destination.particles);
    }
    void accumulate_state_adjoint(const StateAdjoint& source, StateAdjoint& destination) const {
        domain.accumulate(source.particles, destination.particles);
    }

    void forward_step(const State& state, const Control& control, const Parameters& parameters, State& next_state,
                      StepCache& cache) {
        phx_pbf_state s{detail::state_view(const_cast<State&>(state).particles)};
        phx_pbf_state n{detail::state_view(next_state.particles)};
        phx_control c{control.external_accelerations.const_view()};
        const phx_pbf_parameters p = to_c(parameters);
        check(phx_pbf_forward_step(solver_, &s, &c, &p, &n, cache.get()), "PBF::forward_step");
        next_state.particles.step_index = n.particles.step_index;
    }
    // pbf/pbf.cpp:193-215 on the step recorded in `cache`
    void jvp_step(const State&, const Control&, const Parameters& parameters, const State&, const StepCache& cache,
                  const StateTangent& state_tangent, const ControlTangent& control_tangent,
                  const ParameterTangent& parameter_tangent, StateTangent& next_state_tangent) {
        const phx_pbf_parameters p = to_c(parameters);
        auto& st = const_cast<StateTangent&>(state_tangent);
        phx_pbf_state_tangent t{st.particles.positions.view(), st.particles.velocities.view()};
        phx_control_tangent ct{control_tangent.external_accelerations.const_view()};
        phx_pbf_parameter_tangent pt{parameter_tangent.particles.view(),
                                     parameter_tangent.relaxation.data(),
                                     parameter_tangent.artificial_pressure_strength.data(),
                                     parameter_tangent.artificial_pressure_exponent.data(),
                                     parameter_tangent.artificial_pressure_radius.data(),
                                     parameter_tangent.xsph_viscosity.data(),
                                     parameter_tangent.vorticity_confinement.data()};
        phx_pbf_state_tangent nt{next_state_tangent.particles.positions.view(),
                                 next_state_tangent.particles.velocities.view()};
        check(phx_pbf_jvp_step(solver_, &p, cache.get(), &t, &ct, &pt, &nt), "PBF::jvp_step");
    }
    // pbf/pbf.cpp:217-258: adjoints accumulate (+=)
    void vjp_step(const State&, const Control&, const Parameters& parameters, const State&, const StepCache& cache,
                  const StateAdjoint& next_state_adjoint, StateAdjoint& previous_state_adjoint,
                  ControlAdjoint& control_adjoint, ParameterAdjoint& parameter_adjoint) {
        const phx_pbf_parameters p = to_c(parameters);
        auto& na = const_cast<StateAdjoint&>(next_state_adjoint);
        phx_pbf_state_adjoint next{na.particles.positions.view(), na.particles.velocities.view()};
        phx_pbf_state_adjoint prev{previous_state_adjoint.particles.positions.view(),
                                   previous_state_adjoint.particles.velocities.view()};
        phx_control_adjoint ca{control_adjoint.external_accelerations.view()};
        phx_pbf_parameter_adjoint pa{parameter_adjoint.particles.view(),
                                     parameter_adjoint.relaxation.data(),
                                     parameter_adjoint.artificial_pressure_strength.data(),
                                     parameter_adjoint.artificial_pressure_exponent.data(),
                                     parameter_adjoint.artificial_pressure_radius.data(),
                                     parameter_adjoint.xsph_viscosity.data(),
                                     parameter_adjoint.vorticity_confinement.data()};
        check(phx_pbf_vjp_step(solver_, &p, cache.get(), &next, &prev, &ca, &pa), "PBF::vjp_step");
    }

    Domain domain;

  private:
    static phx_pbf_parameters to_c(const Parameters& parameters) {
        return {parameters.particles.view(),
                parameters.relaxation.data(),
                parameters.artificial_pressure_strength.data(),
                parameters.artificial_pressure_exponent.data(),
                parameters.artificial_pressure_radius.data(),
                parameters.xsph_viscosity.data(),
                parameters.vorticity_confinement.data()};
    }
    std::size_t count() const { return domain.configuration.particle_count; }
    phx_pbf* solver_ = nullptr;
};

}  // namespace fluids::liquid::particle
}  // namespace physica_amd
