// C++ host API over the physica_amd C ABI, mirroring the reference module
// physica.fluids.liquid.particle (fluids/liquid/particle/*.ixx): the same
// type and member names, the same "allocate_*, forward_step(state, control,
// parameters, next_state, cache)" flow, HIP device buffers in place of
// ::cuda::device_buffer and hipStream_t in place of ::cuda::stream_ref.
// Errors surface as physica_amd::Error exceptions (the reference's CCCL
// buffers throw on failure as well); the C ABI below never throws.
//
// Header-only; link with libphysica_amd.so and amdhip64.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <span>
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

struct ParticleState {
    VectorField positions;
    VectorField velocities;
    std::uint64_t step_index{};
};

struct Control {
    VectorField external_accelerations;
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
    void clear(ScalarField& f) const { clear(f.values); }
    void clear(VectorField& f) const {
        clear(f.x);
        clear(f.y);
        clear(f.z);
    }
    void clear(device_buffer<float>& b) const {
        check_hip(hipMemsetAsync(b.data(), 0, b.size() * sizeof(float), stream), "hipMemsetAsync");
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

// sph/sph.ixx:195-268
struct DFSPH {
    struct Configuration {
        std::uint32_t divergence_iterations{4u};
        std::uint32_t density_iterations{6u};
        std::uint32_t checkpoint_interval{2u};
        bool pressure_warm_start{true};
    };
    struct State {
        ScalarField warm_divergence_pressure;
        ScalarField warm_density_pressure;
    };
    struct Parameters {
        device_buffer<float> divergence_relaxation;
        device_buffer<float> density_relaxation;
    };
};

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
    struct Parameters {
        ParticleParameters particles;
        DFSPH::Parameters method;
    };
    class StepCache {
      public:
        explicit StepCache(phx_dfsph* solver) { check(phx_dfsph_allocate_step_cache(solver, &cache_), "allocate_step_cache"); }
        StepCache(const StepCache&) = delete;
        StepCache& operator=(const StepCache&) = delete;
        StepCache(StepCache&& o) noexcept : cache_(std::exchange(o.cache_, nullptr)) {}
        ~StepCache() {
            if (cache_ != nullptr) (void)phx_dfsph_free_step_cache(cache_);
        }
        phx_dfsph_step_cache* get() const { return cache_; }

      private:
        phx_dfsph_step_cache* cache_ = nullptr;
    };

    SPH(DomainConfiguration domain_configuration, DFSPH::Configuration method_configuration, ExecutionMode mode,
        hipStream_t stream)
        : domain(std::move(domain_configuration), stream) {
        const phx_domain_configuration d = domain.configuration.to_c();
        phx_dfsph_configuration c{};
        c.divergence_iterations = method_configuration.divergence_iterations;
        c.density_iterations = method_configuration.density_iterations;
        c.checkpoint_interval = method_configuration.checkpoint_interval;
        c.pressure_warm_start = method_configuration.pressure_warm_start ? 1u : 0u;
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
    StepCache allocate_step_cache() const { return StepCache(solver_); }
    void copy_state(const State& source, State& destination) const {
        domain.copy(source.particles, destination.particles);
        domain.copy(source.method.warm_divergence_pressure, destination.method.warm_divergence_pressure);
        domain.copy(source.method.warm_density_pressure, destination.method.warm_density_pressure);
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

    Domain domain;

  private:
    static phx_sph_dfsph_state to_c(State& s) {
        return {{s.particles.positions.view(), s.particles.velocities.view(), s.particles.step_index},
                {s.method.warm_divergence_pressure.values.data(), s.method.warm_density_pressure.values.data()}};
    }
    std::size_t count() const { return domain.configuration.particle_count; }
    phx_dfsph* solver_ = nullptr;
};

// PBF, pbf/pbf.ixx:14-141, over phx_pbf_*.
class PBF {
  public:
    struct Configuration {
        std::uint32_t pressure_iterations{5u};
        std::uint32_t checkpoint_interval{2u};
    };
    struct State {
        ParticleState particles;
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
    class StepCache {
      public:
        explicit StepCache(phx_pbf* solver) { check(phx_pbf_allocate_step_cache(solver, &cache_), "allocate_step_cache"); }
        StepCache(const StepCache&) = delete;
        StepCache& operator=(const StepCache&) = delete;
        StepCache(StepCache&& o) noexcept : cache_(std::exchange(o.cache_, nullptr)) {}
        ~StepCache() {
            if (cache_ != nullptr) (void)phx_pbf_free_step_cache(cache_);
        }
        phx_pbf_step_cache* get() const { return cache_; }

      private:
        phx_pbf_step_cache* cache_ = nullptr;
    };

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
        auto b = [&] { return device_buffer<float>(domain.stream, domain.configuration.particle_count); };
        return {domain.allocate_particle_parameters(), b(), b(), b(), b(), b(), b()};
    }
    StepCache allocate_step_cache() const { return StepCache(solver_); }
    void copy_state(const State& source, State& destination) const { domain.copy(source.particles, destination.particles); }

    void forward_step(const State& state, const Control& control, const Parameters& parameters, State& next_state,
                      StepCache& cache) {
        State& in = const_cast<State&>(state);
        phx_pbf_state s{{in.particles.positions.view(), in.particles.velocities.view(), in.particles.step_index}};
        phx_pbf_state n{{next_state.particles.positions.view(), next_state.particles.velocities.view(),
                         next_state.particles.step_index}};
        phx_control c{control.external_accelerations.const_view()};
        phx_pbf_parameters p{parameters.particles.view(),
                             parameters.relaxation.data(),
                             parameters.artificial_pressure_strength.data(),
                             parameters.artificial_pressure_exponent.data(),
                             parameters.artificial_pressure_radius.data(),
                             parameters.xsph_viscosity.data(),
                             parameters.vorticity_confinement.data()};
        check(phx_pbf_forward_step(solver_, &s, &c, &p, &n, cache.get()), "PBF::forward_step");
        next_state.particles.step_index = n.particles.step_index;
    }

    Domain domain;

  private:
    phx_pbf* solver_ = nullptr;
};

}  // namespace fluids::liquid::particle
}  // namespace physica_amd
