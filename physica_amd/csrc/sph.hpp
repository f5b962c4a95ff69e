// SPH<WCSPH>, SPH<PCISPH>, SPH<IISPH> (sph/sph.ixx:39-193, 273-372) on the
// sorted-order engine.
#pragma once

#include <memory>

#include "dfsph.hpp"
#include "engine.hpp"
#include "masks.hpp"
#include "profiler.hpp"

namespace phx {

// A WCSPH step recorded in differentiable mode, in sorted order: the inputs
// the derivatives read, the exact masks with their flags and the
// integration's collision bits (densities and pressures stay in SphCache).
struct WcsphRecord {
    bool recorded = false;
    uint64_t step_index = 0;
    DeviceBuffer<float4> xm, vmu, aux;  // aux: density, p / rho^2, speed of sound, surface tension
    DeviceBuffer<float> rho0, tait;
    DeviceBuffer<float> btension;  // boundary surface tensions (boundary samples only)
    DeviceBuffer<uint8_t> collisions;
    DeviceBuffer<uint16_t> exact;
    DeviceBuffer<uint8_t> crowded;  // per particle: walks the full stencil
    DeviceBuffer<uint32_t> flags;
};

// Sorted-order tangent and adjoint workspace of the WCSPH derivative passes.
struct WcsphDerivativeWorkspace {
    DeviceBuffer<float4> xt, vt, ut, pt;  // (x', m'), (v', mu'), (u', rho0'), (c', gamma', sigma')
    DeviceBuffer<float> rho_t;
    DeviceBuffer<float2> tt;  // (p / rho^2, its tangent)
    DVecBuffer pos_a, vel_a, total_a;
    DeviceBuffer<double> rho_a, m_a, rest_a, visc_a, tension_a, speed_a, tait_a, btension_a;
    void allocate_tangents(uint32_t n);
    void allocate_adjoints(uint32_t n);
};

// StepCache of SPH<Method> (sph/sph.ixx:299-303) with the method caches used
// here: densities, pressures (WCSPH::Cache::pressures, or the last iteration
// of PCISPH/IISPH) and the non-pressure accelerations (PCISPH/IISPH), all in
// sorted order.
struct SphCache {
    hipStream_t stream = nullptr;
    phx_sph_method method = PHX_SPH_WCSPH;
    Neighborhood neighborhood;
    DeviceBuffer<float> densities;
    DeviceBuffer<float> pressures;
    DeviceBuffer<float4> non_pressure;  // xyz used
    // differentiable mode (PCISPH, IISPH): the step recorded by the
    // projection engine, which then holds every field above
    std::unique_ptr<DfsphCache> record;
    // differentiable mode (WCSPH)
    std::unique_ptr<WcsphRecord> wcsph;
    void allocate(uint32_t n, uint32_t cells);

    const Neighborhood& grid() const { return record ? record->neighborhood : neighborhood; }
    const float* density_data() const { return record ? record->densities.data() : densities.data(); }
    // the last iteration's pressures
    const float* pressure_data() const {
        return record ? record->pressures(1u, record->iterations[1]) : pressures.data();
    }
    const float4* non_pressure_data() const { return record ? record->non_pressure.data() : non_pressure.data(); }
};

enum SphPass : int {
    kSphGrid = 0,
    kSphDensity,     // masks + density (+ Tait pressure for WCSPH)
    kSphForceInit,   // PCISPH/IISPH: non-pressure + first predict
    kSphProjection,  // PCISPH/IISPH: density at x* + pressure update
    kSphForce,       // PCISPH/IISPH: pressure acceleration + predict
    kSphForces,      // WCSPH: pressure, artificial viscosity, surface, external + integrate
    kSphPassCount
};
const char* sph_pass_name(int pass);
// Algorithmic bytes per particle per launch (unique per-particle inputs and
// outputs once, plus the cell offsets).
double sph_pass_bytes(phx_sph_method method, int pass, uint32_t particles, uint32_t cells);

class SphSolver {
  public:
    SphSolver(const phx_domain_configuration* domain, const phx_sph_configuration* configuration,
              phx_execution_mode mode, hipStream_t stream);
    ~SphSolver();
    SphSolver(const SphSolver&) = delete;
    SphSolver& operator=(const SphSolver&) = delete;

    void forward_step(const phx_sph_state& state, const phx_control& control, const phx_sph_parameters& parameters,
                      phx_sph_state& next, SphCache& cache);
    // SPH<PCISPH|IISPH>::jvp_step / vjp_step (sph/sph.ixx:356-365) of a step
    // recorded in differentiable mode
    void jvp_step(const SphCache& cache, const phx_sph_state_tangent& state_t, const phx_control_tangent& control_t,
                  const phx_sph_parameter_tangent& parameter_t, phx_sph_state_tangent& next_t);
    void vjp_step(const SphCache& cache, const phx_sph_state_adjoint& next_a, phx_sph_state_adjoint& previous_a,
                  phx_control_adjoint& control_a, phx_sph_parameter_adjoint& parameter_a);

    const Domain& domain() const { return domain_; }
    hipStream_t stream() const { return stream_; }
    phx_sph_method method() const { return config_.method; }
    float gradient_norm() const { return gradient_norm_; }
    PassProfiler& profiler() { return profiler_; }
    // pressure iterations run by the last PCISPH/IISPH step
    uint32_t last_iterations() const { return last_iterations_; }

  private:
    void forward_wcsph(const phx_sph_state& state, const phx_control& control, const phx_sph_parameters& parameters,
                       phx_sph_state& next, SphCache& cache);
    void forward_pressure(const phx_sph_state& state, const phx_control& control,
                          const phx_sph_parameters& parameters, phx_sph_state& next, SphCache& cache);

    void require_derivatives(const SphCache& cache) const;
    void wcsph_jvp_step(const SphCache& cache, const phx_sph_state_tangent& state_t,
                        const phx_control_tangent& control_t, const phx_sph_parameter_tangent& parameter_t,
                        phx_sph_state_tangent& next_t);
    void wcsph_vjp_step(const SphCache& cache, const phx_sph_state_adjoint& next_a, phx_sph_state_adjoint& previous_a,
                        phx_control_adjoint& control_a, phx_sph_parameter_adjoint& parameter_a);

    Domain domain_;
    phx_sph_configuration config_{};
    phx_execution_mode mode_;
    hipStream_t stream_;
    // differentiable PCISPH/IISPH: the DFSPH projection engine without a
    // divergence phase or warm start (V = m / rho0 for PCISPH), which steps,
    // records and differentiates exactly these steps
    std::unique_ptr<DfsphSolver> engine_;
    DeviceBuffer<float> engine_warm_[2];  // its (unused) warm-start outputs
    WcsphDerivativeWorkspace wdw_;
    float gradient_norm_ = 0.0F;
    KernelConstants constants_{};
    PassProfiler profiler_;
    MaskSet masks_;
    // per-cell skins of the forward PCISPH/IISPH step (as DfsphSolver's):
    // last displacements (original order), their sorted copy, the per-cell
    // bounds and the cells' taint stamps
    DeviceBuffer<float> moved_, moved_sorted_, cell_bound_;
    DeviceBuffer<uint32_t> taint_;
    bool have_moved_ = false;
    uint32_t stamp_base_ = 1u;
    // the force pass's candidate records (x, y, z, t) and the step's mass state
    DeviceBuffer<float4> xt_;
    DeviceBuffer<uint32_t> mass_state_;
    uint32_t mass_stamp_ = 0u;  // this step's mark of unequal masses in mass_state_[0]
    // sorted-order workspace
    DeviceBuffer<float4> xm_, vmu_, xs_, aux_;
    VecBuffer u_;
    DeviceBuffer<float> rho0_, relax_, tait_, t_;
    DeviceBuffer<float> btension_;  // WCSPH boundary_surface_tension in cell order (boundary samples only)
    // tolerance mode (PCISPH, IISPH): per-wave FP64 error sums, the device
    // done word, its pinned host copies and their events
    DeviceBuffer<double> tol_partials_;
    DeviceBuffer<uint32_t> tol_done_;
    PinnedBuffer tol_host_;
    hipEvent_t tol_events_[2] = {nullptr, nullptr};
    uint32_t last_iterations_ = 0;
    BoundarySamples boundary_;      // the domain's boundary samples on the device
};

}  // namespace phx
