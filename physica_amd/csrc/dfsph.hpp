// SPH<DFSPH> solver (sph/sph.ixx:195-268, 273-372) on the sorted-order engine.
#pragma once

#include "engine.hpp"
#include "masks.hpp"
#include "profiler.hpp"

namespace phx {

// StepCache of SPH<DFSPH> (sph/sph.ixx:299-303): neighbourhood, densities and
// the non-pressure accelerations, all in sorted order. A differentiable
// solver also records, in sorted order, the step's inputs and every iterate
// of both pressure solves, so its jvp_step and vjp_step read them instead of
// recomputing iterations from checkpoints (DFSPH::Cache, sph/sph.ixx:222-228).
struct DfsphCache {
    hipStream_t stream = nullptr;
    Neighborhood neighborhood;
    DeviceBuffer<float> densities;
    DeviceBuffer<float4> non_pressure;  // xyz used

    bool record = false;  // set by the solver in differentiable mode
    bool recorded = false;
    uint32_t iterations[2] = {0u, 0u};  // divergence, density
    uint64_t step_index = 0;
    bool warm_start = false;
    bool volume_rest = false;  // V = m / rho0 in the pressure update (PCISPH)
    DeviceBuffer<float4> xm, vmu;  // positions + masses, velocities + viscosities
    DeviceBuffer<float> rho0, relax_div, relax_den;
    // projection pass r = 1 .. nd + ne (divergence first): predicted positions
    // (+ masses) and densities, slot r - 1
    DeviceBuffer<float4> xs;
    DeviceBuffer<float> rs;
    // pressures: divergence p_0 .. p_nd in slots 0 .. nd, density p_0 .. p_ne
    // in slots nd + 1 .. nd + ne + 1
    DeviceBuffer<float> p;
    // box-collision bits of prediction s = 0 .. nd + ne (the last gives the
    // next state)
    DeviceBuffer<uint8_t> collisions;
    DeviceBuffer<uint16_t> exact, skin;  // the step's neighbour masks and flags
    DeviceBuffer<uint8_t> crowded;       // per particle: walks the full stencil
    DeviceBuffer<uint32_t> mask_flags;
    float mask_limit2 = -1.0F;

    void allocate(uint32_t n, uint32_t cells, uint32_t divergence_iterations = 0u, uint32_t density_iterations = 0u);
    size_t pressure_offset(uint32_t phase, uint32_t k) const {
        return (static_cast<size_t>(phase == 0u ? 0u : iterations[0] + 1u) + k) * neighborhood.particles;
    }
    float* pressures(uint32_t phase, uint32_t k) { return p.data() + pressure_offset(phase, k); }
    const float* pressures(uint32_t phase, uint32_t k) const { return p.data() + pressure_offset(phase, k); }
};

// Sorted-order tangent and adjoint workspace of the derivative passes.
struct DfsphDerivativeWorkspace {
    // tangents (FP32)
    DeviceBuffer<float4> xt, vt, ut, anp_t, adiv_t, xs_t;  // .w: mass, viscosity, rest density tangents
    DeviceBuffer<float2> t_div, t_den;                      // (p / rho^2, its tangent)
    DeviceBuffer<float> rho_t, p_t_div, p_t_den, rest_t, relax_div_t, relax_den_t;
    // adjoints (FP64)
    DVecBuffer pos_a, vel_a, anp_a, ap_a[3];
    DeviceBuffer<double> p_a[3], rs_a, rho_a, m_a, rest_a, visc_a, relax_div_a, relax_den_a;
    void allocate_tangents(uint32_t n);
    void allocate_adjoints(uint32_t n);
};

// Passes of one DFSPH step, for the profiler.
enum DfsphPass : int {
    kDfsphGrid = 0,
    kDfsphDensity,
    kDfsphForceInit,
    kDfsphProjectionDivergence,
    kDfsphForceDivergence,
    kDfsphForceDivergenceEnd,
    kDfsphProjectionDensity,
    kDfsphForceDensity,
    kDfsphJvp,
    kDfsphVjp,
    kDfsphPassCount
};
const char* dfsph_pass_name(int pass);
// Algorithmic bytes per particle per launch: the pass's unique per-particle
// inputs and outputs moved once, plus 4 (C + 1) / N for the cell offsets it
// reads (the grid pass writes them and clears the histogram: 8 (C + 1) / N).
// The neighbour masks are an acceleration structure and are not counted.
double dfsph_pass_bytes(int pass, uint32_t particles, uint32_t cells, bool warm_start);

class DfsphSolver {
  public:
    DfsphSolver(const phx_domain_configuration* domain, const phx_dfsph_configuration* configuration,
                phx_execution_mode mode, hipStream_t stream);
    ~DfsphSolver();
    DfsphSolver(const DfsphSolver&) = delete;
    DfsphSolver& operator=(const DfsphSolver&) = delete;

    void forward_step(const phx_sph_dfsph_state& state, const phx_control& control,
                      const phx_sph_dfsph_parameters& parameters, phx_sph_dfsph_state& next, DfsphCache& cache);
    // SPH<DFSPH>::jvp_step (sph/sph.ixx:356-359, DFSPH::jvp sph/sph.cpp:612-659)
    // of the step recorded in `cache`; tangents FP32, overwriting next_t.
    void jvp_step(const DfsphCache& cache, const phx_sph_dfsph_state_tangent& state_t,
                  const phx_control_tangent& control_t, const phx_sph_dfsph_parameter_tangent& parameter_t,
                  phx_sph_dfsph_state_tangent& next_t);
    // SPH<DFSPH>::vjp_step (sph/sph.ixx:361-365, DFSPH::vjp sph/sph.cpp:661-723);
    // FP64 adjoints accumulated into the outputs, warm-start pressure
    // adjoints copied (written only with warm start).
    void vjp_step(const DfsphCache& cache, const phx_sph_dfsph_state_adjoint& next_a,
                  phx_sph_dfsph_state_adjoint& previous_a, phx_control_adjoint& control_a,
                  phx_sph_dfsph_parameter_adjoint& parameter_a);

    const Domain& domain() const { return domain_; }
    hipStream_t stream() const { return stream_; }
    float gradient_norm() const { return gradient_norm_; }
    const uint32_t* last_iterations() const { return last_iterations_; }
    PassProfiler& profiler() { return profiler_; }
    phx_execution_mode mode() const { return mode_; }
    bool warm_start() const { return config_.pressure_warm_start != 0u; }
    float skin() const { return masks_.skin(); }
    // Takes V = m / rho0 instead of m / rho in the density phase's pressure
    // update: with no divergence iterations and no warm start the step is
    // then PCISPH's (pcisph.cu:104-110), as it is IISPH's without.
    void set_volume_rest_density(bool enabled) { volume_rest_ = enabled; }

  private:
    void require_recorded(const DfsphCache& cache) const;

    Domain domain_;
    phx_dfsph_configuration config_{};
    phx_execution_mode mode_;
    bool volume_rest_ = false;
    hipStream_t stream_;
    float gradient_norm_ = 0.0F;
    KernelConstants constants_{};
    uint32_t last_iterations_[2]{};
    // tolerance mode: per-iteration FP64 error sums and arrival counters of
    // each phase, the phases' done words, and their host copies (pinned),
    // read one iteration behind the launches through two events
    DeviceBuffer<double> tol_partials_;
    DeviceBuffer<uint32_t> tol_done_;
    uint32_t* tol_host_ = nullptr;
    uint32_t* tol_host_device_ = nullptr;  // tol_host_ as the device addresses it
    hipEvent_t tol_events_[2] = {nullptr, nullptr};
    PassProfiler profiler_;
    // sorted-order workspace
    DeviceBuffer<float4> xm_, vmu_, xs_, adiv_;
    VecBuffer u_;
    DeviceBuffer<float> rho0_, relax_div_, relax_den_, p_div_, p_den_, t_div_, t_den_;
    MaskSet masks_;
    // Per-cell skins (forward mode, no boundary samples): each particle's
    // displacement in the last forward step (original order), its sorted
    // copy, the step's per-cell bounds, and the cells' taint stamps (one
    // stamp per predicting pass, never reused until the counter wraps).
    DeviceBuffer<float> moved_, moved_sorted_, cell_bound_;
    DeviceBuffer<uint32_t> taint_;
    bool have_moved_ = false;
    uint32_t stamp_base_ = 1u;
    // the force passes' candidate records (x, y, z, t) of each phase and the
    // step's mass state (all masses equal, their bits)
    DeviceBuffer<float4> xt_div_, xt_den_;
    DeviceBuffer<uint32_t> mass_state_;
    uint32_t mass_stamp_ = 0u;  // this step's mark of unequal masses in mass_state_[0]
    BoundarySamples boundary_;  // the domain's boundary samples on the device
    DfsphDerivativeWorkspace dw_;
};

// Constants of the cubic and viscosity kernels for support radius h.
KernelConstants kernel_constants(float h);

}  // namespace phx
