// SPH<DFSPH> solver (sph/sph.ixx:195-268, 273-372) on the sorted-order engine.
#pragma once

#include "engine.hpp"
#include "masks.hpp"
#include "profiler.hpp"

namespace phx {

// StepCache of SPH<DFSPH> (sph/sph.ixx:299-303): neighbourhood, densities and
// the non-pressure accelerations, all in sorted order.
struct DfsphCache {
    hipStream_t stream = nullptr;
    Neighborhood neighborhood;
    DeviceBuffer<float> densities;
    DeviceBuffer<float4> non_pressure;  // xyz used
    void allocate(uint32_t n, uint32_t cells);
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

    const Domain& domain() const { return domain_; }
    hipStream_t stream() const { return stream_; }
    float gradient_norm() const { return gradient_norm_; }
    const uint32_t* last_iterations() const { return last_iterations_; }
    PassProfiler& profiler() { return profiler_; }
    phx_execution_mode mode() const { return mode_; }
    bool warm_start() const { return config_.pressure_warm_start != 0u; }
    float skin() const { return masks_.skin(); }

  private:

    Domain domain_;
    phx_dfsph_configuration config_{};
    phx_execution_mode mode_;
    hipStream_t stream_;
    float gradient_norm_ = 0.0F;
    KernelConstants constants_{};
    uint32_t last_iterations_[2]{};
    PassProfiler profiler_;
    // sorted-order workspace
    DeviceBuffer<float4> xm_, vmu_, xs_, adiv_;
    VecBuffer u_;
    DeviceBuffer<float> rho0_, relax_div_, relax_den_, p_div_, p_den_, t_div_, t_den_;
    MaskSet masks_;
};

// Constants of the cubic and viscosity kernels for support radius h.
KernelConstants kernel_constants(float h);

// Smallest float t with sqrtf(t) >= h: a candidate at squared distance d2
// lies inside the support (sqrtf(d2) < h) exactly when d2 < t.
float sqrt_support_threshold(float h);

}  // namespace phx
