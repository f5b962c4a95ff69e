// PBF solver (pbf/pbf.ixx:14-141) on the sorted-order engine.
#pragma once

#include "engine.hpp"
#include "profiler.hpp"

namespace phx {

// StepCache of PBF (pbf/pbf.ixx:54-70): neighbourhood built on the predicted
// positions, the predicted positions (original order) and the densities and
// lambdas of the last constraint iteration (sorted order).
struct PbfCache {
    Neighborhood neighborhood;
    VecBuffer predicted;  // original order
    DeviceBuffer<float> densities;
    DeviceBuffer<float> lambdas;
    void allocate(uint32_t n, uint32_t cells);
};

enum PbfPass : int {
    kPbfGrid = 0,
    kPbfLambda,
    kPbfCorrection,
    kPbfVorticity,
    kPbfNormal,
    kPbfXsph,
    kPbfPassCount
};
const char* pbf_pass_name(int pass);
double pbf_pass_bytes(int pass, uint32_t particles, uint32_t cells);

class PbfSolver {
  public:
    PbfSolver(const phx_domain_configuration* domain, const phx_pbf_configuration* configuration,
              phx_execution_mode mode, hipStream_t stream);
    void forward_step(const phx_pbf_state& state, const phx_control& control, const phx_pbf_parameters& parameters,
                      phx_pbf_state& next, PbfCache& cache);

    const Domain& domain() const { return domain_; }
    hipStream_t stream() const { return stream_; }
    PassProfiler& profiler() { return profiler_; }
    phx_execution_mode mode() const { return mode_; }

  private:
    Domain domain_;
    phx_pbf_configuration config_{};
    phx_execution_mode mode_;
    hipStream_t stream_;
    float spiky_ = 0.0F;
    PassProfiler profiler_;
    // sorted-order workspace
    VecBuffer xs_, x0_, pa_, pb_, vel_, vort_, vc_;
    DeviceBuffer<float> m_, rho0_, eps_, strength_, exponent_, radius_, xsph_, confinement_, magnitude_;
};

}  // namespace phx
