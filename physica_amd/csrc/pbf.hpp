// PBF solver (pbf/pbf.ixx:14-141) on the sorted-order engine, with the
// forward-mode and reverse-mode derivatives of one step (pbf/pbf.cpp:193-258).
#pragma once

#include "engine.hpp"
#include "masks.hpp"
#include "profiler.hpp"

namespace phx {

// `count` vector fields of n particles in one allocation; field k is
// [3k n, 3k n + 3 n).
struct VecStack {
    DeviceBuffer<float> data;
    size_t n = 0;
    void resize(size_t particles, size_t count) {
        n = particles;
        data.resize(3u * particles * count);
    }
    Vec view(size_t k) {
        float* b = data.data() + 3u * n * k;
        return {b, b + n, b + 2u * n};
    }
    CVec cview(size_t k) const {
        const float* b = data.data() + 3u * n * k;
        return {b, b + n, b + 2u * n};
    }
};

// StepCache of PBF (pbf/pbf.ixx:54-70): neighbourhood built on the predicted
// positions, the predicted positions (original order) and the densities and
// lambdas of the last constraint iteration (sorted order). A differentiable
// solver also records every intermediate of the step in sorted order, so its
// jvp_step and vjp_step read them instead of recomputing constraint
// iterations from checkpoints.
struct PbfCache {
    Neighborhood neighborhood;
    VecBuffer predicted;  // original order
    DeviceBuffer<float> densities;
    DeviceBuffer<float> lambdas;

    bool record = false;  // set by the solver in differentiable mode
    bool recorded = false;
    uint32_t iterations = 0;
    uint64_t step_index = 0;
    VecBuffer xs, x0;                      // x* (topology and first iterate), state positions
    VecStack positions;                    // position after constraint iteration k
    DeviceBuffer<float> rho, lambda, den;  // [iterations][n]
    VecStack gradient_sums;                // sum_j (m_j / rho0_i) grad W_ij per iteration
    DeviceBuffer<uint8_t> masks;           // [iterations][n] box-collision bits of the projection
    VecBuffer velocities, vorticities, normals, confined;
    DeviceBuffer<float> magnitudes;  // vorticity magnitudes (cell order), every mode
    DeviceBuffer<float> normalizers;
    // the step's neighbour masks and mask flags, so the derivative passes
    // walk the same candidates as the forward passes
    DeviceBuffer<uint16_t> masks_skin;
    DeviceBuffer<uint8_t> crowded;  // per particle: walks the full stencil (a stencil cell overflows the masks)
    DeviceBuffer<uint32_t> mask_flags;
    float mask_limit2 = -1.0F;
    // exact masks of every set of positions (iterates 0 .. I - 1, then the
    // final positions), [set][kStencil][n], and one overflow word per set
    DeviceBuffer<uint16_t> exact_masks;
    DeviceBuffer<uint32_t> exact_flags;

    void allocate(uint32_t n, uint32_t cells, uint32_t pressure_iterations);
};

enum PbfPass : int {
    kPbfGrid = 0,
    kPbfMask,
    kPbfLambda,
    kPbfCorrection,
    kPbfVorticity,
    kPbfNormal,
    kPbfXsph,
    kPbfJvp,
    kPbfVjp,
    kPbfPassCount
};
const char* pbf_pass_name(int pass);
double pbf_pass_bytes(int pass, uint32_t particles, uint32_t cells);

// Sorted-order parameter and tangent workspace of the derivative passes.
struct PbfDerivativeWorkspace {
    DeviceBuffer<float> m, rho0, eps, strength, exponent, radius, xsph, confinement;
    // tangents (FP32) of parameters, state and control, sorted
    DeviceBuffer<float> m_t, rho0_t, eps_t, strength_t, exponent_t, radius_t, xsph_t, confinement_t;
    VecBuffer x_t, v_t, u_t, p_t[2], vrec_t, w_t, vc_t;
    DeviceBuffer<float> rho_t, lambda_t, magnitude_t;
    // adjoints (FP64), sorted
    DVecBuffer p_a[2], vrec_a, w_a, g_a, next_x_a, next_v_a, x_a, vel_a;
    DeviceBuffer<double> lambda_a, magnitude_a;
    DeviceBuffer<double> m_a, rho0_a, eps_a, strength_a, exponent_a, radius_a, xsph_a, confinement_a;
    DeviceBuffer<uint8_t> vanishes;  // per particle: the artificial pressure is the signed zero -k
    DeviceBuffer<double> inv_rho0;   // 1 / rho0
    DeviceBuffer<double2> lambda_terms;  // (density adjoint, D_bar) of one iteration
    void allocate(uint32_t n);
};

class PbfSolver {
  public:
    PbfSolver(const phx_domain_configuration* domain, const phx_pbf_configuration* configuration,
              phx_execution_mode mode, hipStream_t stream);
    void forward_step(const phx_pbf_state& state, const phx_control& control, const phx_pbf_parameters& parameters,
                      phx_pbf_state& next, PbfCache& cache);
    void jvp_step(const phx_pbf_parameters& parameters, const PbfCache& cache, const phx_pbf_state_tangent& state_t,
                  const phx_control_tangent& control_t, const phx_pbf_parameter_tangent& parameter_t,
                  phx_pbf_state_tangent& next_t);
    void vjp_step(const phx_pbf_parameters& parameters, const PbfCache& cache, const phx_pbf_state_adjoint& next_a,
                  phx_pbf_state_adjoint& previous_a, phx_control_adjoint& control_a,
                  phx_pbf_parameter_adjoint& parameter_a);

    const Domain& domain() const { return domain_; }
    hipStream_t stream() const { return stream_; }
    PassProfiler& profiler() { return profiler_; }
    phx_execution_mode mode() const { return mode_; }
    uint32_t iterations() const { return config_.pressure_iterations; }

  private:
    void gather_parameters(const phx_pbf_parameters& parameters, const PbfCache& cache);


    Domain domain_;
    phx_pbf_configuration config_{};
    phx_execution_mode mode_;
    hipStream_t stream_;
    float spiky_ = 0.0F;
    PassProfiler profiler_;
    MaskSet masks_;
    BoundarySamples boundary_;  // the domain's boundary samples on the device
    // sorted-order workspace
    VecBuffer xs_, x0_, pa_, pb_, vel_, vort_, vc_;
    DeviceBuffer<float> m_, rho0_, eps_, strength_, exponent_, radius_, xsph_, confinement_;
    DeviceBuffer<uint16_t> exact_;  // exact masks of the current set of positions (forward mode)
    DeviceBuffer<uint8_t> crowded_;  // per particle: walks the full stencil (forward mode)
    DeviceBuffer<uint32_t> exact_flags_;
    PbfDerivativeWorkspace dw_;
};

// Independent PBF scenes stepped together, forward mode (the configured
// uneven batch): the scenes share the support radius, time step and
// iteration count, and each keeps its own box, gravity and particle count.
// Every pass of a substep is one launch over all scenes (PBF::forward_step,
// pbf/pbf.cpp:167-191, per scene), and with static boxes a substep is
// captured once in a HIP graph and replayed.
class PbfBatch {
  public:
    PbfBatch(const phx_domain_configuration* domains, uint32_t scene_count, const phx_pbf_configuration* configuration,
             const phx_pbf_batch_configuration* batch_configuration, hipStream_t stream);
    ~PbfBatch();
    PbfBatch(const PbfBatch&) = delete;
    PbfBatch& operator=(const PbfBatch&) = delete;

    // `steps` substeps of every scene, in place on the concatenated state
    // (scene k holds particles [offsets[k], offsets[k + 1])).
    void run(phx_pbf_state& state, const phx_control& control, const phx_pbf_parameters& parameters, uint32_t steps);

    uint32_t scene_count() const { return static_cast<uint32_t>(scenes_.size()); }
    uint32_t particle_count() const { return n_; }
    const std::vector<uint64_t>& offsets() const { return offsets_; }
    uint32_t kernels_per_step() const { return kernels_per_step_; }
    bool graphed() const { return use_graph_; }

  private:
    void substep(uint32_t from, uint64_t step, const phx_control& control, const phx_pbf_parameters& parameters);
    void capture(const phx_control& control, const phx_pbf_parameters& parameters);
    void release_graphs();

    phx_pbf_configuration config_{};
    hipStream_t caller_;
    hipStream_t stream_ = nullptr;  // the batch's own stream (graphs cannot be captured on the null stream)
    hipEvent_t ready_ = nullptr, done_ = nullptr;
    std::vector<BatchScene> scenes_;
    std::vector<uint64_t> offsets_;
    uint32_t n_ = 0, cells_ = 0, blocks_ = 0;
    float h_ = 0.0F, dt_ = 0.0F, spiky_ = 0.0F, skin_ = 0.3F;
    bool static_boxes_ = true, use_graph_ = false;
    uint32_t kernels_per_step_ = 0;
    DeviceBuffer<BatchBlock> block_table_;
    DeviceBuffer<BatchScene> scene_table_;
    Neighborhood grid_;
    VecBuffer positions_[2], velocities_[2];  // the state in original order, ping-pong
    VecBuffer predicted_, xs_, x0_, pa_, pb_, vel_, vort_, vc_;
    DeviceBuffer<float> m_, rho0_, eps_, strength_, exponent_, radius_, xsph_, confinement_;
    DeviceBuffer<float> rho_, lambda_, magnitude_;
    DeviceBuffer<uint16_t> skin_masks_, exact_;
    DeviceBuffer<uint8_t> crowded_;
    DeviceBuffer<uint32_t> flags_, exact_flags_;
    DeviceBuffer<uint32_t> cell_moved_;  // per cell: displacement bound of the substep (MaskView::cell_moved)
    // captured substeps 0 -> 1 and 1 -> 0, and the control and parameter
    // arrays they read
    hipGraphExec_t graph_[2] = {nullptr, nullptr};
    std::vector<const void*> captured_inputs_;
};

}  // namespace phx
