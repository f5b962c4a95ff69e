// SPH<DFSPH> forward step for gfx950 in sorted (cell-major) particle order.
//
// Reference: SPH<DFSPH>::forward_step (sph/sph.ixx:350-354) and
// DFSPH::forward (sph/sph.cpp:572-610) with kernels from sph/common.cu,
// sph/dfsph.cu and density/density.cu. The reference runs one kernel per
// operation in original particle order; here particles are reordered by cell
// once per step, each per-particle chain is fused into the neighbour pass
// that produces its input, and every neighbour pass walks a precomputed mask
// instead of all 27 cells:
//
//   masks+density  27-cell test once per step -> exact (r < h) and skin masks;
//                  rho = sum m W over the exact mask; t = p / rho^2 (warm p)
//   force (init)   non-pressure + pressure(t_div) + predict -> x*
//   per iteration  A: rho* at x* over the skin mask, projection update -> p, t
//                  B: pressure(t) over the exact mask + predict -> x* (or the
//                     next state)
//
// The end of the divergence phase and the density phase's initial pressure
// are one pass with two accumulators. Every per-particle sum runs in the
// reference order (27 cells, offset_x outer, ascending original index) and a
// masked-out candidate contributes an exact zero in the reference, so results
// equal the reference expression by expression. An A pass uses the skin mask
// only if no particle moved farther than half the skin (read on the device
// from the preceding B pass); otherwise it scans the full stencil. In forward
// mode after the first step the skin is per cell instead: each cell's bound
// comes from its particles' displacements in the previous step, a pair is
// kept while it lies within h plus both cells' bounds, and a B pass marks
// the 27 cells around every particle that moved past its cell's bound, whose
// particles then scan the full stencil in the A pass that reads it.
#include <cmath>
#include <cstring>

#include "dfsph.hpp"
#include "sph_kernels.hpp"

namespace phx {

const char* dfsph_pass_name(int pass) {
    static const char* names[kDfsphPassCount] = {"grid",       "mask_density",     "force_init",
                                                 "projection_divergence", "force_divergence",
                                                 "force_divergence_end",  "projection_density",
                                                 "force_density",         "jvp",
                                                 "vjp"};
    return pass >= 0 && pass < kDfsphPassCount ? names[pass] : "unknown";
}

double dfsph_pass_bytes(int pass, uint32_t particles, uint32_t cells, bool warm_start) {
    const double offsets = particles > 0 ? 4.0 * (static_cast<double>(cells) + 1.0) / particles : 0.0;
    switch (pass) {
        case kDfsphGrid:  // x, v, u (36) + m, rho0, mu, two relaxations (20) [+ two warm p (8)], read and
                          // written in cell order, + the permutation (4)
            return 2.0 * (56.0 + (warm_start ? 8.0 : 0.0)) + 4.0 + 2.0 * offsets;
        case kDfsphDensity:  // x, m, p_div, p_den in; rho, t_div, t_den out
            return 24.0 + 12.0 + offsets;
        case kDfsphForceInit:  // x, v, u, m, mu, rho, t in; a_np, x* out
            return 52.0 + 24.0 + offsets;
        case kDfsphProjectionDivergence:  // x, x*, m, rho, relaxation, p in; p, t out
            return 40.0 + 8.0 + offsets;
        case kDfsphForceDivergence:  // x, v, m, t, a_np in; x* out
            return 44.0 + 12.0 + offsets;
        case kDfsphForceDivergenceEnd:  // x, v, m, t_div, t_den, a_np, p_div in; a_div, x*, warm p_div out
            return 52.0 + 28.0 + offsets;
        case kDfsphProjectionDensity:  // as divergence plus the rest density target
            return 44.0 + 8.0 + offsets;
        case kDfsphForceDensity:  // x, v, m, t, a_np, a_div in; x* out
            return 56.0 + 12.0 + offsets;
        case kDfsphJvp:  // whole-step derivative passes, not a single roofline figure
        case kDfsphVjp:
            return 0.0;
        default:
            return 0.0;
    }
}

void DfsphCache::allocate(uint32_t n, uint32_t cells, uint32_t divergence_iterations, uint32_t density_iterations) {
    neighborhood.allocate(n, cells);
    densities.resize(n);
    non_pressure.resize(n);
    if (!record) return;
    const size_t nn = n;
    const size_t updates = static_cast<size_t>(divergence_iterations) + density_iterations;
    for (DeviceBuffer<float4>* b : {&xm, &vmu}) b->resize(nn);
    for (DeviceBuffer<float>* b : {&rho0, &relax_div, &relax_den}) b->resize(nn);
    xs.resize(updates * nn);
    rs.resize(updates * nn);
    p.resize((updates + 2u) * nn);
    collisions.resize((updates + 1u) * nn);
    exact.resize(static_cast<size_t>(kStencil) * nn);
    skin.resize(static_cast<size_t>(kStencil) * nn);
    crowded.resize(nn);
    mask_flags.resize(kFlagDisplacement + updates + 1u);
}

KernelConstants kernel_constants(float h) {
    KernelConstants k{};
    k.h = h;
    k.cubic = 8.0F / (kPi * h * h * h);
    k.viscosity = kPi * h * h * h * h * h * h;
    k.sqrt_support = sqrt_support_threshold(h);
    k.inv_h = 1.0F / h;
    k.inv_viscosity = 1.0F / k.viscosity;
    return k;
}

DfsphSolver::DfsphSolver(const phx_domain_configuration* domain, const phx_dfsph_configuration* configuration,
                         phx_execution_mode mode, hipStream_t stream)
    : domain_(make_domain(domain)), mode_(mode), stream_(stream) {
    require(configuration != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null DFSPH configuration");
    config_ = *configuration;
    require(config_.checkpoint_interval > 0u, PHX_ERROR_INVALID_ARGUMENT, "checkpoint_interval must be positive");
    require(mode == PHX_MODE_FORWARD || mode == PHX_MODE_DIFFERENTIABLE, PHX_ERROR_INVALID_ARGUMENT,
            "unknown execution mode");
    const float tolerance = config_.density_error_tolerance;
    require(tolerance >= 0.0F && std::isfinite(tolerance), PHX_ERROR_INVALID_ARGUMENT,
            "density_error_tolerance must be finite and non-negative");
    require(tolerance == 0.0F || mode == PHX_MODE_FORWARD, PHX_ERROR_UNSUPPORTED,
            "tolerance mode is supported in forward mode only");
    require(tolerance == 0.0F || (config_.max_iterations >= 1u && config_.max_iterations < 2048u),
            PHX_ERROR_INVALID_ARGUMENT, "tolerance mode needs 1 <= max_iterations < 2048");
    require(config_.divergence_iterations + config_.density_iterations < 4096u, PHX_ERROR_INVALID_ARGUMENT,
            "too many pressure iterations");
    boundary_.upload(domain_.config);
    domain_.config.boundary_particles = nullptr;  // the caller's host array is not kept
    gradient_norm_ = reference_gradient_norm(domain_.config);
    constants_ = kernel_constants(domain_.config.support_radius);
    const uint32_t n = domain_.config.particle_count;
    for (DeviceBuffer<float4>* b : {&xm_, &vmu_, &xs_, &adiv_, &xt_div_, &xt_den_}) b->resize(n);
    mass_state_.resize(2u);
    fill_u32(stream_, mass_state_.data(), 2u, 0u);
    u_.resize(n);
    for (DeviceBuffer<float>* b : {&rho0_, &relax_div_, &relax_den_, &p_div_, &p_den_, &t_div_, &t_den_}) b->resize(n);
    // one displacement slot per predicting pass: the first, then one per
    // iteration of each phase
    const uint32_t predictions = tolerance > 0.0F ? 2u * config_.max_iterations + 1u
                                                  : config_.divergence_iterations + config_.density_iterations + 1u;
    masks_.allocate(n, predictions, domain_.config.support_radius);
    if (mode == PHX_MODE_FORWARD) {
        for (DeviceBuffer<float>* b : {&moved_, &moved_sorted_}) b->resize(n);
        cell_bound_.resize(domain_.cells);
        taint_.resize(domain_.cells);
        fill_u32(stream_, taint_.data(), domain_.cells, 0u);
    }
    if (tolerance > 0.0F) {
        tol_partials_.resize(static_cast<size_t>(pass_blocks(n)) * kWaves);  // one per projection wave
        tol_done_.resize(2u);
        PHX_HIP(hipHostMalloc(reinterpret_cast<void**>(&tol_host_), 4u * sizeof(uint32_t), hipHostMallocDefault));
        PHX_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&tol_host_device_), tol_host_, 0u));
        for (hipEvent_t& e : tol_events_) PHX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
}

DfsphSolver::~DfsphSolver() {
    for (hipEvent_t e : tol_events_)
        if (e != nullptr) (void)hipEventDestroy(e);
    if (tol_host_ != nullptr) (void)hipHostFree(tol_host_);
}

void DfsphSolver::forward_step(const phx_sph_dfsph_state& state, const phx_control& control,
                               const phx_sph_dfsph_parameters& parameters, phx_sph_dfsph_state& next,
                               DfsphCache& cache) {
    const auto& c = domain_.config;
    const uint32_t n = c.particle_count;
    const uint64_t step = state.particles.step_index;
    if (n == 0) {
        next.particles.step_index = step + 1u;
        last_iterations_[0] = config_.divergence_iterations;
        last_iterations_[1] = config_.density_iterations;
        return;
    }
    require_vec(state.particles.positions, "state positions");
    require_vec(state.particles.velocities, "state velocities");
    require_vec(control.external_accelerations, "control");
    require_vec(next.particles.positions, "next positions");
    require_vec(next.particles.velocities, "next velocities");
    const auto& pp = parameters.particles;
    require(pp.masses && pp.rest_densities && pp.viscosities, PHX_ERROR_INVALID_ARGUMENT, "null particle parameters");
    require(parameters.method.divergence_relaxation && parameters.method.density_relaxation,
            PHX_ERROR_INVALID_ARGUMENT, "null DFSPH relaxation parameters");
    const bool warm = config_.pressure_warm_start != 0u;
    if (warm) {
        require(state.method.warm_divergence_pressure && state.method.warm_density_pressure,
                PHX_ERROR_INVALID_ARGUMENT, "null warm-start pressures");
    }
    require(next.method.warm_divergence_pressure && next.method.warm_density_pressure, PHX_ERROR_INVALID_ARGUMENT,
            "null next warm-start pressures");
    const uint32_t nd = config_.divergence_iterations;
    const uint32_t ne = config_.density_iterations;
    cache.record = mode_ == PHX_MODE_DIFFERENTIABLE;
    cache.recorded = false;
    cache.iterations[0] = nd;
    cache.iterations[1] = ne;
    cache.allocate(n, domain_.cells, nd, ne);
    const bool rec = cache.record;
    const size_t nn = n;
    // a recorded step keeps its inputs, masks and iterates in the cache
    float4* xm = rec ? cache.xm.data() : xm_.data();
    float4* vmu = rec ? cache.vmu.data() : vmu_.data();
    float* rho0 = rec ? cache.rho0.data() : rho0_.data();
    float* relax_div = rec ? cache.relax_div.data() : relax_div_.data();
    float* relax_den = rec ? cache.relax_den.data() : relax_den_.data();
    float* p_div0 = rec ? cache.pressures(0u, 0u) : p_div_.data();
    float* p_den0 = rec ? cache.pressures(1u, 0u) : p_den_.data();
    uint16_t* exact = rec ? cache.exact.data() : masks_.exact();
    uint16_t* skin = rec ? cache.skin.data() : masks_.skin_mask();
    uint8_t* crowded = rec ? cache.crowded.data() : masks_.crowded();

    // neighborhood.build(step, x) with every input gathered into cell order
    ChannelList gather;
    gather.add(cvec(state.particles.positions), xm);
    gather.add(pp.masses, xm, 3);
    gather.add(cvec(state.particles.velocities), vmu);
    gather.add(pp.viscosities, vmu, 3);
    gather.add(cvec(control.external_accelerations), u_.view());
    gather.add(pp.rest_densities, rho0);
    gather.add(parameters.method.divergence_relaxation, relax_div);
    gather.add(parameters.method.density_relaxation, relax_den);
    if (warm) {
        gather.add(state.method.warm_divergence_pressure, p_div0);
        gather.add(state.method.warm_density_pressure, p_den0);
    }
    // per-cell skins need the previous step's displacements
    const bool cell_skins = !rec && have_moved_ && boundary_.count() == 0u;
    if (cell_skins) gather.add(moved_.data(), moved_sorted_.data());
    // stamps of this step's predicting passes: stamp_base_ + slot
    const uint32_t predictions = config_.density_error_tolerance > 0.0F ? 2u * config_.max_iterations + 1u
                                                                         : nd + ne + 1u;
    if (cell_skins && stamp_base_ > 0xFFFFFFFFu - predictions) {
        fill_u32(stream_, taint_.data(), domain_.cells, 0u);
        stamp_base_ = 1u;
    }
    {
        PassScope scope(profiler_, stream_, kDfsphGrid);
        cache.neighborhood.build(stream_, domain_, step, cvec(state.particles.positions), nullptr, gather.get(),
                                 &boundary_);
        if (!warm) {
            fill_f32(stream_, p_div0, n, 0.0F);
            fill_f32(stream_, p_den0, n, 0.0F);
        }
        // boundary samples: every pass walks the full stencil
        masks_.begin_step(stream_, boundary_.count() > 0u);
        // the mask pass marks unequal masses with this step's stamp
        if (++mass_stamp_ == 0u) {
            fill_u32(stream_, mass_state_.data(), 1u, 0u);
            mass_stamp_ = 1u;
        }
        if (cell_skins) {
            // 1.2 x the previous step's largest displacement in the cell plus
            // 0.01 h: on the dam break at step 100 this keeps 1.28 x the
            // exact neighbours of the next predictions with no particle
            // past its bound, where the global skin keeps 2.39 x
            const uint32_t cells = domain_.cells;
            cell_bound_kernel<<<blocks_for(cells, kBlock), kBlock, 0, stream_>>>(
                cells, cache.neighborhood.grid().offsets, moved_sorted_.data(), 1.2F, 0.01F * c.support_radius,
                cell_bound_.data());
            PHX_LAUNCH_CHECK();
        }
    }
    const SortedGrid grid = cache.neighborhood.grid();
    const BoundaryGrid boundary = cache.neighborhood.boundary();
    const uint32_t blocks = pass_blocks(n);

    MaskDensityArgs ma{n,
                       constants_,
                       masks_.t_skin(),
                       grid,
                       boundary,
                       xm,
                       exact,
                       skin,
                       p_div0,
                       p_den0,
                       cache.densities.data(),
                       t_div_.data(),
                       t_den_.data(),
                       masks_.flags(),
                       rho0};
    ma.crowded = crowded;
    if (cell_skins) ma.cell_bound = cell_bound_.data();
    ma.xt_div = xt_div_.data();
    ma.xt_den = xt_den_.data();
    ma.mass_state = mass_state_.data();
    ma.mass_stamp = mass_stamp_;
    launch_mask_density<false>(profiler_, stream_, blocks, ma, kDfsphDensity);

    ForceArgs fa{};
    fa.n = n;
    fa.kc = constants_;
    fa.dt = c.time_step;
    fa.grid = grid;
    fa.box = collision_box(c, step);
    fa.gravity = Float3{c.gravity.x, c.gravity.y, c.gravity.z};
    fa.xm = xm;
    fa.vmu = vmu;
    fa.u = u_.cview();
    fa.rho = cache.densities.data();
    fa.exact = exact;
    fa.flags = masks_.flags();
    fa.anp = cache.non_pressure.data();
    fa.adiv = adiv_.data();
    fa.perm = cache.neighborhood.perm.data();
    fa.next_x = vec(next.particles.positions);
    fa.next_v = vec(next.particles.velocities);
    fa.p_div = rec ? cache.pressures(0u, nd) : p_div_.data();
    fa.warm_div_out = next.method.warm_divergence_pressure;
    fa.p_den = rec ? cache.pressures(1u, ne) : p_den_.data();
    fa.warm_den_out = next.method.warm_density_pressure;
    fa.boundary = boundary;
    fa.rest_density = rho0;
    fa.crowded = crowded;
    if (cell_skins) {
        fa.cell_bound = cell_bound_.data();
        fa.taint = taint_.data();
    }
    if (!rec) fa.moved_out = moved_.data();
    fa.mass_state = mass_state_.data();
    fa.mass_stamp = mass_stamp_;
    // prediction `slot` writes x* (recorded per slot) and, recorded, its collision bits
    auto set_slot = [&](uint32_t slot) {
        fa.slot = slot;
        fa.stamp = stamp_base_ + slot;
        fa.xs = rec ? cache.xs.data() + slot * nn : xs_.data();
        fa.collisions = rec ? cache.collisions.data() + slot * nn : nullptr;
    };

    ProjectionArgs pa{};
    pa.n = n;
    pa.kc = constants_;
    pa.dt = c.time_step;
    pa.gradient_norm = gradient_norm_;
    pa.limit2 = masks_.limit2();
    pa.grid = grid;
    pa.xm = xm;
    pa.skin = skin;
    pa.flags = masks_.flags();
    pa.rho = cache.densities.data();
    pa.volume_density = volume_rest_ ? rho0 : cache.densities.data();
    pa.boundary = boundary;
    pa.rest_density = rho0;
    pa.crowded = crowded;
    if (cell_skins) pa.taint = taint_.data();
    // update `it` of phase `phase` reads the x* of prediction `slot`
    auto set_update = [&](uint32_t phase, uint32_t it, uint32_t slot) {
        pa.slot = slot;
        pa.stamp = stamp_base_ + slot;
        pa.xs = rec ? cache.xs.data() + slot * nn : xs_.data();
        pa.rs_out = rec ? cache.rs.data() + slot * nn : nullptr;
        float* base = phase == 0u ? p_div_.data() : p_den_.data();
        pa.p_prev = rec ? cache.pressures(phase, it - 1u) : base;
        pa.p = rec ? cache.pressures(phase, it) : base;
    };

    uint32_t slot = 0;  // displacement slot written by the latest predicting pass

    if (config_.density_error_tolerance > 0.0F) {
        // Tolerance mode: each phase runs until its projection's mean
        // compression error is within the tolerance (or max_iterations).
        // The projection leaves one FP64 error sum per wave; a
        // one-workgroup pass adds them in a fixed order and records the
        // ending iteration on the device; the force pass of that iteration
        // then ends the phase by itself (kDivAuto, or the final state). The
        // host reads each iteration's done word while that iteration's
        // force pass runs and only then launches the next iteration, so
        // the GPU stays busy and no pass runs past the end.
        const uint32_t max_it = config_.max_iterations;
        const double tolerance = static_cast<double>(config_.density_error_tolerance);
        fill_u32(stream_, tol_done_.data(), 2u, 0u);
        // non-pressure + the warm divergence pressure + the first prediction
        fa.t1 = t_div_.data();
        fa.xt1 = xt_div_.data();
        fa.t2 = t_den_.data();
        fa.final_step = 0u;
        set_slot(slot);
        launch_force<true, kDivPredict>(profiler_, stream_, blocks, fa, kDfsphForceInit);
        uint32_t executed[2] = {0u, 0u};
        for (uint32_t phase = 0; phase < 2u; ++phase) {
            uint32_t* done = tol_done_.data() + phase;
            if (phase == 0u) {
                pa.target = cache.densities.data();
                pa.relaxation = relax_div;
                pa.t = t_div_.data();
                pa.xt_out = xt_div_.data();
            } else {
                pa.target = rho0;
                pa.relaxation = relax_den;
                pa.t = t_den_.data();
                pa.xt_out = xt_den_.data();
                fa.t1 = t_den_.data();
                fa.xt1 = xt_den_.data();
                fa.t2 = nullptr;
                fa.warm_div_out = nullptr;
            }
            const uint32_t first_slot = slot;  // the prediction the phase's first projection reads
            auto launch_iteration = [&](uint32_t it) {
                set_update(phase, it, first_slot + it - 1u);
                pa.cv = Convergence{tol_partials_.data(), done, rho0};
                if (phase == 0u) {
                    launch_projection(profiler_, stream_, blocks, pa, kDfsphProjectionDivergence, kProjectDivergence);
                } else {
                    launch_projection(profiler_, stream_, blocks, pa, kDfsphProjectionDensity, kProjectDensity);
                }
                converge_kernel<<<1, kConvergeBlock, 0, stream_>>>(tol_partials_.data(), blocks * kWaves, n,
                                                                   tolerance, it, max_it, done,
                                                                   tol_host_device_ + 2u * phase + (it & 1u));
                PHX_HIP(hipGetLastError());
                PHX_HIP(hipEventRecord(tol_events_[it & 1u], stream_));
                set_slot(first_slot + it);
                fa.done = done;
                fa.iteration = it;
                fa.final_step = 0u;
                if (phase == 0u) {
                    launch_force<false, kDivAuto>(profiler_, stream_, blocks, fa, kDfsphForceDivergence);
                } else {
                    launch_force<false, kDenPredict>(profiler_, stream_, blocks, fa, kDfsphForceDensity);
                }
            };
            launch_iteration(1u);
            uint32_t it = 1u;
            for (;; ++it) {
                // the event follows the converge pass, so the force pass of
                // this iteration runs while the host reads its done word
                PHX_HIP(hipEventSynchronize(tol_events_[it & 1u]));
                // the projection of iteration max_it always ends the phase
                const volatile uint32_t* flag = tol_host_ + 2u * phase + (it & 1u);
                if (*flag != 0u || it >= max_it) break;
                launch_iteration(it + 1u);
            }
            executed[phase] = it;
            slot = first_slot + it;  // the prediction of the pass that ended the phase
        }
        pa.cv = Convergence{};
        fa.done = nullptr;
        masks_.end_step(stream_);
        if (cell_skins) stamp_base_ += predictions;
        have_moved_ = true;  // the final pass wrote every particle's displacement
        next.particles.step_index = step + 1u;
        last_iterations_[0] = executed[0];
        last_iterations_[1] = executed[1];
        return;
    }

    // divergence phase: initial pressure with the warm p_div (sph.cpp:574-578)
    fa.t1 = t_div_.data();
    fa.xt1 = xt_div_.data();
    fa.t2 = t_den_.data();
    set_slot(slot);
    if (nd == 0u) {
        fa.final_step = ne == 0u ? 1u : 0u;
        launch_force<true, kDivEnd>(profiler_, stream_, blocks, fa, kDfsphForceInit);
    } else {
        fa.final_step = 0u;
        launch_force<true, kDivPredict>(profiler_, stream_, blocks, fa, kDfsphForceInit);
        pa.target = cache.densities.data();
        pa.relaxation = relax_div;
        pa.t = t_div_.data();
        pa.xt_out = xt_div_.data();
        for (uint32_t it = 1; it <= nd; ++it) {  // sph.cpp:580-587
            set_update(0u, it, slot);
            launch_projection(profiler_, stream_, blocks, pa, kDfsphProjectionDivergence, kProjectDivergence);
            set_slot(++slot);
            if (it < nd) {
                launch_force<false, kDivPredict>(profiler_, stream_, blocks, fa, kDfsphForceDivergence);
            } else {
                fa.final_step = ne == 0u ? 1u : 0u;
                launch_force<false, kDivEnd>(profiler_, stream_, blocks, fa, kDfsphForceDivergenceEnd);
            }
        }
    }
    // density phase (sph.cpp:591-609)
    pa.target = rho0;
    pa.relaxation = relax_den;
    pa.t = t_den_.data();
    pa.xt_out = xt_den_.data();
    fa.t1 = t_den_.data();
    fa.xt1 = xt_den_.data();
    fa.t2 = nullptr;
    fa.warm_div_out = nullptr;
    for (uint32_t it = 1; it <= ne; ++it) {
        set_update(1u, it, slot);
        launch_projection(profiler_, stream_, blocks, pa, kDfsphProjectionDensity, kProjectDensity);
        set_slot(++slot);
        fa.final_step = it == ne ? 1u : 0u;
        launch_force<false, kDenPredict>(profiler_, stream_, blocks, fa, kDfsphForceDensity);
    }
    if (rec) {
        PHX_HIP(hipMemcpyAsync(cache.mask_flags.data(), masks_.flags(), sizeof(uint32_t) * masks_.flag_count(),
                               hipMemcpyDeviceToDevice, stream_));
        cache.mask_limit2 = masks_.limit2();
        cache.step_index = step;
        cache.warm_start = warm;
        cache.volume_rest = volume_rest_;
        cache.recorded = true;
    }
    // read the step's displacements back for the next step's skin
    masks_.end_step(stream_);
    if (cell_skins) stamp_base_ += predictions;
    if (!rec) have_moved_ = true;  // the final pass wrote every particle's displacement
    next.particles.step_index = step + 1u;
    last_iterations_[0] = nd;
    last_iterations_[1] = ne;
}

}  // namespace phx
