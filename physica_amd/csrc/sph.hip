// SPH<WCSPH>, SPH<PCISPH> and SPH<IISPH> forward steps for gfx950 in sorted
// (cell-major) particle order.
//
// Reference: SPH<Method>::forward_step (sph/sph.ixx:350-354), WCSPH::forward
// (sph/sph.cpp:214-226, kernels sph/wcsph.cu), PCISPH::forward
// (sph/sph.cpp:296-315, sph/pcisph.cu) and IISPH::forward (sph/sph.cpp:409-428,
// sph/iisph.cu). Each step reorders the particles by cell once and fuses every
// per-particle chain into the neighbour pass that produces its input:
//
//   WCSPH          masks + density + Tait pressure -> rho, p, p / rho^2
//                  pressure + artificial viscosity + surface + external +
//                  integrate -> next state
//   PCISPH/IISPH   masks + density; non-pressure + first predict -> x*
//                  per iteration  A: rho* at x* (skin mask) + pressure update
//                                 B: pressure acceleration + predict -> x*
//                                    (the last one writes the next state)
//
// PCISPH and IISPH differ only in the volume of the pressure update
// (V = m / rho0 and V = m / rho). Sums run in the reference order over the
// neighbour masks, so results equal the reference expression by expression
// (the Tait pressure goes through the platform's powf).
#include <cmath>

#include "sph.hpp"
#include "sph_kernels.hpp"

namespace phx {
namespace {

struct WcsphForceArgs {
    uint32_t n;
    KernelConstants kc;
    float dt;
    float diameter;
    SortedGrid grid;
    Box box;
    Float3 gravity;
    const float4* xm;   // position + mass
    const float4* vmu;  // velocity + viscosity
    const float4* aux;  // density, p / rho^2, speed of sound, surface tension
    CVec u;
    const uint16_t* exact;
    const uint32_t* flags;
    const uint32_t* perm;
    Vec next_x, next_v;
    uint8_t* collisions;  // optional (recorded step): box-collision bits of the integration
    BoundaryGrid boundary;          // offsets null: no boundary samples
    const float* rest_density;      // rho0, for the boundary terms
    const float* boundary_tension;  // WCSPH boundary_surface_tension, for the boundary cohesion
    const uint8_t* crowded;         // per particle: walks the full stencil
};

// pressure_forward_kernel (common.cu:165-194), viscosity_forward_kernel
// (wcsph.cu:136-174), surface_forward_kernel (wcsph.cu:306-339) as three
// accumulators over one neighbour walk, then the sums of WCSPH::forward
// (sph.cpp:219-222), external_forward_kernel (wcsph.cu:86-90) and
// integrate_forward_kernel (common.cu:302-320). A boundary sample b, after
// each cell's particles, adds the pressure term grad W (-rho0_i V_b p_i /
// rho_i^2), on approach the artificial viscosity with
// nu = mu_i h c_i / rho_i and mass rho0_i V_b, and within h the cohesion
// -sigma_b,i rho0_i V_b W / m_i d (wcsph.cu:160-170, 326-335).
__global__ void __launch_bounds__(kBlock) wcsph_force_kernel(WcsphForceArgs a) {
    __shared__ uint16_t smask[kStencil * kBlock];
    __shared__ int delta[kStencil];
    fill_slot_deltas(a.grid, delta);
    __syncthreads();
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.n) return;
    const float h = a.kc.h;
    const float4 self = a.xm[i];
    const Float3 p = xyz(self);
    const float4 vm = a.vmu[i];
    const Float3 vel = xyz(vm);
    const float4 own = a.aux[i];
    const float rho_i = own.x;
    const float t_i = own.y;
    const float c_i = own.z;
    const float sigma_i = own.w;
    const float softening = 0.01F * h * h;
    const float contact_weight = cubic_value(pair_distance(Float3{a.diameter, 0.0F, 0.0F}), a.kc);
    Float3 ap{0.0F, 0.0F, 0.0F};
    Float3 av{0.0F, 0.0F, 0.0F};
    Float3 as{0.0F, 0.0F, 0.0F};
    auto visit = [&](uint32_t s) {
        if (s == i) return;
        const float4 q = a.xm[s];
        const float ms = q.w;
        const Float3 d = sub(p, xyz(q));
        const float dist = pair_distance(d);
        const Float3 gradient = scale(d, cubic_gradient_factor(dist, a.kc));
        const float4 other = a.aux[s];
        ap = add(ap, scale(gradient, -ms * (t_i + other.y)));
        const float4 vq = a.vmu[s];
        const float compression = dot(sub(vel, xyz(vq)), d);
        if (!(compression >= 0.0F)) {
            const float alpha = 0.5F * (vm.w + vq.w);
            const float speed = 0.5F * (c_i + other.z);
            const float nu = 2.0F * alpha * h * speed / (rho_i + other.x);
            const float pi_value = -nu * compression / (dot(d, d) + softening);
            av = add(av, scale(gradient, -ms * pi_value));
        }
        if (!(dist >= h)) {
            const float weight = dist > a.diameter ? cubic_value(dist, a.kc) : contact_weight;
            const float factor = -sigma_i * ms * weight / self.w;
            as = add(as, scale(d, factor));
        }
    };
    int cx, cy, cz;
    particle_cell(a.grid, p, cx, cy, cz);
    if (a.boundary.offsets != nullptr) {
        const float rest = a.rest_density[i];
        const float tension = a.boundary_tension[i];
        for_each_candidate(a.grid, a.boundary, cx, cy, cz, visit, [&](uint32_t s) {
            const float4 b = a.boundary.x[s];
            const Float3 d = sub(p, xyz(b));
            const float dist = pair_distance(d);
            const Float3 gradient = scale(d, cubic_gradient_factor(dist, a.kc));
            ap = add(ap, scale(gradient, -rest * b.w * t_i));
            const float compression = dot(sub(vel, xyz(a.boundary.v[s])), d);
            if (!(compression >= 0.0F)) {
                const float nu = vm.w * h * c_i / rho_i;
                const float pi_value = -nu * compression / (dot(d, d) + softening);
                av = add(av, scale(gradient, -rest * b.w * pi_value));
            }
            if (!(dist >= h)) {
                const float weight = dist > a.diameter ? cubic_value(dist, a.kc) : contact_weight;
                const float factor = -tension * rest * b.w * weight / self.w;
                as = add(as, scale(d, factor));
            }
        });
    } else if (a.flags[kFlagOverflow] == 0u && a.crowded[i] == 0u) {
        uint16_t* column = smask + threadIdx.x;
        const uint32_t nonzero = load_lane_column(a.exact, a.n, i, column, kBlock);
        walk_masked(a.grid.offsets, cell_of(a.grid, cx, cy, cz), delta, column, kBlock, nonzero, visit);
    } else {
        for_each_candidate(a.grid, cx, cy, cz, visit);
    }
    Float3 total = add(ap, av);
    total = add(total, as);
    total = add(total, add(a.gravity, load(a.u, i)));
    Float3 nx, nv;
    predict_clamped(a.box, a.dt, p, vel, total, nx, nv);
    if (a.collisions != nullptr) a.collisions[i] = static_cast<uint8_t>(predict_collisions(a.box, a.dt, p, vel, total));
    const uint32_t j = a.perm[i];
    store(a.next_x, j, nx);
    store(a.next_v, j, nv);
}

}  // namespace

const char* sph_pass_name(int pass) {
    static const char* names[kSphPassCount] = {"grid", "mask_density", "force_init", "projection", "force",
                                               "forces"};
    return pass >= 0 && pass < kSphPassCount ? names[pass] : "unknown";
}

double sph_pass_bytes(phx_sph_method method, int pass, uint32_t particles, uint32_t cells) {
    const double offsets = particles > 0 ? 4.0 * (static_cast<double>(cells) + 1.0) / particles : 0.0;
    const bool wcsph = method == PHX_SPH_WCSPH;
    switch (pass) {
        case kSphGrid:  // x, v, u (36) + m, rho0, mu and the method's per-particle parameters, read and written
                        // in cell order, + the permutation (4)
            return 2.0 * (36.0 + (wcsph ? 24.0 : 16.0)) + 4.0 + 2.0 * offsets;
        case kSphDensity:  // x, m in; rho out (WCSPH: + rho0, c, gamma, sigma in; p, p / rho^2 out)
            return wcsph ? 32.0 + 12.0 + offsets : 16.0 + 4.0 + offsets;
        case kSphForceInit:  // x, v, u, m, mu, rho in; a_np, x* out
            return 48.0 + 24.0 + offsets;
        case kSphProjection:  // x, x*, m, rho, rho0, relaxation, p in; p, p / rho^2 out
            return 44.0 + 8.0 + offsets;
        case kSphForce:  // x, v, m, p / rho^2, a_np in; x* out
            return 44.0 + 12.0 + offsets;
        case kSphForces:  // x, v, u, m, mu, rho, p / rho^2, c, sigma in; next x, v out
            return 60.0 + 24.0 + offsets;
        default:
            return 0.0;
    }
}

void SphCache::allocate(uint32_t n, uint32_t cells) {
    neighborhood.allocate(n, cells);
    densities.resize(n);
    pressures.resize(n);
    if (method != PHX_SPH_WCSPH) non_pressure.resize(n);
}

SphSolver::SphSolver(const phx_domain_configuration* domain, const phx_sph_configuration* configuration,
                     phx_execution_mode mode, hipStream_t stream)
    : domain_(make_domain(domain)), mode_(mode), stream_(stream) {
    require(configuration != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null SPH configuration");

    config_ = *configuration;
    require(config_.method == PHX_SPH_WCSPH || config_.method == PHX_SPH_PCISPH || config_.method == PHX_SPH_IISPH,
            PHX_ERROR_INVALID_ARGUMENT, "unknown SPH method");
    require(mode == PHX_MODE_FORWARD || mode == PHX_MODE_DIFFERENTIABLE, PHX_ERROR_INVALID_ARGUMENT,
            "unknown execution mode");
    boundary_.upload(domain_.config);
    domain_.config.boundary_particles = nullptr;  // the caller's host array is not kept
    const float tolerance = config_.density_error_tolerance;
    require(tolerance >= 0.0F && std::isfinite(tolerance), PHX_ERROR_INVALID_ARGUMENT,
            "density_error_tolerance must be finite and non-negative");
    if (config_.method != PHX_SPH_WCSPH) {
        require(config_.checkpoint_interval > 0u, PHX_ERROR_INVALID_ARGUMENT, "checkpoint_interval must be positive");
        require(config_.pressure_iterations < 4096u, PHX_ERROR_INVALID_ARGUMENT, "too many pressure iterations");
        require(tolerance == 0.0F || mode == PHX_MODE_FORWARD, PHX_ERROR_UNSUPPORTED,
                "tolerance mode is supported in forward mode only");
        require(tolerance == 0.0F || (config_.max_iterations >= 1u && config_.max_iterations < 2048u),
                PHX_ERROR_INVALID_ARGUMENT, "tolerance mode needs 1 <= max_iterations < 2048");
    } else {
        require(tolerance == 0.0F, PHX_ERROR_INVALID_ARGUMENT, "tolerance mode applies to PCISPH and IISPH");
        config_.pressure_iterations = 0u;
    }
    gradient_norm_ = reference_gradient_norm(domain_.config);
    constants_ = kernel_constants(domain_.config.support_radius);
    const uint32_t n = domain_.config.particle_count;
    for (DeviceBuffer<float4>* b : {&xm_, &vmu_, &xs_, &aux_}) b->resize(n);
    u_.resize(n);
    for (DeviceBuffer<float>* b : {&rho0_, &relax_, &tait_, &t_}) b->resize(n);
    if (boundary_.count() > 0u && config_.method == PHX_SPH_WCSPH) btension_.resize(n);
    // one displacement slot per predicting pass
    const uint32_t predictions = (tolerance > 0.0F ? config_.max_iterations : config_.pressure_iterations) + 1u;
    masks_.allocate(n, predictions, domain_.config.support_radius);
    if (mode == PHX_MODE_FORWARD && config_.method != PHX_SPH_WCSPH) {
        xt_.resize(n);
        mass_state_.resize(2u);
        fill_u32(stream_, mass_state_.data(), 2u, 0u);
        for (DeviceBuffer<float>* b : {&moved_, &moved_sorted_}) b->resize(n);
        cell_bound_.resize(domain_.cells);
        taint_.resize(domain_.cells);
        fill_u32(stream_, taint_.data(), domain_.cells, 0u);
    }
    if (tolerance > 0.0F) {
        tol_partials_.resize(static_cast<size_t>(pass_blocks(n)) * kWaves);  // one per projection wave
        tol_done_.resize(1u);
        tol_host_.resize(2u * sizeof(uint32_t));
        for (hipEvent_t& e : tol_events_) PHX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    if (mode == PHX_MODE_DIFFERENTIABLE && config_.method != PHX_SPH_WCSPH) {
        // PCISPH::forward (sph.cpp:295-309) and IISPH::forward (sph.cpp:408-421)
        // are the DFSPH step with nd = 0, no warm start and their relaxation in
        // the density phase: the divergence phase then adds a zero
        // acceleration, and the pressure update's volume is m / rho0 (PCISPH)
        // or m / rho (IISPH)
        phx_dfsph_configuration dc{};
        dc.divergence_iterations = 0u;
        dc.density_iterations = config_.pressure_iterations;
        dc.checkpoint_interval = config_.checkpoint_interval;
        dc.pressure_warm_start = 0u;
        dc.exact_no_slip_adjoint = config_.exact_no_slip_adjoint;
        engine_ = std::make_unique<DfsphSolver>(domain, &dc, PHX_MODE_DIFFERENTIABLE, stream);
        engine_->set_volume_rest_density(config_.method == PHX_SPH_PCISPH);
        for (auto& b : engine_warm_) b.resize(n);
    }
}

SphSolver::~SphSolver() {
    for (hipEvent_t e : tol_events_)
        if (e != nullptr) (void)hipEventDestroy(e);
}

void SphSolver::require_derivatives(const SphCache& cache) const {
    require(mode_ == PHX_MODE_DIFFERENTIABLE, PHX_ERROR_UNSUPPORTED,
            "derivatives need a solver created in differentiable mode");
    const bool recorded = config_.method == PHX_SPH_WCSPH ? cache.wcsph != nullptr && cache.wcsph->recorded
                                                          : cache.record != nullptr && cache.record->recorded;
    require(recorded && cache.method == config_.method, PHX_ERROR_INVALID_ARGUMENT,
            "the step cache holds no step recorded by this solver");
    require(cache.grid().particles == domain_.config.particle_count, PHX_ERROR_INVALID_ARGUMENT,
            "the step cache belongs to another particle count");
}

void SphSolver::jvp_step(const SphCache& cache, const phx_sph_state_tangent& state_t,
                         const phx_control_tangent& control_t, const phx_sph_parameter_tangent& parameter_t,
                         phx_sph_state_tangent& next_t) {
    require_derivatives(cache);
    if (config_.method == PHX_SPH_WCSPH) {
        wcsph_jvp_step(cache, state_t, control_t, parameter_t, next_t);
        return;
    }
    const bool iisph = config_.method == PHX_SPH_IISPH;
    phx_sph_dfsph_state_tangent st{state_t.particles, {nullptr, nullptr}};
    phx_sph_dfsph_state_tangent nt{next_t.particles, {nullptr, nullptr}};
    phx_sph_dfsph_parameter_tangent pt{
        parameter_t.particles,
        {nullptr, iisph ? parameter_t.method.jacobi_relaxation : parameter_t.method.pressure_relaxation}};
    engine_->jvp_step(*cache.record, st, control_t, pt, nt);
}

void SphSolver::vjp_step(const SphCache& cache, const phx_sph_state_adjoint& next_a,
                         phx_sph_state_adjoint& previous_a, phx_control_adjoint& control_a,
                         phx_sph_parameter_adjoint& parameter_a) {
    require_derivatives(cache);
    if (config_.method == PHX_SPH_WCSPH) {
        wcsph_vjp_step(cache, next_a, previous_a, control_a, parameter_a);
        return;
    }
    const bool iisph = config_.method == PHX_SPH_IISPH;
    phx_sph_dfsph_state_adjoint na{next_a.particles, {nullptr, nullptr}};
    phx_sph_dfsph_state_adjoint pa{previous_a.particles, {nullptr, nullptr}};
    phx_sph_dfsph_parameter_adjoint pp{
        parameter_a.particles,
        {nullptr, iisph ? parameter_a.method.jacobi_relaxation : parameter_a.method.pressure_relaxation}};
    engine_->vjp_step(*cache.record, na, pa, control_a, pp);
}

void SphSolver::forward_step(const phx_sph_state& state, const phx_control& control,
                             const phx_sph_parameters& parameters, phx_sph_state& next, SphCache& cache) {
    const uint32_t n = domain_.config.particle_count;
    if (n == 0) {
        next.particles.step_index = state.particles.step_index + 1u;
        return;
    }
    require_vec(state.particles.positions, "state positions");
    require_vec(state.particles.velocities, "state velocities");
    require_vec(control.external_accelerations, "control");
    require_vec(next.particles.positions, "next positions");
    require_vec(next.particles.velocities, "next velocities");
    const auto& pp = parameters.particles;
    require(pp.masses && pp.rest_densities && pp.viscosities, PHX_ERROR_INVALID_ARGUMENT, "null particle parameters");
    cache.method = config_.method;
    if (engine_) {
        const float* relaxation = config_.method == PHX_SPH_IISPH ? parameters.method.jacobi_relaxation
                                                                  : parameters.method.pressure_relaxation;
        require(relaxation != nullptr, PHX_ERROR_INVALID_ARGUMENT,
                config_.method == PHX_SPH_IISPH ? "null IISPH jacobi_relaxation" : "null PCISPH pressure_relaxation");
        if (!cache.record) cache.record = std::make_unique<DfsphCache>();
        cache.stream = stream_;
        phx_sph_dfsph_state s{state.particles, {nullptr, nullptr}};
        phx_sph_dfsph_state nx{next.particles, {engine_warm_[0].data(), engine_warm_[1].data()}};
        phx_sph_dfsph_parameters p{parameters.particles, {relaxation, relaxation}};
        engine_->forward_step(s, control, p, nx, *cache.record);
        next.particles.step_index = nx.particles.step_index;
        return;
    }
    cache.record.reset();
    cache.allocate(n, domain_.cells);
    if (config_.method == PHX_SPH_WCSPH) {
        forward_wcsph(state, control, parameters, next, cache);
    } else {
        forward_pressure(state, control, parameters, next, cache);
    }
    next.particles.step_index = state.particles.step_index + 1u;
}

void SphSolver::forward_wcsph(const phx_sph_state& state, const phx_control& control,
                              const phx_sph_parameters& parameters, phx_sph_state& next, SphCache& cache) {
    const auto& c = domain_.config;
    const uint32_t n = c.particle_count;
    const uint64_t step = state.particles.step_index;
    const auto& pp = parameters.particles;
    const auto& mp = parameters.method;
    require(pp.surface_tensions && mp.speed_of_sound && mp.tait_exponent && mp.boundary_surface_tension,
            PHX_ERROR_INVALID_ARGUMENT, "null WCSPH parameters");
    // a recorded step keeps its sorted inputs, masks and collision bits in the cache
    const bool rec = mode_ == PHX_MODE_DIFFERENTIABLE;
    WcsphRecord* r = nullptr;
    if (rec) {
        if (!cache.wcsph) cache.wcsph = std::make_unique<WcsphRecord>();
        r = cache.wcsph.get();
        r->recorded = false;
        for (DeviceBuffer<float4>* b : {&r->xm, &r->vmu, &r->aux}) b->resize(n);
        for (DeviceBuffer<float>* b : {&r->rho0, &r->tait}) b->resize(n);
        r->collisions.resize(n);
        r->exact.resize(static_cast<size_t>(kStencil) * n);
        r->crowded.resize(n);
        r->flags.resize(masks_.flag_count());
        if (boundary_.count() > 0u) r->btension.resize(n);
    }
    float4* xm = rec ? r->xm.data() : xm_.data();
    float4* vmu = rec ? r->vmu.data() : vmu_.data();
    float4* aux = rec ? r->aux.data() : aux_.data();
    float* rho0 = rec ? r->rho0.data() : rho0_.data();
    float* tait = rec ? r->tait.data() : tait_.data();
    uint16_t* exact = rec ? r->exact.data() : masks_.exact();
    uint8_t* crowded = rec ? r->crowded.data() : masks_.crowded();
    ChannelList gather;
    gather.add(cvec(state.particles.positions), xm);
    gather.add(pp.masses, xm, 3);
    gather.add(cvec(state.particles.velocities), vmu);
    gather.add(pp.viscosities, vmu, 3);
    gather.add(cvec(control.external_accelerations), u_.view());
    gather.add(pp.rest_densities, rho0);
    gather.add(mp.speed_of_sound, aux, 2);
    gather.add(pp.surface_tensions, aux, 3);
    gather.add(mp.tait_exponent, tait);
    const bool with_boundary = boundary_.count() > 0u;
    float* btension = rec && with_boundary ? r->btension.data() : btension_.data();
    if (with_boundary) gather.add(mp.boundary_surface_tension, btension);
    {
        PassScope scope(profiler_, stream_, kSphGrid);
        cache.neighborhood.build(stream_, domain_, step, cvec(state.particles.positions), nullptr, gather.get(),
                                 &boundary_);
        masks_.begin_step(stream_, with_boundary);
    }
    const SortedGrid grid = cache.neighborhood.grid();
    const BoundaryGrid boundary = cache.neighborhood.boundary();
    const uint32_t blocks = pass_blocks(n);

    MaskDensityArgs ma{};
    ma.n = n;
    ma.kc = constants_;
    ma.t_skin = masks_.t_skin();
    ma.grid = grid;
    ma.boundary = boundary;
    ma.xm = xm;
    ma.exact = exact;
    ma.skin = masks_.skin_mask();
    ma.rho = cache.densities.data();
    ma.flags = masks_.flags();
    ma.rest_density = rho0;
    ma.tait_exponent = tait;
    ma.pressures = cache.pressures.data();
    ma.aux = aux;
    ma.crowded = crowded;
    launch_mask_density<true>(profiler_, stream_, blocks, ma, kSphDensity);

    WcsphForceArgs wa{};
    wa.n = n;
    wa.kc = constants_;
    wa.dt = c.time_step;
    wa.diameter = 2.0F * c.particle_radius;
    wa.grid = grid;
    wa.box = collision_box(c, step);
    wa.gravity = Float3{c.gravity.x, c.gravity.y, c.gravity.z};
    wa.xm = xm;
    wa.vmu = vmu;
    wa.aux = aux;
    wa.u = u_.cview();
    wa.exact = exact;
    wa.flags = masks_.flags();
    wa.perm = cache.neighborhood.perm.data();
    wa.next_x = vec(next.particles.positions);
    wa.next_v = vec(next.particles.velocities);
    wa.collisions = rec ? r->collisions.data() : nullptr;
    wa.boundary = boundary;
    wa.rest_density = rho0;
    wa.boundary_tension = btension;
    wa.crowded = crowded;
    {
        PassScope scope(profiler_, stream_, kSphForces);
        wcsph_force_kernel<<<blocks, kBlock, 0, stream_>>>(wa);
        PHX_LAUNCH_CHECK();
    }
    if (rec) {
        PHX_HIP(hipMemcpyAsync(r->flags.data(), masks_.flags(), sizeof(uint32_t) * masks_.flag_count(),
                               hipMemcpyDeviceToDevice, stream_));
        r->step_index = step;
        r->recorded = true;
    }
    masks_.end_step(stream_);
}

void SphSolver::forward_pressure(const phx_sph_state& state, const phx_control& control,
                                 const phx_sph_parameters& parameters, phx_sph_state& next, SphCache& cache) {
    const auto& c = domain_.config;
    const uint32_t n = c.particle_count;
    const uint64_t step = state.particles.step_index;
    const auto& pp = parameters.particles;
    const bool iisph = config_.method == PHX_SPH_IISPH;
    const float* relaxation = iisph ? parameters.method.jacobi_relaxation : parameters.method.pressure_relaxation;
    require(relaxation != nullptr, PHX_ERROR_INVALID_ARGUMENT,
            iisph ? "null IISPH jacobi_relaxation" : "null PCISPH pressure_relaxation");
    ChannelList gather;
    gather.add(cvec(state.particles.positions), xm_.data());
    gather.add(pp.masses, xm_.data(), 3);
    gather.add(cvec(state.particles.velocities), vmu_.data());
    gather.add(pp.viscosities, vmu_.data(), 3);
    gather.add(cvec(control.external_accelerations), u_.view());
    gather.add(pp.rest_densities, rho0_.data());
    gather.add(relaxation, relax_.data());
    const bool with_boundary = boundary_.count() > 0u;
    // per-cell skins (DfsphSolver::forward_step) from the previous step's displacements
    const bool cell_skins = have_moved_ && !with_boundary;
    if (cell_skins) gather.add(moved_.data(), moved_sorted_.data());
    const uint32_t predictions =
        (config_.density_error_tolerance > 0.0F ? config_.max_iterations : config_.pressure_iterations) + 1u;
    if (cell_skins && stamp_base_ > 0xFFFFFFFFu - predictions) {
        fill_u32(stream_, taint_.data(), domain_.cells, 0u);
        stamp_base_ = 1u;
    }
    {
        PassScope scope(profiler_, stream_, kSphGrid);
        cache.neighborhood.build(stream_, domain_, step, cvec(state.particles.positions), nullptr, gather.get(),
                                 &boundary_);
        masks_.begin_step(stream_, with_boundary);
        fill_f32(stream_, cache.pressures.data(), n, 0.0F);  // clear_iteration (sph.cpp:298)
        // the mask pass marks unequal masses with this step's stamp
        if (++mass_stamp_ == 0u) {
            fill_u32(stream_, mass_state_.data(), 1u, 0u);
            mass_stamp_ = 1u;
        }
        if (cell_skins) {
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

    MaskDensityArgs ma{};
    ma.n = n;
    ma.kc = constants_;
    ma.t_skin = masks_.t_skin();
    ma.grid = grid;
    ma.boundary = boundary;
    ma.rest_density = rho0_.data();
    ma.xm = xm_.data();
    ma.exact = masks_.exact();
    ma.skin = masks_.skin_mask();
    ma.rho = cache.densities.data();
    ma.flags = masks_.flags();
    ma.crowded = masks_.crowded();
    if (cell_skins) ma.cell_bound = cell_bound_.data();
    ma.mass_state = mass_state_.data();
    ma.mass_stamp = mass_stamp_;
    launch_mask_density<false>(profiler_, stream_, blocks, ma, kSphDensity);

    ForceArgs fa{};
    fa.n = n;
    fa.kc = constants_;
    fa.dt = c.time_step;
    fa.grid = grid;
    fa.box = collision_box(c, step);
    fa.gravity = Float3{c.gravity.x, c.gravity.y, c.gravity.z};
    fa.xm = xm_.data();
    fa.vmu = vmu_.data();
    fa.u = u_.cview();
    fa.rho = cache.densities.data();
    fa.t1 = t_.data();
    fa.exact = masks_.exact();
    fa.flags = masks_.flags();
    fa.anp = cache.non_pressure.data();
    fa.xs = xs_.data();
    fa.perm = cache.neighborhood.perm.data();
    fa.next_x = vec(next.particles.positions);
    fa.next_v = vec(next.particles.velocities);
    fa.boundary = boundary;
    fa.rest_density = rho0_.data();
    fa.crowded = masks_.crowded();
    fa.moved_out = moved_.data();
    fa.xt1 = xt_.data();
    fa.mass_state = mass_state_.data();
    fa.mass_stamp = mass_stamp_;
    if (cell_skins) {
        fa.cell_bound = cell_bound_.data();
        fa.taint = taint_.data();
    }

    ProjectionArgs pa{};
    pa.n = n;
    pa.kc = constants_;
    pa.dt = c.time_step;
    pa.gradient_norm = gradient_norm_;
    pa.limit2 = masks_.limit2();
    pa.grid = grid;
    pa.xm = xm_.data();
    pa.xs = xs_.data();
    pa.skin = masks_.skin_mask();
    pa.flags = masks_.flags();
    pa.rho = cache.densities.data();
    pa.volume_density = iisph ? cache.densities.data() : rho0_.data();
    pa.target = rho0_.data();
    pa.relaxation = relax_.data();
    pa.p_prev = cache.pressures.data();
    pa.p = cache.pressures.data();
    pa.t = t_.data();
    pa.boundary = boundary;
    pa.rest_density = rho0_.data();
    pa.crowded = masks_.crowded();
    pa.xt_out = xt_.data();
    if (cell_skins) pa.taint = taint_.data();
    // the stamps of this step's predictions: stamp_base_ + slot
    auto set_force_slot = [&](uint32_t slot) {
        fa.slot = slot;
        fa.stamp = stamp_base_ + slot;
    };
    auto set_projection_slot = [&](uint32_t slot) {
        pa.slot = slot;
        pa.stamp = stamp_base_ + slot;
    };
    auto end_step = [&] {
        masks_.end_step(stream_);
        if (cell_skins) stamp_base_ += predictions;
        have_moved_ = true;  // the final pass wrote every particle's displacement
    };

    if (config_.density_error_tolerance > 0.0F) {
        // Tolerance mode: iterate until the projection's mean compression
        // error is within the tolerance (or max_iterations); the force pass
        // of that iteration writes the state. As DfsphSolver's tolerance
        // mode, the host reads an iteration's done word while its force
        // pass runs, then launches the next iteration.
        const uint32_t max_it = config_.max_iterations;
        const double tolerance = static_cast<double>(config_.density_error_tolerance);
        uint32_t* done = tol_done_.data();
        const volatile uint32_t* host = static_cast<uint32_t*>(tol_host_.data());
        auto* host_device = static_cast<uint32_t*>(tol_host_.device());
        fill_u32(stream_, done, 1u, 0u);
        set_force_slot(0u);
        fa.final_step = 0u;
        launch_force<true, kInitPredict>(profiler_, stream_, blocks, fa, kSphForceInit);
        auto launch_iteration = [&](uint32_t it) {
            set_projection_slot(it - 1u);  // the prediction of the previous pass
            pa.cv = Convergence{tol_partials_.data(), done, rho0_.data()};
            launch_projection(profiler_, stream_, blocks, pa, kSphProjection, kProjectDensity);
            converge_kernel<<<1, kConvergeBlock, 0, stream_>>>(tol_partials_.data(), blocks * kWaves, n, tolerance,
                                                               it, max_it, done, host_device + (it & 1u));
            PHX_HIP(hipGetLastError());
            PHX_HIP(hipEventRecord(tol_events_[it & 1u], stream_));
            set_force_slot(it);
            fa.done = done;
            fa.iteration = it;
            fa.final_step = 0u;
            launch_force<false, kDivPredict>(profiler_, stream_, blocks, fa, kSphForce);
        };
        launch_iteration(1u);
        uint32_t it = 1u;
        for (;; ++it) {
            // the force pass of this iteration runs while the host reads its done word
            PHX_HIP(hipEventSynchronize(tol_events_[it & 1u]));
            // the projection of iteration max_it always ends the step
            if (host[it & 1u] != 0u || it >= max_it) break;
            launch_iteration(it + 1u);
        }
        last_iterations_ = it;
        pa.cv = Convergence{};
        fa.done = nullptr;
        end_step();
        return;
    }
    const uint32_t iterations = config_.pressure_iterations;
    last_iterations_ = iterations;
    uint32_t slot = 0;  // displacement slot written by the latest predicting pass
    set_force_slot(slot);
    fa.final_step = iterations == 0u ? 1u : 0u;
    launch_force<true, kInitPredict>(profiler_, stream_, blocks, fa, kSphForceInit);
    for (uint32_t it = 1; it <= iterations; ++it) {  // sph.cpp:300-307 / 413-420
        set_projection_slot(slot);
        launch_projection(profiler_, stream_, blocks, pa, kSphProjection, kProjectDensity);
        set_force_slot(++slot);
        fa.final_step = it == iterations ? 1u : 0u;
        launch_force<false, kDivPredict>(profiler_, stream_, blocks, fa, kSphForce);
    }
    end_step();
}

}  // namespace phx
