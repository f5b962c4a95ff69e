// SPH<DFSPH> forward step for gfx950 in sorted (cell-major) particle order.
//
// Reference: SPH<DFSPH>::forward_step (sph/sph.ixx:350-354) and
// DFSPH::forward (sph/sph.cpp:572-610) with kernels from sph/common.cu,
// sph/dfsph.cu and density/density.cu. The reference runs one kernel per
// operation in original particle order; here particles are reordered by cell
// once per step and each per-particle chain is fused into the neighbour pass
// that produces its input:
//
//   density        rho = sum m W(x_i - x_j);  t = p / rho^2 for both warm p
//   force (init)   non-pressure + pressure(t_div) + predict -> x*
//   per iteration  A: rho* at x*, projection update -> p, t
//                  B: pressure(t) + predict -> x*  (or the next state)
//
// The end of the divergence phase and the density phase's initial pressure
// are one pass with two accumulators. Every per-particle sum is evaluated in
// the reference order (27 cells, offset_x outer, ascending original index), so
// results equal the reference expression by expression.
#include "dfsph.hpp"

namespace phx {
namespace {

constexpr uint32_t kBlock = 256;

struct DensityArgs {
    uint32_t n;
    float h;
    SortedGrid grid;
    CVec x;
    const float* m;
    const float* p_div;
    const float* p_den;
    float* rho;
    float* t_div;
    float* t_den;
};

// density_forward_kernel (density.cu:18-40, cubic kernel, topology = positions)
// followed by the first_term p/rho^2 of pressure_forward_kernel (common.cu:171)
// for the warm-started pressures of both phases.
__global__ void __launch_bounds__(kBlock) dfsph_density_kernel(DensityArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const Float3 p = load(a.x, i);
    int cx, cy, cz;
    particle_cell(a.grid, p, cx, cy, cz);
    float result = 0.0F;
    for_each_candidate(a.grid, cx, cy, cz, [&](uint32_t s) { result += a.m[s] * cubic(sub(p, load(a.x, s)), a.h); });
    a.rho[i] = result;
    a.t_div[i] = a.p_div[i] / (result * result);
    a.t_den[i] = a.p_den[i] / (result * result);
}

struct ProjectionArgs {
    uint32_t n;
    float h;
    float dt;
    float gradient_norm;
    SortedGrid grid;
    CVec x;   // topology: state positions
    CVec xs;  // predicted positions
    const float* m;
    const float* rho;
    const float* target;
    const float* relaxation;
    float* p;
    float* t;
    float* rho_star;
};

// density at predicted positions (density.cu:18-40 with topology = x,
// positions = x*) fused with projection_update_forward_kernel
// (dfsph.cu:104-110) and the next pressure pass's first_term.
__global__ void __launch_bounds__(kBlock) dfsph_projection_kernel(ProjectionArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const Float3 ps = load(a.xs, i);
    int cx, cy, cz;
    particle_cell(a.grid, load(a.x, i), cx, cy, cz);
    float rs = 0.0F;
    for_each_candidate(a.grid, cx, cy, cz, [&](uint32_t s) { rs += a.m[s] * cubic(sub(ps, load(a.xs, s)), a.h); });
    if (a.rho_star != nullptr) a.rho_star[i] = rs;
    const float rho = a.rho[i];
    const float volume = a.m[i] / rho;
    const float delta = 1.0F / (2.0F * a.dt * a.dt * volume * volume * a.gradient_norm);
    const float pressure = fmaxf(0.0F, a.p[i] + a.relaxation[i] * delta * (rs - a.target[i]));
    a.p[i] = pressure;
    a.t[i] = pressure / (rho * rho);
}

enum Epilogue : int {
    kDivPredict = 0,  // total = a_p(t1)
    kDivEnd = 1,      // a_div = a_p(t1); total = a_div + a_p(t2)
    kDenPredict = 2,  // total = a_div + a_p(t1)
};

struct ForceArgs {
    uint32_t n;
    float h;
    float dt;
    SortedGrid grid;
    Box box;
    Float3 gravity;
    CVec x, v, u;
    const float* m;
    const float* mu;
    const float* rho;
    const float* t1;
    const float* t2;
    Vec anp;
    Vec adiv;
    Vec xs;
    uint32_t final_step;
    const uint32_t* perm;
    Vec next_x, next_v;
    const float* p_div;
    float* warm_div_out;
    const float* p_den;
    float* warm_den_out;
};

// Pressure acceleration (pressure_forward_kernel, common.cu:165-194),
// optionally fused with the non-pressure acceleration
// (non_pressure_forward_kernel, common.cu:14-47), then the predict of the
// next iteration (dfsph.cu:26-44) or, on the last pass, the next state.
template <bool kNonPressure, int kEpilogue>
__global__ void __launch_bounds__(kBlock) dfsph_force_kernel(ForceArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const Float3 p = load(a.x, i);
    const Float3 vel = load(a.v, i);
    Float3 acc{0.0F, 0.0F, 0.0F};
    float mu_i = 0.0F;
    if constexpr (kNonPressure) {
        acc = add(a.gravity, load(a.u, i));
        mu_i = a.mu[i];
    }
    const float t1_i = a.t1[i];
    float t2_i = 0.0F;
    if constexpr (kEpilogue == kDivEnd) t2_i = a.t2[i];
    Float3 ap1{0.0F, 0.0F, 0.0F};
    Float3 ap2{0.0F, 0.0F, 0.0F};
    int cx, cy, cz;
    particle_cell(a.grid, p, cx, cy, cz);
    for_each_candidate(a.grid, cx, cy, cz, [&](uint32_t s) {
        if (s == i) return;
        const Float3 d = sub(p, load(a.x, s));
        const float ms = a.m[s];
        if constexpr (kNonPressure) {
            const float viscosity = 0.5F * (mu_i + a.mu[s]);
            const float weight = viscosity * ms * viscosity_laplacian(d, a.h) / a.rho[s];
            acc = add(acc, scale(sub(load(a.v, s), vel), weight));
        }
        const Float3 gradient = cubic_gradient(d, a.h);
        ap1 = add(ap1, scale(gradient, -ms * (t1_i + a.t1[s])));
        if constexpr (kEpilogue == kDivEnd) ap2 = add(ap2, scale(gradient, -ms * (t2_i + a.t2[s])));
    });
    if constexpr (kNonPressure) {
        store(a.anp, i, acc);
    } else {
        acc = load(a.anp, i);
    }
    Float3 total;
    if constexpr (kEpilogue == kDivPredict) {
        total = ap1;
    } else if constexpr (kEpilogue == kDivEnd) {
        store(a.adiv, i, ap1);
        total = add(ap1, ap2);
    } else {
        total = add(load(a.adiv, i), ap1);
    }
    Float3 nx, nv;
    predict_clamped(a.box, a.dt, p, vel, add(acc, total), nx, nv);
    if constexpr (kEpilogue == kDivEnd) {
        if (a.warm_div_out != nullptr) a.warm_div_out[a.perm[i]] = a.p_div[i];
    }
    if (a.final_step != 0u) {
        const uint32_t j = a.perm[i];
        store(a.next_x, j, nx);
        store(a.next_v, j, nv);
        if (a.warm_den_out != nullptr) a.warm_den_out[j] = a.p_den[i];
    } else {
        store(a.xs, i, nx);
    }
}

template <bool kNonPressure, int kEpilogue>
void launch_force(PassProfiler& profiler, hipStream_t stream, const ForceArgs& args) {
    const int pass = kNonPressure ? kDfsphForceInit
                     : kEpilogue == kDivPredict ? kDfsphForceDivergence
                     : kEpilogue == kDivEnd     ? kDfsphForceDivergenceEnd
                                                : kDfsphForceDensity;
    PassScope scope(profiler, stream, pass);
    dfsph_force_kernel<kNonPressure, kEpilogue><<<blocks_for(args.n, kBlock), kBlock, 0, stream>>>(args);
    PHX_LAUNCH_CHECK();
}

void launch_projection(PassProfiler& profiler, hipStream_t stream, const ProjectionArgs& args, int pass) {
    PassScope scope(profiler, stream, pass);
    dfsph_projection_kernel<<<blocks_for(args.n, kBlock), kBlock, 0, stream>>>(args);
    PHX_LAUNCH_CHECK();
}

}  // namespace

const char* dfsph_pass_name(int pass) {
    static const char* names[kDfsphPassCount] = {"grid",       "density",          "force_init",
                                                 "projection_divergence", "force_divergence",
                                                 "force_divergence_end",  "projection_density",
                                                 "force_density"};
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
        default:
            return 0.0;
    }
}

void DfsphCache::allocate(uint32_t n, uint32_t cells) {
    neighborhood.allocate(n, cells);
    densities.resize(n);
    non_pressure.resize(n);
}

DfsphSolver::DfsphSolver(const phx_domain_configuration* domain, const phx_dfsph_configuration* configuration,
                         phx_execution_mode mode, hipStream_t stream)
    : domain_(make_domain(domain)), mode_(mode), stream_(stream) {
    require(configuration != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null DFSPH configuration");
    config_ = *configuration;
    require(config_.checkpoint_interval > 0u, PHX_ERROR_INVALID_ARGUMENT, "checkpoint_interval must be positive");
    require(mode == PHX_MODE_FORWARD, PHX_ERROR_UNSUPPORTED, "differentiable mode is not supported yet");
    require(config_.density_error_tolerance == 0.0F, PHX_ERROR_UNSUPPORTED, "tolerance mode is not supported yet");
    gradient_norm_ = reference_gradient_norm(domain_.config);
    const uint32_t n = domain_.config.particle_count;
    x_.resize(n);
    v_.resize(n);
    u_.resize(n);
    xs_.resize(n);
    adiv_.resize(n);
    m_.resize(n);
    rho0_.resize(n);
    mu_.resize(n);
    relax_div_.resize(n);
    relax_den_.resize(n);
    p_div_.resize(n);
    p_den_.resize(n);
    t_div_.resize(n);
    t_den_.resize(n);
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
    cache.allocate(n, domain_.cells);

    // neighborhood.build(step, x) with every input gathered into cell order
    ChannelList gather;
    gather.add(cvec(state.particles.positions), x_.view());
    gather.add(cvec(state.particles.velocities), v_.view());
    gather.add(cvec(control.external_accelerations), u_.view());
    gather.add(pp.masses, m_.data());
    gather.add(pp.rest_densities, rho0_.data());
    gather.add(pp.viscosities, mu_.data());
    gather.add(parameters.method.divergence_relaxation, relax_div_.data());
    gather.add(parameters.method.density_relaxation, relax_den_.data());
    if (warm) {
        gather.add(state.method.warm_divergence_pressure, p_div_.data());
        gather.add(state.method.warm_density_pressure, p_den_.data());
    }
    {
        PassScope scope(profiler_, stream_, kDfsphGrid);
        cache.neighborhood.build(stream_, domain_, step, cvec(state.particles.positions), nullptr, gather.get());
        if (!warm) {
            fill_f32(stream_, p_div_.data(), n, 0.0F);
            fill_f32(stream_, p_den_.data(), n, 0.0F);
        }
    }
    const SortedGrid grid = cache.neighborhood.grid();
    const uint32_t blocks = blocks_for(n, kBlock);
    const float h = c.support_radius;

    DensityArgs da{n, h, grid, x_.cview(), m_.data(), p_div_.data(), p_den_.data(), cache.densities.data(),
                   t_div_.data(), t_den_.data()};
    {
        PassScope scope(profiler_, stream_, kDfsphDensity);
        dfsph_density_kernel<<<blocks, kBlock, 0, stream_>>>(da);
        PHX_LAUNCH_CHECK();
    }

    ForceArgs fa{};
    fa.n = n;
    fa.h = h;
    fa.dt = c.time_step;
    fa.grid = grid;
    fa.box = collision_box(c, step);
    fa.gravity = Float3{c.gravity.x, c.gravity.y, c.gravity.z};
    fa.x = x_.cview();
    fa.v = v_.cview();
    fa.u = u_.cview();
    fa.m = m_.data();
    fa.mu = mu_.data();
    fa.rho = cache.densities.data();
    fa.anp = cache.non_pressure.view();
    fa.adiv = adiv_.view();
    fa.xs = xs_.view();
    fa.perm = cache.neighborhood.perm.data();
    fa.next_x = vec(next.particles.positions);
    fa.next_v = vec(next.particles.velocities);
    fa.p_div = p_div_.data();
    fa.warm_div_out = next.method.warm_divergence_pressure;
    fa.p_den = p_den_.data();
    fa.warm_den_out = next.method.warm_density_pressure;

    ProjectionArgs pa{};
    pa.n = n;
    pa.h = h;
    pa.dt = c.time_step;
    pa.gradient_norm = gradient_norm_;
    pa.grid = grid;
    pa.x = x_.cview();
    pa.xs = xs_.cview();
    pa.m = m_.data();
    pa.rho = cache.densities.data();
    pa.rho_star = nullptr;

    const uint32_t nd = config_.divergence_iterations;
    const uint32_t ne = config_.density_iterations;

    // divergence phase: initial pressure with the warm p_div (sph.cpp:574-578)
    fa.t1 = t_div_.data();
    fa.t2 = t_den_.data();
    if (nd == 0u) {
        fa.final_step = ne == 0u ? 1u : 0u;
        launch_force<true, kDivEnd>(profiler_, stream_, fa);
    } else {
        fa.final_step = 0u;
        launch_force<true, kDivPredict>(profiler_, stream_, fa);
        pa.target = cache.densities.data();
        pa.relaxation = relax_div_.data();
        pa.p = p_div_.data();
        pa.t = t_div_.data();
        for (uint32_t it = 1; it <= nd; ++it) {  // sph.cpp:580-587
            launch_projection(profiler_, stream_, pa, kDfsphProjectionDivergence);
            if (it < nd) {
                launch_force<false, kDivPredict>(profiler_, stream_, fa);
            } else {
                fa.final_step = ne == 0u ? 1u : 0u;
                launch_force<false, kDivEnd>(profiler_, stream_, fa);
            }
        }
    }
    // density phase (sph.cpp:591-609)
    pa.target = rho0_.data();
    pa.relaxation = relax_den_.data();
    pa.p = p_den_.data();
    pa.t = t_den_.data();
    fa.t1 = t_den_.data();
    fa.t2 = nullptr;
    fa.warm_div_out = nullptr;
    for (uint32_t it = 1; it <= ne; ++it) {
        launch_projection(profiler_, stream_, pa, kDfsphProjectionDensity);
        fa.final_step = it == ne ? 1u : 0u;
        launch_force<false, kDenPredict>(profiler_, stream_, fa);
    }
    next.particles.step_index = step + 1u;
    last_iterations_[0] = nd;
    last_iterations_[1] = ne;
}

}  // namespace phx
