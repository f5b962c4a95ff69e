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
// from the preceding B pass); otherwise it scans the full stencil.
#include <cmath>
#include <cstring>

#include "dfsph.hpp"
#include "neighbor_mask.hpp"

namespace phx {
namespace {

__device__ __forceinline__ Float3 xyz(float4 q) { return Float3{q.x, q.y, q.z}; }

constexpr uint32_t kBlock = 256;

__device__ __forceinline__ uint32_t cell_of(const SortedGrid& g, int x, int y, int z) {
    return (static_cast<uint32_t>(z) * g.cy + static_cast<uint32_t>(y)) * g.cx + static_cast<uint32_t>(x);
}

struct MaskDensityArgs {
    uint32_t n;
    KernelConstants kc;
    float t_skin;
    SortedGrid grid;
    const float4* xm;
    uint16_t* exact;
    uint16_t* skin;
    const float* p_div;
    const float* p_den;
    float* rho;
    float* t_div;
    float* t_den;
    uint32_t* flags;
};

// Mask build, then density_forward_kernel (density.cu:18-40, cubic kernel,
// topology = positions) over the exact mask and the first_term p/rho^2 of
// pressure_forward_kernel (common.cu:171) for both warm-started pressures.
__global__ void __launch_bounds__(kBlock) dfsph_mask_density_kernel(MaskDensityArgs a) {
    __shared__ uint16_t smask[kStencil * kBlock];
    __shared__ int delta[kStencil];
    fill_slot_deltas(a.grid, delta);
    __syncthreads();
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.n) return;
    uint16_t* column = smask + threadIdx.x;
    const Float3 p = xyz(a.xm[i]);
    int cx, cy, cz;
    particle_cell(a.grid, p, cx, cy, cz);
    uint32_t nonzero = 0u;
    const bool complete = build_masks(a.grid, a.xm, a.n, i, p, cx, cy, cz, a.kc.sqrt_support, a.t_skin, a.exact,
                                      a.skin, column, kBlock, &nonzero);
    float result = 0.0F;
    if (complete) {
        walk_masked(a.grid.offsets, cell_of(a.grid, cx, cy, cz), delta, column, kBlock, nonzero, [&](uint32_t s) {
            const float4 q = a.xm[s];
            const float dist = length(sub(p, xyz(q)));
            result += q.w * cubic_value(dist, a.kc);
        });
    } else {
        atomicOr(&a.flags[kFlagOverflow], 1u);
        for_each_candidate(a.grid, cx, cy, cz, [&](uint32_t s) {
            const float4 q = a.xm[s];
            result += q.w * cubic_value(length(sub(p, xyz(q))), a.kc);
        });
    }
    a.rho[i] = result;
    a.t_div[i] = a.p_div[i] / (result * result);
    a.t_den[i] = a.p_den[i] / (result * result);
}

struct ProjectionArgs {
    uint32_t n;
    KernelConstants kc;
    float dt;
    float gradient_norm;
    float limit2;  // largest squared displacement for which the skin mask is exact
    uint32_t slot;
    SortedGrid grid;
    const float4* xm;  // topology: state positions (+ masses)
    const float4* xs;  // predicted positions (+ masses)
    const uint16_t* skin;
    const uint32_t* flags;
    const float* rho;
    const float* target;
    const float* relaxation;
    float* p;
    float* t;
};

// density at predicted positions (density.cu:18-40 with topology = x,
// positions = x*) fused with projection_update_forward_kernel
// (dfsph.cu:104-110) and the next pressure pass's first_term.
__global__ void __launch_bounds__(kBlock) dfsph_projection_kernel(ProjectionArgs a) {
    __shared__ uint16_t smask[kStencil * kBlock];
    __shared__ int delta[kStencil];
    fill_slot_deltas(a.grid, delta);
    __syncthreads();
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.n) return;
    const bool masked = a.flags[kFlagOverflow] == 0u &&
                        __uint_as_float(a.flags[kFlagDisplacement + a.slot]) <= a.limit2;
    const float4 self = a.xs[i];
    const Float3 ps = xyz(self);
    int cx, cy, cz;
    particle_cell(a.grid, xyz(a.xm[i]), cx, cy, cz);
    float rs = 0.0F;
    auto visit = [&](uint32_t s) {
        const float4 q = a.xs[s];
        rs += q.w * cubic_value(length(sub(ps, xyz(q))), a.kc);
    };
    if (masked) {
        uint16_t* column = smask + threadIdx.x;
        const uint32_t nonzero = load_lane_column(a.skin, a.n, i, column, kBlock);
        walk_masked(a.grid.offsets, cell_of(a.grid, cx, cy, cz), delta, column, kBlock, nonzero, visit);
    } else {
        for_each_candidate(a.grid, cx, cy, cz, visit);
    }
    const float rho = a.rho[i];
    const float volume = self.w / rho;
    const float delta_p = 1.0F / (2.0F * a.dt * a.dt * volume * volume * a.gradient_norm);
    const float pressure = fmaxf(0.0F, a.p[i] + a.relaxation[i] * delta_p * (rs - a.target[i]));
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
    KernelConstants kc;
    float dt;
    SortedGrid grid;
    Box box;
    Float3 gravity;
    const float4* xm;
    const float4* vmu;  // velocity (+ viscosity)
    CVec u;
    const float* rho;
    const float* t1;
    const float* t2;
    const uint16_t* exact;
    uint32_t* flags;
    uint32_t slot;
    float4* anp;
    float4* adiv;
    float4* xs;
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
// cubic_gradient(d, h) is evaluated as f * d with f its radial factor, the
// same operations as the reference's expression.
template <bool kNonPressure, int kEpilogue>
__global__ void __launch_bounds__(kBlock) dfsph_force_kernel(ForceArgs a) {
    constexpr bool kTwo = kEpilogue == kDivEnd;
    __shared__ uint16_t smask[kStencil * kBlock];
    __shared__ int delta[kStencil];
    fill_slot_deltas(a.grid, delta);
    __syncthreads();
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    const bool active = i < a.n;
    float moved2 = 0.0F;
    if (active) {
        const bool masked = a.flags[kFlagOverflow] == 0u;
        const float4 self = a.xm[i];
        const Float3 p = xyz(self);
        const float4 vm = a.vmu[i];
        const Float3 vel = xyz(vm);
        Float3 acc{0.0F, 0.0F, 0.0F};
        if constexpr (kNonPressure) acc = add(a.gravity, load(a.u, i));
        const float t1_i = a.t1[i];
        float t2_i = 0.0F;
        if constexpr (kTwo) t2_i = a.t2[i];
        Float3 ap1{0.0F, 0.0F, 0.0F};
        Float3 ap2{0.0F, 0.0F, 0.0F};
        int cx, cy, cz;
        particle_cell(a.grid, p, cx, cy, cz);
        auto pair = [&](uint32_t s, float4 q, Float3 d, float f) {
            const float ms = q.w;
            if constexpr (kNonPressure) {
                const float4 vq = a.vmu[s];
                const float viscosity = 0.5F * (vm.w + vq.w);
                const float weight = viscosity * ms * viscosity_laplacian_c(d, a.kc) / a.rho[s];
                acc = add(acc, scale(sub(xyz(vq), vel), weight));
            }
            const Float3 gradient = scale(d, f);
            ap1 = add(ap1, scale(gradient, -ms * (t1_i + a.t1[s])));
            if constexpr (kTwo) ap2 = add(ap2, scale(gradient, -ms * (t2_i + a.t2[s])));
        };
        if (masked) {
            uint16_t* column = smask + threadIdx.x;
            const uint32_t nonzero = load_lane_column(a.exact, a.n, i, column, kBlock);
            walk_masked(a.grid.offsets, cell_of(a.grid, cx, cy, cz), delta, column, kBlock, nonzero,
                        [&](uint32_t s) {
                            if (s == i) return;
                            const float4 q = a.xm[s];
                            const Float3 d = sub(p, xyz(q));
                            pair(s, q, d, cubic_gradient_factor(length(d), a.kc));
                        });
        } else {
            for_each_candidate(a.grid, cx, cy, cz, [&](uint32_t s) {
                if (s == i) return;
                const float4 q = a.xm[s];
                const Float3 d = sub(p, xyz(q));
                pair(s, q, d, cubic_gradient_factor(length(d), a.kc));
            });
        }
        if constexpr (kNonPressure) {
            a.anp[i] = make_float4(acc.x, acc.y, acc.z, 0.0F);
        } else {
            acc = xyz(a.anp[i]);
        }
        Float3 total;
        if constexpr (kEpilogue == kDivPredict) {
            total = ap1;
        } else if constexpr (kTwo) {
            a.adiv[i] = make_float4(ap1.x, ap1.y, ap1.z, 0.0F);
            total = add(ap1, ap2);
        } else {
            total = add(xyz(a.adiv[i]), ap1);
        }
        Float3 nx, nv;
        predict_clamped(a.box, a.dt, p, vel, add(acc, total), nx, nv);
        if constexpr (kTwo) {
            if (a.warm_div_out != nullptr) a.warm_div_out[a.perm[i]] = a.p_div[i];
        }
        if (a.final_step != 0u) {
            const uint32_t j = a.perm[i];
            store(a.next_x, j, nx);
            store(a.next_v, j, nv);
            if (a.warm_den_out != nullptr) a.warm_den_out[j] = a.p_den[i];
        } else {
            a.xs[i] = make_float4(nx.x, nx.y, nx.z, self.w);
            const Float3 moved = sub(nx, p);
            moved2 = dot(moved, moved);
        }
    }
    if (a.final_step == 0u) block_max_to<kBlock>(moved2, &a.flags[kFlagDisplacement + a.slot]);
}

template <bool kNonPressure, int kEpilogue>
void launch_force(PassProfiler& profiler, hipStream_t stream, uint32_t blocks, const ForceArgs& args) {
    const int pass = kNonPressure ? kDfsphForceInit
                     : kEpilogue == kDivPredict ? kDfsphForceDivergence
                     : kEpilogue == kDivEnd     ? kDfsphForceDivergenceEnd
                                                : kDfsphForceDensity;
    PassScope scope(profiler, stream, pass);
    dfsph_force_kernel<kNonPressure, kEpilogue><<<blocks, kBlock, 0, stream>>>(args);
    PHX_LAUNCH_CHECK();
}

void launch_projection(PassProfiler& profiler, hipStream_t stream, uint32_t blocks, const ProjectionArgs& args,
                       int pass) {
    PassScope scope(profiler, stream, pass);
    dfsph_projection_kernel<<<blocks, kBlock, 0, stream>>>(args);
    PHX_LAUNCH_CHECK();
}

}  // namespace

float sqrt_support_threshold(float h) {
    float t = h * h;
    while (std::sqrt(t) < h) t = std::nextafter(t, INFINITY);
    for (;;) {
        const float below = std::nextafter(t, 0.0F);
        if (below <= 0.0F || std::sqrt(below) < h) break;
        t = below;
    }
    return t;
}

const char* dfsph_pass_name(int pass) {
    static const char* names[kDfsphPassCount] = {"grid",       "mask_density",     "force_init",
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

KernelConstants kernel_constants(float h) {
    KernelConstants k{};
    k.h = h;
    k.cubic = 8.0F / (kPi * h * h * h);
    k.viscosity = kPi * h * h * h * h * h * h;
    k.sqrt_support = sqrt_support_threshold(h);
    return k;
}

DfsphSolver::DfsphSolver(const phx_domain_configuration* domain, const phx_dfsph_configuration* configuration,
                         phx_execution_mode mode, hipStream_t stream)
    : domain_(make_domain(domain)), mode_(mode), stream_(stream) {
    require(configuration != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null DFSPH configuration");
    config_ = *configuration;
    require(config_.checkpoint_interval > 0u, PHX_ERROR_INVALID_ARGUMENT, "checkpoint_interval must be positive");
    require(mode == PHX_MODE_FORWARD, PHX_ERROR_UNSUPPORTED, "differentiable mode is not supported yet");
    require(config_.density_error_tolerance == 0.0F, PHX_ERROR_UNSUPPORTED, "tolerance mode is not supported yet");
    require(config_.divergence_iterations + config_.density_iterations < 4096u, PHX_ERROR_INVALID_ARGUMENT,
            "too many pressure iterations");
    gradient_norm_ = reference_gradient_norm(domain_.config);
    constants_ = kernel_constants(domain_.config.support_radius);
    const uint32_t n = domain_.config.particle_count;
    for (DeviceBuffer<float4>* b : {&xm_, &vmu_, &xs_, &adiv_}) b->resize(n);
    u_.resize(n);
    for (DeviceBuffer<float>* b : {&rho0_, &relax_div_, &relax_den_, &p_div_, &p_den_, &t_div_, &t_den_}) b->resize(n);
    mask_exact_.resize(static_cast<size_t>(kStencil) * n);
    mask_skin_.resize(static_cast<size_t>(kStencil) * n);
    flag_count_ = kFlagDisplacement + config_.divergence_iterations + config_.density_iterations + 1u;
    flags_.resize(flag_count_);
    flags_host_.resize(sizeof(uint32_t) * flag_count_);
    PHX_HIP(hipEventCreateWithFlags(&readback_, hipEventDisableTiming));
}

DfsphSolver::~DfsphSolver() {
    if (readback_ != nullptr) {
        (void)hipEventSynchronize(readback_);
        (void)hipEventDestroy(readback_);
    }
}

// Picks the skin of the next step from the largest predicted displacement of
// a previous step (read back without blocking): twice the displacement (both
// particles of a pair move) with a 30 % margin.
void DfsphSolver::adapt_skin() {
    if (!readback_pending_ || hipEventQuery(readback_) != hipSuccess) return;
    readback_pending_ = false;
    const auto* flags = static_cast<const uint32_t*>(flags_host_.data());
    float moved2 = 0.0F;
    for (uint32_t k = kFlagDisplacement; k < flag_count_; ++k) {
        float v;
        std::memcpy(&v, &flags[k], sizeof(float));
        moved2 = std::fmax(moved2, v);
    }
    const float moved = std::sqrt(moved2) / domain_.config.support_radius;
    skin_ = std::fmin(0.6F, std::fmax(0.02F, 2.6F * moved + 0.01F));
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
    adapt_skin();

    // neighborhood.build(step, x) with every input gathered into cell order
    ChannelList gather;
    gather.add(cvec(state.particles.positions), xm_.data());
    gather.add(pp.masses, xm_.data(), 3);
    gather.add(cvec(state.particles.velocities), vmu_.data());
    gather.add(pp.viscosities, vmu_.data(), 3);
    gather.add(cvec(control.external_accelerations), u_.view());
    gather.add(pp.rest_densities, rho0_.data());
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
        fill_u32(stream_, flags_.data(), flag_count_, 0u);
    }
    const SortedGrid grid = cache.neighborhood.grid();
    // whole blocks of kBlock particles, padded to a multiple of the XCD count
    const uint32_t blocks = (blocks_for(n, kBlock) + 7u) / 8u * 8u;
    const float h = c.support_radius;
    // skin radius R = h (1 + skin): the mask keeps d2 < R^2 (with margin), and
    // an A pass may use it while every particle moved at most (R - h) / 2
    const float radius = h * (1.0F + skin_);
    const float t_skin = radius * radius * (1.0F + 1e-5F);
    const float limit = 0.5F * (radius - h * (1.0F + 2e-5F));
    const float limit2 = limit > 0.0F ? limit * limit : -1.0F;

    MaskDensityArgs ma{n,
                       constants_,
                       t_skin,
                       grid,
                       xm_.data(),
                       mask_exact_.data(),
                       mask_skin_.data(),
                       p_div_.data(),
                       p_den_.data(),
                       cache.densities.data(),
                       t_div_.data(),
                       t_den_.data(),
                       flags_.data()};
    {
        PassScope scope(profiler_, stream_, kDfsphDensity);
        dfsph_mask_density_kernel<<<blocks, kBlock, 0, stream_>>>(ma);
        PHX_LAUNCH_CHECK();
    }

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
    fa.exact = mask_exact_.data();
    fa.flags = flags_.data();
    fa.anp = cache.non_pressure.data();
    fa.adiv = adiv_.data();
    fa.xs = xs_.data();
    fa.perm = cache.neighborhood.perm.data();
    fa.next_x = vec(next.particles.positions);
    fa.next_v = vec(next.particles.velocities);
    fa.p_div = p_div_.data();
    fa.warm_div_out = next.method.warm_divergence_pressure;
    fa.p_den = p_den_.data();
    fa.warm_den_out = next.method.warm_density_pressure;

    ProjectionArgs pa{};
    pa.n = n;
    pa.kc = constants_;
    pa.dt = c.time_step;
    pa.gradient_norm = gradient_norm_;
    pa.limit2 = limit2;
    pa.grid = grid;
    pa.xm = xm_.data();
    pa.xs = xs_.data();
    pa.skin = mask_skin_.data();
    pa.flags = flags_.data();
    pa.rho = cache.densities.data();

    const uint32_t nd = config_.divergence_iterations;
    const uint32_t ne = config_.density_iterations;
    uint32_t slot = 0;  // displacement slot written by the latest predicting pass

    // divergence phase: initial pressure with the warm p_div (sph.cpp:574-578)
    fa.t1 = t_div_.data();
    fa.t2 = t_den_.data();
    fa.slot = slot;
    if (nd == 0u) {
        fa.final_step = ne == 0u ? 1u : 0u;
        launch_force<true, kDivEnd>(profiler_, stream_, blocks, fa);
    } else {
        fa.final_step = 0u;
        launch_force<true, kDivPredict>(profiler_, stream_, blocks, fa);
        pa.target = cache.densities.data();
        pa.relaxation = relax_div_.data();
        pa.p = p_div_.data();
        pa.t = t_div_.data();
        for (uint32_t it = 1; it <= nd; ++it) {  // sph.cpp:580-587
            pa.slot = slot;
            launch_projection(profiler_, stream_, blocks, pa, kDfsphProjectionDivergence);
            fa.slot = ++slot;
            if (it < nd) {
                launch_force<false, kDivPredict>(profiler_, stream_, blocks, fa);
            } else {
                fa.final_step = ne == 0u ? 1u : 0u;
                launch_force<false, kDivEnd>(profiler_, stream_, blocks, fa);
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
        pa.slot = slot;
        launch_projection(profiler_, stream_, blocks, pa, kDfsphProjectionDensity);
        fa.slot = ++slot;
        fa.final_step = it == ne ? 1u : 0u;
        launch_force<false, kDenPredict>(profiler_, stream_, blocks, fa);
    }
    // read the step's displacements back for the next step's skin
    PHX_HIP(hipMemcpyAsync(flags_host_.data(), flags_.data(), sizeof(uint32_t) * flag_count_, hipMemcpyDeviceToHost,
                           stream_));
    PHX_HIP(hipEventRecord(readback_, stream_));
    readback_pending_ = true;
    next.particles.step_index = step + 1u;
    last_iterations_[0] = nd;
    last_iterations_[1] = ne;
}

}  // namespace phx
