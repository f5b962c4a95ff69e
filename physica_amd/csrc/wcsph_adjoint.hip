// SPH<WCSPH>::jvp_step and vjp_step (sph/sph.ixx:356-365, WCSPH::jvp and
// WCSPH::vjp sph/sph.cpp:228-249, kernels sph/wcsph.cu:102-435,
// sph/common.cu:196-294, 322-366, density/density.cu:42-100) in sorted order,
// over the step recorded by a differentiable forward step: its sorted inputs,
// densities, pressures, exact neighbour masks and the integration's collision
// bits, so nothing is recomputed.
//
// JVP (FP32, as the reference): the state densities' tangent with the Tait
// pressure's, then one pass over the exact masks for the tangents of the
// pressure, artificial-viscosity and cohesion accelerations and of the
// integration. VJP (FP64 adjoints of the FP32 primal): the integration's
// transpose, one gather pass for the three accelerations (each particle
// collects the pair terms where it is the centre and those where it is the
// neighbour, so nothing is written twice) ending in the Tait pressure's VJP,
// the state densities' VJP, and a scatter into the caller's arrays.
#include "adjoint_math.hpp"
#include "neighbor_mask.hpp"
#include "sph.hpp"

namespace phx {
namespace {

constexpr uint32_t kBlock = 256;

#define PHX_MASK_LDS(grid)                          \
    __shared__ uint16_t smask[kStencil * kBlock];   \
    __shared__ int delta[kStencil];                 \
    fill_slot_deltas(grid, delta);                  \
    __syncthreads()

// The recorded step in sorted order.
struct WcsphStep {
    uint32_t n;
    KernelConstants kc;
    float dt;
    float diameter;  // 2 r_p, below which the cohesion weight is clamped
    Box box;
    uint32_t exact_no_slip;  // 1: exact transpose of a no-slip collision, 0: the reference's VJP
    SortedGrid grid;
    const float4* xm;   // position + mass
    const float4* vmu;  // velocity + viscosity
    const float4* aux;  // density, p / rho^2, speed of sound, surface tension
    const float* p;     // pressures
    const float* rho0;
    const float* tait;
    const uint16_t* exact;
    const uint32_t* flags;
    const uint8_t* collisions;
    BoundaryGrid boundary;   // offsets null: no boundary samples
    const float* btension;   // boundary surface tensions (boundary samples only)
    const uint8_t* crowded;  // per particle: walks the full stencil
};

__device__ __forceinline__ MaskView exact_view(const WcsphStep& s) {
    return MaskView{s.exact, s.flags, s.n, 0.0F, 0u, s.crowded};
}
// Whether particle i walks the step's exact masks (not crowded, no boundary samples).
__device__ __forceinline__ bool exact_masked(const WcsphStep& s, uint32_t i) {
    return s.flags[kFlagOverflow] == 0u && s.crowded[i] == 0u;
}

// Visits the fluid candidates of particle i through its exact masks or, with
// boundary samples, every fluid particle and every sample of the 27 cells in
// the reference order: each cell's fluid, then its samples.
template <class Visit, class VisitBoundary>
__device__ __forceinline__ void walk(const WcsphStep& s, uint32_t i, int cx, int cy, int cz, const int* delta,
                                     uint16_t* column, Visit&& visit, VisitBoundary&& visit_boundary) {
    if (s.boundary.offsets != nullptr) {
        for_each_candidate(s.grid, s.boundary, cx, cy, cz, visit, visit_boundary);
    } else {
        visit_candidates(s.grid, exact_view(s), exact_masked(s, i), i, cx, cy, cz, delta, column, kBlock, visit);
    }
}

// The cohesion weight of surface_forward_kernel (wcsph.cu:324): W at the
// distance, clamped below at the particle diameter.
__device__ __forceinline__ float cohesion_weight(float dist, const WcsphStep& s) {
    return dist > s.diameter ? cubic_value(dist, s.kc) : cubic_value(length(Float3{s.diameter, 0.0F, 0.0F}), s.kc);
}

// ----------------------------------------------------------------- JVP
struct WcsphGatherArgs {
    uint32_t n;
    const uint32_t* perm;
    CVec x_t, v_t, u_t;
    const float *m_t, *rest_t, *visc_t, *tension_t, *speed_t, *tait_t, *btension_t;  // null: zero
    float4* xt;  // .w: mass tangent
    float4* vt;  // .w: viscosity tangent
    float4* ut;  // .w: rest-density tangent
    float4* pt;  // speed of sound, Tait exponent, surface tension, boundary surface tension tangents
};

__global__ void __launch_bounds__(kBlock) wcsph_gather_tangent_kernel(WcsphGatherArgs a) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t j = a.perm[i];
    a.xt[i] = make_float4(a.x_t.x[j], a.x_t.y[j], a.x_t.z[j], or_zero(a.m_t, j));
    a.vt[i] = make_float4(a.v_t.x[j], a.v_t.y[j], a.v_t.z[j], or_zero(a.visc_t, j));
    a.ut[i] = make_float4(a.u_t.x[j], a.u_t.y[j], a.u_t.z[j], or_zero(a.rest_t, j));
    a.pt[i] = make_float4(or_zero(a.speed_t, j), or_zero(a.tait_t, j), or_zero(a.tension_t, j),
                          or_zero(a.btension_t, j));
}

struct WcsphDensityJvpArgs {
    WcsphStep s;
    const float4* xt;
    const float4* ut;
    const float4* pt;
    float* rho_t;
    float2* tt;  // (p / rho^2, its tangent)
};

// density_jvp_kernel (density.cu:42-67, topology = positions = x) +
// eos_jvp_kernel (wcsph.cu:102-117) + the first term's tangent.
__global__ void __launch_bounds__(kBlock) wcsph_jvp_density_kernel(WcsphDensityJvpArgs a) {
    const WcsphStep& s = a.s;
    PHX_MASK_LDS(s.grid);
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= s.n) return;
    const Float3 p = xyz(s.xm[i]);
    const Float3 pti = xyz(a.xt[i]);
    int cx, cy, cz;
    particle_cell(s.grid, p, cx, cy, cz);
    float rho_t = 0.0F;
    const float rest = s.rho0[i];
    const float rest_t = a.ut[i].w;
    walk(
        s, i, cx, cy, cz, delta, smask + threadIdx.x,
        [&](uint32_t j) {
            const float4 q = s.xm[j];
            const float4 qt = a.xt[j];
            const Float3 d = sub(p, xyz(q));
            const float dist = length(d);
            rho_t += qt.w * cubic_value(dist, s.kc) +
                     q.w * dot(scale(d, cubic_gradient_factor(dist, s.kc)), sub(pti, xyz(qt)));
        },
        [&](uint32_t b) {
            const float4 q = s.boundary.x[b];
            const Float3 d = sub(p, xyz(q));
            const float dist = length(d);
            rho_t += q.w * (rest_t * cubic_value(dist, s.kc) +
                            rest * dot(scale(d, cubic_gradient_factor(dist, s.kc)), pti));
        });
    a.rho_t[i] = rho_t;
    const float4 own = s.aux[i];
    const float density = own.x;
    const float speed = own.z;
    const float exponent = s.tait[i];
    const float4 pt = a.pt[i];
    const float ratio = density / rest;
    const float powered = powf(ratio, exponent);
    const float difference = powered - 1.0F;
    const float d_density = speed * speed * powf(ratio, exponent - 1.0F);
    const float d_rest = speed * speed * difference / exponent - speed * speed * powered;
    const float d_speed = 2.0F * rest * speed * difference / exponent;
    const float d_exponent = rest * speed * speed * (exponent * powered * logf(ratio) - difference) / (exponent * exponent);
    const float p_t = d_density * rho_t + d_rest * rest_t + d_speed * pt.x + d_exponent * pt.y;
    a.tt[i] = first_term(s.p[i], p_t, density, rho_t);
}

struct WcsphForceJvpArgs {
    WcsphStep s;
    const float4* xt;
    const float4* vt;
    const float4* ut;
    const float4* pt;
    const float* rho_t;
    const float2* tt;
    const uint32_t* perm;
    Vec next_x_t, next_v_t;
};

// pressure_jvp_kernel (common.cu:196-237), viscosity_jvp_kernel
// (wcsph.cu:176-237) and surface_jvp_kernel (wcsph.cu:341-386) over one walk,
// the sums of WCSPH::jvp (sph.cpp:234-236) and integrate_jvp_kernel
// (common.cu:322-342) with the recorded collision bits.
__global__ void __launch_bounds__(kBlock) wcsph_jvp_force_kernel(WcsphForceJvpArgs a) {
    const WcsphStep& s = a.s;
    PHX_MASK_LDS(s.grid);
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= s.n) return;
    const float h = s.kc.h;
    const float4 self = s.xm[i];
    const Float3 x = xyz(self);
    const float4 xt_i = a.xt[i];
    const Float3 xti = xyz(xt_i);
    const float4 vm = s.vmu[i];
    const Float3 vel = xyz(vm);
    const float4 vt_i = a.vt[i];
    const Float3 vti = xyz(vt_i);
    const float4 own = s.aux[i];
    const float4 pt_i = a.pt[i];
    const float2 tt_i = a.tt[i];
    const float rho_t_i = a.rho_t[i];
    const float m_i = self.w;
    const float m_t_i = xt_i.w;
    Float3 ap{0.0F, 0.0F, 0.0F};
    Float3 av{0.0F, 0.0F, 0.0F};
    Float3 as{0.0F, 0.0F, 0.0F};
    const float rest_i = s.rho0[i];
    const float rest_t_i = a.ut[i].w;
    int cx, cy, cz;
    particle_cell(s.grid, x, cx, cy, cz);
    walk(s, i, cx, cy, cz, delta, smask + threadIdx.x,
                     [&](uint32_t j) {
                         if (j == i) return;
                         const float4 q = s.xm[j];
                         const float4 qt = a.xt[j];
                         const float mass = q.w;
                         const float mass_t = qt.w;
                         const Float3 d = sub(x, xyz(q));
                         const Float3 dd = sub(xti, xyz(qt));
                         const float dist = length(d);
                         const Float3 gradient = scale(d, cubic_gradient_factor(dist, s.kc));
                         const Float3 gradient_t = cubic_gradient_tangent(d, dd, s.kc);
                         // pressure
                         const float2 tt_j = a.tt[j];
                         const float factor = -mass * (tt_i.x + tt_j.x);
                         const float factor_t = -mass_t * (tt_i.x + tt_j.x) - mass * (tt_i.y + tt_j.y);
                         ap = add(ap, add(scale(gradient, factor_t), scale(gradient_t, factor)));
                         // artificial viscosity on approaching pairs
                         const float4 vq = s.vmu[j];
                         const Float3 dv = sub(vel, xyz(vq));
                         const float compression = dot(dv, d);
                         if (!(compression >= 0.0F)) {
                             const float4 vqt = a.vt[j];
                             const float4 other = s.aux[j];
                             const float4 pt_j = a.pt[j];
                             const Float3 dvd = sub(vti, xyz(vqt));
                             const float alpha = 0.5F * (vm.w + vq.w);
                             const float alpha_t = 0.5F * (vt_i.w + vqt.w);
                             const float speed = 0.5F * (own.z + other.z);
                             const float speed_t = 0.5F * (pt_i.x + pt_j.x);
                             const float rho_sum = own.x + other.x;
                             const float rho_sum_t = rho_t_i + a.rho_t[j];
                             const float nu = 2.0F * alpha * h * speed / rho_sum;
                             const float nu_t =
                                 2.0F * h * (alpha_t * speed + alpha * speed_t) / rho_sum - nu * rho_sum_t / rho_sum;
                             const float compression_t = dot(dvd, d) + dot(dv, dd);
                             const float denominator = dot(d, d) + 0.01F * h * h;
                             const float denominator_t = 2.0F * dot(d, dd);
                             const float pi = -nu * compression / denominator;
                             const float pi_t = -(nu_t * compression + nu * compression_t) / denominator +
                                                nu * compression * denominator_t / (denominator * denominator);
                             av = add(av, add(scale(gradient, -(mass_t * pi + mass * pi_t)),
                                              scale(gradient_t, -mass * pi)));
                         }
                         // cohesion inside the support
                         if (!(dist >= h)) {
                             const float w = cohesion_weight(dist, s);
                             const float w_t = dist > s.diameter ? dot(gradient, dd) : 0.0F;
                             const float sigma = own.w;
                             const float f = -sigma * mass * w / m_i;
                             const float f_t = -pt_i.z * mass * w / m_i - sigma * mass_t * w / m_i -
                                               sigma * mass * w_t / m_i + sigma * mass * w * m_t_i / (m_i * m_i);
                             as = add(as, add(scale(d, f_t), scale(dd, f)));
                         }
                     },
                     [&](uint32_t b) {
                         const float4 q = s.boundary.x[b];
                         const float vol = q.w;
                         const Float3 d = sub(x, xyz(q));
                         const float dist = length(d);
                         const Float3 gradient = scale(d, cubic_gradient_factor(dist, s.kc));
                         const Float3 gradient_t = cubic_gradient_tangent(d, xti, s.kc);
                         // pressure: -rho0_i V (p_i / rho_i^2) grad W
                         const float factor = -rest_i * vol * tt_i.x;
                         const float factor_t = -vol * (rest_t_i * tt_i.x + rest_i * tt_i.y);
                         ap = add(ap, add(scale(gradient, factor_t), scale(gradient_t, factor)));
                         // artificial viscosity: nu = mu_i h c_i / rho_i, mass rho0_i V
                         const Float3 dv = sub(vel, xyz(s.boundary.v[b]));
                         const float compression = dot(dv, d);
                         if (!(compression >= 0.0F)) {
                             const float nu = vm.w * h * own.z / own.x;
                             const float nu_t = h * (vt_i.w * own.z + vm.w * pt_i.x) / own.x - nu * rho_t_i / own.x;
                             const float mass = rest_i * vol;
                             const float mass_t = rest_t_i * vol;
                             const float compression_t = dot(vti, d) + dot(dv, xti);
                             const float denominator = dot(d, d) + 0.01F * h * h;
                             const float denominator_t = 2.0F * dot(d, xti);
                             const float pi = -nu * compression / denominator;
                             const float pi_t = -(nu_t * compression + nu * compression_t) / denominator +
                                                nu * compression * denominator_t / (denominator * denominator);
                             av = add(av, add(scale(gradient, -(mass_t * pi + mass * pi_t)),
                                              scale(gradient_t, -mass * pi)));
                         }
                         // cohesion with the boundary surface tension, mass rho0_i V
                         if (!(dist >= h)) {
                             const float w = cohesion_weight(dist, s);
                             const float w_t = dist > s.diameter ? dot(gradient, xti) : 0.0F;
                             const float sigma = s.btension[i];
                             const float mb = rest_i * vol;
                             const float mb_t = rest_t_i * vol;
                             const float f = -sigma * mb * w / m_i;
                             const float f_t = -pt_i.w * mb * w / m_i - sigma * mb_t * w / m_i -
                                               sigma * mb * w_t / m_i + sigma * mb * w * m_t_i / (m_i * m_i);
                             as = add(as, add(scale(d, f_t), scale(xti, f)));
                         }
                     });
    Float3 total = add(ap, av);
    total = add(total, as);
    total = add(total, xyz(a.ut[i]));
    Float3 nx, nv;
    predict_tangent(s.collisions[i], s.box.no_slip, s.dt, xti, vti, total, nx, nv);
    const uint32_t j = a.perm[i];
    store(a.next_x_t, j, nx);
    store(a.next_v_t, j, nv);
}

// ----------------------------------------------------------------- VJP
struct WcsphFinalVjpArgs {
    WcsphStep s;
    const uint32_t* perm;
    phx_const_adjoint_vector_view next_x_a, next_v_a;
    DVec pos_a, vel_a, total_a;  // written
};

// integrate_vjp_kernel (common.cu:344-366) with the recorded collision bits
// (under no-slip collisions the reference's zeroed velocity adjoint, or the
// exact transpose of the tangent when the solver asks for it).
__global__ void __launch_bounds__(kBlock) wcsph_vjp_final_kernel(WcsphFinalVjpArgs a) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.s.n) return;
    const uint32_t j = a.perm[i];
    const D3 xa{a.next_x_a.x[j], a.next_x_a.y[j], a.next_x_a.z[j]};
    const D3 va = a.next_v_a.x != nullptr ? D3{a.next_v_a.x[j], a.next_v_a.y[j], a.next_v_a.z[j]} : D3{0.0, 0.0, 0.0};
    D3 position, velocity;
    predict_adjoint(a.s.collisions[i], a.s.box.no_slip, a.s.exact_no_slip, static_cast<double>(a.s.dt), xa, va, position,
                    velocity);
    store(a.pos_a, i, position);
    store(a.vel_a, i, velocity);
    store(a.total_a, i, scale(velocity, static_cast<double>(a.s.dt)));
}

// viscosity_reverse (wcsph.cu:40-72): the adjoint of one artificial-viscosity
// term -m_b pi W'(d) of an approaching pair for the acceleration adjoint w of
// its centre a.
struct ViscosityAdjoint {
    D3 displacement, velocity;
    double mass_b, rho, mu, speed;  // rho, mu, speed: the adjoint of each end's value
};

__device__ __forceinline__ ViscosityAdjoint viscosity_reverse(Float3 d, Float3 dv, float rho_a, float rho_b, float mu_a,
                                                              float mu_b, float c_a, float c_b, float mass_b, float h,
                                                              const KernelConstants& kc, D3 w) {
    ViscosityAdjoint r{};
    const double compression = dot(dv, d);
    if (compression >= 0.0) return r;
    const double denominator = static_cast<double>(dot(d, d)) + 0.01 * h * h;
    const double alpha = 0.5 * static_cast<double>(mu_a + mu_b);
    const double speed = 0.5 * static_cast<double>(c_a + c_b);
    const double rho_sum = static_cast<double>(rho_a + rho_b);
    const double nu = 2.0 * alpha * h * speed / rho_sum;
    const double pi = -nu * compression / denominator;
    const D3 gradient = d3(scale(d, cubic_gradient_factor(length(d), kc)));
    r.mass_b = -pi * dot(w, gradient);
    const double pi_a = -static_cast<double>(mass_b) * dot(w, gradient);
    r.displacement = cubic_hessian_product(d, scale(w, -static_cast<double>(mass_b) * pi), h);
    const double nu_a = -compression * pi_a / denominator;
    const double compression_a = -nu * pi_a / denominator;
    const double denominator_a = nu * compression * pi_a / (denominator * denominator);
    r.velocity = scale(d3(d), compression_a);
    r.displacement = add(r.displacement, add(scale(d3(dv), compression_a), scale(d3(d), 2.0 * denominator_a)));
    r.mu = 0.5 * (2.0 * h * speed * nu_a / rho_sum);
    r.speed = 0.5 * (2.0 * alpha * h * nu_a / rho_sum);
    r.rho = -nu * nu_a / rho_sum;
    return r;
}

// surface_reverse (wcsph.cu:74-88): the adjoint of one cohesion term
// -s m_b w(d) / m_a d for the acceleration adjoint w of its centre a.
struct SurfaceAdjoint {
    D3 displacement;
    double mass_a, mass_b, tension;
};

__device__ __forceinline__ SurfaceAdjoint surface_reverse(Float3 d, float mass_a, float mass_b, float tension,
                                                          const WcsphStep& s, D3 w) {
    SurfaceAdjoint r{};
    const float dist = length(d);
    if (dist >= s.kc.h) return r;
    const float weight = cohesion_weight(dist, s);
    const double factor = -static_cast<double>(tension) * mass_b * weight / mass_a;
    const double factor_a = dot(w, d3(d));
    r.displacement = scale(w, factor);
    if (dist > s.diameter)
        r.displacement = add(r.displacement, scale(d3(scale(d, cubic_gradient_factor(dist, s.kc))),
                                                   -static_cast<double>(tension) * mass_b * factor_a / mass_a));
    r.tension = -static_cast<double>(mass_b) * weight * factor_a / mass_a;
    r.mass_b = -static_cast<double>(tension) * weight * factor_a / mass_a;
    r.mass_a = static_cast<double>(tension) * mass_b * weight * factor_a / (static_cast<double>(mass_a) * mass_a);
    return r;
}

struct WcsphForceVjpArgs {
    WcsphStep s;
    DVec total_a;  // acceleration adjoints (complete)
    DVec pos_a, vel_a;  // accumulated
    double *rho_a, *m_a, *rest_a, *visc_a, *tension_a, *speed_a, *tait_a, *btension_a;  // written
};

// pressure_vjp_kernel (common.cu:239-294), viscosity_vjp_kernel
// (wcsph.cu:239-304) and surface_vjp_kernel (wcsph.cu:388-435) over one walk,
// then eos_vjp_kernel (wcsph.cu:119-134).
__global__ void __launch_bounds__(kBlock) wcsph_vjp_force_kernel(WcsphForceVjpArgs a) {
    const WcsphStep& s = a.s;
    PHX_MASK_LDS(s.grid);
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= s.n) return;
    const float h = s.kc.h;
    const float4 self = s.xm[i];
    const Float3 x = xyz(self);
    const float m_i = self.w;
    const float4 vm = s.vmu[i];
    const Float3 vel = xyz(vm);
    const float4 own = s.aux[i];
    const float rho_i = own.x;
    const D3 local = load(a.total_a, i);
    const double first = static_cast<double>(s.p[i]) / (rho_i * rho_i);
    D3 position{0.0, 0.0, 0.0};
    D3 velocity{0.0, 0.0, 0.0};
    double first_a = 0.0, mass_c = 0.0, rho_c = 0.0, mu_c = 0.0, speed_c = 0.0, tension_c = 0.0;
    double rest_c = 0.0, btension_c = 0.0;
    const float rest_i = s.rho0[i];
    int cx, cy, cz;
    particle_cell(s.grid, x, cx, cy, cz);
    walk(s, i, cx, cy, cz, delta, smask + threadIdx.x,
                     [&](uint32_t j) {
                         if (j == i) return;
                         const float4 q = s.xm[j];
                         const Float3 d = sub(x, xyz(q));
                         const float mass = q.w;
                         const D3 cross = load(a.total_a, j);
                         // pressure
                         const D3 gradient = d3(scale(d, cubic_gradient_factor(length(d), s.kc)));
                         const double second = static_cast<double>(s.p[j]) / (s.aux[j].x * s.aux[j].x);
                         const double sum = first + second;
                         first_a -= mass * dot(local, gradient);
                         position = add(position, cubic_hessian_product(d, scale(local, -mass * sum), h));
                         const double projection = dot(cross, gradient);
                         mass_c += sum * projection;
                         first_a += m_i * projection;
                         position = add(position, cubic_hessian_product(d, scale(cross, m_i * sum), h));
                         // artificial viscosity: as centre, then as neighbour
                         const float4 vq = s.vmu[j];
                         const float4 other = s.aux[j];
                         const Float3 dv = sub(vel, xyz(vq));
                         const ViscosityAdjoint mine = viscosity_reverse(d, dv, rho_i, other.x, vm.w, vq.w, own.z,
                                                                         other.z, mass, h, s.kc, local);
                         position = add(position, mine.displacement);
                         velocity = add(velocity, mine.velocity);
                         rho_c += mine.rho;
                         mu_c += mine.mu;
                         speed_c += mine.speed;
                         const ViscosityAdjoint theirs =
                             viscosity_reverse(scale(d, -1.0F), scale(dv, -1.0F), other.x, rho_i, vq.w, vm.w, other.z,
                                               own.z, m_i, h, s.kc, cross);
                         position = add(position, scale(theirs.displacement, -1.0));
                         velocity = add(velocity, scale(theirs.velocity, -1.0));
                         mass_c += theirs.mass_b;
                         rho_c += theirs.rho;
                         mu_c += theirs.mu;
                         speed_c += theirs.speed;
                         // cohesion: as centre, then as neighbour
                         const SurfaceAdjoint smine = surface_reverse(d, m_i, mass, own.w, s, local);
                         position = add(position, smine.displacement);
                         mass_c += smine.mass_a;
                         tension_c += smine.tension;
                         const SurfaceAdjoint stheirs = surface_reverse(scale(d, -1.0F), mass, m_i, other.w, s, cross);
                         position = add(position, scale(stheirs.displacement, -1.0));
                         mass_c += stheirs.mass_b;
                     },
                     [&](uint32_t b) {
                         // a sample is not differentiated: its terms reach particle i only
                         const float4 q = s.boundary.x[b];
                         const float vol = q.w;
                         const Float3 d = sub(x, xyz(q));
                         // pressure: -rho0_i V (p_i / rho_i^2) grad W
                         const D3 gradient = d3(scale(d, cubic_gradient_factor(length(d), s.kc)));
                         const double factor = static_cast<double>(-rest_i * vol) * first;
                         const double factor_a = dot(local, gradient);
                         rest_c -= factor_a * vol * first;
                         first_a -= factor_a * rest_i * vol;
                         position = add(position, cubic_hessian_product(d, scale(local, factor), h));
                         // artificial viscosity: the fluid form with both ends particle i, mass rho0_i V
                         const Float3 dv = sub(vel, xyz(s.boundary.v[b]));
                         const ViscosityAdjoint mine = viscosity_reverse(d, dv, rho_i, rho_i, vm.w, vm.w, own.z, own.z,
                                                                         rest_i * vol, h, s.kc, local);
                         position = add(position, mine.displacement);
                         velocity = add(velocity, mine.velocity);
                         rest_c += mine.mass_b * vol;
                         rho_c += mine.rho + mine.rho;
                         mu_c += mine.mu + mine.mu;
                         speed_c += mine.speed + mine.speed;
                         // cohesion with the boundary surface tension, mass rho0_i V
                         const SurfaceAdjoint smine = surface_reverse(d, m_i, rest_i * vol, s.btension[i], s, local);
                         position = add(position, smine.displacement);
                         mass_c += smine.mass_a;
                         rest_c += smine.mass_b * vol;
                         btension_c += smine.tension;
                     });
    accumulate(a.pos_a, i, position);
    accumulate(a.vel_a, i, velocity);
    const double p_a = first_a / (rho_i * rho_i);
    double rho_a = rho_c - 2.0 * first_a * s.p[i] / (rho_i * rho_i * rho_i);
    // eos_vjp_kernel (wcsph.cu:119-134)
    const double rest = s.rho0[i];
    const double speed = own.z;
    const double exponent = s.tait[i];
    const double ratio = static_cast<double>(rho_i) / rest;
    const double powered = pow(ratio, exponent);
    const double difference = powered - 1.0;
    rho_a += p_a * speed * speed * pow(ratio, exponent - 1.0);
    a.rest_a[i] = rest_c + p_a * (speed * speed * difference / exponent - speed * speed * powered);
    a.speed_a[i] = speed_c + p_a * 2.0 * rest * speed * difference / exponent;
    a.tait_a[i] = p_a * rest * speed * speed * (exponent * powered * log(ratio) - difference) / (exponent * exponent);
    a.rho_a[i] = rho_a;
    a.m_a[i] = mass_c;
    a.visc_a[i] = mu_c;
    a.tension_a[i] = tension_c;
    a.btension_a[i] = btension_c;
}

struct WcsphDensityVjpArgs {
    WcsphStep s;
    const double* rho_a;
    DVec pos_a;
    double* m_a;
    double* rest_a;  // accumulated: the boundary terms' rest-density adjoint
};

// density_vjp_kernel (density.cu:69-100, topology = positions = x), gather
// form: particle i collects the terms of its own density and of its
// neighbours' densities (sph.ixx:364).
__global__ void __launch_bounds__(kBlock) wcsph_vjp_density_kernel(WcsphDensityVjpArgs a) {
    const WcsphStep& s = a.s;
    PHX_MASK_LDS(s.grid);
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= s.n) return;
    const float4 self = s.xm[i];
    const Float3 x = xyz(self);
    const double out_a = a.rho_a[i];
    D3 position{0.0, 0.0, 0.0};
    double mass_c = 0.0;
    double rest_c = 0.0;
    const float rest_i = s.rho0[i];
    int cx, cy, cz;
    particle_cell(s.grid, x, cx, cy, cz);
    walk(
        s, i, cx, cy, cz, delta, smask + threadIdx.x,
        [&](uint32_t j) {
            const float4 q = s.xm[j];
            const Float3 d = sub(x, xyz(q));
            const float dist = length(d);
            const double rho_a_j = a.rho_a[j];
            position = add(position, scale(d3(scale(d, cubic_gradient_factor(dist, s.kc))),
                                           out_a * q.w + rho_a_j * self.w));
            mass_c += rho_a_j * cubic_value(dist, s.kc);
        },
        [&](uint32_t b) {
            const float4 q = s.boundary.x[b];
            const Float3 d = sub(x, xyz(q));
            const float dist = length(d);
            position = add(position, scale(d3(scale(d, cubic_gradient_factor(dist, s.kc))), out_a * rest_i * q.w));
            rest_c += out_a * q.w * cubic_value(dist, s.kc);
        });
    accumulate(a.pos_a, i, position);
    a.m_a[i] += mass_c;
    if (s.boundary.offsets != nullptr) a.rest_a[i] += rest_c;
}

struct WcsphScatterArgs {
    uint32_t n;
    const uint32_t* perm;
    DVec pos_a, vel_a, total_a;
    const double *m_a, *rest_a, *visc_a, *tension_a, *speed_a, *tait_a, *btension_a;
    phx_adjoint_vector_view out_pos, out_vel, out_ctl;  // null: skipped
    double *out_m, *out_rest, *out_visc, *out_tension, *out_speed, *out_tait, *out_btension;
};

__device__ __forceinline__ void accumulate_view(phx_adjoint_vector_view v, uint32_t j, D3 a) {
    if (v.x == nullptr) return;
    v.x[j] += a.x;
    v.y[j] += a.y;
    v.z[j] += a.z;
}

__global__ void __launch_bounds__(kBlock) wcsph_vjp_scatter_kernel(WcsphScatterArgs a) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t j = a.perm[i];
    accumulate_view(a.out_pos, j, load(a.pos_a, i));
    accumulate_view(a.out_vel, j, load(a.vel_a, i));
    // external_forward_kernel: the control's adjoint is the acceleration's
    accumulate_view(a.out_ctl, j, load(a.total_a, i));
    accumulate_to(a.out_m, j, a.m_a[i]);
    accumulate_to(a.out_rest, j, a.rest_a[i]);
    accumulate_to(a.out_visc, j, a.visc_a[i]);
    accumulate_to(a.out_tension, j, a.tension_a[i]);
    accumulate_to(a.out_btension, j, a.btension_a[i]);
    accumulate_to(a.out_speed, j, a.speed_a[i]);
    accumulate_to(a.out_tait, j, a.tait_a[i]);
}

uint32_t derivative_blocks(uint32_t n) { return (blocks_for(n, kBlock) + 7u) / 8u * 8u; }

}  // namespace

void WcsphDerivativeWorkspace::allocate_tangents(uint32_t n) {
    for (DeviceBuffer<float4>* b : {&xt, &vt, &ut, &pt}) b->resize(n);
    rho_t.resize(n);
    tt.resize(n);
}

void WcsphDerivativeWorkspace::allocate_adjoints(uint32_t n) {
    for (DVecBuffer* b : {&pos_a, &vel_a, &total_a}) b->resize(n);
    for (DeviceBuffer<double>* b : {&rho_a, &m_a, &rest_a, &visc_a, &tension_a, &speed_a, &tait_a, &btension_a})
        b->resize(n);
}

namespace {

WcsphStep make_wcsph_step(const Domain& domain, const SphCache& cache, const KernelConstants& kc, bool exact_no_slip) {
    const auto& c = domain.config;
    const WcsphRecord& r = *cache.wcsph;
    WcsphStep s{};
    s.n = c.particle_count;
    s.kc = kc;
    s.dt = c.time_step;
    s.diameter = 2.0F * c.particle_radius;
    s.box = collision_box(c, r.step_index);
    s.exact_no_slip = exact_no_slip ? 1u : 0u;
    s.grid = cache.neighborhood.grid();
    s.xm = r.xm.data();
    s.vmu = r.vmu.data();
    s.aux = r.aux.data();
    s.p = cache.pressures.data();
    s.rho0 = r.rho0.data();
    s.tait = r.tait.data();
    s.exact = r.exact.data();
    s.flags = r.flags.data();
    s.collisions = r.collisions.data();
    s.boundary = cache.neighborhood.boundary();
    s.btension = r.btension.data();
    s.crowded = r.crowded.data();
    return s;
}

}  // namespace

void SphSolver::wcsph_jvp_step(const SphCache& cache, const phx_sph_state_tangent& state_t,
                               const phx_control_tangent& control_t, const phx_sph_parameter_tangent& parameter_t,
                               phx_sph_state_tangent& next_t) {
    const uint32_t n = domain_.config.particle_count;
    if (n == 0) return;
    require_vec(state_t.particles.positions, "position tangent");
    require_vec(state_t.particles.velocities, "velocity tangent");
    require_vec(control_t.external_accelerations, "control tangent");
    require_vec(next_t.particles.positions, "next position tangent");
    require_vec(next_t.particles.velocities, "next velocity tangent");
    wdw_.allocate_tangents(n);
    const WcsphStep s = make_wcsph_step(domain_, cache, constants_, config_.exact_no_slip_adjoint != 0u);
    const uint32_t* perm = cache.neighborhood.perm.data();
    WcsphGatherArgs ga{};
    ga.n = n;
    ga.perm = perm;
    ga.x_t = cvec(state_t.particles.positions);
    ga.v_t = cvec(state_t.particles.velocities);
    ga.u_t = cvec(control_t.external_accelerations);
    ga.m_t = parameter_t.particles.masses;
    ga.rest_t = parameter_t.particles.rest_densities;
    ga.visc_t = parameter_t.particles.viscosities;
    ga.tension_t = parameter_t.particles.surface_tensions;
    ga.speed_t = parameter_t.method.speed_of_sound;
    ga.tait_t = parameter_t.method.tait_exponent;
    ga.btension_t = parameter_t.method.boundary_surface_tension;
    ga.xt = wdw_.xt.data();
    ga.vt = wdw_.vt.data();
    ga.ut = wdw_.ut.data();
    ga.pt = wdw_.pt.data();
    wcsph_gather_tangent_kernel<<<blocks_for(n, kBlock), kBlock, 0, stream_>>>(ga);
    PHX_LAUNCH_CHECK();
    const uint32_t blocks = derivative_blocks(n);
    WcsphDensityJvpArgs da{s, wdw_.xt.data(), wdw_.ut.data(), wdw_.pt.data(), wdw_.rho_t.data(), wdw_.tt.data()};
    wcsph_jvp_density_kernel<<<blocks, kBlock, 0, stream_>>>(da);
    PHX_LAUNCH_CHECK();
    WcsphForceJvpArgs fa{};
    fa.s = s;
    fa.xt = wdw_.xt.data();
    fa.vt = wdw_.vt.data();
    fa.ut = wdw_.ut.data();
    fa.pt = wdw_.pt.data();
    fa.rho_t = wdw_.rho_t.data();
    fa.tt = wdw_.tt.data();
    fa.perm = perm;
    fa.next_x_t = vec(next_t.particles.positions);
    fa.next_v_t = vec(next_t.particles.velocities);
    wcsph_jvp_force_kernel<<<blocks, kBlock, 0, stream_>>>(fa);
    PHX_LAUNCH_CHECK();
}

void SphSolver::wcsph_vjp_step(const SphCache& cache, const phx_sph_state_adjoint& next_a,
                               phx_sph_state_adjoint& previous_a, phx_control_adjoint& control_a,
                               phx_sph_parameter_adjoint& parameter_a) {
    const uint32_t n = domain_.config.particle_count;
    if (n == 0) return;
    const auto& nx = next_a.particles.positions;
    require(nx.x && nx.y && nx.z, PHX_ERROR_INVALID_ARGUMENT, "null next-state adjoint");
    const auto& nv = next_a.particles.velocities;
    require((nv.x == nullptr) == (nv.y == nullptr) && (nv.x == nullptr) == (nv.z == nullptr),
            PHX_ERROR_INVALID_ARGUMENT, "next velocity adjoint components must be all null or all set");
    auto all_or_none = [](const phx_adjoint_vector_view& v, const char* what) {
        require((v.x == nullptr) == (v.y == nullptr) && (v.x == nullptr) == (v.z == nullptr),
                PHX_ERROR_INVALID_ARGUMENT, std::string(what) + " components must be all null or all set");
    };
    all_or_none(previous_a.particles.positions, "position adjoint");
    all_or_none(previous_a.particles.velocities, "velocity adjoint");
    all_or_none(control_a.external_accelerations, "control adjoint");
    wdw_.allocate_adjoints(n);
    const WcsphStep s = make_wcsph_step(domain_, cache, constants_, config_.exact_no_slip_adjoint != 0u);
    const uint32_t* perm = cache.neighborhood.perm.data();
    WcsphFinalVjpArgs fa{};
    fa.s = s;
    fa.perm = perm;
    fa.next_x_a = {nx.x, nx.y, nx.z};
    fa.next_v_a = {nv.x, nv.y, nv.z};
    fa.pos_a = wdw_.pos_a.view();
    fa.vel_a = wdw_.vel_a.view();
    fa.total_a = wdw_.total_a.view();
    wcsph_vjp_final_kernel<<<blocks_for(n, kBlock), kBlock, 0, stream_>>>(fa);
    PHX_LAUNCH_CHECK();
    const uint32_t blocks = derivative_blocks(n);
    WcsphForceVjpArgs va{};
    va.s = s;
    va.total_a = wdw_.total_a.view();
    va.pos_a = wdw_.pos_a.view();
    va.vel_a = wdw_.vel_a.view();
    va.rho_a = wdw_.rho_a.data();
    va.m_a = wdw_.m_a.data();
    va.rest_a = wdw_.rest_a.data();
    va.visc_a = wdw_.visc_a.data();
    va.tension_a = wdw_.tension_a.data();
    va.speed_a = wdw_.speed_a.data();
    va.tait_a = wdw_.tait_a.data();
    va.btension_a = wdw_.btension_a.data();
    wcsph_vjp_force_kernel<<<blocks, kBlock, 0, stream_>>>(va);
    PHX_LAUNCH_CHECK();
    WcsphDensityVjpArgs dv{s, wdw_.rho_a.data(), wdw_.pos_a.view(), wdw_.m_a.data(), wdw_.rest_a.data()};
    wcsph_vjp_density_kernel<<<blocks, kBlock, 0, stream_>>>(dv);
    PHX_LAUNCH_CHECK();
    WcsphScatterArgs sa{};
    sa.n = n;
    sa.perm = perm;
    sa.pos_a = wdw_.pos_a.view();
    sa.vel_a = wdw_.vel_a.view();
    sa.total_a = wdw_.total_a.view();
    sa.m_a = wdw_.m_a.data();
    sa.rest_a = wdw_.rest_a.data();
    sa.visc_a = wdw_.visc_a.data();
    sa.tension_a = wdw_.tension_a.data();
    sa.speed_a = wdw_.speed_a.data();
    sa.tait_a = wdw_.tait_a.data();
    sa.btension_a = wdw_.btension_a.data();
    sa.out_pos = previous_a.particles.positions;
    sa.out_vel = previous_a.particles.velocities;
    sa.out_ctl = control_a.external_accelerations;
    sa.out_m = parameter_a.particles.masses;
    sa.out_rest = parameter_a.particles.rest_densities;
    sa.out_visc = parameter_a.particles.viscosities;
    sa.out_tension = parameter_a.particles.surface_tensions;
    sa.out_speed = parameter_a.method.speed_of_sound;
    sa.out_tait = parameter_a.method.tait_exponent;
    sa.out_btension = parameter_a.method.boundary_surface_tension;
    wcsph_vjp_scatter_kernel<<<blocks_for(n, kBlock), kBlock, 0, stream_>>>(sa);
    PHX_LAUNCH_CHECK();
}

}  // namespace phx
