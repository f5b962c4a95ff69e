// Derivatives of one recorded DFSPH step for gfx950, in sorted order.
//
// Reference: SPH<DFSPH>::jvp_step / vjp_step (sph/sph.ixx:356-365) around
// DFSPH::jvp (sph/sph.cpp:612-659) and DFSPH::vjp (sph/sph.cpp:661-723),
// with the derivative kernels of sph/common.cu (non_pressure_*, pressure_*),
// sph/dfsph.cu (predict_*, projection_update_*) and density/density.cu. The
// neighbour graph, the box-collision bits of every prediction and the active
// set of every max(0, .) pressure update are those of the recorded forward
// step. Tangents are FP32 like the reference's; adjoints are FP64 with the
// FP32 primal values promoted in the reference's operand order. The
// reference recomputes pressure iterations from checkpoints; here the
// forward step recorded every iterate in the step cache.
//
// Forward-mode passes (the forward step's pass structure):
//   density    tangent of the state densities and (p / rho^2)' of both
//              initial pressures
//   force      non-pressure tangent (first pass only) + pressure tangent of
//              one or two pressure fields + prediction tangent
//   projection tangent of the predicted densities + pressure update tangent
// Reverse-mode passes (gather form: particle i collects the pair terms in
// which it is the centre and those in which it is the neighbour, so no output
// is written by two particles and no atomics are needed):
//   final       prediction VJP of the next state; clears the accumulators
//   per iteration, last to first, density phase then divergence phase:
//     pressure  pressure VJP + pressure update VJP
//     density   predicted-density VJP + prediction VJP
//   pressure    pressure VJP of the initial density-phase pressures
//   state       pressure VJP of the initial divergence-phase pressures +
//               non-pressure VJP
//   scatter     state-density VJP, then every adjoint accumulated into the
//               caller's original-order arrays
//
// The prediction VJP is the exact transpose of the prediction JVP: under a
// no-slip collision the positions of the components that did not collide
// keep their dt * xbar term in the velocity and acceleration adjoints. The
// reference's predict_vjp_kernel (sph/dfsph.cu:81) zeroes it; everywhere
// else the two agree.
#include <cmath>

#include "adjoint_math.hpp"
#include "dfsph.hpp"
#include "neighbor_mask.hpp"

namespace phx {
namespace {

constexpr uint32_t kBlock = 256;

#define PHX_MASK_LDS(grid)                          \
    __shared__ uint16_t smask[kStencil * kBlock];   \
    __shared__ int delta[kStencil];                 \
    fill_slot_deltas(grid, delta);                  \
    __syncthreads()

// The recorded step in sorted order.
struct Step {
    uint32_t n;
    KernelConstants kc;
    float dt;
    float gradient_norm;
    double pi_h6;  // double(pi) * pow(double(h), 6)
    Box box;
    uint32_t exact_no_slip;  // 1: exact transpose of a no-slip collision, 0: the reference's VJP
    SortedGrid grid;
    const float4* xm;
    const float4* vmu;
    const float* rho;
    const uint16_t* exact;
    const uint16_t* skin;
    const uint32_t* flags;
    float limit2;
    BoundaryGrid boundary;   // offsets null: no boundary samples
    const float* rest;       // rho0 (the boundary terms)
    const uint8_t* crowded;  // per particle: walks the full stencil
};

// Whether particle i walks the step's exact masks (not crowded, no boundary samples).
__device__ __forceinline__ bool exact_masked(const Step& s, uint32_t i) {
    return s.flags[kFlagOverflow] == 0u && s.crowded[i] == 0u;
}
__device__ __forceinline__ MaskView exact_view(const Step& s) {
    return MaskView{s.exact, s.flags, s.n, s.limit2, 0u, s.crowded};
}
__device__ __forceinline__ MaskView skin_view(const Step& s) {
    return MaskView{s.skin, s.flags, s.n, s.limit2, 0u, s.crowded};
}

// Visits the fluid candidates of particle i (through the masks where they
// apply) or, with boundary samples, every fluid particle and every sample of
// the 27 cells in the reference order: each cell's fluid, then its samples.
template <class Visit, class VisitBoundary>
__device__ __forceinline__ void walk(const Step& s, const MaskView& mv, bool masked, uint32_t i, int cx, int cy,
                                     int cz, const int* delta, uint16_t* column, Visit&& visit,
                                     VisitBoundary&& visit_boundary) {
    if (s.boundary.offsets != nullptr) {
        for_each_candidate(s.grid, s.boundary, cx, cy, cz, visit, visit_boundary);
    } else {
        visit_candidates(s.grid, mv, masked, i, cx, cy, cz, delta, column, kBlock, visit);
    }
}

// ----------------------------------------------------------------- JVP
struct GatherTangentArgs {
    uint32_t n;
    const uint32_t* perm;
    CVec x_t, v_t, u_t;
    const float *m_t, *rest_t, *visc_t, *relax_div_t, *relax_den_t, *warm_div_t, *warm_den_t;  // null: zero
    float4* xt;
    float4* vt;
    float4* ut;
    float* rest_sorted;
    float* relax_div_sorted;
    float* relax_den_sorted;
    float* p_t_div;
    float* p_t_den;
};


__global__ void __launch_bounds__(kBlock) dfsph_gather_tangent_kernel(GatherTangentArgs a) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t j = a.perm[i];
    const float rest_t = or_zero(a.rest_t, j);
    a.xt[i] = make_float4(a.x_t.x[j], a.x_t.y[j], a.x_t.z[j], or_zero(a.m_t, j));
    a.vt[i] = make_float4(a.v_t.x[j], a.v_t.y[j], a.v_t.z[j], or_zero(a.visc_t, j));
    a.ut[i] = make_float4(a.u_t.x[j], a.u_t.y[j], a.u_t.z[j], rest_t);
    a.rest_sorted[i] = rest_t;
    a.relax_div_sorted[i] = or_zero(a.relax_div_t, j);
    a.relax_den_sorted[i] = or_zero(a.relax_den_t, j);
    a.p_t_div[i] = or_zero(a.warm_div_t, j);
    a.p_t_den[i] = or_zero(a.warm_den_t, j);
}

struct DensityJvpArgs {
    Step s;
    const float4* xt;
    const float* p_div;  // initial pressures of both phases
    const float* p_den;
    const float* p_t_div;
    const float* p_t_den;
    float* rho_t;
    float2* t_div;
    float2* t_den;
    const float* rest_t;  // rest-density tangents (the boundary terms)
};

// density_jvp_kernel (density.cu:42-67, topology = positions = x) and the
// first terms of the two initial pressure fields.
__global__ void __launch_bounds__(kBlock) dfsph_jvp_density_kernel(DensityJvpArgs a) {
    const Step& s = a.s;
    PHX_MASK_LDS(s.grid);
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= s.n) return;
    const Float3 p = xyz(s.xm[i]);
    const Float3 pt = xyz(a.xt[i]);
    int cx, cy, cz;
    particle_cell(s.grid, p, cx, cy, cz);
    float result = 0.0F;
    const float rest_i = s.rest[i];
    const float rest_t_i = a.rest_t[i];
    walk(
        s, exact_view(s), exact_masked(s, i), i, cx, cy, cz, delta, smask + threadIdx.x,
        [&](uint32_t j) {
            const float4 q = s.xm[j];
            const float4 qt = a.xt[j];
            const Float3 d = sub(p, xyz(q));
            const float dist = length(d);
            result += qt.w * cubic_value(dist, s.kc) +
                      q.w * dot(scale(d, cubic_gradient_factor(dist, s.kc)), sub(pt, xyz(qt)));
        },
        [&](uint32_t b) {
            const float4 q = s.boundary.x[b];
            const Float3 d = sub(p, xyz(q));
            const float dist = length(d);
            result += q.w * (rest_t_i * cubic_value(dist, s.kc) +
                             rest_i * dot(scale(d, cubic_gradient_factor(dist, s.kc)), pt));
        });
    a.rho_t[i] = result;
    const float rho = s.rho[i];
    a.t_div[i] = first_term(a.p_div[i], a.p_t_div[i], rho, result);
    a.t_den[i] = first_term(a.p_den[i], a.p_t_den[i], rho, result);
}

enum JvpEpilogue : int {
    kJvpDivPredict = 0,  // total = a_p'(t1)
    kJvpDivEnd = 1,      // a_div' = a_p'(t1); total = a_div' + a_p'(t2)
    kJvpDenPredict = 2,  // total = a_div' + a_p'(t1)
};

struct ForceJvpArgs {
    Step s;
    const uint8_t* collisions;  // bits of the prediction this pass differentiates
    const float4* xt;
    const float4* vt;
    const float4* ut;
    const float* rho_t;
    const float2* t1;
    const float2* t2;
    float4* anp_t;
    float4* adiv_t;
    float4* xs_t;
    uint32_t final_step;
    const uint32_t* perm;
    Vec next_x_t, next_v_t;
    const float* p_t_div;
    float* warm_div_t_out;
    const float* p_t_den;
    float* warm_den_t_out;
};

// non_pressure_jvp_kernel (common.cu:49-96, first pass only) +
// pressure_jvp_kernel (common.cu:196-237) for one or two pressure fields +
// predict_jvp_kernel (dfsph.cu:46-66) of the next prediction or the next
// state.
template <bool kNonPressure, int kEpilogue>
__global__ void __launch_bounds__(kBlock) dfsph_jvp_force_kernel(ForceJvpArgs a) {
    constexpr bool kTwo = kEpilogue == kJvpDivEnd;
    const Step& s = a.s;
    PHX_MASK_LDS(s.grid);
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= s.n) return;
    const float4 self = s.xm[i];
    const Float3 x = xyz(self);
    const float4 xt_i = a.xt[i];
    const Float3 xti = xyz(xt_i);
    const float4 vm = s.vmu[i];
    const float4 vt_i = a.vt[i];
    Float3 acc_t{0.0F, 0.0F, 0.0F};
    if constexpr (kNonPressure) acc_t = xyz(a.ut[i]);
    const float2 t1_i = a.t1[i];
    float2 t2_i = make_float2(0.0F, 0.0F);
    if constexpr (kTwo) t2_i = a.t2[i];
    Float3 ap1{0.0F, 0.0F, 0.0F};
    Float3 ap2{0.0F, 0.0F, 0.0F};
    const float rest_i = s.rest[i];
    const float rest_t_i = a.ut[i].w;
    int cx, cy, cz;
    particle_cell(s.grid, x, cx, cy, cz);
    // the pressure terms of sample b (pressure_jvp_kernel's boundary half)
    auto boundary_pressure = [&](const Float3& gradient, const Float3& gradient_t, float vol, float2 t, Float3& ap) {
        const float factor = -rest_i * vol * t.x;
        const float factor_t = -vol * (rest_t_i * t.x + rest_i * t.y);
        ap = add(ap, add(scale(gradient, factor_t), scale(gradient_t, factor)));
    };
    walk(s, exact_view(s), exact_masked(s, i), i, cx, cy, cz, delta, smask + threadIdx.x,
                     [&](uint32_t j) {
                         if (j == i) return;
                         const float4 q = s.xm[j];
                         const float4 qt = a.xt[j];
                         const float mass = q.w;
                         const Float3 d = sub(x, xyz(q));
                         const Float3 dd = sub(xti, xyz(qt));
                         if constexpr (kNonPressure) {
                             const float4 vq = s.vmu[j];
                             const float4 vqt = a.vt[j];
                             const float lap = viscosity_laplacian_c(d, s.kc);
                             const float lap_t = viscosity_laplacian_tangent(d, dd, s.kc);
                             const float nu = 0.5F * (vm.w + vq.w);
                             const float nu_t = 0.5F * (vt_i.w + vqt.w);
                             const float dens = s.rho[j];
                             const float w = nu * mass * lap / dens;
                             const float w_t = nu_t * mass * lap / dens + nu * qt.w * lap / dens +
                                               nu * mass * lap_t / dens - nu * mass * lap * a.rho_t[j] / (dens * dens);
                             const Float3 diff = sub(xyz(vq), xyz(vm));
                             const Float3 diff_t = sub(xyz(vqt), xyz(vt_i));
                             acc_t = add(acc_t, add(scale(diff, w_t), scale(diff_t, w)));
                         }
                         const Float3 gradient = scale(d, cubic_gradient_factor(length(d), s.kc));
                         const Float3 gradient_t = cubic_gradient_tangent(d, dd, s.kc);
                         const float2 t1_j = a.t1[j];
                         const float factor = -mass * (t1_i.x + t1_j.x);
                         const float factor_t = -qt.w * (t1_i.x + t1_j.x) - mass * (t1_i.y + t1_j.y);
                         ap1 = add(ap1, add(scale(gradient, factor_t), scale(gradient_t, factor)));
                         if constexpr (kTwo) {
                             const float2 t2_j = a.t2[j];
                             const float f2 = -mass * (t2_i.x + t2_j.x);
                             const float f2_t = -qt.w * (t2_i.x + t2_j.x) - mass * (t2_i.y + t2_j.y);
                             ap2 = add(ap2, add(scale(gradient, f2_t), scale(gradient_t, f2)));
                         }
                     },
                     [&](uint32_t b) {
                         const float4 q = s.boundary.x[b];
                         const float vol = q.w;
                         const Float3 d = sub(x, xyz(q));
                         if constexpr (kNonPressure) {
                             // non_pressure_jvp_kernel's boundary half: mu_i V lap (v_b - v_i)
                             const float lap = viscosity_laplacian_c(d, s.kc);
                             const float w = vm.w * vol * lap;
                             const float w_t = vol * (vt_i.w * lap + vm.w * viscosity_laplacian_tangent(d, xti, s.kc));
                             const Float3 diff = sub(xyz(s.boundary.v[b]), xyz(vm));
                             acc_t = add(acc_t, add(scale(diff, w_t), scale(xyz(vt_i), -w)));
                         }
                         const Float3 gradient = scale(d, cubic_gradient_factor(length(d), s.kc));
                         const Float3 gradient_t = cubic_gradient_tangent(d, xti, s.kc);
                         boundary_pressure(gradient, gradient_t, vol, t1_i, ap1);
                         if constexpr (kTwo) boundary_pressure(gradient, gradient_t, vol, t2_i, ap2);
                     });
    if constexpr (kNonPressure) {
        a.anp_t[i] = make_float4(acc_t.x, acc_t.y, acc_t.z, 0.0F);
    } else {
        acc_t = xyz(a.anp_t[i]);
    }
    Float3 total;
    if constexpr (kEpilogue == kJvpDivPredict) {
        total = ap1;
    } else if constexpr (kTwo) {
        a.adiv_t[i] = make_float4(ap1.x, ap1.y, ap1.z, 0.0F);
        total = add(ap1, ap2);
    } else {
        total = add(xyz(a.adiv_t[i]), ap1);
    }
    Float3 nx, nv;
    predict_tangent(a.collisions[i], s.box.no_slip, s.dt, xti, xyz(vt_i), add(acc_t, total), nx, nv);
    if constexpr (kTwo) {
        if (a.warm_div_t_out != nullptr) a.warm_div_t_out[a.perm[i]] = a.p_t_div[i];
    }
    if (a.final_step != 0u) {
        const uint32_t j = a.perm[i];
        store(a.next_x_t, j, nx);
        store(a.next_v_t, j, nv);
        if (a.warm_den_t_out != nullptr) a.warm_den_t_out[j] = a.p_t_den[i];
    } else {
        a.xs_t[i] = make_float4(nx.x, nx.y, nx.z, 0.0F);
    }
}

struct ProjectionJvpArgs {
    Step s;
    uint32_t slot;     // displacement slot of the predicted positions
    const float4* xs;  // recorded predicted positions (+ masses)
    const float* rs;   // recorded predicted densities
    const float* p_prev;
    const float* p;
    const float* target;
    const float* target_t;
    const float* relax;
    const float* relax_t;
    const float4* xt;  // .w: mass tangents
    const float4* xs_t;
    const float* rho_t;
    const float* volume;    // density of V = m / density: rho, or rho0 (PCISPH)
    const float* volume_t;  // its tangent
    float* p_t;  // in: tangent of p_prev; out: tangent of p
    float2* t_out;
    const float* rest_t;  // rest-density tangents (the boundary terms)
};

// density_jvp_kernel at the predicted positions (topology x) +
// projection_update_jvp_kernel (dfsph.cu:112-121).
__global__ void __launch_bounds__(kBlock) dfsph_jvp_projection_kernel(ProjectionJvpArgs a) {
    const Step& s = a.s;
    PHX_MASK_LDS(s.grid);
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= s.n) return;
    const float4 self = a.xs[i];
    const Float3 ps = xyz(self);
    const Float3 pst = xyz(a.xs_t[i]);
    int cx, cy, cz;
    particle_cell(s.grid, xyz(s.xm[i]), cx, cy, cz);
    float rs_t = 0.0F;
    const float rest_i = s.rest[i];
    const float rest_t_i = a.rest_t[i];
    const MaskView mv = skin_view(s);
    walk(
        s, mv, masks_apply(mv, a.slot, i), i, cx, cy, cz, delta, smask + threadIdx.x,
        [&](uint32_t j) {
            const float4 q = a.xs[j];
            const Float3 d = sub(ps, xyz(q));
            const float dist = length(d);
            rs_t += a.xt[j].w * cubic_value(dist, s.kc) +
                    q.w * dot(scale(d, cubic_gradient_factor(dist, s.kc)), sub(pst, xyz(a.xs_t[j])));
        },
        [&](uint32_t b) {
            const float4 q = s.boundary.x[b];
            const Float3 d = sub(ps, xyz(q));
            const float dist = length(d);
            rs_t += q.w * (rest_t_i * cubic_value(dist, s.kc) +
                           rest_i * dot(scale(d, cubic_gradient_factor(dist, s.kc)), pst));
        });
    const float m = self.w;
    const float m_t = a.xt[i].w;
    const float rho = s.rho[i];
    const float rho_t = a.rho_t[i];
    const float error = a.rs[i] - a.target[i];
    const float density = a.volume[i];
    const float volume = m / density;
    const float dp = 1.0F / (2.0F * s.dt * s.dt * volume * volume * s.gradient_norm);
    const float dp_t = -2.0F * dp * (m_t / m - a.volume_t[i] / density);
    const float relax = a.relax[i];
    const float candidate = a.p_prev[i] + relax * dp * error;
    const float pt = candidate > 0.0F ? a.p_t[i] + a.relax_t[i] * dp * error + relax * dp_t * error +
                                            relax * dp * (rs_t - a.target_t[i])
                                      : 0.0F;
    a.p_t[i] = pt;
    a.t_out[i] = first_term(a.p[i], pt, rho, rho_t);
}

// ----------------------------------------------------------------- VJP
struct FinalVjpArgs {
    Step s;
    const uint8_t* collisions;  // the last prediction
    const uint32_t* perm;
    phx_const_adjoint_vector_view next_x_a, next_v_a;
    const double* next_warm_div_a;  // null: zero
    const double* next_warm_den_a;
    DVec pos_a, vel_a, anp_a, adiv_a, cur_ap;
    double* cur_p;            // next warm density-pressure adjoint, sorted
    double* warm_div_sorted;  // next warm divergence-pressure adjoint, sorted
    double *rho_a, *m_a, *rest_a, *visc_a, *relax_div_a, *relax_den_a;
};

// predict_vjp_kernel (dfsph.cu:68-102) of the next state (sph.cpp:671); the
// total pressure acceleration adjoint seeds both the divergence result and
// the density phase (sph.cpp:672-674). Starts every sorted accumulator.
__global__ void __launch_bounds__(kBlock) dfsph_vjp_final_kernel(FinalVjpArgs a) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.s.n) return;
    const uint32_t j = a.perm[i];
    const D3 xa{a.next_x_a.x[j], a.next_x_a.y[j], a.next_x_a.z[j]};
    const D3 va{a.next_v_a.x[j], a.next_v_a.y[j], a.next_v_a.z[j]};
    D3 position, velocity;
    predict_adjoint(a.collisions[i], a.s.box.no_slip, a.s.exact_no_slip, static_cast<double>(a.s.dt), xa, va, position,
                    velocity);
    const D3 acc = scale(velocity, static_cast<double>(a.s.dt));
    store(a.pos_a, i, position);
    store(a.vel_a, i, velocity);
    store(a.anp_a, i, acc);
    store(a.adiv_a, i, acc);
    store(a.cur_ap, i, acc);
    a.cur_p[i] = a.next_warm_den_a != nullptr ? a.next_warm_den_a[j] : 0.0;
    a.warm_div_sorted[i] = a.next_warm_div_a != nullptr ? a.next_warm_div_a[j] : 0.0;
    a.rho_a[i] = 0.0;
    a.m_a[i] = 0.0;
    a.rest_a[i] = 0.0;
    a.visc_a[i] = 0.0;
    a.relax_div_a[i] = 0.0;
    a.relax_den_a[i] = 0.0;
}

struct PressureVjpArgs {
    Step s;
    const float* p;  // the pressures whose accelerations are differentiated
    DVec ap_a;       // adjoint of those accelerations (complete)
    double* p_a;     // adjoint of p: accumulated here, then consumed
    // pressure update that produced p (kUpdate)
    const float* p_prev;
    const float* rs;
    const float* target;
    const float* relax;
    double* p_prev_a;  // written
    double* rs_a;      // written
    double* target_a;  // accumulated: rest densities (density phase); null: the
                       // target is the state density (divergence phase)
    const float* volume;  // density of V = m / density; null: the state density
    double* volume_a;     // accumulated when volume is set (rest densities, PCISPH)
    double* relax_a;
    // warm-start adjoint of an initial pressure field (!kUpdate), copied into
    // original order; null: dropped
    const uint32_t* perm;
    double* warm_out;
    // non-pressure VJP (kNonPressure)
    DVec anp_a;
    DVec vel_a;
    double* visc_a;
    DVec pos_a;
    double* rho_a;
    double* m_a;
    double* rest_a;  // accumulated: the boundary terms' rest-density adjoint
};

// pressure_vjp_kernel (common.cu:239-294) + projection_update_vjp_kernel
// (dfsph.cu:123-139), or + non_pressure_vjp_kernel (common.cu:98-163).
template <bool kUpdate, bool kNonPressure>
__global__ void __launch_bounds__(kBlock) dfsph_vjp_pressure_kernel(PressureVjpArgs a) {
    const Step& s = a.s;
    PHX_MASK_LDS(s.grid);
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= s.n) return;
    const float h = s.kc.h;
    const float4 self = s.xm[i];
    const Float3 x = xyz(self);
    const float m_i = self.w;
    const float rho_i = s.rho[i];
    const D3 local = load(a.ap_a, i);
    const double first = static_cast<double>(a.p[i]) / (rho_i * rho_i);
    D3 position{0.0, 0.0, 0.0};
    double first_a = 0.0;
    double mass_c = 0.0;
    // non-pressure terms
    const float4 vm = s.vmu[i];
    const Float3 vel = xyz(vm);
    D3 lnp{0.0, 0.0, 0.0};
    if constexpr (kNonPressure) lnp = load(a.anp_a, i);
    D3 velocity{0.0, 0.0, 0.0};
    double dens_c = 0.0;
    double visc_c = 0.0;
    double rest_c = 0.0;
    const float rest_i = s.rest[i];
    int cx, cy, cz;
    particle_cell(s.grid, x, cx, cy, cz);
    walk(s, exact_view(s), exact_masked(s, i), i, cx, cy, cz, delta, smask + threadIdx.x,
                     [&](uint32_t j) {
                         if (j == i) return;
                         const float4 q = s.xm[j];
                         const Float3 d = sub(x, xyz(q));
                         const float mass = q.w;
                         if constexpr (kNonPressure) {
                             const float4 vq = s.vmu[j];
                             const Float3 diff = sub(xyz(vq), vel);
                             const float lap = viscosity_laplacian_c(d, s.kc);
                             const float nu = 0.5F * (vm.w + vq.w);
                             const float dens = s.rho[j];
                             const float w = nu * mass * lap / dens;
                             const double w_a = dot(lnp, d3(diff));
                             velocity = add(velocity, scale(lnp, static_cast<double>(-w)));
                             const double nu_a = w_a * mass * lap / dens;
                             visc_c += 0.5 * nu_a;
                             position = add(position, viscosity_laplacian_gradient(d, h, s.pi_h6, w_a * nu * mass / dens));
                             const D3 cross = load(a.anp_a, j);
                             const float cross_w = nu * m_i * lap / rho_i;
                             const double cross_w_a = dot(cross, d3(scale(diff, -1.0F)));
                             velocity = add(velocity, scale(cross, static_cast<double>(cross_w)));
                             mass_c += cross_w_a * nu * lap / rho_i;
                             dens_c -= cross_w_a * nu * m_i * lap / (rho_i * rho_i);
                             visc_c += 0.5 * cross_w_a * m_i * lap / rho_i;
                             position = add(position,
                                            viscosity_laplacian_gradient(d, h, s.pi_h6, cross_w_a * nu * m_i / rho_i));
                         }
                         const D3 gradient = d3(scale(d, cubic_gradient_factor(length(d), s.kc)));
                         const double second = static_cast<double>(a.p[j]) / (s.rho[j] * s.rho[j]);
                         const double sum = first + second;
                         const double local_factor = -mass * sum;
                         first_a -= mass * dot(local, gradient);
                         position = add(position, cubic_hessian_product(d, scale(local, local_factor), h));
                         const D3 cross = load(a.ap_a, j);
                         const double projection = dot(cross, gradient);
                         mass_c += sum * projection;
                         first_a += m_i * projection;
                         position = add(position, cubic_hessian_product(d, scale(cross, m_i * sum), h));
                     },
                     [&](uint32_t b) {
                         const float4 q = s.boundary.x[b];
                         const float vol = q.w;
                         const Float3 d = sub(x, xyz(q));
                         if constexpr (kNonPressure) {
                             // non_pressure_vjp_kernel's boundary half
                             const Float3 diff = sub(xyz(s.boundary.v[b]), vel);
                             const float lap = viscosity_laplacian_c(d, s.kc);
                             const float w = vm.w * vol * lap;
                             const double w_a = dot(lnp, d3(diff));
                             velocity = add(velocity, scale(lnp, static_cast<double>(-w)));
                             visc_c += w_a * vol * lap;
                             position = add(position, viscosity_laplacian_gradient(d, h, s.pi_h6, w_a * vm.w * vol));
                         }
                         // pressure_vjp_kernel's boundary half: -rho0_i V (p_i / rho_i^2) grad W
                         const D3 gradient = d3(scale(d, cubic_gradient_factor(length(d), s.kc)));
                         const double factor = static_cast<double>(-rest_i * vol) * first;
                         const double factor_a = dot(local, gradient);
                         rest_c -= factor_a * vol * first;
                         first_a -= factor_a * rest_i * vol;
                         position = add(position, cubic_hessian_product(d, scale(local, factor), h));
                     });
    if (s.boundary.offsets != nullptr) a.rest_a[i] += rest_c;
    accumulate(a.pos_a, i, position);
    double m_a = a.m_a[i] + mass_c;
    double rho_a = a.rho_a[i] - 2.0 * first_a * a.p[i] / (rho_i * rho_i * rho_i);
    const double p_a = a.p_a[i] + first_a / (rho_i * rho_i);
    if constexpr (kNonPressure) {
        accumulate(a.vel_a, i, velocity);
        rho_a += dens_c;
        a.visc_a[i] += visc_c;
    }
    if constexpr (kUpdate) {
        const double error = static_cast<double>(a.rs[i]) - a.target[i];
        const float density = a.volume != nullptr ? a.volume[i] : rho_i;
        const double volume = static_cast<double>(m_i) / density;
        const double dp = 1.0 / (2.0 * s.dt * s.dt * volume * volume * s.gradient_norm);
        const float relax = a.relax[i];
        const double candidate = a.p_prev[i] + relax * dp * error;
        if (candidate > 0.0) {
            a.p_prev_a[i] = p_a;
            a.rs_a[i] = relax * dp * p_a;
            if (a.target_a != nullptr) {
                a.target_a[i] -= relax * dp * p_a;
            } else {
                rho_a -= relax * dp * p_a;
            }
            m_a -= 2.0 * relax * dp * error * p_a / m_i;
            if (a.volume_a != nullptr) {
                a.volume_a[i] += 2.0 * relax * dp * error * p_a / density;
            } else {
                rho_a += 2.0 * relax * dp * error * p_a / rho_i;
            }
            a.relax_a[i] += dp * error * p_a;
        } else {
            a.p_prev_a[i] = 0.0;
            a.rs_a[i] = 0.0;
        }
    } else {
        if (a.warm_out != nullptr) a.warm_out[a.perm[i]] = p_a;
    }
    a.m_a[i] = m_a;
    a.rho_a[i] = rho_a;
}

struct DensityVjpArgs {
    Step s;
    uint32_t slot;  // displacement slot of the predicted positions
    const float4* xs;
    const double* rs_a;
    const uint8_t* collisions;  // the prediction that produced xs
    DVec pos_a, vel_a, anp_a;
    DVec prev_ap;  // written: adjoint of the previous iterate's pressure accelerations
    DVec adiv_a;   // accumulated in the density phase, unused in the divergence phase
    double* m_a;
    double* rest_a;  // accumulated: the boundary terms' rest-density adjoint
};

// density_vjp_kernel at the predicted positions (density.cu:69-100, topology
// x) + predict_vjp_kernel of the prediction that produced them; in the
// density phase the prediction read a_div + a_p, and its adjoint goes to
// both (add_adjoint_kernel, common.cu:376-388).
template <bool kDensityPhase>
__global__ void __launch_bounds__(kBlock) dfsph_vjp_density_kernel(DensityVjpArgs a) {
    const Step& s = a.s;
    PHX_MASK_LDS(s.grid);
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= s.n) return;
    const float4 self = a.xs[i];
    const Float3 ps = xyz(self);
    const double out_a = a.rs_a[i];
    const float m_i = self.w;
    int cx, cy, cz;
    particle_cell(s.grid, xyz(s.xm[i]), cx, cy, cz);
    D3 position{0.0, 0.0, 0.0};
    double mass_c = 0.0;
    double rest_c = 0.0;
    const float rest_i = s.rest[i];
    const MaskView mv = skin_view(s);
    walk(
        s, mv, masks_apply(mv, a.slot, i), i, cx, cy, cz, delta, smask + threadIdx.x,
        [&](uint32_t j) {
            const float4 q = a.xs[j];
            const Float3 d = sub(ps, xyz(q));
            const float dist = length(d);
            const double rs_a_j = a.rs_a[j];
            position = add(position, scale(d3(scale(d, cubic_gradient_factor(dist, s.kc))),
                                           out_a * q.w + rs_a_j * m_i));
            mass_c += rs_a_j * cubic_value(dist, s.kc);
        },
        [&](uint32_t b) {
            const float4 q = s.boundary.x[b];
            const Float3 d = sub(ps, xyz(q));
            const float dist = length(d);
            position = add(position, scale(d3(scale(d, cubic_gradient_factor(dist, s.kc))), out_a * rest_i * q.w));
            rest_c += out_a * q.w * cubic_value(dist, s.kc);
        });
    a.m_a[i] += mass_c;
    if (s.boundary.offsets != nullptr) a.rest_a[i] += rest_c;
    D3 pos, velocity;
    predict_adjoint(a.collisions[i], s.box.no_slip, s.exact_no_slip, static_cast<double>(s.dt), position,
                    D3{0.0, 0.0, 0.0}, pos, velocity);
    const D3 acc = scale(velocity, static_cast<double>(s.dt));
    accumulate(a.pos_a, i, pos);
    accumulate(a.vel_a, i, velocity);
    accumulate(a.anp_a, i, acc);
    store(a.prev_ap, i, acc);
    if constexpr (kDensityPhase) accumulate(a.adiv_a, i, acc);
}

struct ScatterVjpArgs {
    Step s;
    const uint32_t* perm;
    const double* rho_a;
    DVec pos_a, vel_a, anp_a;
    const double *m_a, *rest_a, *visc_a, *relax_div_a, *relax_den_a;
    DVec out_pos, out_vel, out_ctl;  // null components: skipped
    double *out_m, *out_rest, *out_visc, *out_relax_div, *out_relax_den;
};

// density_vjp_kernel of the state densities (sph.ixx:364, topology =
// positions = x), then every sorted adjoint accumulated into the caller's
// original-order arrays.
__global__ void __launch_bounds__(kBlock) dfsph_vjp_scatter_kernel(ScatterVjpArgs a) {
    const Step& s = a.s;
    PHX_MASK_LDS(s.grid);
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= s.n) return;
    const float4 self = s.xm[i];
    const Float3 x = xyz(self);
    const double out_a = a.rho_a[i];
    const float m_i = self.w;
    int cx, cy, cz;
    particle_cell(s.grid, x, cx, cy, cz);
    D3 position{0.0, 0.0, 0.0};
    double mass_c = 0.0;
    double rest_c = 0.0;
    const float rest_i = s.rest[i];
    walk(
        s, exact_view(s), exact_masked(s, i), i, cx, cy, cz, delta, smask + threadIdx.x,
        [&](uint32_t j) {
            const float4 q = s.xm[j];
            const Float3 d = sub(x, xyz(q));
            const float dist = length(d);
            const double rho_a_j = a.rho_a[j];
            position = add(position, scale(d3(scale(d, cubic_gradient_factor(dist, s.kc))),
                                           out_a * q.w + rho_a_j * m_i));
            mass_c += rho_a_j * cubic_value(dist, s.kc);
        },
        [&](uint32_t b) {
            const float4 q = s.boundary.x[b];
            const Float3 d = sub(x, xyz(q));
            const float dist = length(d);
            position = add(position, scale(d3(scale(d, cubic_gradient_factor(dist, s.kc))), out_a * rest_i * q.w));
            rest_c += out_a * q.w * cubic_value(dist, s.kc);
        });
    const uint32_t j = a.perm[i];
    const D3 pos = add(load(a.pos_a, i), position);
    if (a.out_pos.x != nullptr) accumulate(a.out_pos, j, pos);
    if (a.out_vel.x != nullptr) accumulate(a.out_vel, j, load(a.vel_a, i));
    if (a.out_ctl.x != nullptr) accumulate(a.out_ctl, j, load(a.anp_a, i));
    accumulate_to(a.out_m, j, a.m_a[i] + mass_c);
    accumulate_to(a.out_rest, j, a.rest_a[i] + rest_c);
    accumulate_to(a.out_visc, j, a.visc_a[i]);
    accumulate_to(a.out_relax_div, j, a.relax_div_a[i]);
    accumulate_to(a.out_relax_den, j, a.relax_den_a[i]);
}

uint32_t derivative_blocks(uint32_t n) { return (blocks_for(n, kBlock) + 7u) / 8u * 8u; }

}  // namespace

void DfsphDerivativeWorkspace::allocate_tangents(uint32_t n) {
    for (DeviceBuffer<float4>* b : {&xt, &vt, &ut, &anp_t, &adiv_t, &xs_t}) b->resize(n);
    for (DeviceBuffer<float2>* b : {&t_div, &t_den}) b->resize(n);
    for (DeviceBuffer<float>* b : {&rho_t, &p_t_div, &p_t_den, &rest_t, &relax_div_t, &relax_den_t}) b->resize(n);
}

void DfsphDerivativeWorkspace::allocate_adjoints(uint32_t n) {
    for (DVecBuffer* b : {&pos_a, &vel_a, &anp_a, &ap_a[0], &ap_a[1], &ap_a[2]}) b->resize(n);
    for (DeviceBuffer<double>* b : {&p_a[0], &p_a[1], &p_a[2], &rs_a, &rho_a, &m_a, &rest_a, &visc_a,
                                    &relax_div_a, &relax_den_a})
        b->resize(n);
}

void DfsphSolver::require_recorded(const DfsphCache& cache) const {
    require(mode_ == PHX_MODE_DIFFERENTIABLE, PHX_ERROR_UNSUPPORTED,
            "derivatives need a solver created in differentiable mode");
    require(cache.recorded, PHX_ERROR_INVALID_ARGUMENT, "the step cache holds no recorded step");
    require(cache.neighborhood.particles == domain_.config.particle_count, PHX_ERROR_INVALID_ARGUMENT,
            "the step cache belongs to another particle count");
}

namespace {

Step make_step(const Domain& domain, const DfsphCache& cache, const KernelConstants& kc, float gradient_norm,
               bool exact_no_slip) {
    const auto& c = domain.config;
    Step s{};
    s.n = c.particle_count;
    s.kc = kc;
    s.dt = c.time_step;
    s.gradient_norm = gradient_norm;
    s.pi_h6 = static_cast<double>(kPi) * std::pow(static_cast<double>(c.support_radius), 6.0);
    s.box = collision_box(c, cache.step_index);
    s.exact_no_slip = exact_no_slip ? 1u : 0u;
    s.grid = cache.neighborhood.grid();
    s.xm = cache.xm.data();
    s.vmu = cache.vmu.data();
    s.rho = cache.densities.data();
    s.exact = cache.exact.data();
    s.skin = cache.skin.data();
    s.flags = cache.mask_flags.data();
    s.limit2 = cache.mask_limit2;
    s.boundary = cache.neighborhood.boundary();
    s.rest = cache.rho0.data();
    s.crowded = cache.crowded.data();
    return s;
}

template <bool kNonPressure, int kEpilogue>
void launch_force_jvp(hipStream_t stream, uint32_t blocks, const ForceJvpArgs& args) {
    dfsph_jvp_force_kernel<kNonPressure, kEpilogue><<<blocks, kBlock, 0, stream>>>(args);
    PHX_LAUNCH_CHECK();
}

}  // namespace

void DfsphSolver::jvp_step(const DfsphCache& cache, const phx_sph_dfsph_state_tangent& state_t,
                           const phx_control_tangent& control_t, const phx_sph_dfsph_parameter_tangent& parameter_t,
                           phx_sph_dfsph_state_tangent& next_t) {
    const uint32_t n = domain_.config.particle_count;
    if (n == 0) return;
    require_recorded(cache);
    require_vec(state_t.particles.positions, "position tangent");
    require_vec(state_t.particles.velocities, "velocity tangent");
    require_vec(control_t.external_accelerations, "control tangent");
    require_vec(next_t.particles.positions, "next position tangent");
    require_vec(next_t.particles.velocities, "next velocity tangent");
    PassScope scope(profiler_, stream_, kDfsphJvp);
    dw_.allocate_tangents(n);
    const uint32_t nd = cache.iterations[0];
    const uint32_t ne = cache.iterations[1];
    const bool warm = cache.warm_start;
    const uint32_t* perm = cache.neighborhood.perm.data();
    const uint32_t blocks = derivative_blocks(n);
    const Step s = make_step(domain_, cache, constants_, gradient_norm_, config_.exact_no_slip_adjoint != 0u);

    GatherTangentArgs ga{};
    ga.n = n;
    ga.perm = perm;
    ga.x_t = cvec(state_t.particles.positions);
    ga.v_t = cvec(state_t.particles.velocities);
    ga.u_t = cvec(control_t.external_accelerations);
    ga.m_t = parameter_t.particles.masses;
    ga.rest_t = parameter_t.particles.rest_densities;
    ga.visc_t = parameter_t.particles.viscosities;
    ga.relax_div_t = parameter_t.method.divergence_relaxation;
    ga.relax_den_t = parameter_t.method.density_relaxation;
    // without warm start the solves start from cleared pressures (sph.cpp:616-621)
    ga.warm_div_t = warm ? state_t.method.warm_divergence_pressure : nullptr;
    ga.warm_den_t = warm ? state_t.method.warm_density_pressure : nullptr;
    ga.xt = dw_.xt.data();
    ga.vt = dw_.vt.data();
    ga.ut = dw_.ut.data();
    ga.rest_sorted = dw_.rest_t.data();
    ga.relax_div_sorted = dw_.relax_div_t.data();
    ga.relax_den_sorted = dw_.relax_den_t.data();
    ga.p_t_div = dw_.p_t_div.data();
    ga.p_t_den = dw_.p_t_den.data();
    dfsph_gather_tangent_kernel<<<blocks_for(n, kBlock), kBlock, 0, stream_>>>(ga);
    PHX_LAUNCH_CHECK();

    // sph.ixx:357: tangent of the state densities
    DensityJvpArgs da{s,
                      dw_.xt.data(),
                      cache.pressures(0u, 0u),
                      cache.pressures(1u, 0u),
                      dw_.p_t_div.data(),
                      dw_.p_t_den.data(),
                      dw_.rho_t.data(),
                      dw_.t_div.data(),
                      dw_.t_den.data(),
                      dw_.rest_t.data()};
    dfsph_jvp_density_kernel<<<blocks, kBlock, 0, stream_>>>(da);
    PHX_LAUNCH_CHECK();

    ForceJvpArgs fa{};
    fa.s = s;
    fa.xt = dw_.xt.data();
    fa.vt = dw_.vt.data();
    fa.ut = dw_.ut.data();
    fa.rho_t = dw_.rho_t.data();
    fa.anp_t = dw_.anp_t.data();
    fa.adiv_t = dw_.adiv_t.data();
    fa.xs_t = dw_.xs_t.data();
    fa.perm = perm;
    fa.next_x_t = vec(next_t.particles.positions);
    fa.next_v_t = vec(next_t.particles.velocities);
    fa.p_t_div = dw_.p_t_div.data();
    fa.warm_div_t_out = next_t.method.warm_divergence_pressure;
    fa.p_t_den = dw_.p_t_den.data();
    fa.warm_den_t_out = next_t.method.warm_density_pressure;

    ProjectionJvpArgs pa{};
    pa.s = s;
    pa.xt = dw_.xt.data();
    pa.xs_t = dw_.xs_t.data();
    pa.rho_t = dw_.rho_t.data();
    pa.volume = cache.volume_rest ? cache.rho0.data() : cache.densities.data();
    pa.volume_t = cache.volume_rest ? dw_.rest_t.data() : dw_.rho_t.data();
    pa.rest_t = dw_.rest_t.data();
    const size_t nn = n;
    uint32_t slot = 0;  // prediction index
    auto collisions = [&](uint32_t k) { return cache.collisions.data() + k * nn; };

    fa.t1 = dw_.t_div.data();
    fa.t2 = dw_.t_den.data();
    fa.collisions = collisions(slot);
    if (nd == 0u) {
        fa.final_step = ne == 0u ? 1u : 0u;
        launch_force_jvp<true, kJvpDivEnd>(stream_, blocks, fa);
    } else {
        fa.final_step = 0u;
        launch_force_jvp<true, kJvpDivPredict>(stream_, blocks, fa);
        pa.target = cache.densities.data();
        pa.target_t = dw_.rho_t.data();
        pa.relax = cache.relax_div.data();
        pa.relax_t = dw_.relax_div_t.data();
        pa.p_t = dw_.p_t_div.data();
        pa.t_out = dw_.t_div.data();
        for (uint32_t it = 1; it <= nd; ++it) {
            pa.slot = slot;
            pa.xs = cache.xs.data() + slot * nn;
            pa.rs = cache.rs.data() + slot * nn;
            pa.p_prev = cache.pressures(0u, it - 1u);
            pa.p = cache.pressures(0u, it);
            dfsph_jvp_projection_kernel<<<blocks, kBlock, 0, stream_>>>(pa);
            PHX_LAUNCH_CHECK();
            fa.collisions = collisions(++slot);
            if (it < nd) {
                launch_force_jvp<false, kJvpDivPredict>(stream_, blocks, fa);
            } else {
                fa.final_step = ne == 0u ? 1u : 0u;
                launch_force_jvp<false, kJvpDivEnd>(stream_, blocks, fa);
            }
        }
    }
    pa.target = cache.rho0.data();
    pa.target_t = dw_.rest_t.data();
    pa.relax = cache.relax_den.data();
    pa.relax_t = dw_.relax_den_t.data();
    pa.p_t = dw_.p_t_den.data();
    pa.t_out = dw_.t_den.data();
    fa.t1 = dw_.t_den.data();
    fa.t2 = nullptr;
    fa.warm_div_t_out = nullptr;
    for (uint32_t it = 1; it <= ne; ++it) {
        pa.slot = slot;
        pa.xs = cache.xs.data() + slot * nn;
        pa.rs = cache.rs.data() + slot * nn;
        pa.p_prev = cache.pressures(1u, it - 1u);
        pa.p = cache.pressures(1u, it);
        dfsph_jvp_projection_kernel<<<blocks, kBlock, 0, stream_>>>(pa);
        PHX_LAUNCH_CHECK();
        fa.collisions = collisions(++slot);
        fa.final_step = it == ne ? 1u : 0u;
        launch_force_jvp<false, kJvpDenPredict>(stream_, blocks, fa);
    }
}

void DfsphSolver::vjp_step(const DfsphCache& cache, const phx_sph_dfsph_state_adjoint& next_a,
                           phx_sph_dfsph_state_adjoint& previous_a, phx_control_adjoint& control_a,
                           phx_sph_dfsph_parameter_adjoint& parameter_a) {
    const uint32_t n = domain_.config.particle_count;
    if (n == 0) return;
    require_recorded(cache);
    const auto& nx = next_a.particles.positions;
    const auto& nv = next_a.particles.velocities;
    require(nx.x && nx.y && nx.z && nv.x && nv.y && nv.z, PHX_ERROR_INVALID_ARGUMENT, "null next-state adjoint");
    auto all_or_none = [](const phx_adjoint_vector_view& v, const char* what) {
        require((v.x == nullptr) == (v.y == nullptr) && (v.x == nullptr) == (v.z == nullptr),
                PHX_ERROR_INVALID_ARGUMENT, std::string(what) + " components must be all null or all set");
    };
    all_or_none(previous_a.particles.positions, "position adjoint");
    all_or_none(previous_a.particles.velocities, "velocity adjoint");
    all_or_none(control_a.external_accelerations, "control adjoint");
    PassScope scope(profiler_, stream_, kDfsphVjp);
    dw_.allocate_adjoints(n);
    const uint32_t nd = cache.iterations[0];
    const uint32_t ne = cache.iterations[1];
    const bool warm = cache.warm_start;
    const uint32_t* perm = cache.neighborhood.perm.data();
    const uint32_t blocks = derivative_blocks(n);
    const Step s = make_step(domain_, cache, constants_, gradient_norm_, config_.exact_no_slip_adjoint != 0u);
    const size_t nn = n;
    auto collisions = [&](uint32_t k) { return cache.collisions.data() + k * nn; };

    // Buffer roles: ap_a[2] holds a_div's adjoint through the density phase
    // and starts the divergence phase; p_a[0] and p_a[2] start with the next
    // warm density / divergence pressure adjoints. Each iteration writes the
    // previous iterate's adjoints into a buffer that holds nothing still needed.
    DVec adiv_a = dw_.ap_a[2].view();
    FinalVjpArgs fa{s,
                    collisions(nd + ne),
                    perm,
                    {nx.x, nx.y, nx.z},
                    {nv.x, nv.y, nv.z},
                    next_a.method.warm_divergence_pressure,
                    next_a.method.warm_density_pressure,
                    dw_.pos_a.view(),
                    dw_.vel_a.view(),
                    dw_.anp_a.view(),
                    adiv_a,
                    dw_.ap_a[0].view(),
                    dw_.p_a[0].data(),
                    dw_.p_a[2].data(),
                    dw_.rho_a.data(),
                    dw_.m_a.data(),
                    dw_.rest_a.data(),
                    dw_.visc_a.data(),
                    dw_.relax_div_a.data(),
                    dw_.relax_den_a.data()};
    dfsph_vjp_final_kernel<<<blocks_for(n, kBlock), kBlock, 0, stream_>>>(fa);
    PHX_LAUNCH_CHECK();

    PressureVjpArgs pv{};
    pv.s = s;
    pv.perm = perm;
    pv.anp_a = dw_.anp_a.view();
    pv.vel_a = dw_.vel_a.view();
    pv.visc_a = dw_.visc_a.data();
    pv.pos_a = dw_.pos_a.view();
    pv.rho_a = dw_.rho_a.data();
    pv.m_a = dw_.m_a.data();
    pv.rest_a = dw_.rest_a.data();
    pv.rs_a = dw_.rs_a.data();
    pv.volume = cache.volume_rest ? cache.rho0.data() : nullptr;
    pv.volume_a = cache.volume_rest ? dw_.rest_a.data() : nullptr;
    DensityVjpArgs dv{};
    dv.s = s;
    dv.pos_a = dw_.pos_a.view();
    dv.vel_a = dw_.vel_a.view();
    dv.anp_a = dw_.anp_a.view();
    dv.adiv_a = adiv_a;
    dv.m_a = dw_.m_a.data();
    dv.rest_a = dw_.rest_a.data();
    dv.rs_a = dw_.rs_a.data();
    // the buffer after `cur`, skipping buffer 2 while it is still needed
    auto next_buffer = [](int cur, bool keep2) {
        int k = (cur + 1) % 3;
        if (keep2 && k == 2) k = (cur + 2) % 3;
        return k;
    };

    // reverse_phase (sph.cpp:675-714): density phase, then divergence phase
    int ca = 0;  // current pressure-acceleration adjoint
    int cp = 0;  // current pressure adjoint
    for (uint32_t phase = 2u; phase-- > 0u;) {
        const bool density_phase = phase == 1u;
        if (!density_phase) ca = cp = 2;
        const uint32_t iterations = density_phase ? ne : nd;
        pv.target = density_phase ? cache.rho0.data() : cache.densities.data();
        pv.target_a = density_phase ? dw_.rest_a.data() : nullptr;
        pv.relax = density_phase ? cache.relax_den.data() : cache.relax_div.data();
        pv.relax_a = density_phase ? dw_.relax_den_a.data() : dw_.relax_div_a.data();
        for (uint32_t it = iterations; it >= 1u; --it) {
            const uint32_t slot = (density_phase ? nd : 0u) + it - 1u;  // prediction read by this update
            const int pa = next_buffer(ca, density_phase);
            const int pp = next_buffer(cp, density_phase);
            pv.p = cache.pressures(phase, it);
            pv.ap_a = dw_.ap_a[ca].view();
            pv.p_a = dw_.p_a[cp].data();
            pv.p_prev = cache.pressures(phase, it - 1u);
            pv.rs = cache.rs.data() + slot * nn;
            pv.p_prev_a = dw_.p_a[pp].data();
            dfsph_vjp_pressure_kernel<true, false><<<blocks, kBlock, 0, stream_>>>(pv);
            PHX_LAUNCH_CHECK();
            dv.slot = slot;
            dv.xs = cache.xs.data() + slot * nn;
            dv.collisions = collisions(slot);
            dv.prev_ap = dw_.ap_a[pa].view();
            if (density_phase) {
                dfsph_vjp_density_kernel<true><<<blocks, kBlock, 0, stream_>>>(dv);
            } else {
                dfsph_vjp_density_kernel<false><<<blocks, kBlock, 0, stream_>>>(dv);
            }
            PHX_LAUNCH_CHECK();
            ca = pa;
            cp = pp;
        }
        // initial pressures p_0 (sph.cpp:712-713); the divergence phase's pass
        // also carries the non-pressure VJP (sph.cpp:722)
        pv.p = cache.pressures(phase, 0u);
        pv.ap_a = dw_.ap_a[ca].view();
        pv.p_a = dw_.p_a[cp].data();
        pv.warm_out = !warm ? nullptr
                            : (density_phase ? previous_a.method.warm_density_pressure
                                             : previous_a.method.warm_divergence_pressure);
        if (density_phase) {
            dfsph_vjp_pressure_kernel<false, false><<<blocks, kBlock, 0, stream_>>>(pv);
        } else {
            dfsph_vjp_pressure_kernel<false, true><<<blocks, kBlock, 0, stream_>>>(pv);
        }
        PHX_LAUNCH_CHECK();
        pv.warm_out = nullptr;
    }

    ScatterVjpArgs sa{};
    sa.s = s;
    sa.perm = perm;
    sa.rho_a = dw_.rho_a.data();
    sa.pos_a = dw_.pos_a.view();
    sa.vel_a = dw_.vel_a.view();
    sa.anp_a = dw_.anp_a.view();
    sa.m_a = dw_.m_a.data();
    sa.rest_a = dw_.rest_a.data();
    sa.visc_a = dw_.visc_a.data();
    sa.relax_div_a = dw_.relax_div_a.data();
    sa.relax_den_a = dw_.relax_den_a.data();
    const auto& pp = previous_a.particles;
    sa.out_pos = DVec{pp.positions.x, pp.positions.y, pp.positions.z};
    sa.out_vel = DVec{pp.velocities.x, pp.velocities.y, pp.velocities.z};
    const auto& ctl = control_a.external_accelerations;
    sa.out_ctl = DVec{ctl.x, ctl.y, ctl.z};
    sa.out_m = parameter_a.particles.masses;
    sa.out_rest = parameter_a.particles.rest_densities;
    sa.out_visc = parameter_a.particles.viscosities;
    sa.out_relax_div = parameter_a.method.divergence_relaxation;
    sa.out_relax_den = parameter_a.method.density_relaxation;
    dfsph_vjp_scatter_kernel<<<blocks, kBlock, 0, stream_>>>(sa);
    PHX_LAUNCH_CHECK();
}

}  // namespace phx
