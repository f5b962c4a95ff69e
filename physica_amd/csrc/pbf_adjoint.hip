// Derivatives of one recorded PBF step for gfx950, in sorted order.
//
// Reference: PBF::jvp_step (pbf/pbf.cpp:193-215) and PBF::vjp_step
// (pbf/pbf.cpp:217-258) with the *_jvp_kernel / *_vjp_kernel bodies of
// pbf/pbf.cu and density/density.cu. The neighbour graph, the box-collision
// masks and every branch are those of the recorded forward step; tangents are
// FP32 like the reference's, adjoints FP64 with the FP32 primal values
// promoted. The reference recomputes constraint iterations from checkpoints;
// here the forward step recorded every iterate in the step cache.
//
// Reverse passes (gather form: particle a collects the pair terms where it is
// the centre and those where it is the neighbour, so nothing is written by
// two particles and no atomics are needed):
//   post        XSPH VJP + confinement VJP + normal-gradient adjoint
//   normal      normal VJP + magnitude VJP
//   vorticity   vorticity VJP + reconstruct VJP
//   per iteration, last to first:
//     correction  projection VJP (masks) + correction VJP
//     lambda      lambda VJP + density VJP (a neighbour's density adjoint is
//                 rebuilt from its lambda adjoint)
//   scatter     predict VJP, accumulated into the caller's original-order
//               adjoints
// Forward passes: predict, per iteration density+lambda then correction +
// projection, vorticity (+ reconstruct), normal + confinement, XSPH.
#include <cmath>

#include "neighbor_mask.hpp"
#include "pbf.hpp"

namespace phx {
namespace {

constexpr uint32_t kBlock = 256;
// records a derivative pass's walks fetch ahead (two were no faster: the
// deeper ring costs occupancy)
constexpr int kVjpDepth = 1;

// LDS of a masked neighbour pass: the lanes' mask columns and the slot table
// (every thread reaches the barrier before any returns).
#define PHX_MASK_LDS(grid)                          \
    __shared__ uint16_t smask[kStencil * kBlock];   \
    __shared__ int delta[kStencil];                 \
    fill_slot_deltas(grid, delta);                  \
    __syncthreads()

struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 d3(Float3 a) {
    return {static_cast<double>(a.x), static_cast<double>(a.y), static_cast<double>(a.z)};
}
__device__ __forceinline__ D3 add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 scale(D3 v, double f) { return {f * v.x, f * v.y, f * v.z}; }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 cross(D3 a, D3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ D3 load(DVec f, uint32_t i) { return {f.x[i], f.y[i], f.z[i]}; }
__device__ __forceinline__ void store(DVec f, uint32_t i, D3 v) {
    f.x[i] = v.x;
    f.y[i] = v.y;
    f.z[i] = v.z;
}

// Derivative of the spiky gradient G(d) = f(r) d, f = k (h - r)^2 / r, along
// w: H(d) w = f w + (f'(r) / r) (d . w) d, f' = -k (h - r)(h + r) / r^2; zero
// where G is zero (r = 0 or r >= h). H is symmetric and even in d
// (spiky_hessian_product, spiky_gradient_tangent, pbf.cu:34-54).
// The FP64 form (reverse passes) is spiky_hessian_one_division, below.
__device__ __forceinline__ D3 spiky_hessian_one_division(Float3 d, D3 w, double h, double k);
__device__ __forceinline__ D3 spiky_hessian(Float3 d, D3 w, double h, double k) {
    return spiky_hessian_one_division(d, w, h, k);
}

__device__ __forceinline__ Float3 spiky_hessian(Float3 d, Float3 w, float h, float k) {
    const float r = length(d);
    if (r == 0.0F || r >= h) return {0.0F, 0.0F, 0.0F};
    const float f = k * (h - r) * (h - r) / r;
    const float fp = -k * (h - r) * (h + r) / (r * r);
    return add(scale(w, f), scale(d, fp / r * dot(d, w)));
}

// artificial_pressure (pbf.cu:56-71): s(d) = -k (W(d) / W(dq))^n with its
// gradient in d and its partial derivatives in k, n and dq.
struct Artificial {
    float value;
    Float3 gradient;
    float d_strength, d_exponent, d_radius;
};

// With `vanishes` (artificial_vanishes of the particle) the value is -k and
// its gradient, exponent and radius derivatives are signed zeros that no sum
// can notice; only the strength derivative -(W / W(dq))^n needs a powf, and
// only when `strength_wanted`.
// `kernel` is poly6(d, h), which a pair's two directions share (d and -d
// give the same value).
__device__ __forceinline__ Artificial artificial_at(float kernel, Float3 d, float h, float strength, float exponent,
                                                    float radius, bool vanishes = false, bool strength_wanted = true) {
    Artificial a{0.0F, {0.0F, 0.0F, 0.0F}, 0.0F, 0.0F, 0.0F};
    if (kernel == 0.0F) return a;
    if (vanishes) {
        a.value = -strength;
        if (strength_wanted) a.d_strength = -powf(kernel / poly6(Float3{radius, 0.0F, 0.0F}, h), exponent);
        return a;
    }
    const float reference = poly6(Float3{radius, 0.0F, 0.0F}, h);
    const float ratio = kernel / reference;
    const float power = powf(ratio, exponent);
    a.value = -strength * power;
    a.d_strength = -power;
    a.d_exponent = a.value * logf(ratio);
    const float hr = h * h;
    const float diff = hr - radius * radius;
    const float radial =
        radius >= h ? 0.0F : -6.0F * 315.0F / (64.0F * kPi * hr * hr * hr * h * h * h) * radius * diff * diff;
    a.d_radius = -a.value * exponent * radial / reference;
    a.gradient = scale(poly6_gradient(d, h), -strength * exponent * powf(ratio, exponent - 1.0F) / reference);
    return a;
}

__device__ __forceinline__ Artificial artificial_full(Float3 d, float h, float strength, float exponent, float radius,
                                                      bool vanishes = false, bool strength_wanted = true) {
    return artificial_at(poly6(d, h), d, h, strength, exponent, radius, vanishes, strength_wanted);
}

// Sorted-order views of a recorded step and the sorted parameters.
struct Step {
    uint32_t n;
    float h, spiky, dt, inv_dt;
    SortedGrid grid;
    CVec xs, x0;
    const float *m, *rho0, *eps, *strength, *exponent, *radius, *xsph, *confinement;
    MaskView mv;  // the recorded step's masks and flags
    const double* inv_rho0;  // 1 / rho0 (reverse passes)
    BoundaryGrid boundary;   // the step's sorted boundary samples; offsets null: none
};

// Visits particle i's fluid candidates through its masks where they apply or,
// with boundary samples, every particle and every sample of the 27 cells,
// each cell's samples after its particles (the order of the reference's
// loops, e.g. pbf.cu:143-168). Samples are inputs: their pair terms reach
// only particle i's values and adjoints.
template <class Visit, class VisitBoundary>
__device__ __forceinline__ void walk_with_samples(const Step& s, bool masked, uint32_t i, int cx, int cy, int cz,
                                                  const int* delta, uint16_t* column, Visit&& visit,
                                                  VisitBoundary&& visit_boundary) {
    if (s.boundary.offsets != nullptr) {
        for_each_candidate(s.grid, s.boundary, cx, cy, cz, visit, visit_boundary);
    } else {
        visit_candidates(s.grid, s.mv, masked, i, cx, cy, cz, delta, column, kBlock, visit);
    }
}

// The same with fluid candidate j visited as pair(load(j), j) (visit_records
// without samples).
template <class Load, class Pair, class VisitBoundary>
__device__ __forceinline__ void walk_with_samples(const Step& s, bool masked, uint32_t i, int cx, int cy, int cz,
                                                  const int* delta, uint16_t* column, Load&& load, Pair&& pair,
                                                  VisitBoundary&& visit_boundary) {
    if (s.boundary.offsets != nullptr) {
        for_each_candidate(
            s.grid, s.boundary, cx, cy, cz, [&](uint32_t j) { pair(load(j), j); }, visit_boundary);
    } else {
        visit_records<kVjpDepth>(s.grid, s.mv, masked, i, cx, cy, cz, delta, column, kBlock, load, pair);
    }
}

__device__ __forceinline__ Float3 sample_position(const Step& s, uint32_t b) {
    const float4 q = s.boundary.x[b];
    return Float3{q.x, q.y, q.z};
}

__device__ __forceinline__ void cell_of_topology(const Step& s, uint32_t i, int& cx, int& cy, int& cz) {
    particle_cell(s.grid, load(s.xs, i), cx, cy, cz);
}

__device__ __forceinline__ D3 unclamped(D3 b, uint8_t m) {
    if (m & 1u) b.x = 0.0;
    if (m & 2u) b.y = 0.0;
    if (m & 4u) b.z = 0.0;
    return b;
}

// ----------------------------------------------------------------------- VJP
struct PostVjpArgs {
    Step s;
    uint32_t slot;  // displacement slot of the positions read
    CVec pos, vc, w, nrm;
    const float* nz;
    DVec next_x_a, next_v_a;  // adjoints of the next state, sorted
    DVec p_a, vrec_a, w_a, g_a;
    double* xsph_a;
    double* confinement_a;
};

// xsph_vjp_kernel (pbf.cu:756-794), confinement_vjp_kernel (pbf.cu:688-698)
// and the adjoint of the normalised gradient (pbf.cu:630-633).
__global__ void __launch_bounds__(kBlock) pbf_vjp_post_kernel(PostVjpArgs a) {
    PHX_MASK_LDS(a.s.grid);
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.s.n) return;
    const bool masked = masks_apply(a.s.mv, a.slot, i);
    const Step& s = a.s;
    const Float3 pi = load(a.pos, i);
    const D3 out = load(a.next_v_a, i);
    const D3 vci = d3(load(a.vc, i));
    const double ci = static_cast<double>(s.xsph[i]);
    D3 pcon{0.0, 0.0, 0.0};
    D3 vcon = out;
    double ccon = 0.0;
    int cx, cy, cz;
    cell_of_topology(s, i, cx, cy, cz);
    struct Record {
        Float3 x, vc;
        D3 vb;
        float c;
    };
    auto fetch = [&](uint32_t j) { return Record{load(a.pos, j), load(a.vc, j), load(a.next_v_a, j), s.xsph[j]}; };
    visit_records<kVjpDepth>(s.grid, s.mv, masked, i, cx, cy, cz, delta, smask + threadIdx.x, kBlock, fetch,
                  [&](const Record& r, uint32_t j) {
        if (j == i) return;
        const Float3 dd = sub(pi, r.x);
        const double wgt = static_cast<double>(poly6(dd, s.h));
        const D3 gw = d3(poly6_gradient(dd, s.h));
        const D3 diff = sub(d3(r.vc), vci);
        const D3 vb = r.vb;
        const double cj = static_cast<double>(r.c);
        vcon = add(vcon, scale(out, -ci * wgt));
        vcon = add(vcon, scale(vb, cj * wgt));
        pcon = add(pcon, scale(gw, ci * dot(out, diff) - cj * dot(vb, diff)));
        ccon += wgt * dot(out, diff);
    });
    store(a.p_a, i, add(load(a.next_x_a, i), pcon));
    a.xsph_a[i] += ccon;
    // confinement: v_c = v + dt eps (n x w)
    const D3 wv = d3(load(a.w, i));
    const D3 nv = d3(load(a.nrm, i));
    const double f = static_cast<double>(s.dt) * static_cast<double>(s.confinement[i]);
    store(a.vrec_a, i, vcon);
    store(a.w_a, i, scale(cross(vcon, nv), f));
    const D3 n_bar = scale(cross(wv, vcon), f);
    a.confinement_a[i] += static_cast<double>(s.dt) * dot(vcon, cross(nv, wv));
    const float nz = a.nz[i];
    store(a.g_a, i,
          nz == 0.0F ? D3{0.0, 0.0, 0.0}
                     : scale(sub(n_bar, scale(nv, dot(nv, n_bar))), 1.0 / static_cast<double>(nz)));
}

struct NormalVjpArgs {
    Step s;
    uint32_t slot;  // displacement slot of the positions read
    CVec pos, w;
    const float* mag;
    DVec p_a, w_a, g_a;
};

// normal_vjp_kernel (pbf.cu:635-672) and the magnitude adjoint into omega.
__global__ void __launch_bounds__(kBlock) pbf_vjp_normal_kernel(NormalVjpArgs a) {
    PHX_MASK_LDS(a.s.grid);
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.s.n) return;
    const bool masked = masks_apply(a.s.mv, a.slot, i);
    const Step& s = a.s;
    const Float3 pi = load(a.pos, i);
    const D3 gi = load(a.g_a, i);
    const double mi = static_cast<double>(a.mag[i]);
    D3 pcon{0.0, 0.0, 0.0};
    double mcon = 0.0;
    int cx, cy, cz;
    cell_of_topology(s, i, cx, cy, cz);
    struct Record {
        Float3 x;
        float mag;
        D3 g;
    };
    auto fetch = [&](uint32_t j) { return Record{load(a.pos, j), a.mag[j], load(a.g_a, j)}; };
    visit_records<kVjpDepth>(s.grid, s.mv, masked, i, cx, cy, cz, delta, smask + threadIdx.x, kBlock, fetch,
                  [&](const Record& r, uint32_t j) {
        if (j == i) return;
        const Float3 dd = sub(pi, r.x);
        const D3 gr = d3(spiky_gradient(dd, s.h, s.spiky));
        const D3 both = add(gi, r.g);
        pcon = add(pcon, spiky_hessian(dd, scale(both, static_cast<double>(r.mag) - mi), static_cast<double>(s.h),
                                       static_cast<double>(s.spiky)));
        mcon -= dot(gr, both);
    });
    store(a.p_a, i, add(load(a.p_a, i), pcon));
    if (a.mag[i] > 0.0F) store(a.w_a, i, add(load(a.w_a, i), scale(d3(load(a.w, i)), mcon / mi)));
}

struct VorticityVjpArgs {
    Step s;
    uint32_t slot;  // displacement slot of the positions read
    CVec pos, vrec;
    DVec p_a, vrec_a, w_a, x_a;
};

// vorticity_vjp_kernel (pbf.cu:529-563) and reconstruct_vjp_kernel
// (pbf.cu:468-474).
__global__ void __launch_bounds__(kBlock) pbf_vjp_vorticity_kernel(VorticityVjpArgs a) {
    PHX_MASK_LDS(a.s.grid);
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.s.n) return;
    const bool masked = masks_apply(a.s.mv, a.slot, i);
    const Step& s = a.s;
    const Float3 pi = load(a.pos, i);
    const D3 vi = d3(load(a.vrec, i));
    const D3 wb = load(a.w_a, i);
    D3 pcon{0.0, 0.0, 0.0};
    D3 vcon{0.0, 0.0, 0.0};
    int cx, cy, cz;
    cell_of_topology(s, i, cx, cy, cz);
    struct Record {
        Float3 x, v;
        D3 w;
    };
    auto fetch = [&](uint32_t j) { return Record{load(a.pos, j), load(a.vrec, j), load(a.w_a, j)}; };
    visit_records<kVjpDepth>(s.grid, s.mv, masked, i, cx, cy, cz, delta, smask + threadIdx.x, kBlock, fetch,
                  [&](const Record& r, uint32_t j) {
        if (j == i) return;
        const Float3 dd = sub(pi, r.x);
        const D3 gr = d3(spiky_gradient(dd, s.h, s.spiky));
        const D3 both = add(wb, r.w);
        const D3 diff = sub(d3(r.v), vi);
        vcon = sub(vcon, cross(gr, both));
        pcon = add(pcon, spiky_hessian(dd, cross(both, diff), static_cast<double>(s.h), static_cast<double>(s.spiky)));
    });
    const D3 va = add(load(a.vrec_a, i), vcon);
    const D3 cb = scale(va, static_cast<double>(s.inv_dt));
    store(a.x_a, i, scale(cb, -1.0));
    store(a.p_a, i, add(add(load(a.p_a, i), pcon), cb));
}

struct CorrectionVjpArgs {
    Step s;
    uint32_t slot;  // displacement slot of the positions read
    const uint8_t* vanishes;  // per particle: artificial_vanishes
    uint32_t strength_wanted;
    CVec pos;  // positions entering the iteration
    const float* lambda;
    const uint8_t* mask;
    DVec p_in;   // adjoint of the positions leaving the iteration
    DVec p_out;  // adjoint of the positions entering it (written)
    double* lambda_a;
    double *m_a, *rho0_a, *strength_a, *exponent_a, *radius_a;
    // the iteration's densities and denominators, for the per-particle terms
    // of the lambda VJP
    const float *rho, *den;
    double2* lambda_terms;  // (density adjoint, D_bar) per particle (written)
};

// 1 / sqrt(x) for a positive normal x: the hardware estimate refined by two
// Newton steps, y <- y + y (1 - x y^2) / 2, to within a few units of the
// last place (the derivative passes compare with the FP64 oracle to 1e-10;
// the square root and the division they replace cost about 30 instructions).
__device__ __forceinline__ double inverse_sqrt(double x) {
    double y = __builtin_amdgcn_rsq(x);
#pragma unroll
    for (int step = 0; step < 2; ++step) {
        const double e = fma(-x * y, y, 1.0);
        y = fma(0.5 * y, e, y);
    }
    return y;
}

// Derivative of G(d) = f(r) d (spiky gradient) along w, with 1 / r from
// inverse_sqrt instead of a square root and a division.
__device__ __forceinline__ D3 spiky_hessian_one_division(Float3 d, D3 w, double h, double k) {
    const D3 dd = d3(d);
    const double r2 = dot(dd, dd);
    if (r2 == 0.0) return {0.0, 0.0, 0.0};
    const double inv_r = inverse_sqrt(r2);
    const double r = r2 * inv_r;
    if (r >= h) return {0.0, 0.0, 0.0};
    const double hr = h - r;
    const double f = k * hr * hr * inv_r;
    const double fpr = -k * hr * (h + r) * inv_r * inv_r * inv_r;  // f'(r) / r
    return add(scale(w, f), scale(dd, fpr * dot(dd, w)));
}

// project_vjp_kernel (pbf.cu:445-454) and correction_vjp_kernel
// (pbf.cu:335-415). The Hessian is even in d, so the terms where i is the
// centre and where it is the neighbour share one product; the neighbour's
// spiky gradient is the exact negation of the centre's.
__global__ void __launch_bounds__(kBlock) pbf_vjp_correction_kernel(CorrectionVjpArgs a) {
    PHX_MASK_LDS(a.s.grid);
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.s.n) return;
    const bool masked = masks_apply(a.s.mv, a.slot, i);
    const Step& s = a.s;
    const double h = static_cast<double>(s.h), k = static_cast<double>(s.spiky);
    const Float3 pi = load(a.pos, i);
    const double inv_rho0 = s.inv_rho0[i];
    const D3 oa = unclamped(load(a.p_in, i), a.mask[i]);
    const float lam_i = a.lambda[i];
    const double m_i = static_cast<double>(s.m[i]);
    D3 pcon{0.0, 0.0, 0.0};
    double lcon = 0.0, mcon = 0.0, rcon = 0.0, scon = 0.0, econ = 0.0, qcon = 0.0;
    int cx, cy, cz;
    cell_of_topology(s, i, cx, cy, cz);
    // candidate j: position, lambda, mass, artificial-pressure parameters and
    // vanishing flag, 1 / rho0 and its unclamped position adjoint
    struct Record {
        Float3 x;
        float lambda, m, strength, exponent, radius;
        uint32_t vanishes;
        double inv_rho0;
        D3 ob;
    };
    auto fetch = [&](uint32_t j) {
        return Record{load(a.pos, j),    a.lambda[j],  s.m[j],
                      s.strength[j],     s.exponent[j], s.radius[j],
                      a.vanishes[j],     s.inv_rho0[j], unclamped(load(a.p_in, j), a.mask[j])};
    };
    const float strength_i = s.strength[i], exponent_i = s.exponent[i], radius_i = s.radius[i];
    const bool vanishes_i = a.vanishes[i] != 0u;
    walk_with_samples(s, masked, i, cx, cy, cz, delta, smask + threadIdx.x, fetch, [&](const Record& r, uint32_t j) {
        if (j == i) return;
        const Float3 dd = sub(pi, r.x);
        const D3 gr = d3(spiky_gradient(dd, s.h, s.spiky));
        const float lam_j = r.lambda;
        // i as the centre
        const float kernel = poly6(dd, s.h);
        const Artificial art =
            artificial_at(kernel, dd, s.h, strength_i, exponent_i, radius_i, vanishes_i, a.strength_wanted != 0u);
        const double cf = static_cast<double>(r.m) * inv_rho0;
        const double mu = static_cast<double>(lam_i) + static_cast<double>(lam_j) + static_cast<double>(art.value);
        const double oa_gr = dot(oa, gr);
        const double sa = cf * oa_gr;
        lcon += sa;
        rcon -= cf * mu * oa_gr * inv_rho0;
        scon += sa * static_cast<double>(art.d_strength);
        econ += sa * static_cast<double>(art.d_exponent);
        qcon += sa * static_cast<double>(art.d_radius);
        // i as a neighbour of j: d_ji = -d_ij, gradient -gr
        const Artificial rart = artificial_at(kernel, scale(dd, -1.0F), s.h, r.strength, r.exponent, r.radius,
                                              r.vanishes != 0u, a.strength_wanted != 0u);
        const double inv_rrho0 = r.inv_rho0;
        const double rcf = m_i * inv_rrho0;
        const double rmu = static_cast<double>(lam_j) + static_cast<double>(lam_i) + static_cast<double>(rart.value);
        const D3 ob = r.ob;
        const double ob_rgr = -dot(ob, gr);
        const double rsa = rcf * ob_rgr;
        lcon += rsa;
        mcon += rmu * inv_rrho0 * ob_rgr;
        pcon = add(pcon, spiky_hessian_one_division(dd, sub(scale(oa, cf * mu), scale(ob, rcf * rmu)), h, k));
        pcon = add(pcon, sub(scale(d3(art.gradient), sa), scale(d3(rart.gradient), rsa)));
    }, [&](uint32_t b) {
        // a boundary sample: V_b (lambda_i + s_corr) grad W (pbf.cu:390-397)
        const Float3 dd = sub(pi, sample_position(s, b));
        const D3 gr = d3(spiky_gradient(dd, s.h, s.spiky));
        const Artificial art =
            artificial_full(dd, s.h, strength_i, exponent_i, radius_i, vanishes_i, a.strength_wanted != 0u);
        const double vol = static_cast<double>(s.boundary.x[b].w);
        const double mu = static_cast<double>(lam_i) + static_cast<double>(art.value);
        const double sa = vol * dot(oa, gr);
        lcon += sa;
        scon += sa * static_cast<double>(art.d_strength);
        econ += sa * static_cast<double>(art.d_exponent);
        qcon += sa * static_cast<double>(art.d_radius);
        pcon = add(pcon, spiky_hessian_one_division(dd, scale(oa, vol * mu), h, k));
        pcon = add(pcon, scale(d3(art.gradient), sa));
    });
    store(a.p_out, i, add(oa, pcon));
    a.lambda_a[i] = lcon;
    a.m_a[i] += mcon;
    a.rho0_a[i] += rcon;
    a.strength_a[i] += scon;
    a.exponent_a[i] += econ;
    a.radius_a[i] += qcon;
    // lambda_vjp_kernel's per-particle terms: the density adjoint
    // -lambda_bar / (D rho0) and D_bar = lambda_bar C / D^2 (pbf.cu:207-214)
    const double rho0 = static_cast<double>(s.rho0[i]);
    const double den = static_cast<double>(a.den[i]);
    const double con = static_cast<double>(a.rho[i]) / rho0 - 1.0;
    a.lambda_terms[i] = make_double2(-(lcon / (den * rho0)), lcon * con / (den * den));
}

struct LambdaVjpArgs {
    Step s;
    uint32_t slot;  // displacement slot of the positions read
    CVec pos;
    const float *rho, *den;
    CVec gsum;
    const double* lambda_a;
    const double2* lambda_terms;
    DVec p_out;
    double *m_a, *rho0_a, *eps_a;
};

// lambda_vjp_kernel (pbf.cu:178-252) and density_vjp_kernel (density.cu:69-98,
// poly6) with the neighbours' density adjoints and D_bar read from the
// per-particle terms; the centre and neighbour Hessian terms share one
// product.
__global__ void __launch_bounds__(kBlock) pbf_vjp_lambda_kernel(LambdaVjpArgs a) {
    PHX_MASK_LDS(a.s.grid);
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.s.n) return;
    const bool masked = masks_apply(a.s.mv, a.slot, i);
    const Step& s = a.s;
    const double h = static_cast<double>(s.h), k = static_cast<double>(s.spiky);
    const Float3 pi = load(a.pos, i);
    const double rho0 = static_cast<double>(s.rho0[i]);
    const double inv_rho0 = s.inv_rho0[i];
    const double den = static_cast<double>(a.den[i]);
    const double lam_a = a.lambda_a[i];
    const double2 terms = a.lambda_terms[i];
    const double rho_a = terms.x;
    const double dbar = terms.y;
    const D3 sgi = d3(load(a.gsum, i));
    const double mi = static_cast<double>(s.m[i]);
    D3 pl{0.0, 0.0, 0.0}, pd{0.0, 0.0, 0.0};
    double ml = 0.0, md = 0.0;
    double rcon = lam_a * static_cast<double>(a.rho[i]) / (den * rho0 * rho0);
    double rd = 0.0;  // the density's rest-density adjoint (boundary samples)
    int cx, cy, cz;
    cell_of_topology(s, i, cx, cy, cz);
    // candidate j: position, mass, per-particle lambda terms, 1 / rho0 and
    // gradient sum
    struct Record {
        Float3 x;
        float m;
        double2 terms;
        double inv_rho0;
        Float3 gsum;
    };
    auto fetch = [&](uint32_t j) {
        return Record{load(a.pos, j), s.m[j], a.lambda_terms[j], s.inv_rho0[j], load(a.gsum, j)};
    };
    walk_with_samples(s, masked, i, cx, cy, cz, delta, smask + threadIdx.x, fetch, [&](const Record& r, uint32_t j) {
        const Float3 dd = sub(pi, r.x);
        const double2 jterms = r.terms;
        // density term (the self term has a zero gradient)
        const D3 gw = d3(poly6_gradient(dd, s.h));
        pd = add(pd, scale(gw, rho_a * static_cast<double>(r.m) + jterms.x * mi));
        md += jterms.x * static_cast<double>(poly6(dd, s.h));
        if (j == i) return;
        const D3 gr = d3(spiky_gradient(dd, s.h, s.spiky));
        const double gr2 = dot(gr, gr);
        // i as the centre: g_bar = 2 D_bar c (c g + S), c_bar = 2 D_bar (c |g|^2 + g . S)
        const double cf = static_cast<double>(r.m) * inv_rho0;
        const D3 gbar = scale(add(scale(gr, cf), sgi), 2.0 * dbar * cf);
        const double cbar = 2.0 * dbar * (cf * gr2 + dot(gr, sgi));
        rcon -= cbar * cf * inv_rho0;
        // i as a neighbour of j: gradient -gr
        const double inv_jrho0 = r.inv_rho0;
        const D3 sgj = d3(r.gsum);
        const double rcf = mi * inv_jrho0;
        const D3 rgbar = scale(sub(sgj, scale(gr, rcf)), 2.0 * jterms.y * rcf);
        const double rcbar = 2.0 * jterms.y * (rcf * gr2 - dot(gr, sgj));
        ml += rcbar * inv_jrho0;
        pl = add(pl, spiky_hessian_one_division(dd, sub(gbar, rgbar), h, k));
    }, [&](uint32_t b) {
        // a boundary sample: the density term rho0_i V_b W (density.cu:88-93)
        // and V_b grad W in the gradient sum only (pbf.cu:226-229)
        const Float3 dd = sub(pi, sample_position(s, b));
        const double vol = static_cast<double>(s.boundary.x[b].w);
        pd = add(pd, scale(d3(poly6_gradient(dd, s.h)), rho_a * rho0 * vol));
        rd += rho_a * vol * static_cast<double>(poly6(dd, s.h));
        pl = add(pl, spiky_hessian_one_division(dd, scale(sgi, 2.0 * dbar * vol), h, k));
    });
    store(a.p_out, i, add(add(load(a.p_out, i), pl), pd));
    a.m_a[i] += ml;
    a.m_a[i] += md;
    a.rho0_a[i] += rcon;
    a.rho0_a[i] += rd;
    a.eps_a[i] += dbar;
}

struct ScatterVjpArgs {
    uint32_t n;
    double dt, dt2;
    const uint32_t* perm;
    DVec p_a, x_a;
    DVec prev_x, prev_v, ctl;  // original order, accumulated; any component may be null
    const double* sorted[8];
    double* original[8];
};

// predict_vjp_kernel (pbf.cu:86-93) and the accumulation of every adjoint
// into the caller's original-order arrays.
__global__ void __launch_bounds__(kBlock) pbf_vjp_scatter_kernel(ScatterVjpArgs a) {
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t o = a.perm[i];
    const D3 b = load(a.p_a, i);
    const D3 xb = add(load(a.x_a, i), b);
    if (a.prev_x.x != nullptr) {
        a.prev_x.x[o] += xb.x;
        a.prev_x.y[o] += xb.y;
        a.prev_x.z[o] += xb.z;
    }
    if (a.prev_v.x != nullptr) {
        a.prev_v.x[o] += a.dt * b.x;
        a.prev_v.y[o] += a.dt * b.y;
        a.prev_v.z[o] += a.dt * b.z;
    }
    if (a.ctl.x != nullptr) {
        a.ctl.x[o] += a.dt2 * b.x;
        a.ctl.y[o] += a.dt2 * b.y;
        a.ctl.z[o] += a.dt2 * b.z;
    }
#pragma unroll
    for (int c = 0; c < 8; ++c)
        if (a.original[c] != nullptr) a.original[c][o] += a.sorted[c][i];
}

// Per-particle constants of the reverse passes: whether the artificial
// pressure vanishes, and 1 / rho0.
__global__ void __launch_bounds__(kBlock) pbf_vjp_setup_kernel(uint32_t n, float h, const float* strength,
                                                               const float* exponent, const float* radius,
                                                               const float* rho0, uint8_t* vanishes,
                                                               double* inv_rho0) {
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    const float reference = poly6(Float3{radius[i], 0.0F, 0.0F}, h);
    vanishes[i] = artificial_vanishes(h, strength[i], exponent[i], reference) ? 1u : 0u;
    inv_rho0[i] = 1.0 / static_cast<double>(rho0[i]);
}

struct GatherAdjointArgs {
    uint32_t n;
    const uint32_t* perm;
    const double *xx, *xy, *xz, *vx, *vy, *vz;
    DVec x_a, v_a;
};

__global__ void __launch_bounds__(kBlock) pbf_gather_adjoint_kernel(GatherAdjointArgs a) {
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t o = a.perm[i];
    store(a.x_a, i, D3{a.xx[o], a.xy[o], a.xz[o]});
    store(a.v_a, i, D3{a.vx[o], a.vy[o], a.vz[o]});
}

// ----------------------------------------------------------------------- JVP
struct PredictJvpArgs {
    uint32_t n;
    float dt;
    const uint32_t* perm;
    CVec x_t, v_t, u_t;  // original order
    Vec xs_t, p_t;       // sorted: state position tangent, first iterate tangent
};

// predict_jvp_kernel (pbf.cu:80-84), gathered into sorted order.
__global__ void __launch_bounds__(kBlock) pbf_jvp_predict_kernel(PredictJvpArgs a) {
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t o = a.perm[i];
    const Float3 xt = load(a.x_t, o);
    store(a.xs_t, i, xt);
    store(a.p_t, i, add(xt, scale(add(load(a.v_t, o), scale(load(a.u_t, o), a.dt)), a.dt)));
}

struct TangentParams {
    const float *m, *rho0, *eps, *strength, *exponent, *radius, *xsph, *confinement;
};

struct LambdaJvpArgs {
    Step s;
    TangentParams t;
    uint32_t slot;  // displacement slot of the positions read
    CVec pos, p_t;
    const float *rho, *den;
    float* lambda_t;
};

// density_jvp_kernel (density.cu:42-67, poly6) and lambda_jvp_kernel
// (pbf.cu:129-176); the density tangent is the particle's own.
__global__ void __launch_bounds__(kBlock) pbf_jvp_lambda_kernel(LambdaJvpArgs a) {
    PHX_MASK_LDS(a.s.grid);
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.s.n) return;
    const bool masked = masks_apply(a.s.mv, a.slot, i);
    const Step& s = a.s;
    const Float3 pi = load(a.pos, i), ti = load(a.p_t, i);
    const float rho0 = s.rho0[i], rho0_t = a.t.rho0[i];
    float rho_t = 0.0F;
    Float3 sg{0.0F, 0.0F, 0.0F}, sg_t{0.0F, 0.0F, 0.0F};
    float den_t = a.t.eps[i];
    int cx, cy, cz;
    cell_of_topology(s, i, cx, cy, cz);
    walk_with_samples(s, masked, i, cx, cy, cz, delta, smask + threadIdx.x, [&](uint32_t j) {
        const Float3 dd = sub(pi, load(a.pos, j));
        const Float3 dd_t = sub(ti, load(a.p_t, j));
        rho_t += a.t.m[j] * poly6(dd, s.h) + s.m[j] * dot(poly6_gradient(dd, s.h), dd_t);
        if (j == i) return;
        const Float3 gr = spiky_gradient(dd, s.h, s.spiky);
        const Float3 gr_t = spiky_hessian(dd, dd_t, s.h, s.spiky);
        const float cf = s.m[j] / rho0;
        const float cf_t = a.t.m[j] / rho0 - s.m[j] * rho0_t / (rho0 * rho0);
        const Float3 ng = scale(gr, cf);
        const Float3 ng_t = add(scale(gr, cf_t), scale(gr_t, cf));
        sg = add(sg, ng);
        sg_t = add(sg_t, ng_t);
        den_t += 2.0F * dot(ng, ng_t);
    }, [&](uint32_t b) {
        // a boundary sample (density.cu:59-63, pbf.cu:164-168), itself fixed
        const Float3 dd = sub(pi, sample_position(s, b));
        const float vol = s.boundary.x[b].w;
        rho_t += vol * (rho0_t * poly6(dd, s.h) + rho0 * dot(poly6_gradient(dd, s.h), ti));
        sg = add(sg, scale(spiky_gradient(dd, s.h, s.spiky), vol));
        sg_t = add(sg_t, scale(spiky_hessian(dd, ti, s.h, s.spiky), vol));
    });
    den_t += 2.0F * dot(sg, sg_t);
    const float den = a.den[i];
    const float con = a.rho[i] / rho0 - 1.0F;
    const float con_t = rho_t / rho0 - a.rho[i] * rho0_t / (rho0 * rho0);
    a.lambda_t[i] = -con_t / den + con * den_t / (den * den);
}

struct CorrectionJvpArgs {
    Step s;
    TangentParams t;
    uint32_t slot;  // displacement slot of the positions read
    CVec pos, p_t;
    const float *lambda, *lambda_t;
    const uint8_t* mask;
    Vec p_t_out;
};

// correction_jvp_kernel (pbf.cu:290-333) and project_jvp_kernel
// (pbf.cu:433-443).
__global__ void __launch_bounds__(kBlock) pbf_jvp_correction_kernel(CorrectionJvpArgs a) {
    PHX_MASK_LDS(a.s.grid);
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.s.n) return;
    const bool masked = masks_apply(a.s.mv, a.slot, i);
    const Step& s = a.s;
    const Float3 pi = load(a.pos, i), ti = load(a.p_t, i);
    const float rho0 = s.rho0[i], rho0_t = a.t.rho0[i];
    Float3 res{0.0F, 0.0F, 0.0F};
    int cx, cy, cz;
    cell_of_topology(s, i, cx, cy, cz);
    walk_with_samples(s, masked, i, cx, cy, cz, delta, smask + threadIdx.x, [&](uint32_t j) {
        if (j == i) return;
        const Float3 dd = sub(pi, load(a.pos, j));
        const Float3 dd_t = sub(ti, load(a.p_t, j));
        const Artificial art = artificial_full(dd, s.h, s.strength[i], s.exponent[i], s.radius[i]);
        const float a_t = dot(art.gradient, dd_t) + art.d_strength * a.t.strength[i] +
                          art.d_exponent * a.t.exponent[i] + art.d_radius * a.t.radius[i];
        const float cf = s.m[j] / rho0;
        const float cf_t = a.t.m[j] / rho0 - s.m[j] * rho0_t / (rho0 * rho0);
        const float mu = a.lambda[i] + a.lambda[j] + art.value;
        const float mu_t = a.lambda_t[i] + a.lambda_t[j] + a_t;
        res = add(res, add(scale(spiky_gradient(dd, s.h, s.spiky), cf_t * mu + cf * mu_t),
                           scale(spiky_hessian(dd, dd_t, s.h, s.spiky), cf * mu)));
    }, [&](uint32_t b) {
        // a boundary sample: V_b (lambda_i + s_corr) grad W (pbf.cu:321-329)
        const Float3 dd = sub(pi, sample_position(s, b));
        const Artificial art = artificial_full(dd, s.h, s.strength[i], s.exponent[i], s.radius[i]);
        const float a_t = dot(art.gradient, ti) + art.d_strength * a.t.strength[i] +
                          art.d_exponent * a.t.exponent[i] + art.d_radius * a.t.radius[i];
        const float mu = a.lambda[i] + art.value;
        res = add(res, scale(add(scale(spiky_gradient(dd, s.h, s.spiky), a.lambda_t[i] + a_t),
                                 scale(spiky_hessian(dd, ti, s.h, s.spiky), mu)),
                             s.boundary.x[b].w));
    });
    Float3 out = add(ti, res);
    const uint8_t m = a.mask[i];
    if (m & 1u) out.x = 0.0F;
    if (m & 2u) out.y = 0.0F;
    if (m & 4u) out.z = 0.0F;
    store(a.p_t_out, i, out);
}

struct PostJvpArgs {
    Step s;
    TangentParams t;
    uint32_t slot;  // displacement slot of the positions read
    CVec pos, p_t, x_t, vrec, w, nrm, vc;
    const float *mag, *nz;
    Vec w_t, vc_t;
    float* mag_t;
    const uint32_t* perm;
    Vec next_x_t, next_v_t;  // original order
};

__device__ __forceinline__ Float3 vrec_tangent(const PostJvpArgs& a, uint32_t j) {
    return scale(sub(load(a.p_t, j), load(a.x_t, j)), a.s.inv_dt);
}

// reconstruct_jvp_kernel (pbf.cu:462-466), vorticity_jvp_kernel
// (pbf.cu:499-527) and magnitude_jvp_kernel (pbf.cu:595-599).
__global__ void __launch_bounds__(kBlock) pbf_jvp_vorticity_kernel(PostJvpArgs a) {
    PHX_MASK_LDS(a.s.grid);
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.s.n) return;
    const bool masked = masks_apply(a.s.mv, a.slot, i);
    const Step& s = a.s;
    const Float3 pi = load(a.pos, i), ti = load(a.p_t, i);
    const Float3 vi = load(a.vrec, i), vti = vrec_tangent(a, i);
    Float3 res{0.0F, 0.0F, 0.0F};
    int cx, cy, cz;
    cell_of_topology(s, i, cx, cy, cz);
    visit_candidates(s.grid, s.mv, masked, i, cx, cy, cz, delta, smask + threadIdx.x, kBlock, [&](uint32_t j) {
        if (j == i) return;
        const Float3 dd = sub(pi, load(a.pos, j));
        const Float3 dd_t = sub(ti, load(a.p_t, j));
        const Float3 diff = sub(load(a.vrec, j), vi);
        const Float3 diff_t = sub(vrec_tangent(a, j), vti);
        res = add(res, add(cross(diff_t, spiky_gradient(dd, s.h, s.spiky)),
                           cross(diff, spiky_hessian(dd, dd_t, s.h, s.spiky))));
    });
    store(a.w_t, i, res);
    a.mag_t[i] = a.mag[i] > 0.0F ? dot(load(a.w, i), res) / a.mag[i] : 0.0F;
}

// normal_jvp_kernel (pbf.cu:600-628) and confinement_jvp_kernel
// (pbf.cu:680-686).
__global__ void __launch_bounds__(kBlock) pbf_jvp_normal_kernel(PostJvpArgs a) {
    PHX_MASK_LDS(a.s.grid);
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.s.n) return;
    const bool masked = masks_apply(a.s.mv, a.slot, i);
    const Step& s = a.s;
    const Float3 pi = load(a.pos, i), ti = load(a.p_t, i);
    Float3 gt{0.0F, 0.0F, 0.0F};
    int cx, cy, cz;
    cell_of_topology(s, i, cx, cy, cz);
    visit_candidates(s.grid, s.mv, masked, i, cx, cy, cz, delta, smask + threadIdx.x, kBlock, [&](uint32_t j) {
        if (j == i) return;
        const Float3 dd = sub(pi, load(a.pos, j));
        const Float3 dd_t = sub(ti, load(a.p_t, j));
        gt = add(gt, add(scale(spiky_gradient(dd, s.h, s.spiky), a.mag_t[j] - a.mag_t[i]),
                         scale(spiky_hessian(dd, dd_t, s.h, s.spiky), a.mag[j] - a.mag[i])));
    });
    const Float3 nv = load(a.nrm, i);
    const Float3 ntan = a.nz[i] == 0.0F ? Float3{0.0F, 0.0F, 0.0F}
                                        : scale(add(gt, scale(nv, -dot(nv, gt))), 1.0F / a.nz[i]);
    const Float3 wv = load(a.w, i);
    const Float3 force = cross(nv, wv);
    const Float3 force_t = add(cross(ntan, wv), cross(nv, load(CVec{a.w_t.x, a.w_t.y, a.w_t.z}, i)));
    store(a.vc_t, i,
          add(vrec_tangent(a, i),
              scale(add(scale(force, a.t.confinement[i]), scale(force_t, s.confinement[i])), s.dt)));
}

// xsph_jvp_kernel (pbf.cu:723-754), written to original order together with
// the position tangent.
__global__ void __launch_bounds__(kBlock) pbf_jvp_xsph_kernel(PostJvpArgs a) {
    PHX_MASK_LDS(a.s.grid);
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.s.n) return;
    const bool masked = masks_apply(a.s.mv, a.slot, i);
    const Step& s = a.s;
    const CVec vct{a.vc_t.x, a.vc_t.y, a.vc_t.z};
    const Float3 pi = load(a.pos, i), ti = load(a.p_t, i);
    const Float3 vi = load(a.vc, i), vti = load(vct, i);
    Float3 sm{0.0F, 0.0F, 0.0F}, sm_t{0.0F, 0.0F, 0.0F};
    int cx, cy, cz;
    cell_of_topology(s, i, cx, cy, cz);
    visit_candidates(s.grid, s.mv, masked, i, cx, cy, cz, delta, smask + threadIdx.x, kBlock, [&](uint32_t j) {
        if (j == i) return;
        const Float3 dd = sub(pi, load(a.pos, j));
        const Float3 dd_t = sub(ti, load(a.p_t, j));
        const Float3 diff = sub(load(a.vc, j), vi);
        const Float3 diff_t = sub(load(vct, j), vti);
        const float wgt = poly6(dd, s.h);
        const float wgt_t = dot(poly6_gradient(dd, s.h), dd_t);
        sm = add(sm, scale(diff, wgt));
        sm_t = add(sm_t, add(scale(diff_t, wgt), scale(diff, wgt_t)));
    });
    const uint32_t o = a.perm[i];
    store(a.next_v_t, o, add(vti, add(scale(sm, a.t.xsph[i]), scale(sm_t, s.xsph[i]))));
    store(a.next_x_t, o, ti);
}

}  // namespace

void PbfDerivativeWorkspace::allocate(uint32_t n) {
    for (DeviceBuffer<float>* b : {&m, &rho0, &eps, &strength, &exponent, &radius, &xsph, &confinement, &m_t, &rho0_t,
                                   &eps_t, &strength_t, &exponent_t, &radius_t, &xsph_t, &confinement_t, &rho_t,
                                   &lambda_t, &magnitude_t})
        b->resize(n);
    for (VecBuffer* b : {&x_t, &v_t, &u_t, &p_t[0], &p_t[1], &vrec_t, &w_t, &vc_t}) b->resize(n);
    for (DVecBuffer* b : {&p_a[0], &p_a[1], &vrec_a, &w_a, &g_a, &next_x_a, &next_v_a, &x_a, &vel_a}) b->resize(n);
    for (DeviceBuffer<double>* b : {&lambda_a, &magnitude_a, &m_a, &rho0_a, &eps_a, &strength_a, &exponent_a, &radius_a,
                                    &xsph_a, &confinement_a, &inv_rho0})
        b->resize(n);
    lambda_terms.resize(n);
    vanishes.resize(n);
}

void PbfSolver::gather_parameters(const phx_pbf_parameters& parameters, const PbfCache& cache) {
    const auto& pp = parameters.particles;
    require(pp.masses && pp.rest_densities, PHX_ERROR_INVALID_ARGUMENT, "null particle parameters");
    require(parameters.relaxation && parameters.artificial_pressure_strength && parameters.artificial_pressure_exponent &&
                parameters.artificial_pressure_radius && parameters.xsph_viscosity && parameters.vorticity_confinement,
            PHX_ERROR_INVALID_ARGUMENT, "null PBF parameters");
    ChannelList ch;
    ch.add(pp.masses, dw_.m.data());
    ch.add(pp.rest_densities, dw_.rho0.data());
    ch.add(parameters.relaxation, dw_.eps.data());
    ch.add(parameters.artificial_pressure_strength, dw_.strength.data());
    ch.add(parameters.artificial_pressure_exponent, dw_.exponent.data());
    ch.add(parameters.artificial_pressure_radius, dw_.radius.data());
    ch.add(parameters.xsph_viscosity, dw_.xsph.data());
    ch.add(parameters.vorticity_confinement, dw_.confinement.data());
    gather_to_sorted(stream_, cache.neighborhood.perm.data(), cache.neighborhood.particles, ch.get());
}

namespace {

void require_recorded(const PbfCache& cache, phx_execution_mode mode, uint32_t n) {
    require(mode == PHX_MODE_DIFFERENTIABLE, PHX_ERROR_INVALID_ARGUMENT,
            "derivatives need a solver created in differentiable mode");
    require(cache.recorded && cache.neighborhood.particles == n, PHX_ERROR_INVALID_ARGUMENT,
            "the step cache holds no recorded forward step of this solver");
}

}  // namespace

namespace {
// The recorded exact masks of position set k (iterate k; the final positions
// are set I), which apply to every particle that is not crowded unless the
// set's overflow word is raised.
MaskView exact_view(const PbfCache& cache, uint32_t k, uint32_t n) {
    require(k <= cache.iterations, PHX_ERROR_INVALID_ARGUMENT, "exact mask set out of range");
    return MaskView{cache.exact_masks.data() + static_cast<size_t>(k) * kStencil * n, cache.exact_flags.data() + k, n,
                    0.0F, 1u, cache.crowded.data()};
}

// The recorded skin masks of the step and their displacement flags.
MaskView skin_view(const PbfCache& cache, uint32_t n) {
    return MaskView{cache.masks_skin.data(), cache.mask_flags.data(), n, cache.mask_limit2, 0u,
                    cache.crowded.data()};
}
}  // namespace

void PbfSolver::jvp_step(const phx_pbf_parameters& parameters, const PbfCache& cache,
                         const phx_pbf_state_tangent& state_t, const phx_control_tangent& control_t,
                         const phx_pbf_parameter_tangent& parameter_t, phx_pbf_state_tangent& next_t) {
    const auto& c = domain_.config;
    const uint32_t n = c.particle_count;
    if (n == 0) return;
    require_recorded(cache, mode_, n);
    require_vec(state_t.positions, "state position tangent");
    require_vec(state_t.velocities, "state velocity tangent");
    require_vec(control_t.external_accelerations, "control tangent");
    require_vec(next_t.positions, "next position tangent");
    require_vec(next_t.velocities, "next velocity tangent");
    PassScope scope(profiler_, stream_, kPbfJvp);
    dw_.allocate(n);
    gather_parameters(parameters, cache);
    // parameter tangents (null = zero)
    const float* tangents[8] = {parameter_t.particles.masses,         parameter_t.particles.rest_densities,
                                parameter_t.relaxation,               parameter_t.artificial_pressure_strength,
                                parameter_t.artificial_pressure_exponent, parameter_t.artificial_pressure_radius,
                                parameter_t.xsph_viscosity,           parameter_t.vorticity_confinement};
    float* sorted_t[8] = {dw_.m_t.data(),        dw_.rho0_t.data(),   dw_.eps_t.data(),  dw_.strength_t.data(),
                          dw_.exponent_t.data(), dw_.radius_t.data(), dw_.xsph_t.data(), dw_.confinement_t.data()};
    ChannelList ch;
    for (int k = 0; k < 8; ++k) {
        if (tangents[k] != nullptr) {
            ch.add(tangents[k], sorted_t[k]);
        } else {
            fill_f32(stream_, sorted_t[k], n, 0.0F);
        }
    }
    if (ch.get().count > 0) gather_to_sorted(stream_, cache.neighborhood.perm.data(), n, ch.get());

    const uint32_t blocks = (blocks_for(n, kBlock) + 7u) / 8u * 8u;
    const uint32_t iterations = cache.iterations;
    Step s{n,
           c.support_radius,
           spiky_,
           c.time_step,
           1.0F / c.time_step,
           cache.neighborhood.grid(),
           cache.xs.cview(),
           cache.x0.cview(),
           dw_.m.data(),
           dw_.rho0.data(),
           dw_.eps.data(),
           dw_.strength.data(),
           dw_.exponent.data(),
           dw_.radius.data(),
           dw_.xsph.data(),
           dw_.confinement.data(),
           skin_view(cache, n),
           nullptr,
           cache.neighborhood.boundary()};
    const TangentParams tp{dw_.m_t.data(),        dw_.rho0_t.data(),   dw_.eps_t.data(),  dw_.strength_t.data(),
                           dw_.exponent_t.data(), dw_.radius_t.data(), dw_.xsph_t.data(), dw_.confinement_t.data()};
    PredictJvpArgs pa{n,
                      c.time_step,
                      cache.neighborhood.perm.data(),
                      cvec(state_t.positions),
                      cvec(state_t.velocities),
                      cvec(control_t.external_accelerations),
                      dw_.x_t.view(),
                      dw_.p_t[0].view()};
    pbf_jvp_predict_kernel<<<blocks, kBlock, 0, stream_>>>(pa);
    PHX_LAUNCH_CHECK();
    int cur = 0;
    Step se = s;  // the passes at one set of positions walk its exact masks
    for (uint32_t it = 0; it < iterations; ++it) {
        const CVec pos = it == 0u ? cache.xs.cview() : cache.positions.cview(it - 1u);
        const size_t off = static_cast<size_t>(it) * n;
        se.mv = exact_view(cache, it, n);
        LambdaJvpArgs la{se, tp, it, pos, dw_.p_t[cur].cview(), cache.rho.data() + off, cache.den.data() + off,
                         dw_.lambda_t.data()};
        pbf_jvp_lambda_kernel<<<blocks, kBlock, 0, stream_>>>(la);
        PHX_LAUNCH_CHECK();
        CorrectionJvpArgs ca{se,
                             tp,
                             it,
                             pos,
                             dw_.p_t[cur].cview(),
                             cache.lambda.data() + off,
                             dw_.lambda_t.data(),
                             cache.masks.data() + off,
                             dw_.p_t[cur ^ 1].view()};
        pbf_jvp_correction_kernel<<<blocks, kBlock, 0, stream_>>>(ca);
        PHX_LAUNCH_CHECK();
        cur ^= 1;
    }
    PostJvpArgs post{};
    post.t = tp;
    post.slot = iterations;
    post.pos = iterations == 0u ? cache.xs.cview() : cache.positions.cview(iterations - 1u);
    se.mv = exact_view(cache, iterations, n);
    post.s = se;
    post.p_t = dw_.p_t[cur].cview();
    post.x_t = dw_.x_t.cview();
    post.vrec = cache.velocities.cview();
    post.w = cache.vorticities.cview();
    post.nrm = cache.normals.cview();
    post.vc = cache.confined.cview();
    post.mag = cache.magnitudes.data();
    post.nz = cache.normalizers.data();
    post.w_t = dw_.w_t.view();
    post.vc_t = dw_.vc_t.view();
    post.mag_t = dw_.magnitude_t.data();
    post.perm = cache.neighborhood.perm.data();
    post.next_x_t = vec(next_t.positions);
    post.next_v_t = vec(next_t.velocities);
    pbf_jvp_vorticity_kernel<<<blocks, kBlock, 0, stream_>>>(post);
    PHX_LAUNCH_CHECK();
    pbf_jvp_normal_kernel<<<blocks, kBlock, 0, stream_>>>(post);
    PHX_LAUNCH_CHECK();
    pbf_jvp_xsph_kernel<<<blocks, kBlock, 0, stream_>>>(post);
    PHX_LAUNCH_CHECK();
}

void PbfSolver::vjp_step(const phx_pbf_parameters& parameters, const PbfCache& cache,
                         const phx_pbf_state_adjoint& next_a, phx_pbf_state_adjoint& previous_a,
                         phx_control_adjoint& control_a, phx_pbf_parameter_adjoint& parameter_a) {
    const auto& c = domain_.config;
    const uint32_t n = c.particle_count;
    if (n == 0) return;
    require_recorded(cache, mode_, n);
    const auto& nx = next_a.positions;
    const auto& nv = next_a.velocities;
    require(nx.x && nx.y && nx.z && nv.x && nv.y && nv.z, PHX_ERROR_INVALID_ARGUMENT, "null next-state adjoint");
    PassScope scope(profiler_, stream_, kPbfVjp);
    dw_.allocate(n);
    gather_parameters(parameters, cache);
    const uint32_t blocks = (blocks_for(n, kBlock) + 7u) / 8u * 8u;
    const uint32_t iterations = cache.iterations;
    const uint32_t* perm = cache.neighborhood.perm.data();
    Step s{n,
           c.support_radius,
           spiky_,
           c.time_step,
           1.0F / c.time_step,
           cache.neighborhood.grid(),
           cache.xs.cview(),
           cache.x0.cview(),
           dw_.m.data(),
           dw_.rho0.data(),
           dw_.eps.data(),
           dw_.strength.data(),
           dw_.exponent.data(),
           dw_.radius.data(),
           dw_.xsph.data(),
           dw_.confinement.data(),
           skin_view(cache, n),
           dw_.inv_rho0.data(),
           cache.neighborhood.boundary()};
    // sorted parameter adjoints start at zero (+0.0 is the all-zero bit pattern)
    for (DeviceBuffer<double>* b : {&dw_.m_a, &dw_.rho0_a, &dw_.eps_a, &dw_.strength_a, &dw_.exponent_a,
                                    &dw_.radius_a, &dw_.xsph_a, &dw_.confinement_a})
        fill_u32(stream_, reinterpret_cast<uint32_t*>(b->data()), 2u * static_cast<size_t>(n), 0u);
    pbf_vjp_setup_kernel<<<blocks, kBlock, 0, stream_>>>(n, c.support_radius, dw_.strength.data(),
                                                          dw_.exponent.data(), dw_.radius.data(), dw_.rho0.data(),
                                                          dw_.vanishes.data(), dw_.inv_rho0.data());
    PHX_LAUNCH_CHECK();
    GatherAdjointArgs ga{n, perm, nx.x, nx.y, nx.z, nv.x, nv.y, nv.z, dw_.next_x_a.view(), dw_.next_v_a.view()};
    pbf_gather_adjoint_kernel<<<blocks, kBlock, 0, stream_>>>(ga);
    PHX_LAUNCH_CHECK();

    const CVec final_pos = iterations == 0u ? cache.xs.cview() : cache.positions.cview(iterations - 1u);
    int cur = 0;
    Step se = s;  // the passes at one set of positions walk its exact masks
    se.mv = exact_view(cache, iterations, n);
    PostVjpArgs pv{se,
                   iterations,
                   final_pos,
                   cache.confined.cview(),
                   cache.vorticities.cview(),
                   cache.normals.cview(),
                   cache.normalizers.data(),
                   dw_.next_x_a.view(),
                   dw_.next_v_a.view(),
                   dw_.p_a[cur].view(),
                   dw_.vrec_a.view(),
                   dw_.w_a.view(),
                   dw_.g_a.view(),
                   dw_.xsph_a.data(),
                   dw_.confinement_a.data()};
    pbf_vjp_post_kernel<<<blocks, kBlock, 0, stream_>>>(pv);
    PHX_LAUNCH_CHECK();
    NormalVjpArgs nva{se, iterations, final_pos, cache.vorticities.cview(), cache.magnitudes.data(), dw_.p_a[cur].view(),
                      dw_.w_a.view(), dw_.g_a.view()};
    pbf_vjp_normal_kernel<<<blocks, kBlock, 0, stream_>>>(nva);
    PHX_LAUNCH_CHECK();
    VorticityVjpArgs vva{se, iterations, final_pos, cache.velocities.cview(), dw_.p_a[cur].view(), dw_.vrec_a.view(),
                         dw_.w_a.view(), dw_.x_a.view()};
    pbf_vjp_vorticity_kernel<<<blocks, kBlock, 0, stream_>>>(vva);
    PHX_LAUNCH_CHECK();
    for (uint32_t it = iterations; it-- > 0u;) {
        const CVec pos = it == 0u ? cache.xs.cview() : cache.positions.cview(it - 1u);
        const size_t off = static_cast<size_t>(it) * n;
        se.mv = exact_view(cache, it, n);
        CorrectionVjpArgs cva{se,
                              it,
                              dw_.vanishes.data(),
                              parameter_a.artificial_pressure_strength != nullptr ? 1u : 0u,
                              pos,
                              cache.lambda.data() + off,
                              cache.masks.data() + off,
                              dw_.p_a[cur].view(),
                              dw_.p_a[cur ^ 1].view(),
                              dw_.lambda_a.data(),
                              dw_.m_a.data(),
                              dw_.rho0_a.data(),
                              dw_.strength_a.data(),
                              dw_.exponent_a.data(),
                              dw_.radius_a.data(),
                              cache.rho.data() + off,
                              cache.den.data() + off,
                              dw_.lambda_terms.data()};
        pbf_vjp_correction_kernel<<<blocks, kBlock, 0, stream_>>>(cva);
        PHX_LAUNCH_CHECK();
        LambdaVjpArgs lva{se,
                          it,
                          pos,
                          cache.rho.data() + off,
                          cache.den.data() + off,
                          cache.gradient_sums.cview(it),
                          dw_.lambda_a.data(),
                          dw_.lambda_terms.data(),
                          dw_.p_a[cur ^ 1].view(),
                          dw_.m_a.data(),
                          dw_.rho0_a.data(),
                          dw_.eps_a.data()};
        pbf_vjp_lambda_kernel<<<blocks, kBlock, 0, stream_>>>(lva);
        PHX_LAUNCH_CHECK();
        cur ^= 1;
    }
    ScatterVjpArgs sa{};
    sa.n = n;
    sa.dt = static_cast<double>(c.time_step);
    sa.dt2 = static_cast<double>(c.time_step * c.time_step);
    sa.perm = perm;
    sa.p_a = dw_.p_a[cur].view();
    sa.x_a = dw_.x_a.view();
    sa.prev_x = DVec{previous_a.positions.x, previous_a.positions.y, previous_a.positions.z};
    sa.prev_v = DVec{previous_a.velocities.x, previous_a.velocities.y, previous_a.velocities.z};
    sa.ctl = DVec{control_a.external_accelerations.x, control_a.external_accelerations.y,
                  control_a.external_accelerations.z};
    require((sa.prev_x.x == nullptr) == (sa.prev_x.y == nullptr) && (sa.prev_x.x == nullptr) == (sa.prev_x.z == nullptr),
            PHX_ERROR_INVALID_ARGUMENT, "position adjoint components must be all null or all set");
    require((sa.prev_v.x == nullptr) == (sa.prev_v.y == nullptr) && (sa.prev_v.x == nullptr) == (sa.prev_v.z == nullptr),
            PHX_ERROR_INVALID_ARGUMENT, "velocity adjoint components must be all null or all set");
    require((sa.ctl.x == nullptr) == (sa.ctl.y == nullptr) && (sa.ctl.x == nullptr) == (sa.ctl.z == nullptr),
            PHX_ERROR_INVALID_ARGUMENT, "control adjoint components must be all null or all set");
    const double* sorted[8] = {dw_.m_a.data(),        dw_.rho0_a.data(),   dw_.eps_a.data(),  dw_.strength_a.data(),
                               dw_.exponent_a.data(), dw_.radius_a.data(), dw_.xsph_a.data(), dw_.confinement_a.data()};
    double* original[8] = {parameter_a.particles.masses,
                           parameter_a.particles.rest_densities,
                           parameter_a.relaxation,
                           parameter_a.artificial_pressure_strength,
                           parameter_a.artificial_pressure_exponent,
                           parameter_a.artificial_pressure_radius,
                           parameter_a.xsph_viscosity,
                           parameter_a.vorticity_confinement};
    for (int k = 0; k < 8; ++k) {
        sa.sorted[k] = sorted[k];
        sa.original[k] = original[k];
    }
    pbf_vjp_scatter_kernel<<<blocks, kBlock, 0, stream_>>>(sa);
    PHX_LAUNCH_CHECK();
}

}  // namespace phx
