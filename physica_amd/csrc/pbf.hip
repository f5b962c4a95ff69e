// PBF forward step for gfx950 in sorted (cell-major) particle order.
//
// Reference: PBF::forward_step (pbf/pbf.cpp:167-191) and pbf/pbf.cu. Fusion:
//   grid       predict x* fused into the cell-key pass (pbf.cu:73-78)
//   mask       skin masks at x* (radius h (1 + skin)); every later pass walks
//              them while no particle moved more than half the skin from x*
//              (checked on the device from the correction passes' largest
//              displacement), and the full 27-cell stencil otherwise
//   per iter   A: density + lambda in one neighbour loop (density.cu:18-40,
//                 pbf.cu:95-127; the two sums keep their own order)
//              B: correction + project (pbf.cu:254-288, 417-431), written to
//                 a second position buffer so neighbours read the old values;
//                 the last iteration also reconstructs v = (p - x)/dt
//   post       vorticity + magnitude (pbf.cu:476-497, 565-569)
//              normal + confinement (pbf.cu:571-593, 674-678)
//              XSPH + scatter of the next state (pbf.cu:700-721)
// Per-particle sums run in the reference order, so each value is the IEEE
// result of the reference expression.
#include "pbf.hpp"

#include "neighbor_mask.hpp"

namespace phx {
namespace {

constexpr uint32_t kBlock = 256;
// Records a neighbour walk fetches ahead: two in a single scene's launch
// (too few waves to hide the loads of long crowded walks), none in a batch
// launch, which fills the GPU.
template <bool kBatch>
constexpr int kPbfWalkDepth = kBatch ? 0 : 2;

// LDS of a masked neighbour pass: the lanes' mask columns and the slot table.
#define PHX_MASK_LDS(grid)                          \
    __shared__ uint16_t smask[kStencil * kBlock];   \
    __shared__ int delta[kStencil];                 \
    fill_slot_deltas(grid, delta);                  \
    __syncthreads()

// The particle, grid and collision box of a lane: one scene (kBatch false),
// or in a batch the scene of the workgroup's block, whose grid starts at
// the scene's cell base in the batch's offsets.
struct Lane {
    uint32_t i;
    bool valid;
    SortedGrid grid;
    Box box;
};

template <bool kBatch>
__device__ __forceinline__ Lane lane_of(uint32_t n, const SortedGrid& grid, const Box& box, const BatchView& batch) {
    Lane l;
    if constexpr (kBatch) {
        const BatchBlock blk = batch.blocks[xcd_block()];
        const BatchScene& sc = batch.scenes[blk.scene];
        l.i = blk.first + threadIdx.x;
        l.valid = threadIdx.x < blk.count;
        // the lookup origin min + t v_box and the collision box of the step
        // (neighborhood.cpp:83-88, pbf.cpp:17-31), as the host computes them
        const float t = batch.time;
        l.grid = SortedGrid{grid.offsets + sc.cell_base,
                            sc.res[0],
                            sc.res[1],
                            sc.res[2],
                            sc.minimum[0] + t * sc.velocity[0],
                            sc.minimum[1] + t * sc.velocity[1],
                            sc.minimum[2] + t * sc.velocity[2],
                            batch.h};
        const float rp = sc.particle_radius;
        for (int a = 0; a < 3; ++a) {
            l.box.lo[a] = sc.minimum[a] + t * sc.velocity[a] + rp;
            l.box.hi[a] = sc.maximum[a] + t * sc.velocity[a] - rp;
            l.box.vel[a] = sc.velocity[a];
        }
        l.box.no_slip = sc.no_slip;
    } else {
        l.i = xcd_block() * kBlock + threadIdx.x;
        l.valid = l.i < n;
        l.grid = grid;
        l.box = box;
    }
    return l;
}

// The lane's mask view: in a batch the per-cell displacement bounds of the
// lane's scene start at its cell base, like its grid's offsets.
__device__ __forceinline__ MaskView lane_masks(const MaskView& mv, const SortedGrid& launch, const SortedGrid& lane) {
    MaskView v = mv;
    if (v.cell_moved != nullptr) v.cell_moved += lane.offsets - launch.offsets;
    return v;
}

struct MaskArgs {
    uint32_t n;
    SortedGrid grid;
    CVec xs;
    float t_skin;
    uint16_t* masks;
    uint8_t* crowded;
};

// Skin masks at x*; a particle with a stencil cell of more than kMaskBits
// particles is marked crowded and walks the full stencil in every pass of
// the step (the others keep their masks).
template <bool kBatch>
__global__ void __launch_bounds__(kBlock) pbf_mask_kernel(MaskArgs a, BatchView batch) {
    const Lane lane = lane_of<kBatch>(a.n, a.grid, Box{}, batch);
    const uint32_t i = lane.i;
    if (!lane.valid) return;
    const SortedGrid& grid = lane.grid;
    const Float3 p = load(a.xs, i);
    int cx, cy, cz;
    particle_cell(grid, p, cx, cy, cz);
    a.crowded[i] = build_skin_mask(grid, a.xs, a.n, i, p, cx, cy, cz, a.t_skin, a.masks) ? 0u : 1u;
}

// A particle that is not crowded always gets complete exact masks; the set's
// overflow word is raised only if that ever fails.
__device__ __forceinline__ void flag_incomplete(bool complete, const MaskView& mv, uint32_t i, uint32_t* set_flag) {
    if (!complete && (mv.crowded == nullptr || mv.crowded[i] == 0u)) atomicOr(set_flag, 1u);
}

struct LambdaArgs {
    uint32_t n;
    float h;
    float spiky;
    SortedGrid grid;
    CVec topo;
    CVec pos;
    const float* m;
    const float* rho0;
    const float* eps;
    float* rho;
    float* lambda;
    float* den;  // optional (recorded step)
    Vec gsum;    // optional (recorded step)
    MaskView mv;
    uint32_t slot;  // displacement slot of the positions read
    // exact masks of the positions read (pairs inside exact_mask_threshold),
    // for the correction pass and the derivatives (crowded particles walk the
    // full stencil instead)
    float t_exact;
    uint16_t* exact_out;
    uint32_t* exact_flags;
    BoundaryGrid boundary;  // offsets null: no boundary samples
};

// density_forward_kernel (density.cu:18-40, poly6) and lambda_forward_kernel
// (pbf.cu:95-127); the walk also writes the exact masks of the positions.
// With boundary samples the walk is the full stencil, each cell's samples b
// after its particles adding rho0_i V_b W to the density and V_b grad W to
// the gradient sum (pbf.cu:118-121); no exact masks are written (the host
// marks those sets unusable).
template <bool kBatch>
__global__ void __launch_bounds__(kBlock) pbf_lambda_kernel(LambdaArgs a, BatchView batch) {
    const Lane lane = lane_of<kBatch>(a.n, a.grid, Box{}, batch);
    PHX_MASK_LDS(lane.grid);
    const uint32_t i = lane.i;
    if (!lane.valid) return;
    const SortedGrid& grid = lane.grid;
    const MaskView mv = lane_masks(a.mv, a.grid, grid);
    const Float3 p = load(a.pos, i);
    const float rest = a.rho0[i];
    const Float3 x_star = load(a.topo, i);
    int cx, cy, cz;
    particle_cell(grid, x_star, cx, cy, cz);
    const bool masked =
        masks_apply(mv, a.slot, i, grid, cx, cy, cz, [&] { return displacement_bound(sub(p, x_star)); });
    float density = 0.0F;
    Float3 gradient_sum{0.0F, 0.0F, 0.0F};
    float denominator = a.eps[i];
    auto pair = [&](Float3 q, float ms, uint32_t s) {
        const Float3 d = sub(p, q);
        density += ms * poly6(d, a.h);
        if (s != i) {
            const Float3 g = spiky_gradient(d, a.h, a.spiky);
            const float coefficient = ms / rest;
            gradient_sum = add(gradient_sum, scale(g, coefficient));
            denominator += coefficient * coefficient * dot(g, g);
        }
        return dot(d, d) < a.t_exact;
    };
    auto visit = [&](uint32_t s) { return pair(load(a.pos, s), a.m[s], s); };
    if (a.boundary.offsets != nullptr) {
        for_each_candidate(
            grid, a.boundary, cx, cy, cz, [&](uint32_t s) { (void)visit(s); },
            [&](uint32_t s) {
                const float4 b = a.boundary.x[s];
                const Float3 d = sub(p, Float3{b.x, b.y, b.z});
                density += rest * b.w * poly6(d, a.h);
                gradient_sum = add(gradient_sum, scale(spiky_gradient(d, a.h, a.spiky), b.w));
            });
    } else {
        const bool complete = visit_records_and_mask<kPbfWalkDepth<kBatch>>(
            grid, mv, masked, i, cx, cy, cz, delta, smask + threadIdx.x, kBlock, a.exact_out, a.n,
            [&](uint32_t s) { return PointMass{load(a.pos, s), a.m[s]}; },
            [&](const PointMass& c, uint32_t s) { return pair(c.x, c.m, s); });
        flag_incomplete(complete, mv, i, &a.exact_flags[kFlagOverflow]);
    }
    denominator += dot(gradient_sum, gradient_sum);
    a.rho[i] = density;
    a.lambda[i] = -(density / rest - 1.0F) / denominator;
    if (a.den != nullptr) {
        a.den[i] = denominator;
        store(a.gsum, i, gradient_sum);
    }
}

struct CorrectionArgs {
    uint32_t n;
    float h;
    float spiky;
    SortedGrid grid;
    Box box;
    CVec topo;
    CVec pos;
    const float* m;
    const float* rho0;
    const float* lambda;
    const float* strength;
    const float* exponent;
    const float* radius;
    Vec out;
    uint8_t* mask;  // optional (recorded step): box-collision bits of the projection
    uint32_t reconstruct;
    CVec x0;
    float inv_dt;
    Vec vel;
    MaskView mv;
    uint32_t slot;           // displacement slot of the positions read
    uint32_t* moved_to;      // largest squared displacement of the output from x*
    uint32_t* cell_moved;    // optional: per-cell displacement bounds of the output (MaskView::cell_moved)
    BoundaryGrid boundary;   // offsets null: no boundary samples
};

// correction_forward_kernel (pbf.cu:254-288) and project_forward_kernel
// (pbf.cu:417-431), plus the largest displacement of the projected positions
// from x* for the mask checks of later passes. A boundary sample b adds
// V_b (lambda_i + s_corr(d)) grad W after each cell's particles
// (pbf.cu:278-284).
template <bool kBatch>
__global__ void __launch_bounds__(kBlock) pbf_correction_kernel(CorrectionArgs a, BatchView batch) {
    const Lane lane = lane_of<kBatch>(a.n, a.grid, a.box, batch);
    PHX_MASK_LDS(lane.grid);
    const uint32_t i = lane.i;
    const SortedGrid& grid = lane.grid;
    const Box& box = lane.box;
    float moved2 = 0.0F;
    uint32_t moved_cell = ~0u;
    float moved_bound = 0.0F;
    if (lane.valid) {
        const bool masked = masks_apply(a.mv, a.slot, i);
        const Float3 p = load(a.pos, i);
        const float rest = a.rho0[i];
        const float lambda_i = a.lambda[i];
        const float strength = a.strength[i];
        const float exponent = a.exponent[i];
        const float reference = poly6(Float3{a.radius[i], 0.0F, 0.0F}, a.h);
        // with k = +-0 the term is -k wherever W(d) > 0 (no powf per pair)
        const bool vanishes = artificial_vanishes(a.h, strength, exponent, reference);
        int cx, cy, cz;
        particle_cell(grid, load(a.topo, i), cx, cy, cz);
        Float3 correction{0.0F, 0.0F, 0.0F};
        auto artificial_at = [&](Float3 d) {
            return vanishes ? (poly6(d, a.h) == 0.0F ? 0.0F : -strength)
                            : artificial_value(d, a.h, strength, exponent, reference);
        };
        auto pair = [&](Float3 q, float ms, float lambda_s, uint32_t s) {
            if (s == i) return;
            const Float3 d = sub(p, q);
            const float artificial = artificial_at(d);
            const float coefficient = ms / rest;
            correction = add(correction, scale(spiky_gradient(d, a.h, a.spiky),
                                               coefficient * (lambda_i + lambda_s + artificial)));
        };
        auto visit = [&](uint32_t s) { pair(load(a.pos, s), a.m[s], a.lambda[s], s); };
        if (a.boundary.offsets != nullptr) {
            for_each_candidate(grid, a.boundary, cx, cy, cz, visit, [&](uint32_t s) {
                const float4 b = a.boundary.x[s];
                const Float3 d = sub(p, Float3{b.x, b.y, b.z});
                const float artificial = artificial_at(d);
                correction = add(correction, scale(spiky_gradient(d, a.h, a.spiky), b.w * (lambda_i + artificial)));
            });
        } else {
            visit_records<kPbfWalkDepth<kBatch>>(
                grid, a.mv, masked, i, cx, cy, cz, delta, smask + threadIdx.x, kBlock,
                [&](uint32_t s) { return PointMassLambda{load(a.pos, s), a.m[s], a.lambda[s]}; },
                [&](const PointMassLambda& c, uint32_t s) { pair(c.x, c.m, c.lambda, s); });
        }
        const Float3 candidate = add(p, correction);
        if (a.mask != nullptr) {
            uint32_t bits = 0u;
            if (candidate.x < box.lo[0] || candidate.x > box.hi[0]) bits |= 1u;
            if (candidate.y < box.lo[1] || candidate.y > box.hi[1]) bits |= 2u;
            if (candidate.z < box.lo[2] || candidate.z > box.hi[2]) bits |= 4u;
            a.mask[i] = static_cast<uint8_t>(bits);
        }
        const Float3 projected{fminf(box.hi[0], fmaxf(box.lo[0], candidate.x)),
                               fminf(box.hi[1], fmaxf(box.lo[1], candidate.y)),
                               fminf(box.hi[2], fmaxf(box.lo[2], candidate.z))};
        store(a.out, i, projected);
        if (a.reconstruct != 0u) store(a.vel, i, scale(sub(projected, load(a.x0, i)), a.inv_dt));
        const Float3 moved = sub(projected, load(a.topo, i));
        moved2 = dot(moved, moved);
        if (a.cell_moved != nullptr) {
            moved_cell = static_cast<uint32_t>(grid.offsets - a.grid.offsets) + cell_of(grid, cx, cy, cz);
            moved_bound = displacement_bound(moved);
        }
    }
    if (a.cell_moved != nullptr) raise_cell_bound(a.cell_moved, moved_cell, moved_bound);
    block_max_to<kBlock>(moved2, a.moved_to);
}

__global__ void __launch_bounds__(kBlock) pbf_reconstruct_kernel(uint32_t n, CVec pos, CVec x0, float inv_dt, Vec vel) {
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= n) return;
    store(vel, i, scale(sub(load(pos, i), load(x0, i)), inv_dt));
}

struct PostArgs {
    uint32_t n;
    float h;
    float spiky;
    float dt;
    SortedGrid grid;
    CVec topo;
    CVec pos;
    CVec vel;
    Vec vort;
    float* magnitude;
    const float* confinement;
    Vec vc;
    Vec normals;         // optional (recorded step)
    float* normalizers;  // optional (recorded step)
    const float* xsph;
    const uint32_t* perm;
    Vec next_x, next_v;
    MaskView mv;
    uint32_t slot;  // displacement slot of the final positions
    // exact masks of the final positions: written by the vorticity pass and
    // walked by the normal and XSPH passes (through mv, slot 0)
    float t_exact;
    uint16_t* exact_out;
    uint32_t* exact_flags;
};

template <bool kBatch>
__global__ void __launch_bounds__(kBlock) pbf_vorticity_kernel(PostArgs a, BatchView batch) {
    const Lane lane = lane_of<kBatch>(a.n, a.grid, Box{}, batch);
    PHX_MASK_LDS(lane.grid);
    const uint32_t i = lane.i;
    if (!lane.valid) return;
    const SortedGrid& grid = lane.grid;
    const MaskView mv = lane_masks(a.mv, a.grid, grid);
    const Float3 p = load(a.pos, i);
    const Float3 v = load(a.vel, i);
    const Float3 x_star = load(a.topo, i);
    int cx, cy, cz;
    particle_cell(grid, x_star, cx, cy, cz);
    const bool masked =
        masks_apply(mv, a.slot, i, grid, cx, cy, cz, [&] { return displacement_bound(sub(p, x_star)); });
    Float3 w{0.0F, 0.0F, 0.0F};
    auto pair = [&](Float3 q, Float3 vq, uint32_t s) {
        const Float3 d = sub(p, q);
        if (s != i) w = add(w, cross(sub(vq, v), spiky_gradient(d, a.h, a.spiky)));
        return dot(d, d) < a.t_exact;
    };
    const bool complete = visit_records_and_mask<kPbfWalkDepth<kBatch>>(
        grid, mv, masked, i, cx, cy, cz, delta, smask + threadIdx.x, kBlock, a.exact_out, a.n,
        [&](uint32_t s) { return PointVector{load(a.pos, s), load(a.vel, s)}; },
        [&](const PointVector& c, uint32_t s) { return pair(c.x, c.v, s); });
    flag_incomplete(complete, mv, i, &a.exact_flags[kFlagOverflow]);
    store(a.vort, i, w);
    a.magnitude[i] = length(w);
}

template <bool kBatch>
__global__ void __launch_bounds__(kBlock) pbf_normal_kernel(PostArgs a, BatchView batch) {
    const Lane lane = lane_of<kBatch>(a.n, a.grid, Box{}, batch);
    PHX_MASK_LDS(lane.grid);
    const uint32_t i = lane.i;
    if (!lane.valid) return;
    const SortedGrid& grid = lane.grid;
    const bool masked = masks_apply(a.mv, a.slot, i);
    const Float3 p = load(a.pos, i);
    const float mag_i = a.magnitude[i];
    int cx, cy, cz;
    particle_cell(grid, load(a.topo, i), cx, cy, cz);
    Float3 g{0.0F, 0.0F, 0.0F};
    auto pair = [&](Float3 q, float mag_s, uint32_t s) {
        if (s == i) return;
        g = add(g, scale(spiky_gradient(sub(p, q), a.h, a.spiky), mag_s - mag_i));
    };
    visit_records<kPbfWalkDepth<kBatch>>(
        grid, a.mv, masked, i, cx, cy, cz, delta, smask + threadIdx.x, kBlock,
        [&](uint32_t s) { return PointMass{load(a.pos, s), a.magnitude[s]}; },
        [&](const PointMass& c, uint32_t s) { pair(c.x, c.m, s); });
    const float normalizer = length(g);
    const Float3 normal = normalizer > 0.0F ? scale(g, 1.0F / normalizer) : Float3{0.0F, 0.0F, 0.0F};
    if (a.normalizers != nullptr) {
        a.normalizers[i] = normalizer;
        store(a.normals, i, normal);
    }
    const Float3 w = load(CVec{a.vort.x, a.vort.y, a.vort.z}, i);
    store(a.vc, i, add(load(a.vel, i), scale(cross(normal, w), a.dt * a.confinement[i])));
}

template <bool kBatch>
__global__ void __launch_bounds__(kBlock) pbf_xsph_kernel(PostArgs a, BatchView batch) {
    const Lane lane = lane_of<kBatch>(a.n, a.grid, Box{}, batch);
    PHX_MASK_LDS(lane.grid);
    const uint32_t i = lane.i;
    if (!lane.valid) return;
    const SortedGrid& grid = lane.grid;
    const bool masked = masks_apply(a.mv, a.slot, i);
    const CVec vc{a.vc.x, a.vc.y, a.vc.z};
    const Float3 p = load(a.pos, i);
    const Float3 v = load(vc, i);
    int cx, cy, cz;
    particle_cell(grid, load(a.topo, i), cx, cy, cz);
    Float3 smoothing{0.0F, 0.0F, 0.0F};
    auto pair = [&](Float3 q, Float3 vq, uint32_t s) {
        if (s == i) return;
        smoothing = add(smoothing, scale(sub(vq, v), poly6(sub(p, q), a.h)));
    };
    visit_records<kPbfWalkDepth<kBatch>>(
        grid, a.mv, masked, i, cx, cy, cz, delta, smask + threadIdx.x, kBlock,
        [&](uint32_t s) { return PointVector{load(a.pos, s), load(vc, s)}; },
        [&](const PointVector& c, uint32_t s) { pair(c.x, c.v, s); });
    const uint32_t j = a.perm[i];
    store(a.next_v, j, add(v, scale(smoothing, a.xsph[i])));
    store(a.next_x, j, p);
}

}  // namespace

const char* pbf_pass_name(int pass) {
    static const char* names[kPbfPassCount] = {"grid",      "mask",               "density_lambda",
                                               "correction_project", "vorticity", "normal_confinement",
                                               "xsph",      "jvp",                "vjp"};
    return pass >= 0 && pass < kPbfPassCount ? names[pass] : "unknown";
}

double pbf_pass_bytes(int pass, uint32_t particles, uint32_t cells) {
    const double offsets = particles > 0 ? 4.0 * (static_cast<double>(cells) + 1.0) / particles : 0.0;
    switch (pass) {
        case kPbfGrid:  // x, v, u in, x* out (original order); x*, x + 8 parameters gathered; permutation
            return 36.0 + 12.0 + 2.0 * (24.0 + 32.0) + 4.0 + 2.0 * offsets;
        case kPbfMask:  // x* in (the masks are an acceleration structure)
            return 12.0 + offsets;
        case kPbfLambda:  // x*, p, m, rho0, eps in; rho, lambda out
            return 36.0 + 8.0 + offsets;
        case kPbfCorrection:  // x*, p, m, rho0, lambda, s_corr k, n, dq in; p' out (+ x in, v out on the last)
            return 48.0 + 12.0 + offsets;
        case kPbfVorticity:  // x*, p, v in; omega, |omega| out
            return 36.0 + 16.0 + offsets;
        case kPbfNormal:  // x*, p, |omega|, omega, v, confinement in; v_c out
            return 56.0 + 12.0 + offsets;
        case kPbfXsph:  // x*, p, v_c, xsph in; next x, v out + permutation
            return 40.0 + 24.0 + 4.0 + offsets;
        case kPbfJvp:  // whole-step passes, not a single roofline figure
        case kPbfVjp:
            return 0.0;
        default:
            return 0.0;
    }
}

void PbfCache::allocate(uint32_t n, uint32_t cells, uint32_t pressure_iterations) {
    neighborhood.allocate(n, cells);
    predicted.resize(n);
    densities.resize(n);
    lambdas.resize(n);
    magnitudes.resize(n);  // StepCache::vorticity_magnitudes (pbf/pbf.ixx:67)
    if (!record) return;
    const size_t per = static_cast<size_t>(pressure_iterations) * n;
    for (VecBuffer* b : {&xs, &x0, &velocities, &vorticities, &normals, &confined}) b->resize(n);
    positions.resize(n, pressure_iterations);
    gradient_sums.resize(n, pressure_iterations);
    rho.resize(per);
    lambda.resize(per);
    den.resize(per);
    masks.resize(per);
    normalizers.resize(n);
    masks_skin.resize(static_cast<size_t>(kStencil) * n);
    crowded.resize(n);
    mask_flags.resize(kFlagDisplacement + pressure_iterations + 1u);
    exact_masks.resize(static_cast<size_t>(pressure_iterations + 1u) * kStencil * n);
    exact_flags.resize(pressure_iterations + 1u);
}

PbfSolver::PbfSolver(const phx_domain_configuration* domain, const phx_pbf_configuration* configuration,
                     phx_execution_mode mode, hipStream_t stream)
    : domain_(make_domain(domain)), mode_(mode), stream_(stream) {
    require(configuration != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null PBF configuration");

    config_ = *configuration;
    require(config_.checkpoint_interval > 0u, PHX_ERROR_INVALID_ARGUMENT, "checkpoint_interval must be positive");
    require(mode == PHX_MODE_FORWARD || mode == PHX_MODE_DIFFERENTIABLE, PHX_ERROR_INVALID_ARGUMENT,
            "unknown execution mode");
    boundary_.upload(domain_.config);
    domain_.config.boundary_particles = nullptr;  // the caller's host array is not kept
    spiky_ = spiky_coefficient(domain_.config.support_radius);
    const uint32_t n = domain_.config.particle_count;
    for (VecBuffer* b : {&xs_, &x0_, &pa_, &pb_, &vel_, &vort_, &vc_}) b->resize(n);
    for (DeviceBuffer<float>* b : {&m_, &rho0_, &eps_, &strength_, &exponent_, &radius_, &xsph_, &confinement_})
        b->resize(n);
    exact_.resize(static_cast<size_t>(kStencil) * n);
    crowded_.resize(n);
    exact_flags_.resize(config_.pressure_iterations + 1u);
    // displacement slots: one per iterate (the first, x* itself, stays zero)
    masks_.allocate(n, config_.pressure_iterations + 1u, domain_.config.support_radius);
}

void PbfSolver::forward_step(const phx_pbf_state& state, const phx_control& control,
                             const phx_pbf_parameters& parameters, phx_pbf_state& next, PbfCache& cache) {
    const auto& c = domain_.config;
    const uint32_t n = c.particle_count;
    const uint64_t step = state.particles.step_index;
    if (n == 0) {
        next.particles.step_index = step + 1u;
        return;
    }
    require_vec(state.particles.positions, "state positions");
    require_vec(state.particles.velocities, "state velocities");
    require_vec(control.external_accelerations, "control");
    require_vec(next.particles.positions, "next positions");
    require_vec(next.particles.velocities, "next velocities");
    const auto& pp = parameters.particles;
    require(pp.masses && pp.rest_densities, PHX_ERROR_INVALID_ARGUMENT, "null particle parameters");
    require(parameters.relaxation && parameters.artificial_pressure_strength && parameters.artificial_pressure_exponent &&
                parameters.artificial_pressure_radius && parameters.xsph_viscosity && parameters.vorticity_confinement,
            PHX_ERROR_INVALID_ARGUMENT, "null PBF parameters");
    cache.record = mode_ == PHX_MODE_DIFFERENTIABLE;
    cache.recorded = false;
    cache.allocate(n, domain_.cells, config_.pressure_iterations);
    const bool rec = cache.record;
    // recorded quantities live in the cache, the others in the workspace
    VecBuffer& xs = rec ? cache.xs : xs_;
    VecBuffer& x0 = rec ? cache.x0 : x0_;
    VecBuffer& vel = rec ? cache.velocities : vel_;
    VecBuffer& vort = rec ? cache.vorticities : vort_;
    VecBuffer& vc = rec ? cache.confined : vc_;
    float* magnitude = cache.magnitudes.data();

    // predict (pbf.cu:73-78) fused into the key pass of build(step + 1, x*)
    PredictInput predict{};
    predict.velocities = cvec(state.particles.velocities);
    predict.controls = cvec(control.external_accelerations);
    predict.gravity = Float3{c.gravity.x, c.gravity.y, c.gravity.z};
    predict.dt = c.time_step;
    predict.predicted = cache.predicted.view();
    ChannelList gather;
    gather.add(cache.predicted.cview(), xs.view());
    gather.add(cvec(state.particles.positions), x0.view());
    gather.add(pp.masses, m_.data());
    gather.add(pp.rest_densities, rho0_.data());
    gather.add(parameters.relaxation, eps_.data());
    gather.add(parameters.artificial_pressure_strength, strength_.data());
    gather.add(parameters.artificial_pressure_exponent, exponent_.data());
    gather.add(parameters.artificial_pressure_radius, radius_.data());
    gather.add(parameters.xsph_viscosity, xsph_.data());
    gather.add(parameters.vorticity_confinement, confinement_.data());
    {
        PassScope scope(profiler_, stream_, kPbfGrid);
        cache.neighborhood.build(stream_, domain_, step + 1u, cvec(state.particles.positions), &predict,
                                 gather.get(), &boundary_);
        masks_.begin_step(stream_);
    }

    const SortedGrid grid = cache.neighborhood.grid();
    const BoundaryGrid boundary = cache.neighborhood.boundary();
    const bool with_boundary = boundary.offsets != nullptr;
    // whole blocks of kBlock particles, padded to a multiple of the XCD count
    const uint32_t blocks = (blocks_for(n, kBlock) + 7u) / 8u * 8u;
    const float h = c.support_radius;
    const float inv_dt = 1.0F / c.time_step;
    // a recorded step keeps its masks and flags for its derivative passes
    uint16_t* skin = rec ? cache.masks_skin.data() : masks_.skin_mask();
    uint8_t* crowded = rec ? cache.crowded.data() : crowded_.data();
    {
        PassScope scope(profiler_, stream_, kPbfMask);
        MaskArgs ma{n, grid, xs.cview(), masks_.t_skin(), skin, crowded};
        pbf_mask_kernel<false><<<blocks, kBlock, 0, stream_>>>(ma, BatchView{});
        PHX_LAUNCH_CHECK();
    }
    // the skin adapts to the largest displacement of the previous step, so
    // the step-wide limit holds nearly always: no per-cell bounds (the
    // batch's fixed skin uses them)
    MaskView mv{skin, masks_.flags(), n, masks_.limit2()};
    mv.crowded = crowded;
    // exact masks of position set k (iterate k, the final positions last):
    // written by the pass that first walks those positions, walked by the
    // later passes at them; a recorded step keeps every set. Each set has one
    // overflow word.
    const float t_exact = exact_mask_threshold(h);
    const uint32_t sets = config_.pressure_iterations + 1u;
    uint32_t* exact_flags = rec ? cache.exact_flags.data() : exact_flags_.data();
    fill_u32(stream_, exact_flags, sets, 0u);
    // with boundary samples the iterations' passes walk the full stencil and
    // write no exact masks; the final positions' set (vorticity, normal,
    // XSPH: no boundary terms) is built as usual
    if (with_boundary && sets > 1u) fill_u32(stream_, exact_flags, sets - 1u, 1u);
    auto exact_set = [&](uint32_t k) {
        require(k < sets, PHX_ERROR_INVALID_ARGUMENT, "exact mask set out of range");
        return rec ? cache.exact_masks.data() + static_cast<size_t>(k) * kStencil * n : exact_.data();
    };
    auto exact_view = [&](uint32_t k) { return MaskView{exact_set(k), exact_flags + k, n, 0.0F, 1u, crowded}; };

    LambdaArgs la{n,           h,        spiky_,     grid, xs.cview(), xs.cview(), m_.data(), rho0_.data(),
                  eps_.data(), cache.densities.data(), cache.lambdas.data(), nullptr, Vec{}, mv, 0u};
    la.t_exact = t_exact;
    la.boundary = boundary;
    CorrectionArgs ca{};
    ca.n = n;
    ca.h = h;
    ca.spiky = spiky_;
    ca.grid = grid;
    ca.box = collision_box(c, step + 1u);
    ca.topo = xs.cview();
    ca.m = m_.data();
    ca.rho0 = rho0_.data();
    ca.lambda = cache.lambdas.data();
    ca.strength = strength_.data();
    ca.exponent = exponent_.data();
    ca.radius = radius_.data();
    ca.x0 = x0.cview();
    ca.inv_dt = inv_dt;
    ca.vel = vel.view();
    ca.mv = mv;
    ca.boundary = boundary;

    // positions start at x* (pbf.cpp:171) and ping-pong between two buffers
    // (a recorded step keeps every iterate)
    CVec current = xs.cview();
    VecBuffer* targets[2] = {&pa_, &pb_};
    const uint32_t iterations = config_.pressure_iterations;
    for (uint32_t it = 0; it < iterations; ++it) {
        la.pos = current;
        la.slot = it;
        la.exact_out = exact_set(it);
        la.exact_flags = exact_flags + it;
        ca.mv = exact_view(it);
        ca.slot = 0u;
        ca.moved_to = masks_.flags() + kFlagDisplacement + it + 1u;
        if (rec) {
            const size_t off = static_cast<size_t>(it) * n;
            la.rho = cache.rho.data() + off;
            la.lambda = cache.lambda.data() + off;
            la.den = cache.den.data() + off;
            la.gsum = cache.gradient_sums.view(it);
            ca.lambda = la.lambda;
            ca.mask = cache.masks.data() + off;
        }
        {
            PassScope scope(profiler_, stream_, kPbfLambda);
            pbf_lambda_kernel<false><<<blocks, kBlock, 0, stream_>>>(la, BatchView{});
            PHX_LAUNCH_CHECK();
        }
        const Vec target = rec ? cache.positions.view(it) : targets[it & 1u]->view();
        ca.pos = current;
        ca.out = target;
        ca.reconstruct = it + 1u == iterations ? 1u : 0u;
        {
            PassScope scope(profiler_, stream_, kPbfCorrection);
            pbf_correction_kernel<false><<<blocks, kBlock, 0, stream_>>>(ca, BatchView{});
            PHX_LAUNCH_CHECK();
        }
        current = CVec{target.x, target.y, target.z};
    }
    if (rec && iterations > 0u) {
        // the cache's densities and lambdas are those of the last iteration
        const size_t off = static_cast<size_t>(iterations - 1u) * n;
        PHX_HIP(hipMemcpyAsync(cache.densities.data(), cache.rho.data() + off, sizeof(float) * n,
                               hipMemcpyDeviceToDevice, stream_));
        PHX_HIP(hipMemcpyAsync(cache.lambdas.data(), cache.lambda.data() + off, sizeof(float) * n,
                               hipMemcpyDeviceToDevice, stream_));
    }
    if (iterations == 0u) {
        pbf_reconstruct_kernel<<<blocks, kBlock, 0, stream_>>>(n, current, x0.cview(), inv_dt, vel.view());
        PHX_LAUNCH_CHECK();
    }

    PostArgs post{};
    post.n = n;
    post.h = h;
    post.spiky = spiky_;
    post.dt = c.time_step;
    post.grid = grid;
    post.topo = xs.cview();
    post.pos = current;
    post.vel = vel.cview();
    post.vort = vort.view();
    post.magnitude = magnitude;
    post.confinement = confinement_.data();
    post.vc = vc.view();
    if (rec) {
        post.normals = cache.normals.view();
        post.normalizers = cache.normalizers.data();
    }
    post.xsph = xsph_.data();
    post.perm = cache.neighborhood.perm.data();
    post.mv = mv;
    post.slot = iterations;
    post.t_exact = t_exact;
    post.exact_out = exact_set(iterations);
    post.exact_flags = exact_flags + iterations;
    post.next_x = vec(next.particles.positions);
    post.next_v = vec(next.particles.velocities);
    // the normal and XSPH passes: every field of `post`, with the exact masks
    PostArgs post_exact = post;
    post_exact.mv = exact_view(iterations);
    post_exact.slot = 0u;
    {
        PassScope scope(profiler_, stream_, kPbfVorticity);
        pbf_vorticity_kernel<false><<<blocks, kBlock, 0, stream_>>>(post, BatchView{});
        PHX_LAUNCH_CHECK();
    }
    {
        PassScope scope(profiler_, stream_, kPbfNormal);
        pbf_normal_kernel<false><<<blocks, kBlock, 0, stream_>>>(post_exact, BatchView{});
        PHX_LAUNCH_CHECK();
    }
    {
        PassScope scope(profiler_, stream_, kPbfXsph);
        pbf_xsph_kernel<false><<<blocks, kBlock, 0, stream_>>>(post_exact, BatchView{});
        PHX_LAUNCH_CHECK();
    }
    if (rec) {
        PHX_HIP(hipMemcpyAsync(cache.mask_flags.data(), masks_.flags(), sizeof(uint32_t) * masks_.flag_count(),
                               hipMemcpyDeviceToDevice, stream_));
        cache.mask_limit2 = masks_.limit2();
        cache.recorded = true;
        cache.iterations = iterations;
        cache.step_index = step;
    }
    masks_.end_step(stream_);
    next.particles.step_index = step + 1u;
}

}  // namespace phx

namespace phx {

// ------------------------------------------------------------------ batch
PbfBatch::PbfBatch(const phx_domain_configuration* domains, uint32_t scene_count,
                   const phx_pbf_configuration* configuration, const phx_pbf_batch_configuration* batch_configuration,
                   hipStream_t stream)
    : caller_(stream) {
    require(domains != nullptr && scene_count > 0u, PHX_ERROR_INVALID_ARGUMENT, "a batch needs at least one scene");
    require(configuration != nullptr, PHX_ERROR_INVALID_ARGUMENT, "null PBF configuration");
    config_ = *configuration;
    require(config_.pressure_iterations < 1024u, PHX_ERROR_INVALID_ARGUMENT, "too many pressure iterations");
    if (batch_configuration != nullptr) {
        const float skin = batch_configuration->mask_skin;
        require(skin == 0.0F || (skin >= 0.01F && skin <= 1.0F), PHX_ERROR_INVALID_ARGUMENT,
                "mask_skin must be 0 (default) or in [0.01, 1]");
        if (skin > 0.0F) skin_ = skin;
        use_graph_ = batch_configuration->use_graph != 0u;
    }
    uint64_t particles = 0, cells = 0;
    offsets_.push_back(0u);
    for (uint32_t k = 0; k < scene_count; ++k) {
        const Domain d = make_domain(&domains[k]);
        const auto& c = d.config;
        require(c.boundary_particle_count == 0u, PHX_ERROR_UNSUPPORTED, "batched scenes take no boundary particles");
        if (k == 0u) {
            h_ = c.support_radius;
            dt_ = c.time_step;
        }
        require(c.support_radius == h_ && c.time_step == dt_, PHX_ERROR_INVALID_ARGUMENT,
                "the scenes of a batch share the support radius and time step");
        BatchScene s{};
        s.cell_base = static_cast<uint32_t>(cells);
        for (int a = 0; a < 3; ++a) s.res[a] = d.resolution[a];
        s.minimum[0] = c.boundary.minimum.x;
        s.minimum[1] = c.boundary.minimum.y;
        s.minimum[2] = c.boundary.minimum.z;
        s.maximum[0] = c.boundary.maximum.x;
        s.maximum[1] = c.boundary.maximum.y;
        s.maximum[2] = c.boundary.maximum.z;
        s.velocity[0] = c.boundary.velocity.x;
        s.velocity[1] = c.boundary.velocity.y;
        s.velocity[2] = c.boundary.velocity.z;
        s.particle_radius = c.particle_radius;
        s.gravity[0] = c.gravity.x;
        s.gravity[1] = c.gravity.y;
        s.gravity[2] = c.gravity.z;
        s.no_slip = c.boundary.no_slip != 0u ? 1u : 0u;
        if (s.velocity[0] != 0.0F || s.velocity[1] != 0.0F || s.velocity[2] != 0.0F) static_boxes_ = false;
        scenes_.push_back(s);
        cells += d.cells;
        particles += c.particle_count;
        offsets_.push_back(particles);
    }
    require(particles < (1ull << 31) && cells < (1ull << 31), PHX_ERROR_CAPACITY, "batch too large");
    n_ = static_cast<uint32_t>(particles);
    cells_ = static_cast<uint32_t>(cells);
    // a captured substep bakes in the step time, which moving boxes change
    use_graph_ = use_graph_ && static_boxes_;
    spiky_ = spiky_coefficient(h_);

    // workgroup table: each scene's particles in blocks of kBlock, padded to
    // a multiple of the XCD count
    std::vector<BatchBlock> table;
    for (uint32_t k = 0; k < scene_count; ++k) {
        for (uint64_t first = offsets_[k]; first < offsets_[k + 1]; first += kBlock) {
            const uint64_t count = std::min<uint64_t>(kBlock, offsets_[k + 1] - first);
            table.push_back(BatchBlock{k, static_cast<uint32_t>(first), static_cast<uint32_t>(count), 0u});
        }
    }
    while (table.empty() || table.size() % 8u != 0u) table.push_back(BatchBlock{0u, 0u, 0u, 0u});
    blocks_ = static_cast<uint32_t>(table.size());
    block_table_.resize(table.size());
    scene_table_.resize(scenes_.size());
    PHX_HIP(hipMemcpy(block_table_.data(), table.data(), sizeof(BatchBlock) * table.size(), hipMemcpyHostToDevice));
    PHX_HIP(hipMemcpy(scene_table_.data(), scenes_.data(), sizeof(BatchScene) * scenes_.size(),
                      hipMemcpyHostToDevice));

    const uint32_t n = n_;
    grid_.allocate(n, cells_);
    cell_moved_.resize(cells_);
    for (VecBuffer* b : {&positions_[0], &positions_[1], &velocities_[0], &velocities_[1], &predicted_, &xs_, &x0_,
                         &pa_, &pb_, &vel_, &vort_, &vc_})
        b->resize(n);
    for (DeviceBuffer<float>* b :
         {&m_, &rho0_, &eps_, &strength_, &exponent_, &radius_, &xsph_, &confinement_, &rho_, &lambda_, &magnitude_})
        b->resize(n);
    skin_masks_.resize(static_cast<size_t>(kStencil) * n);
    exact_.resize(static_cast<size_t>(kStencil) * n);
    crowded_.resize(n);
    flags_.resize(kFlagDisplacement + config_.pressure_iterations + 1u);
    exact_flags_.resize(config_.pressure_iterations + 1u);
    // grid build 9, flag and cell-bound fills 3, masks 1, two per iteration, three post passes
    const uint32_t it = config_.pressure_iterations;
    kernels_per_step_ = 9u + 3u + 1u + 2u * it + (it == 0u ? 1u : 0u) + 3u;
    PHX_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    PHX_HIP(hipEventCreateWithFlags(&ready_, hipEventDisableTiming));
    PHX_HIP(hipEventCreateWithFlags(&done_, hipEventDisableTiming));
}

PbfBatch::~PbfBatch() {
    if (stream_ != nullptr) (void)hipStreamSynchronize(stream_);
    release_graphs();
    if (ready_ != nullptr) (void)hipEventDestroy(ready_);
    if (done_ != nullptr) (void)hipEventDestroy(done_);
    if (stream_ != nullptr) (void)hipStreamDestroy(stream_);
}

void PbfBatch::release_graphs() {
    for (hipGraphExec_t& g : graph_) {
        if (g != nullptr) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    captured_inputs_.clear();
}

// One forward step of every scene from state_[from] into state_[1 - from]:
// the pass sequence of PbfSolver::forward_step in forward mode, each pass
// one launch over the batch's workgroup table.
void PbfBatch::substep(uint32_t from, uint64_t step, const phx_control& control,
                       const phx_pbf_parameters& parameters) {
    const uint32_t n = n_;
    const hipStream_t s = stream_;
    const VecBuffer& x = positions_[from];
    const VecBuffer& v = velocities_[from];
    VecBuffer& next_x = positions_[1u - from];
    VecBuffer& next_v = velocities_[1u - from];
    const auto& pp = parameters.particles;
    BatchView bv{block_table_.data(), scene_table_.data(), static_cast<float>(step + 1u) * dt_, h_};
    kernels_per_step_ = 0u;

    PredictInput predict{};
    predict.velocities = v.cview();
    predict.controls = cvec(control.external_accelerations);
    predict.dt = dt_;
    predict.predicted = predicted_.view();
    ChannelList gather;
    gather.add(predicted_.cview(), xs_.view());
    gather.add(x.cview(), x0_.view());
    gather.add(pp.masses, m_.data());
    gather.add(pp.rest_densities, rho0_.data());
    gather.add(parameters.relaxation, eps_.data());
    gather.add(parameters.artificial_pressure_strength, strength_.data());
    gather.add(parameters.artificial_pressure_exponent, exponent_.data());
    gather.add(parameters.artificial_pressure_radius, radius_.data());
    gather.add(parameters.xsph_viscosity, xsph_.data());
    gather.add(parameters.vorticity_confinement, confinement_.data());
    build_batch_grid(s, n, cells_, blocks_, bv, x.cview(), predict, gather.get(), grid_.scratch, grid_.perm.data(),
                     grid_.offsets.data());
    kernels_per_step_ += 9u;  // counts fill, key, three scan kernels, scatter, rank, big-cell sort, gather
    const uint32_t sets = config_.pressure_iterations + 1u;
    fill_u32(s, flags_.data(), kFlagDisplacement + sets, 0u);
    fill_u32(s, exact_flags_.data(), sets, 0u);
    fill_u32(s, cell_moved_.data(), cells_, 0u);
    kernels_per_step_ += 3u;

    // the batch's offsets; each lane's grid starts at its scene's cell base
    const SortedGrid grid{grid_.offsets.data(), 0u, 0u, 0u, 0.0F, 0.0F, 0.0F, h_};
    const float radius = h_ * (1.0F + skin_);
    const float t_skin = radius * radius * (1.0F + 1e-5F);
    const float limit = 0.5F * (radius - h_ * (1.0F + 2e-5F));
    const float limit2 = limit > 0.0F ? limit * limit : -1.0F;
    MaskArgs ma{n, grid, xs_.cview(), t_skin, skin_masks_.data(), crowded_.data()};
    pbf_mask_kernel<true><<<blocks_, kBlock, 0, s>>>(ma, bv);
    PHX_LAUNCH_CHECK();
    ++kernels_per_step_;
    MaskView mv{skin_masks_.data(), flags_.data(), n, limit2};
    mv.crowded = crowded_.data();
    mv.cell_moved = cell_moved_.data();
    mv.pair_limit = 2.0F * limit;
    const MaskView exact_view{exact_.data(), exact_flags_.data(), n, 0.0F, 1u, crowded_.data()};
    const float t_exact = exact_mask_threshold(h_);

    LambdaArgs la{n,           h_,         spiky_,        grid, xs_.cview(), xs_.cview(), m_.data(), rho0_.data(),
                  eps_.data(), rho_.data(), lambda_.data(), nullptr, Vec{},       mv,          0u};
    la.t_exact = t_exact;
    la.exact_out = exact_.data();
    CorrectionArgs ca{};
    ca.n = n;
    ca.h = h_;
    ca.spiky = spiky_;
    ca.grid = grid;
    ca.topo = xs_.cview();
    ca.m = m_.data();
    ca.rho0 = rho0_.data();
    ca.lambda = lambda_.data();
    ca.strength = strength_.data();
    ca.exponent = exponent_.data();
    ca.radius = radius_.data();
    ca.x0 = x0_.cview();
    ca.inv_dt = 1.0F / dt_;
    ca.vel = vel_.view();
    ca.mv = exact_view;
    ca.slot = 0u;
    ca.cell_moved = cell_moved_.data();
    // the exact masks of each set of positions are rebuilt by the pass that
    // first walks them, so one buffer serves every set in forward mode
    CVec current = xs_.cview();
    VecBuffer* targets[2] = {&pa_, &pb_};
    const uint32_t iterations = config_.pressure_iterations;
    for (uint32_t it = 0; it < iterations; ++it) {
        la.pos = current;
        la.slot = it;
        la.exact_flags = exact_flags_.data() + it;
        ca.mv.flags = exact_flags_.data() + it;
        ca.moved_to = flags_.data() + kFlagDisplacement + it + 1u;
        pbf_lambda_kernel<true><<<blocks_, kBlock, 0, s>>>(la, bv);
        PHX_LAUNCH_CHECK();
        const Vec target = targets[it & 1u]->view();
        ca.pos = current;
        ca.out = target;
        ca.reconstruct = it + 1u == iterations ? 1u : 0u;
        pbf_correction_kernel<true><<<blocks_, kBlock, 0, s>>>(ca, bv);
        PHX_LAUNCH_CHECK();
        kernels_per_step_ += 2u;
        current = CVec{target.x, target.y, target.z};
    }
    if (iterations == 0u) {
        pbf_reconstruct_kernel<<<blocks_for(n, kBlock), kBlock, 0, s>>>(n, current, x0_.cview(), 1.0F / dt_,
                                                                      vel_.view());
        PHX_LAUNCH_CHECK();
        ++kernels_per_step_;
    }
    PostArgs post{};
    post.n = n;
    post.h = h_;
    post.spiky = spiky_;
    post.dt = dt_;
    post.grid = grid;
    post.topo = xs_.cview();
    post.pos = current;
    post.vel = vel_.cview();
    post.vort = vort_.view();
    post.magnitude = magnitude_.data();
    post.confinement = confinement_.data();
    post.vc = vc_.view();
    post.xsph = xsph_.data();
    post.perm = grid_.perm.data();
    post.mv = mv;
    post.slot = iterations;
    post.t_exact = t_exact;
    post.exact_out = exact_.data();
    post.exact_flags = exact_flags_.data() + iterations;
    post.next_x = next_x.view();
    post.next_v = next_v.view();
    PostArgs post_exact = post;
    post_exact.mv = exact_view;
    post_exact.mv.flags = exact_flags_.data() + iterations;
    post_exact.slot = 0u;
    pbf_vorticity_kernel<true><<<blocks_, kBlock, 0, s>>>(post, bv);
    PHX_LAUNCH_CHECK();
    pbf_normal_kernel<true><<<blocks_, kBlock, 0, s>>>(post_exact, bv);
    PHX_LAUNCH_CHECK();
    pbf_xsph_kernel<true><<<blocks_, kBlock, 0, s>>>(post_exact, bv);
    PHX_LAUNCH_CHECK();
    kernels_per_step_ += 3u;
}
}  // namespace phx

namespace phx {

void PbfBatch::capture(const phx_control& control, const phx_pbf_parameters& parameters) {
    release_graphs();
    // one eager substep first, so every kernel is loaded before the capture
    // (it writes only the batch's own state buffers, which the replays overwrite)
    substep(0u, 0u, control, parameters);
    for (uint32_t from = 0; from < 2u; ++from) {
        hipGraph_t graph = nullptr;
        PHX_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        try {
            substep(from, 0u, control, parameters);
        } catch (...) {
            (void)hipStreamEndCapture(stream_, &graph);
            if (graph != nullptr) (void)hipGraphDestroy(graph);
            throw;
        }
        PHX_HIP(hipStreamEndCapture(stream_, &graph));
        hipGraphExec_t exec = nullptr;
        const hipError_t status = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        PHX_HIP(status);
        graph_[from] = exec;
    }
}

void PbfBatch::run(phx_pbf_state& state, const phx_control& control, const phx_pbf_parameters& parameters,
                   uint32_t steps) {
    require_vec(state.particles.positions, "state positions");
    require_vec(state.particles.velocities, "state velocities");
    require_vec(control.external_accelerations, "control");
    const auto& pp = parameters.particles;
    require(pp.masses && pp.rest_densities, PHX_ERROR_INVALID_ARGUMENT, "null particle parameters");
    require(parameters.relaxation && parameters.artificial_pressure_strength && parameters.artificial_pressure_exponent &&
                parameters.artificial_pressure_radius && parameters.xsph_viscosity && parameters.vorticity_confinement,
            PHX_ERROR_INVALID_ARGUMENT, "null PBF parameters");
    if (steps == 0u || n_ == 0u) {
        state.particles.step_index += steps;
        return;
    }
    const size_t bytes = sizeof(float) * n_;
    // the batch's stream follows the caller's work on the state and inputs
    PHX_HIP(hipEventRecord(ready_, caller_));
    PHX_HIP(hipStreamWaitEvent(stream_, ready_, 0));
    const float* in[6] = {state.particles.positions.x,  state.particles.positions.y,  state.particles.positions.z,
                          state.particles.velocities.x, state.particles.velocities.y, state.particles.velocities.z};
    float* first[6] = {positions_[0].x.data(),  positions_[0].y.data(),  positions_[0].z.data(),
                       velocities_[0].x.data(), velocities_[0].y.data(), velocities_[0].z.data()};
    for (int k = 0; k < 6; ++k) PHX_HIP(hipMemcpyAsync(first[k], in[k], bytes, hipMemcpyDeviceToDevice, stream_));
    if (use_graph_) {
        // a captured substep reads the control and parameter arrays it was captured with
        const std::vector<const void*> inputs = {
            control.external_accelerations.x, control.external_accelerations.y, control.external_accelerations.z,
            pp.masses, pp.rest_densities, parameters.relaxation, parameters.artificial_pressure_strength,
            parameters.artificial_pressure_exponent, parameters.artificial_pressure_radius,
            parameters.xsph_viscosity, parameters.vorticity_confinement};
        if (graph_[0] == nullptr || inputs != captured_inputs_) {
            capture(control, parameters);
            captured_inputs_ = inputs;
        }
    }
    const uint64_t step = state.particles.step_index;
    uint32_t from = 0u;
    for (uint32_t k = 0; k < steps; ++k) {
        if (use_graph_) {
            PHX_HIP(hipGraphLaunch(graph_[from], stream_));
        } else {
            substep(from, step + k, control, parameters);
        }
        from = 1u - from;
    }
    float* out[6] = {state.particles.positions.x,  state.particles.positions.y,  state.particles.positions.z,
                     state.particles.velocities.x, state.particles.velocities.y, state.particles.velocities.z};
    const float* last[6] = {positions_[from].x.data(),  positions_[from].y.data(),  positions_[from].z.data(),
                            velocities_[from].x.data(), velocities_[from].y.data(), velocities_[from].z.data()};
    for (int k = 0; k < 6; ++k) PHX_HIP(hipMemcpyAsync(out[k], last[k], bytes, hipMemcpyDeviceToDevice, stream_));
    PHX_HIP(hipEventRecord(done_, stream_));
    PHX_HIP(hipStreamWaitEvent(caller_, done_, 0));
    state.particles.step_index = step + steps;
}

}  // namespace phx
