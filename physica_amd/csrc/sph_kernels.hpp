// Neighbour-pass kernels shared by the SPH solvers (DFSPH, PCISPH, IISPH,
// WCSPH) in sorted (cell-major) particle order.
//
// Every per-particle sum runs in the reference order (27 cells, offset_x
// outer, ascending original index) over the precomputed neighbour masks, and
// every masked-out candidate contributes an exact zero in the reference, so
// results equal the reference expressions operation by operation. The kernels
// live in an unnamed namespace: each solver translation unit instantiates the
// ones it launches.
#pragma once

#include "engine.hpp"
#include "neighbor_mask.hpp"
#include "profiler.hpp"

namespace phx {
namespace {

__device__ __forceinline__ Float3 xyz(float4 q) { return Float3{q.x, q.y, q.z}; }

constexpr uint32_t kBlock = 256;
constexpr uint32_t kWaves = kBlock / 64;  // waves of a pass workgroup
// candidates a masked walk fetches ahead of the one it processes
constexpr int kWalkDepth = 2;

// Marks the 27 cells around (cx, cy, cz) with `stamp` (plain stores: every
// writer writes the same value).
__device__ __forceinline__ void taint_neighbour_cells(const SortedGrid& g, int cx, int cy, int cz, uint32_t* taint,
                                                      uint32_t stamp) {
    for (int z = max(cz - 1, 0); z <= min(cz + 1, static_cast<int>(g.cz) - 1); ++z)
        for (int y = max(cy - 1, 0); y <= min(cy + 1, static_cast<int>(g.cy) - 1); ++y)
            for (int x = max(cx - 1, 0); x <= min(cx + 1, static_cast<int>(g.cx) - 1); ++x)
                taint[cell_of(g, x, y, z)] = stamp;
}

// Per-cell displacement bound of a step: scale * the largest displacement
// of the previous step among the cell's particles (sorted order), plus
// `allowance`; 0 for an empty cell.
[[maybe_unused]] __global__ void __launch_bounds__(kBlock)
    cell_bound_kernel(uint32_t cells, const uint32_t* __restrict__ offsets, const float* __restrict__ moved,
                      float scale, float allowance, float* __restrict__ bound) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cells) return;
    const uint32_t first = offsets[c];
    const uint32_t last = offsets[c + 1u];
    float largest = 0.0F;
    for (uint32_t s = first; s < last; ++s) largest = fmaxf(largest, moved[s]);
    bound[c] = first == last ? 0.0F : scale * largest + allowance;
}

struct MaskDensityArgs {
    uint32_t n;
    KernelConstants kc;
    float t_skin;
    SortedGrid grid;
    BoundaryGrid boundary;  // offsets null: no boundary samples
    const float4* xm;
    uint16_t* exact;
    uint16_t* skin;
    const float* p_div;  // optional: t_div = p_div / rho^2
    const float* p_den;  // optional: t_den = p_den / rho^2
    float* rho;
    float* t_div;
    float* t_den;
    uint32_t* flags;
    // rho0 (boundary terms), and the equation of state (WCSPH): p = EOS(rho),
    // aux = (rho, p / rho^2, c, sigma)
    const float* rest_density;
    const float* tait_exponent;
    float* pressures;
    float4* aux;  // z (speed of sound) and w (surface tension) are inputs
    // per particle: set where a stencil cell holds more than kMaskBits
    // particles (that particle walks the full stencil in every pass)
    uint8_t* crowded;
    // per cell: a bound on how far its particles move this step; non-null:
    // the skin mask keeps a pair while d2 < (h + B_ci + B_cj)^2 (plus a
    // rounding margin) instead of d2 < t_skin
    const float* cell_bound = nullptr;
    // optional: (x, y, z, t_div) and (x, y, z, t_den) per particle, the
    // candidate records of the force passes that read these t
    float4* xt_div = nullptr;
    float4* xt_den = nullptr;
    // optional: the step's mass state (see ForceArgs::mass_state): [1] the
    // bits of sorted particle 0's mass; [0] set to mass_stamp by every
    // particle whose mass has other bits
    uint32_t* mass_state = nullptr;
    uint32_t mass_stamp = 0u;
};

// Mask build, then density_forward_kernel (density.cu:18-40, cubic kernel,
// topology = positions) over the exact mask and the first_term p/rho^2 of
// pressure_forward_kernel (common.cu:171) for the given pressures, or with
// kEos the Tait pressure of eos_forward_kernel (wcsph.cu:92-100) and its
// first_term. With boundary samples no masks are built (the host marks the
// step's masks unusable) and the density walks the full stencil, each cell's
// samples b after its particles adding rho0_i V_b W (density.cu:34-37). A
// particle with a stencil cell of more than kMaskBits particles is marked
// crowded and walks the full stencil here and in every later pass.
template <bool kEos>
__global__ void __launch_bounds__(kBlock) mask_density_kernel(MaskDensityArgs a) {
    __shared__ uint16_t smask[kStencil * kBlock];
    __shared__ int delta[kStencil];
    fill_slot_deltas(a.grid, delta);
    __syncthreads();
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    if (i >= a.n) return;
    uint16_t* column = smask + threadIdx.x;
    const float4 self = a.xm[i];
    const Float3 p = xyz(self);
    if (a.mass_state != nullptr) {
        const uint32_t bits = __float_as_uint(self.w);
        if (i == 0u) a.mass_state[1] = bits;
        if (bits != __float_as_uint(a.xm[0].w)) a.mass_state[0] = a.mass_stamp;  // one value for every writer
    }
    int cx, cy, cz;
    particle_cell(a.grid, p, cx, cy, cz);
    uint32_t nonzero = 0u;
    float result = 0.0F;
    if (a.boundary.offsets != nullptr) {
        a.crowded[i] = 1u;  // no masks are built with boundary samples
        const float rest = a.rest_density[i];
        for_each_candidate(
            a.grid, a.boundary, cx, cy, cz,
            [&](uint32_t s) {
                const float4 q = a.xm[s];
                result += q.w * cubic_value(pair_distance(sub(p, xyz(q))), a.kc);
            },
            [&](uint32_t s) {
                const float4 b = a.boundary.x[s];
                result += rest * b.w * cubic_value(pair_distance(sub(p, xyz(b))), a.kc);
            });
    } else if (build_masks(a.grid, a.xm, a.n, i, p, cx, cy, cz, a.kc.sqrt_support, a.t_skin, a.exact, a.skin, column,
                           kBlock, &nonzero, a.cell_bound, a.kc.h)) {
        a.crowded[i] = 0u;
        walk_masked(a.grid.offsets, cell_of(a.grid, cx, cy, cz), delta, column, kBlock, nonzero, [&](uint32_t s) {
            const float4 q = a.xm[s];
            const float dist = pair_distance(sub(p, xyz(q)));
            result += q.w * cubic_value(dist, a.kc);
        });
    } else {
        a.crowded[i] = 1u;
        for_each_candidate(a.grid, cx, cy, cz, [&](uint32_t s) {
            const float4 q = a.xm[s];
            result += q.w * cubic_value(pair_distance(sub(p, xyz(q))), a.kc);
        });
    }
    a.rho[i] = result;
    if constexpr (kEos) {
        const float4 aux = a.aux[i];
        const float rest_density = a.rest_density[i];
        const float speed = aux.z;
        const float exponent = a.tait_exponent[i];
        const float ratio = result / rest_density;
        const float pressure = rest_density * speed * speed * (powf(ratio, exponent) - 1.0F) / exponent;
        a.pressures[i] = pressure;
        a.aux[i] = make_float4(result, pressure / (result * result), speed, aux.w);
    } else {
        if (a.p_div != nullptr) {
            const float t = a.p_div[i] / (result * result);
            a.t_div[i] = t;
            if (a.xt_div != nullptr) a.xt_div[i] = make_float4(p.x, p.y, p.z, t);
        }
        if (a.p_den != nullptr) {
            const float t = a.p_den[i] / (result * result);
            a.t_den[i] = t;
            if (a.xt_den != nullptr) a.xt_den[i] = make_float4(p.x, p.y, p.z, t);
        }
    }
}

// Tolerance mode: the convergence state of a pressure phase. The projection
// adds each wave's compression errors max(0, rho* - target) / rho0 in FP64
// (shuffles, no block barrier, so every wave retires as soon as its own walk
// ends) and stores the sum in partials[blockIdx.x * kWaves + wave];
// converge_kernel then
// marks the phase done at this iteration when the mean is within the
// tolerance or the iteration is the last allowed one. Every pass of the
// phase reads `done`: passes of later iterations (launched ahead by the
// host) then return at once, and the force pass of the marked iteration ends
// the phase.
struct Convergence {
    double* partials;      // null: fixed iteration counts; else one error sum per wave
    const uint32_t* done;  // iteration that ended the phase; 0 while it runs
    const float* rest;  // rho0, the scale of the error
};

// Block-wide FP64 sum reaching thread 0 (other threads get partial sums).
// The order of the additions is fixed, so the result is reproducible.
template <uint32_t kBlockSize>
__device__ __forceinline__ double block_sum(double value) {
    __shared__ double wave_sum[kBlockSize / 64];
    double v = value;
    for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_xor(v, offset, 64);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = threadIdx.x >> 6;
    if (lane == 0u) wave_sum[wave] = v;
    __syncthreads();
    double total = 0.0;
    if (threadIdx.x == 0u)
        for (uint32_t w = 0; w < kBlockSize / 64; ++w) total += wave_sum[w];
    return total;
}

constexpr uint32_t kConvergeBlock = 1024;

// The value *done has after this iteration, for the host: a store to
// host-visible (pinned) memory at system scope.
__device__ __forceinline__ void publish_done(uint32_t* host_flag, uint32_t value) {
    __hip_atomic_store(host_flag, value, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Ends a tolerance-mode phase: sums the projection's per-wave error sums in
// a fixed order (thread t takes partials t, t + 1024, ...), and
// records the iteration in *done when the mean error is within the
// tolerance or the iteration is the last allowed. One workgroup; the kernel
// boundaries order it after the projection and before the force pass, so
// no atomics or fences are needed on the device. The resulting *done is also
// stored to the host's flag of this iteration (pinned memory mapped for the
// device), which the host reads once the iteration's event has completed,
// with no copy in between.
[[maybe_unused]] __global__ void __launch_bounds__(kConvergeBlock)
    converge_kernel(const double* partials, uint32_t count, uint32_t n, double tolerance, uint32_t iteration,
                    uint32_t max_iterations, uint32_t* done, uint32_t* host_flag) {
    const uint32_t ended = *done;
    if (ended != 0u) {  // the phase ended at an earlier iteration
        if (threadIdx.x == 0u) publish_done(host_flag, ended);
        return;
    }
    double sum = 0.0;
    for (uint32_t k = threadIdx.x; k < count; k += kConvergeBlock) sum += partials[k];
    const double total = block_sum<kConvergeBlock>(sum);
    if (threadIdx.x == 0u) {
        const double mean = total / static_cast<double>(n);
        const uint32_t now = mean <= tolerance || iteration >= max_iterations ? iteration : 0u;
        if (now != 0u) *done = now;
        publish_done(host_flag, now);
    }
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
    const float* volume_density;  // rho (DFSPH, IISPH) or rho0 (PCISPH) in V = m / density
    const float* target;
    const float* relaxation;
    const float* p_prev;  // pressures before the update (may equal p)
    float* p;
    float* t;
    float* rs_out;  // optional (recorded step): the predicted densities
    Convergence cv;
    BoundaryGrid boundary;     // offsets null: no boundary samples
    const float* rest_density;  // rho0, for the boundary terms
    const uint8_t* crowded;     // per particle: walks the full stencil
    // per-cell skins: the skin mask applies to particle i unless its cell
    // holds `stamp` (a particle of a neighbouring cell moved past its cell's
    // bound in the prediction this pass reads); null: the global rule
    // (every particle moved at most sqrt(limit2))
    const uint32_t* taint = nullptr;
    uint32_t stamp = 0u;
    float4* xt_out = nullptr;  // optional: (x, y, z, t) per particle, the force pass's candidate record
};

// The pressure phase a projection serves. The kernel is the same for both;
// the two instantiations only give profiles one kernel name per phase.
enum ProjectionPhase : int { kProjectDivergence = 0, kProjectDensity = 1 };

// density at predicted positions (density.cu:18-40 with topology = x,
// positions = x*) fused with the pressure update and the next pressure
// pass's first_term: projection_update_forward_kernel (dfsph.cu:104-110),
// pressure_update_forward_kernel (pcisph.cu:104-110) or
// jacobi_update_forward_kernel (iisph.cu:104-110), which differ only in the
// density of the volume V = m / density.
template <int kPhase>
__global__ void __launch_bounds__(kBlock) projection_kernel(ProjectionArgs a) {
    __shared__ uint16_t smask[kStencil * kBlock];
    __shared__ int delta[kStencil];
    fill_slot_deltas(a.grid, delta);
    __syncthreads();
    const bool tolerance = a.cv.partials != nullptr;
    if (tolerance && *a.cv.done != 0u) return;  // the phase ended at an earlier iteration
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    double error = 0.0;
    if (i < a.n) {
        const float4 topology = a.xm[i];
        int cx, cy, cz;
        particle_cell(a.grid, xyz(topology), cx, cy, cz);
        const bool moved_within = a.taint != nullptr
                                      ? a.taint[cell_of(a.grid, cx, cy, cz)] != a.stamp
                                      : __uint_as_float(a.flags[kFlagDisplacement + a.slot]) <= a.limit2;
        const bool masked =
            a.flags[kFlagOverflow] == 0u && a.boundary.offsets == nullptr && a.crowded[i] == 0u && moved_within;
        const float4 self = a.xs[i];
        const Float3 ps = xyz(self);
        float rs = 0.0F;
        auto visit = [&](uint32_t s) {
            const float4 q = a.xs[s];
            rs += q.w * cubic_value(pair_distance(sub(ps, xyz(q))), a.kc);
        };
        if (masked) {
            uint16_t* column = smask + threadIdx.x;
            const uint32_t nonzero = load_lane_column(a.skin, a.n, i, column, kBlock);
            walk_masked_prefetch<kWalkDepth>(
                a.grid.offsets, cell_of(a.grid, cx, cy, cz), delta, column, kBlock, nonzero, i,
                [&](uint32_t s) { return a.xs[s]; },
                [&](float4 q) { rs += q.w * cubic_value(pair_distance(sub(ps, xyz(q))), a.kc); });
        } else if (a.boundary.offsets != nullptr) {
            // the predicted density's boundary terms rho0_i V_b W(x*_i - x_b)
            const float rest = a.rest_density[i];
            for_each_candidate(a.grid, a.boundary, cx, cy, cz, visit, [&](uint32_t s) {
                const float4 b = a.boundary.x[s];
                rs += rest * b.w * cubic_value(pair_distance(sub(ps, xyz(b))), a.kc);
            });
        } else {
            for_each_candidate(a.grid, cx, cy, cz, visit);
        }
        const float rho = a.rho[i];
        const float target = a.target[i];
        const float volume = self.w / a.volume_density[i];
        const float delta_p = 1.0F / (2.0F * a.dt * a.dt * volume * volume * a.gradient_norm);
        const float pressure = fmaxf(0.0F, a.p_prev[i] + a.relaxation[i] * delta_p * (rs - target));
        a.p[i] = pressure;
        const float t = pressure / (rho * rho);
        a.t[i] = t;
        if (a.xt_out != nullptr) a.xt_out[i] = make_float4(topology.x, topology.y, topology.z, t);
        if (a.rs_out != nullptr) a.rs_out[i] = rs;
        if (tolerance)
            error = fmax(0.0, static_cast<double>(rs) - static_cast<double>(target)) /
                    static_cast<double>(a.cv.rest[i]);
    }
    if (tolerance) {
        // every lane of the wave is here (no early return above)
        double wave_error = error;
        for (int offset = 32; offset > 0; offset >>= 1) wave_error += __shfl_xor(wave_error, offset, 64);
        if ((threadIdx.x & 63u) == 0u) a.cv.partials[blockIdx.x * kWaves + (threadIdx.x >> 6)] = wave_error;
    }
}

enum Epilogue : int {
    kDivPredict = 0,   // total = a_p(t1)
    kDivEnd = 1,       // a_div = a_p(t1); total = a_div + a_p(t2)
    kDenPredict = 2,   // total = a_div + a_p(t1)
    kInitPredict = 3,  // total = 0 (cleared pressure acceleration), no pressure pass
    kDivAuto = 4,      // tolerance mode: kDivEnd at the iteration that ends the phase, else kDivPredict
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
    uint8_t* collisions;  // optional (recorded step): box-collision bits of the prediction
    uint32_t final_step;
    const uint32_t* perm;
    Vec next_x, next_v;
    const float* p_div;
    float* warm_div_out;
    const float* p_den;
    float* warm_den_out;
    // tolerance mode: the phase's done word and this pass's iteration; a
    // pass of a later iteration than the one that ended the phase returns
    const uint32_t* done;
    uint32_t iteration;
    BoundaryGrid boundary;      // offsets null: no boundary samples
    const float* rest_density;  // rho0, for the boundary pressure terms
    const uint8_t* crowded;     // per particle: walks the full stencil
    // per-cell skins: a predicting pass marks the 27 cells around every
    // particle that moved farther than its cell's bound with `stamp`
    const float* cell_bound = nullptr;
    uint32_t* taint = nullptr;
    uint32_t stamp = 0u;
    // optional: |x_next - x| per particle (original order), written with the next state
    float* moved_out = nullptr;
    // optional: (x, y, z, t1) per particle and the step's mass state {
    // mass_stamp if some mass differs from particle 0's, particle 0's mass
    // bits} (written by the mask pass): with equal masses a candidate is one
    // 16-byte record instead of its position, mass and t1
    const float4* xt1 = nullptr;
    const uint32_t* mass_state = nullptr;
    uint32_t mass_stamp = 0u;
};

// Pressure acceleration (pressure_forward_kernel, common.cu:165-194),
// optionally fused with the non-pressure acceleration
// (non_pressure_forward_kernel, common.cu:14-47), then the predict of the
// next iteration (dfsph.cu:26-44) or, on the last pass, the next state.
// cubic_gradient(d, h) is evaluated as f * d with f its radial factor, the
// same operations as the reference's expression. Boundary samples add, after
// each cell's particles, mu_i V_b lap W (v_b - v_i) (common.cu:38-43) and
// grad W (-rho0_i V_b p_i / rho_i^2) (common.cu:185-190).
template <bool kNonPressure, int kEpilogue>
__global__ void __launch_bounds__(kBlock) force_kernel(ForceArgs a) {
    constexpr bool kAuto = kEpilogue == kDivAuto;
    constexpr bool kPressure = kEpilogue != kInitPredict;
    __shared__ uint16_t smask[kStencil * kBlock];
    __shared__ int delta[kStencil];
    fill_slot_deltas(a.grid, delta);
    __syncthreads();
    // the iteration that ends its phase (tolerance mode), uniform over the grid
    const uint32_t ended = a.done != nullptr ? *a.done : 0u;
    if (ended != 0u && ended < a.iteration) return;
    const bool ends = ended != 0u && ended == a.iteration;
    // a second pressure field: the divergence phase's end
    const bool kTwo = kEpilogue == kDivEnd || (kAuto && ends);
    // tolerance mode: the iteration that ends the last phase writes the state
    const uint32_t final_step =
        a.final_step != 0u || ((kEpilogue == kDenPredict || kEpilogue == kDivPredict) && ends) ? 1u : 0u;
    const uint32_t i = xcd_block() * kBlock + threadIdx.x;
    const bool active = i < a.n;
    float moved2 = 0.0F;
    if (active) {
        const bool masked = a.flags[kFlagOverflow] == 0u && a.boundary.offsets == nullptr && a.crowded[i] == 0u;
        const float4 self = a.xm[i];
        const Float3 p = xyz(self);
        const float4 vm = a.vmu[i];
        const Float3 vel = xyz(vm);
        Float3 acc{0.0F, 0.0F, 0.0F};
        if constexpr (kNonPressure) acc = add(a.gravity, load(a.u, i));
        const float t1_i = kPressure ? a.t1[i] : 0.0F;
        float t2_i = 0.0F;
        if (kTwo) t2_i = a.t2[i];
        Float3 ap1{0.0F, 0.0F, 0.0F};
        Float3 ap2{0.0F, 0.0F, 0.0F};
        int cx, cy, cz;
        particle_cell(a.grid, p, cx, cy, cz);
        // candidates as (x, y, z, t1) records with one mass (uniform over the launch)
        const bool packed = kPressure && a.xt1 != nullptr && a.mass_state[0] != a.mass_stamp;
        const float mass = packed ? __uint_as_float(a.mass_state[1]) : 0.0F;
        auto pair = [&](uint32_t s, float4 q, float t1_s, float t2_s, Float3 d, float f) {
            const float ms = q.w;
            if constexpr (kNonPressure) {
                const float4 vq = a.vmu[s];
                const float viscosity = 0.5F * (vm.w + vq.w);
                const float weight = viscosity * ms * viscosity_laplacian_c(d, a.kc) / a.rho[s];
                acc = add(acc, scale(sub(xyz(vq), vel), weight));
            }
            if constexpr (kPressure) {
                const Float3 gradient = scale(d, f);
                ap1 = add(ap1, scale(gradient, -ms * (t1_i + t1_s)));
                if (kTwo) ap2 = add(ap2, scale(gradient, -ms * (t2_i + t2_s)));
            }
        };
        if (masked) {
            uint16_t* column = smask + threadIdx.x;
            const uint32_t center = cell_of(a.grid, cx, cy, cz);
            uint32_t nonzero = load_lane_column(a.exact, a.n, i, column, kBlock);
            nonzero = drop_self(i, a.grid.offsets[center], column, kBlock, nonzero);
            if (packed) {
                // equal masses: (x, y, z, t1) is the whole candidate but t2
                if constexpr (kNonPressure) {
                    walk_masked(a.grid.offsets, center, delta, column, kBlock, nonzero, [&](uint32_t s) {
                        const float4 r = a.xt1[s];
                        const float t2_s = kTwo ? a.t2[s] : 0.0F;
                        const Float3 d = sub(p, xyz(r));
                        pair(s, make_float4(r.x, r.y, r.z, mass), r.w, t2_s, d,
                             kPressure ? cubic_gradient_factor(pair_distance(d), a.kc) : 0.0F);
                    });
                } else {
                    struct Record {
                        float4 r;
                        float t2;
                    };
                    walk_masked_prefetch<kWalkDepth>(
                        a.grid.offsets, center, delta, column, kBlock, nonzero, i,
                        [&](uint32_t s) { return Record{a.xt1[s], kTwo ? a.t2[s] : 0.0F}; },
                        [&](const Record& c) {
                            const Float3 d = sub(p, xyz(c.r));
                            pair(0u, make_float4(c.r.x, c.r.y, c.r.z, mass), c.r.w, c.t2, d,
                                 cubic_gradient_factor(pair_distance(d), a.kc));
                        });
                }
            } else if constexpr (kNonPressure) {
                walk_masked(a.grid.offsets, center, delta, column, kBlock, nonzero, [&](uint32_t s) {
                    const float4 q = a.xm[s];
                    const float t1_s = kPressure ? a.t1[s] : 0.0F;
                    const float t2_s = kTwo ? a.t2[s] : 0.0F;
                    const Float3 d = sub(p, xyz(q));
                    pair(s, q, t1_s, t2_s, d, kPressure ? cubic_gradient_factor(pair_distance(d), a.kc) : 0.0F);
                });
            } else {
                struct Candidate {
                    float4 q;
                    float t1, t2;
                };
                walk_masked_prefetch<kWalkDepth>(
                    a.grid.offsets, center, delta, column, kBlock, nonzero, i,
                    [&](uint32_t s) { return Candidate{a.xm[s], a.t1[s], kTwo ? a.t2[s] : 0.0F}; },
                    [&](const Candidate& c) {
                        const Float3 d = sub(p, xyz(c.q));
                        pair(0u, c.q, c.t1, c.t2, d, cubic_gradient_factor(pair_distance(d), a.kc));
                    });
            }
        } else {
            auto visit = [&](uint32_t s) {
                if (s == i) return;
                const float4 q = a.xm[s];
                const float t1_s = kPressure ? a.t1[s] : 0.0F;
                const float t2_s = kTwo ? a.t2[s] : 0.0F;
                const Float3 d = sub(p, xyz(q));
                pair(s, q, t1_s, t2_s, d, kPressure ? cubic_gradient_factor(pair_distance(d), a.kc) : 0.0F);
            };
            if (a.boundary.offsets != nullptr) {
                const float rest = a.rest_density[i];
                for_each_candidate(a.grid, a.boundary, cx, cy, cz, visit, [&](uint32_t s) {
                    const float4 b = a.boundary.x[s];
                    const Float3 d = sub(p, xyz(b));
                    if constexpr (kNonPressure) {
                        const float weight = vm.w * b.w * viscosity_laplacian_c(d, a.kc);
                        acc = add(acc, scale(sub(xyz(a.boundary.v[s]), vel), weight));
                    }
                    if constexpr (kPressure) {
                        const Float3 gradient = scale(d, cubic_gradient_factor(pair_distance(d), a.kc));
                        ap1 = add(ap1, scale(gradient, -rest * b.w * t1_i));
                        if (kTwo) ap2 = add(ap2, scale(gradient, -rest * b.w * t2_i));
                    }
                });
            } else {
                for_each_candidate(a.grid, cx, cy, cz, visit);
            }
        }
        if constexpr (kNonPressure) {
            a.anp[i] = make_float4(acc.x, acc.y, acc.z, 0.0F);
        } else {
            acc = xyz(a.anp[i]);
        }
        Float3 total;
        if constexpr (kEpilogue == kInitPredict) {
            total = Float3{0.0F, 0.0F, 0.0F};
        } else if constexpr (kEpilogue == kDenPredict) {
            total = add(xyz(a.adiv[i]), ap1);
        } else if (kTwo) {
            a.adiv[i] = make_float4(ap1.x, ap1.y, ap1.z, 0.0F);
            total = add(ap1, ap2);
        } else {
            total = ap1;
        }
        Float3 nx, nv;
        const Float3 a_total = add(acc, total);
        predict_clamped(a.box, a.dt, p, vel, a_total, nx, nv);
        if (a.collisions != nullptr)
            a.collisions[i] = static_cast<uint8_t>(predict_collisions(a.box, a.dt, p, vel, a_total));
        if (kTwo && a.warm_div_out != nullptr) a.warm_div_out[a.perm[i]] = a.p_div[i];
        const Float3 moved = sub(nx, p);
        if (final_step != 0u) {
            const uint32_t j = a.perm[i];
            store(a.next_x, j, nx);
            store(a.next_v, j, nv);
            if (a.warm_den_out != nullptr) a.warm_den_out[j] = a.p_den[i];
            if (a.moved_out != nullptr) a.moved_out[j] = sqrtf(dot(moved, moved));
        } else {
            a.xs[i] = make_float4(nx.x, nx.y, nx.z, self.w);
            moved2 = dot(moved, moved);
            if (a.taint != nullptr) {
                const float bound = a.cell_bound[cell_of(a.grid, cx, cy, cz)];
                if (moved2 > bound * bound) taint_neighbour_cells(a.grid, cx, cy, cz, a.taint, a.stamp);
            }
        }
    }
    if (final_step == 0u) block_max_to<kBlock>(moved2, &a.flags[kFlagDisplacement + a.slot]);
}

template <bool kNonPressure, int kEpilogue>
void launch_force(PassProfiler& profiler, hipStream_t stream, uint32_t blocks, const ForceArgs& args, int pass) {
    PassScope scope(profiler, stream, pass);
    force_kernel<kNonPressure, kEpilogue><<<blocks, kBlock, 0, stream>>>(args);
    PHX_LAUNCH_CHECK();
}

void launch_projection(PassProfiler& profiler, hipStream_t stream, uint32_t blocks, const ProjectionArgs& args,
                       int pass, ProjectionPhase phase) {
    PassScope scope(profiler, stream, pass);
    if (phase == kProjectDivergence) {
        projection_kernel<kProjectDivergence><<<blocks, kBlock, 0, stream>>>(args);
    } else {
        projection_kernel<kProjectDensity><<<blocks, kBlock, 0, stream>>>(args);
    }
    PHX_LAUNCH_CHECK();
}

template <bool kEos>
void launch_mask_density(PassProfiler& profiler, hipStream_t stream, uint32_t blocks, const MaskDensityArgs& args,
                         int pass) {
    PassScope scope(profiler, stream, pass);
    mask_density_kernel<kEos><<<blocks, kBlock, 0, stream>>>(args);
    PHX_LAUNCH_CHECK();
}

// Whole blocks of kBlock particles, padded to a multiple of the XCD count.
inline uint32_t pass_blocks(uint32_t n) { return (blocks_for(n, kBlock) + 7u) / 8u * 8u; }

}  // namespace
}  // namespace phx
