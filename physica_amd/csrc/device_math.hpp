// Device-side vector math and SPH kernel functions for gfx950.
//
// Every function keeps the operand order of the reference's device helpers so
// that, compiled with -ffp-contract=off (no FMA formation) and HIP's default
// correctly rounded FP32 division and square root, each result is the IEEE
// value of the same expression. Reference: fluids/liquid/particle/
// domain/device.cuh:8-82, density/device.cuh:10-88, pbf/pbf.cu:18-71.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace phx {

constexpr float kPi = 3.14159265358979323846F;

struct Float3 {
    float x, y, z;
};

__device__ __forceinline__ Float3 make3(float x, float y, float z) { return Float3{x, y, z}; }
__device__ __forceinline__ Float3 add(Float3 a, Float3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ Float3 sub(Float3 a, Float3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ Float3 scale(Float3 v, float f) { return {f * v.x, f * v.y, f * v.z}; }
__device__ __forceinline__ float dot(Float3 a, Float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ Float3 cross(Float3 a, Float3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float length(Float3 a) { return sqrtf(dot(a, a)); }

// SoA field of three float arrays (domain/device.h:20-36).
struct Vec {
    float* x;
    float* y;
    float* z;
};
struct CVec {
    const float* x;
    const float* y;
    const float* z;
};

__device__ __forceinline__ Float3 load(CVec f, uint32_t i) { return {f.x[i], f.y[i], f.z[i]}; }
__device__ __forceinline__ Float3 load(Vec f, uint32_t i) { return {f.x[i], f.y[i], f.z[i]}; }
__device__ __forceinline__ void store(Vec f, uint32_t i, Float3 v) {
    f.x[i] = v.x;
    f.y[i] = v.y;
    f.z[i] = v.z;
}

// ---------------------------------------------------------------- kernels
// poly6, density/device.cuh:10-16
__device__ __forceinline__ float poly6(Float3 d, float h) {
    const float sq = dot(d, d);
    const float hr = h * h;
    if (sq >= hr) return 0.0F;
    const float diff = hr - sq;
    return 315.0F * diff * diff * diff / (64.0F * kPi * hr * hr * hr * h * h * h);
}

// poly6 gradient, density/device.cuh:18-24
__device__ __forceinline__ Float3 poly6_gradient(Float3 d, float h) {
    const float sq = dot(d, d);
    const float hr = h * h;
    if (sq >= hr) return {0.0F, 0.0F, 0.0F};
    const float diff = hr - sq;
    return scale(d, -945.0F * diff * diff / (32.0F * kPi * hr * hr * hr * h * h * h));
}

// cubic spline, density/device.cuh:26-34
__device__ __forceinline__ float cubic(Float3 d, float h) {
    const float dist = length(d);
    if (dist >= h) return 0.0F;
    const float q = dist / h;
    const float coef = 8.0F / (kPi * h * h * h);
    if (q <= 0.5F) return coef * (6.0F * q * q * q - 6.0F * q * q + 1.0F);
    const float diff = 1.0F - q;
    return coef * 2.0F * diff * diff * diff;
}

// cubic spline gradient, density/device.cuh:36-43
__device__ __forceinline__ Float3 cubic_gradient(Float3 d, float h) {
    const float dist = length(d);
    if (dist == 0.0F || dist >= h) return {0.0F, 0.0F, 0.0F};
    const float q = dist / h;
    const float coef = 8.0F / (kPi * h * h * h);
    const float deriv = q <= 0.5F ? 18.0F * q * q - 12.0F * q : -6.0F * (1.0F - q) * (1.0F - q);
    return scale(d, coef * deriv / (h * dist));
}

// Kernel constants evaluated once on the host with the reference's FP32
// expressions (density/device.cuh:30,40,75): cubic 8/(pi h^3) and the
// viscosity Laplacian's pi h^6.
struct KernelConstants {
    float h;
    float cubic;        // 8 / (pi h h h)
    float viscosity;    // pi h h h h h h
    float sqrt_support; // smallest float t with sqrtf(t) >= h
    float inv_h;        // RN(1 / h)
    float inv_viscosity; // RN(1 / viscosity)
};

// Correctly rounded x / b from inv_b = RN(1 / b), for b > 0: the estimate
// q0 = x inv_b is refined once, q1 = q0 + (x - b q0) inv_b, which leaves it
// within an ulp of x / b; the remainder x - b q1 is then exact and
// q1 + (x - b q1) inv_b rounds to RN(x / b) (P. Markstein, IBM J. Res. Dev.
// 34(1), 1990). Each step is one IEEE multiply or fused multiply-add, so the
// host and the device agree. Exact for x = 0 and for x >= 2^-100 while the
// quotient stays normal (below 2^-105 the remainders underflow and it can
// miss); the callers' arguments are distances, at least 2^-74.5 when
// nonzero, and multiples of them. Checked against x / b for every float x
// of two binades for many b and for every x in [2^-100, h) on the GPU.
__host__ __device__ inline float div_rn(float x, float b, float inv_b) {
    const float q0 = x * inv_b;
    const float r0 = fmaf(-q0, b, x);
    const float q1 = fmaf(r0, inv_b, q0);
    const float r1 = fmaf(-q1, b, x);
    return fmaf(r1, inv_b, q1);
}

// Correctly rounded square root of 0 <= x < 2^64: the hardware estimate s of
// sqrt(y), y = x 2^64 (a normal number or zero for every such x), moved to
// the float neighbour that the signs of y - s' s select (the correction of
// the AMDGPU backend's sqrtf), then scaled back by 2^-32, which is exact
// since every nonzero result is at least 2^-74.5. It equals sqrtf(x) on that
// range (checked on the GPU for every float below 2^64) without the tests
// sqrtf makes for denormal and special arguments; from 2^64 on it returns
// +inf, which the pair kernels only compare with the support radius.
__device__ __forceinline__ float sqrt_rn(float x) {
    const float y = x * 0x1p64F;
    const float s = __builtin_amdgcn_sqrtf(y);
    const float down = __uint_as_float(__float_as_uint(s) - 1u);
    const float up = __uint_as_float(__float_as_uint(s) + 1u);
    const float below = fmaf(-down, s, y);
    const float above = fmaf(-up, s, y);
    float r = below <= 0.0F ? down : s;
    r = above > 0.0F ? up : r;
    return r * 0x1p-32F;
}

// length(d) of a neighbour pair (the same value; any |d| >= 2^32 gives +inf,
// which lies outside every support).
__device__ __forceinline__ float pair_distance(Float3 d) { return sqrt_rn(dot(d, d)); }

// cubic(d, h) as a function of dist = length(d), with the constants
// supplied: the same operations as cubic(), both branches evaluated and
// selected, so that a wave walking neighbours does not diverge.
__device__ __forceinline__ float cubic_value(float dist, const KernelConstants& k) {
    const float q = div_rn(dist, k.h, k.inv_h);
    const float near = k.cubic * (6.0F * q * q * q - 6.0F * q * q + 1.0F);
    const float diff = 1.0F - q;
    const float far = k.cubic * 2.0F * diff * diff * diff;
    const float w = q <= 0.5F ? near : far;
    return dist >= k.h ? 0.0F : w;
}

// The radial factor f of cubic_gradient(d, h) = f * d (zero outside the
// support and at d = 0), with the constants supplied; branch-free like
// cubic_value().
__device__ __forceinline__ float cubic_gradient_factor(float dist, const KernelConstants& k) {
    const float q = div_rn(dist, k.h, k.inv_h);
    const float near = 18.0F * q * q - 12.0F * q;
    const float diff = 1.0F - q;
    const float far = -6.0F * diff * diff;
    const float deriv = q <= 0.5F ? near : far;
    const float f = k.cubic * deriv / (k.h * dist);
    return dist == 0.0F || dist >= k.h ? 0.0F : f;
}

__device__ __forceinline__ float viscosity_laplacian_c(Float3 d, const KernelConstants& k) {
    const float dist = pair_distance(d);
    const float l = div_rn(45.0F * (k.h - dist), k.viscosity, k.inv_viscosity);
    return dist >= k.h ? 0.0F : l;
}

// viscosity Laplacian, density/device.cuh:72-76
__device__ __forceinline__ float viscosity_laplacian(Float3 d, float h) {
    const float dist = length(d);
    if (dist >= h) return 0.0F;
    return 45.0F * (h - dist) / (kPi * h * h * h * h * h * h);
}

// spiky gradient, pbf.cu:26-32. The coefficient -45/(pi h^6) is computed once
// on the host (the reference evaluates the same powf per pair).
__device__ __forceinline__ Float3 spiky_gradient(Float3 d, float h, float coefficient) {
    const float dist = length(d);
    if (dist == 0.0F || dist >= h) return {0.0F, 0.0F, 0.0F};
    const float diff = h - dist;
    return scale(d, coefficient * diff * diff / dist);
}

// artificial pressure value, pbf.cu:56-71 (value only). `reference` is
// poly6((radius,0,0), h) of the particle.
__device__ __forceinline__ float artificial_value(Float3 d, float h, float strength, float exponent, float reference) {
    const float kernel = poly6(d, h);
    if (kernel == 0.0F) return 0.0F;
    const float ratio = kernel / reference;
    const float power = powf(ratio, exponent);
    return -strength * power;
}

// True when the artificial pressure -k (W(d) / W(dq))^n of a particle is the
// signed zero -k for every pair with W(d) > 0: k is +-0 and the power stays
// finite for every ratio up to W(0) / W(dq) (powf is non-decreasing in the
// base for n >= 0). `reference` is W(dq).
__device__ __forceinline__ bool artificial_vanishes(float h, float strength, float exponent, float reference) {
    if (strength != 0.0F || !(reference > 0.0F) || !(exponent >= 0.0F)) return false;
    const float largest = poly6(Float3{0.0F, 0.0F, 0.0F}, h) / reference;
    return powf(largest, exponent) < 1.0e37F;
}

// min(C-1, max(0, floor(v))) with NaN and negatives mapped to 0, as the
// reference's __float2int_rd + clamp (neighborhood.cu:34-36).
__device__ __forceinline__ int cell_clamp(float v, uint32_t count) {
    const float f = floorf(v);
    if (!(f > 0.0F)) return 0;
    if (f >= static_cast<float>(count - 1u)) return static_cast<int>(count) - 1;
    return static_cast<int>(f);
}

// Collision box of the solvers (sph.cpp:21-35, pbf.cpp:17-31).
struct Box {
    float lo[3];
    float hi[3];
    float vel[3];
    uint32_t no_slip;
};

// Semi-implicit Euler with box clamp: integrate/predict body of
// sph/common.cu:302-320 and sph/dfsph.cu:26-44.
__device__ __forceinline__ void predict_clamped(const Box& box, float dt, Float3 x, Float3 v, Float3 a, Float3& out_x,
                                                Float3& out_v) {
    Float3 vel = add(v, scale(a, dt));
    Float3 pos = add(x, scale(vel, dt));
    const bool cx = pos.x < box.lo[0] || pos.x > box.hi[0];
    const bool cy = pos.y < box.lo[1] || pos.y > box.hi[1];
    const bool cz = pos.z < box.lo[2] || pos.z > box.hi[2];
    pos.x = fminf(box.hi[0], fmaxf(box.lo[0], pos.x));
    pos.y = fminf(box.hi[1], fmaxf(box.lo[1], pos.y));
    pos.z = fminf(box.hi[2], fmaxf(box.lo[2], pos.z));
    if (box.no_slip != 0u && (cx || cy || cz)) {
        vel = {box.vel[0], box.vel[1], box.vel[2]};
    } else {
        if (cx) vel.x = box.vel[0];
        if (cy) vel.y = box.vel[1];
        if (cz) vel.z = box.vel[2];
    }
    out_x = pos;
    out_v = vel;
}

// Box-collision bits (x: 1, y: 2, z: 4) of the unclamped prediction
// x + dt (v + dt a) of predict_clamped.
__device__ __forceinline__ uint32_t predict_collisions(const Box& box, float dt, Float3 x, Float3 v, Float3 a) {
    const Float3 pos = add(x, scale(add(v, scale(a, dt)), dt));
    return (pos.x < box.lo[0] || pos.x > box.hi[0] ? 1u : 0u) | (pos.y < box.lo[1] || pos.y > box.hi[1] ? 2u : 0u) |
           (pos.z < box.lo[2] || pos.z > box.hi[2] ? 4u : 0u);
}

// Uniform grid in sorted (cell-major) particle order: the cell range of cell c
// is [offsets[c], offsets[c+1]) of the sorted arrays.
struct SortedGrid {
    const uint32_t* offsets;
    uint32_t cx, cy, cz;
    float ox, oy, oz;
    float cell;
};

// particle_cell, neighborhood/device.cuh:44-48
__device__ __forceinline__ void particle_cell(const SortedGrid& g, Float3 p, int& x, int& y, int& z) {
    x = cell_clamp((p.x - g.ox) / g.cell, g.cx);
    y = cell_clamp((p.y - g.oy) / g.cell, g.cy);
    z = cell_clamp((p.z - g.oz) / g.cell, g.cz);
}

// Sorted boundary samples of a grid build, the boundary half of
// NeighborhoodView and BoundaryView (neighborhood/device.h:8-36): per cell
// the samples in ascending original index, each stored as its position at
// the step's time t, x0 + t v (boundary_position, neighborhood/device.cuh:
// 16-22), with its volume, and its velocity.
struct BoundaryGrid {
    const uint32_t* offsets;  // C + 1 cell starts; null when the domain has no boundary samples
    const float4* x;          // position at t (xyz), volume (w)
    const float4* v;          // velocity (xyz)
};

// Calls cell(c) for the in-grid cells of the 27 around (cx, cy, cz) in the
// reference order: offset_x outer, offset_y, offset_z inner (cell_range,
// neighborhood/device.cuh:32-42, skips cells outside the grid).
template <class Cell>
__device__ __forceinline__ void for_each_cell(const SortedGrid& g, int cx, int cy, int cz, Cell&& cell) {
    // not unrolled: 27 inlined copies of a visitor cost registers (occupancy)
    // and instruction cache for no gain, the visit loop being dynamic anyway
#pragma unroll 1
    for (int ox = -1; ox <= 1; ++ox) {
        const int X = cx + ox;
        if (X < 0 || X >= static_cast<int>(g.cx)) continue;
#pragma unroll 1
        for (int oy = -1; oy <= 1; ++oy) {
            const int Y = cy + oy;
            if (Y < 0 || Y >= static_cast<int>(g.cy)) continue;
            const uint32_t row = (static_cast<uint32_t>(Y)) * g.cx + static_cast<uint32_t>(X);
#pragma unroll 1
            for (int oz = -1; oz <= 1; ++oz) {
                const int Z = cz + oz;
                if (Z < 0 || Z >= static_cast<int>(g.cz)) continue;
                cell(static_cast<uint32_t>(Z) * g.cy * g.cx + row);
            }
        }
    }
}

// Visits the sorted candidates of the 27 cells around (cx, cy, cz) in the
// reference order; ascending sorted position (= ascending original index)
// within a cell.
template <class Visit>
__device__ __forceinline__ void for_each_candidate(const SortedGrid& g, int cx, int cy, int cz, Visit&& visit) {
    for_each_cell(g, cx, cy, cz, [&](uint32_t c) {
        const uint32_t first = g.offsets[c];
        const uint32_t last = g.offsets[c + 1u];
        for (uint32_t s = first; s < last; ++s) visit(s);
    });
}

// The same walk with each cell's boundary samples visited after its
// particles, the order of every reference neighbour loop (e.g.
// density/density.cu:28-37).
template <class Visit, class VisitBoundary>
__device__ __forceinline__ void for_each_candidate(const SortedGrid& g, const BoundaryGrid& b, int cx, int cy, int cz,
                                                   Visit&& visit, VisitBoundary&& visit_boundary) {
    for_each_cell(g, cx, cy, cz, [&](uint32_t c) {
        const uint32_t first = g.offsets[c];
        const uint32_t last = g.offsets[c + 1u];
        for (uint32_t s = first; s < last; ++s) visit(s);
        const uint32_t bfirst = b.offsets[c];
        const uint32_t blast = b.offsets[c + 1u];
        for (uint32_t s = bfirst; s < blast; ++s) visit_boundary(s);
    });
}

}  // namespace phx
