// Device helpers shared by the SPH derivative passes: FP64 vectors, the
// tangents and Hessian products of the SPH kernel functions, and the tangent
// and transpose of the prediction / integration with recorded collision bits.
// Reference: density/device.cuh:45-88, sph/common.cu:201-202, 322-366,
// sph/dfsph.cu:46-102.
#pragma once

#include "device_math.hpp"
#include "engine.hpp"

namespace phx {
namespace {

__device__ __forceinline__ Float3 xyz(float4 q) { return Float3{q.x, q.y, q.z}; }

struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 d3(Float3 a) {
    return {static_cast<double>(a.x), static_cast<double>(a.y), static_cast<double>(a.z)};
}
__device__ __forceinline__ D3 add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 scale(D3 v, double f) { return {f * v.x, f * v.y, f * v.z}; }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 load(DVec f, uint32_t i) { return {f.x[i], f.y[i], f.z[i]}; }
__device__ __forceinline__ void store(DVec f, uint32_t i, D3 v) {
    f.x[i] = v.x;
    f.y[i] = v.y;
    f.z[i] = v.z;
}
__device__ __forceinline__ void accumulate(DVec f, uint32_t i, D3 v) {
    f.x[i] += v.x;
    f.y[i] += v.y;
    f.z[i] += v.z;
}

// cubic_gradient_tangent (density/device.cuh:45-57): derivative of the cubic
// gradient at d along dd.
__device__ __forceinline__ Float3 cubic_gradient_tangent(Float3 d, Float3 dd, const KernelConstants& k) {
    const float h = k.h;
    const float dist = length(d);
    const float coef = k.cubic;
    if (dist == 0.0F) return scale(dd, -12.0F * coef / (h * h));
    if (dist >= h) return Float3{0.0F, 0.0F, 0.0F};
    const float q = dist / h;
    const float deriv = q <= 0.5F ? 18.0F * q * q - 12.0F * q : -6.0F * (1.0F - q) * (1.0F - q);
    const float second = q <= 0.5F ? 36.0F * q - 12.0F : 12.0F * (1.0F - q);
    const float radial = dot(d, dd) / dist;
    const float rs = coef * deriv / (h * dist);
    const float rsd = coef * (second / (h * h * dist) - deriv / (h * dist * dist));
    return add(scale(dd, rs), scale(d, rsd * radial));
}

// viscosity_laplacian_tangent (density/device.cuh:78-82)
__device__ __forceinline__ float viscosity_laplacian_tangent(Float3 d, Float3 dd, const KernelConstants& k) {
    const float dist = length(d);
    if (dist == 0.0F || dist >= k.h) return 0.0F;
    return -45.0F * dot(d, dd) / (k.viscosity * dist);
}

// cubic_hessian_product (density/device.cuh:59-70): H(d) w in FP64.
__device__ __forceinline__ D3 cubic_hessian_product(Float3 d, D3 w, float h) {
    const double dist = sqrt(static_cast<double>(dot(d, d)));
    const double coef = 8.0 / (static_cast<double>(kPi) * h * h * h);
    if (dist == 0.0) return scale(w, -12.0 * coef / (h * h));
    if (dist >= h) return {0.0, 0.0, 0.0};
    const double q = dist / h;
    const double deriv = q <= 0.5 ? 18.0 * q * q - 12.0 * q : -6.0 * (1.0 - q) * (1.0 - q);
    const double second = q <= 0.5 ? 36.0 * q - 12.0 : 12.0 * (1.0 - q);
    const double rs = coef * deriv / (h * dist);
    const double rsd = coef * (second / (h * h * dist) - deriv / (h * dist * dist));
    const D3 dd = d3(d);
    return add(scale(w, rs), scale(dd, rsd * dot(w, dd) / dist));
}

// viscosity_laplacian_gradient (density/device.cuh:84-88) with pow(h, 6)
// evaluated on the host.
__device__ __forceinline__ D3 viscosity_laplacian_gradient(Float3 d, float h, double pi_h6, double s) {
    const double dist = sqrt(static_cast<double>(dot(d, d)));
    if (dist == 0.0 || dist >= h) return {0.0, 0.0, 0.0};
    return scale(d3(d), -45.0 * s / (pi_h6 * dist));
}

// (p / rho^2, its tangent), pressure_jvp_kernel (common.cu:201-202)
__device__ __forceinline__ float2 first_term(float p, float p_t, float rho, float rho_t) {
    return make_float2(p / (rho * rho), p_t / (rho * rho) - 2.0F * p * rho_t / (rho * rho * rho));
}

// Tangent of the prediction (predict_jvp_kernel, dfsph.cu:46-66) with the
// recorded collision bits.
__device__ __forceinline__ void predict_tangent(uint32_t bits, uint32_t no_slip, float dt, Float3 xt, Float3 vt,
                                                Float3 at, Float3& out_x, Float3& out_v) {
    Float3 vd = add(vt, scale(at, dt));
    Float3 xd = add(xt, scale(vd, dt));
    if (bits & 1u) xd.x = 0.0F;
    if (bits & 2u) xd.y = 0.0F;
    if (bits & 4u) xd.z = 0.0F;
    if (no_slip != 0u && bits != 0u) {
        vd = Float3{0.0F, 0.0F, 0.0F};
    } else {
        if (bits & 1u) vd.x = 0.0F;
        if (bits & 2u) vd.y = 0.0F;
        if (bits & 4u) vd.z = 0.0F;
    }
    out_x = xd;
    out_v = vd;
}

// VJP of the prediction: position and velocity adjoints from the adjoints
// xa, va of the predicted position and velocity (dfsph.cu:68-102). Under a
// no-slip collision the reference zeroes the whole velocity adjoint
// (dfsph.cu:81, common.cu:357); with `exact` the dt * xa terms of the
// components that did not collide are kept, which makes it the transpose
// of predict_tangent.
__device__ __forceinline__ void predict_adjoint(uint32_t bits, uint32_t no_slip, uint32_t exact, double dt, D3 xa,
                                                D3 va, D3& position, D3& velocity) {
    position = D3{bits & 1u ? 0.0 : xa.x, bits & 2u ? 0.0 : xa.y, bits & 4u ? 0.0 : xa.z};
    if (no_slip != 0u && bits != 0u) {
        velocity = exact != 0u ? scale(position, dt) : D3{0.0, 0.0, 0.0};
    } else {
        velocity = add(va, scale(position, dt));
        if (bits & 1u) velocity.x = 0.0;
        if (bits & 2u) velocity.y = 0.0;
        if (bits & 4u) velocity.z = 0.0;
    }
}

__device__ __forceinline__ float or_zero(const float* f, uint32_t j) { return f != nullptr ? f[j] : 0.0F; }
// out[j] += v unless the output is absent
__device__ __forceinline__ void accumulate_to(double* out, uint32_t j, double v) {
    if (out != nullptr) out[j] += v;
}

}  // namespace
}  // namespace phx
