// Drives the C++ API (include/physica_amd/particle.hpp) through every solver
// and derivative entry point on small scenes:
//   - SPH<DFSPH|PCISPH|IISPH|WCSPH> and PBF: forward_step of a jittered block
//     (positions written to <out>/<name>_positions.bin), then in
//     differentiable mode jvp_step and vjp_step with seeded directions t, w
//     over every state, control and parameter field, printing
//     "<name> dot <left> <right>" for <J t, w> and <t, J^T w> (FP64 sums);
//   - SPH<DFSPH> in tolerance mode, printing the executed iterations;
//   - NeighborhoodSearch + density::sph_forward with boundary samples
//     (densities written to <out>/density.bin), density::sph_jvp (tangent
//     written to <out>/density_tangent.bin) and sph_vjp ("density dot");
//   - Domain::accumulate (prints "accumulate ok").
// Usage: api_coverage <out_dir>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "physica_amd/particle.hpp"

namespace pp = physica_amd::fluids::liquid::particle;

namespace {

constexpr int kSide = 6;
constexpr std::uint32_t kCount = kSide * kSide * kSide;
constexpr float kSpacing = 0.02F;

// Deterministic uniform values in [-1, 1) (64-bit LCG).
struct Random {
    std::uint64_t state;
    float next() {
        state = state * 6364136223846793005ull + 1442695040888963407ull;
        return static_cast<float>((state >> 40) & 0xFFFFFFu) / 8388608.0F - 1.0F;
    }
};

pp::DomainConfiguration domain(float dt) {
    pp::DomainConfiguration d{};
    d.particle_count = kCount;
    d.time_step = dt;
    d.support_radius = 2.0F * kSpacing;
    d.particle_radius = 0.5F * kSpacing;
    d.gravity = {0.0F, -9.81F, 0.0F};
    d.boundary = {{0.0F, 0.0F, 0.0F}, {0.4F, 0.4F, 0.4F}, {0.0F, 0.0F, 0.0F}, true};
    return d;
}

std::vector<pp::Vector3> block_positions() {
    std::vector<pp::Vector3> x(kCount);
    Random r{7};
    for (std::uint32_t k = 0; k < kCount; ++k) {
        const float i = static_cast<float>(k % kSide), j = static_cast<float>((k / kSide) % kSide),
                    l = static_cast<float>(k / (kSide * kSide));
        // 0.95 spacing: the block starts compressed, so pressures act
        x[k] = {0.1F + 0.95F * kSpacing * i + 0.001F * r.next(), 0.05F + 0.95F * kSpacing * j + 0.001F * r.next(),
                0.1F + 0.95F * kSpacing * l + 0.001F * r.next()};
    }
    return x;
}

std::vector<pp::Vector3> swirl_velocities() {
    std::vector<pp::Vector3> v(kCount);
    Random r{11};
    for (auto& e : v) e = {0.2F * r.next(), 0.2F * r.next(), 0.2F * r.next()};
    return v;
}

template <class T>
std::vector<T> download(const physica_amd::device_buffer<T>& b) {
    std::vector<T> h(b.size());
    physica_amd::check_hip(hipMemcpy(h.data(), b.data(), b.size() * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
    return h;
}

// the host vector ends with this call, so the copy is waited for
void upload(const pp::Domain& d, std::vector<float> values, physica_amd::device_buffer<float>& b) {
    d.upload(values, b);
    physica_amd::check_hip(hipStreamSynchronize(d.stream), "hipStreamSynchronize");
}

void fill(const pp::Domain& d, physica_amd::device_buffer<float>& b, float value) {
    upload(d, std::vector<float>(b.size(), value), b);
}

void random_fill(const pp::Domain& d, physica_amd::device_buffer<float>& b, Random& r, float scale) {
    std::vector<float> v(b.size());
    for (auto& e : v) e = scale * r.next();
    upload(d, v, b);
}

void random_fill(const pp::Domain& d, pp::VectorField& f, Random& r, float scale) {
    random_fill(d, f.x, r, scale);
    random_fill(d, f.y, r, scale);
    random_fill(d, f.z, r, scale);
}

void write_positions(const std::string& path, const pp::VectorField& f) {
    std::ofstream out(path, std::ios::binary);
    for (const auto* b : {&f.x, &f.y, &f.z}) {
        const std::vector<float> h = download(*b);
        out.write(reinterpret_cast<const char*>(h.data()), static_cast<std::streamsize>(h.size() * sizeof(float)));
    }
}

// FP64 <a, b> of a float tangent and a double adjoint
double dot(const physica_amd::device_buffer<float>& a, const physica_amd::device_buffer<double>& b) {
    const std::vector<float> x = download(a);
    const std::vector<double> y = download(b);
    double s = 0.0;
    for (std::size_t k = 0; k < x.size(); ++k) s += static_cast<double>(x[k]) * y[k];
    return s;
}
double dot(const pp::VectorField& a, const pp::VectorAdjointField& b) { return dot(a.x, b.x) + dot(a.y, b.y) + dot(a.z, b.z); }
double dot(const pp::ParticleParameterTangent& a, const pp::ParticleParameterAdjoint& b) {
    return dot(a.masses, b.masses) + dot(a.rest_densities, b.rest_densities) + dot(a.viscosities, b.viscosities) +
           dot(a.surface_tensions, b.surface_tensions);
}

void upload_adjoint(const pp::Domain& d, pp::VectorAdjointField& f, Random& r) {
    for (auto* b : {&f.x, &f.y, &f.z}) {
        std::vector<double> h(b->size());
        for (auto& e : h) e = r.next();
        physica_amd::check_hip(hipMemcpyAsync(b->data(), h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                                              d.stream),
                               "hipMemcpyAsync");
        physica_amd::check_hip(hipStreamSynchronize(d.stream), "hipStreamSynchronize");
    }
}

// Particle parameters of a water block; tangents perturb each by a few percent.
void water(const pp::Domain& d, pp::ParticleParameters& p) {
    fill(d, p.masses, 1000.0F * kSpacing * kSpacing * kSpacing);
    fill(d, p.rest_densities, 1000.0F);
    fill(d, p.viscosities, 0.01F);
    fill(d, p.surface_tensions, 0.0F);
}

void particle_tangent(const pp::Domain& d, pp::ParticleParameterTangent& t, Random& r, bool surface) {
    random_fill(d, t.masses, r, 1e-3F * kSpacing * kSpacing * kSpacing * 1000.0F);
    random_fill(d, t.rest_densities, r, 5.0F);
    random_fill(d, t.viscosities, r, 1e-3F);
    if (surface) random_fill(d, t.surface_tensions, r, 1e-3F);
}

// Forward step in forward mode, then the derivative identity in differentiable mode
// (the SPH solvers with the exact transpose of a no-slip collision, for which
// the identity holds).
template <class Solver, class Configuration, class Setup, class Tangent, class MethodDot>
void run(const std::string& name, const std::string& out_dir, float dt, Configuration configuration, Setup&& setup,
         Tangent&& tangent, MethodDot&& method_dot) {
    const auto x0 = block_positions();
    const auto v0 = swirl_velocities();
    {
        // the initial state, for the comparison with the oracle
        pp::Domain d(domain(dt), nullptr);
        pp::VectorField x = d.allocate_vector_field(kCount), v = d.allocate_vector_field(kCount);
        d.upload(x0, x);
        d.upload(v0, v);
        write_positions(out_dir + "/initial_positions.bin", x);
        write_positions(out_dir + "/initial_velocities.bin", v);
    }
    {
        Solver solver(domain(dt), configuration, pp::ExecutionMode::forward, nullptr);
        auto state = solver.allocate_state();
        auto next = solver.allocate_state();
        auto control = solver.allocate_control();
        auto parameters = solver.allocate_parameters();
        auto cache = solver.allocate_step_cache();
        solver.domain.upload(x0, state.particles.positions);
        solver.domain.upload(v0, state.particles.velocities);
        setup(solver.domain, parameters);
        solver.forward_step(state, control, parameters, next, cache);
        physica_amd::check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
        write_positions(out_dir + "/" + name + "_positions.bin", next.particles.positions);
    }
    Solver solver(domain(dt), configuration, pp::ExecutionMode::differentiable, nullptr);
    auto state = solver.allocate_state();
    auto next = solver.allocate_state();
    auto control = solver.allocate_control();
    auto parameters = solver.allocate_parameters();
    auto cache = solver.allocate_step_cache();
    solver.domain.upload(x0, state.particles.positions);
    solver.domain.upload(v0, state.particles.velocities);
    setup(solver.domain, parameters);
    solver.forward_step(state, control, parameters, next, cache);

    Random r{23};
    auto t = solver.allocate_state_tangent();
    random_fill(solver.domain, t.particles.positions, r, 1e-4F);
    random_fill(solver.domain, t.particles.velocities, r, 1e-2F);
    auto ct = solver.allocate_control_tangent();
    random_fill(solver.domain, ct.external_accelerations, r, 1e-1F);
    auto pt = solver.allocate_parameter_tangent();
    tangent(solver.domain, t, pt, r);
    auto nt = solver.allocate_state_tangent();
    solver.jvp_step(state, control, parameters, next, cache, t, ct, pt, nt);

    auto w = solver.allocate_state_adjoint();
    upload_adjoint(solver.domain, w.particles.positions, r);
    upload_adjoint(solver.domain, w.particles.velocities, r);
    auto prev = solver.allocate_state_adjoint();
    auto ca = solver.allocate_control_adjoint();
    auto pa = solver.allocate_parameter_adjoint();
    solver.vjp_step(state, control, parameters, next, cache, w, prev, ca, pa);
    physica_amd::check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");

    const double left = dot(nt.particles.positions, w.particles.positions) +
                        dot(nt.particles.velocities, w.particles.velocities);
    const double right = dot(t.particles.positions, prev.particles.positions) +
                         dot(t.particles.velocities, prev.particles.velocities) +
                         dot(ct.external_accelerations, ca.external_accelerations) +
                         dot(pt.particles, pa.particles) + method_dot(t, pt, prev, pa);
    std::printf("%s dot %.17g %.17g\n", name.c_str(), left, right);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s <out_dir>\n", argv[0]);
        return 2;
    }
    const std::string out = argv[1];
    try {
        run<pp::SPH<pp::DFSPH>>(
            "dfsph", out, 1e-3F, pp::DFSPH::Configuration{3u, 4u, 2u, false, 0.0F, 100u, true},
            [](const pp::Domain& d, auto& p) {
                water(d, p.particles);
                fill(d, p.method.divergence_relaxation, 0.5F);
                fill(d, p.method.density_relaxation, 0.5F);
            },
            [](const pp::Domain& d, auto&, auto& pt, Random& r) {
                particle_tangent(d, pt.particles, r, false);
                random_fill(d, pt.method.divergence_relaxation, r, 0.05F);
                random_fill(d, pt.method.density_relaxation, r, 0.05F);
            },
            [](auto&, auto& pt, auto&, auto& pa) {
                return dot(pt.method.divergence_relaxation, pa.method.divergence_relaxation) +
                       dot(pt.method.density_relaxation, pa.method.density_relaxation);
            });
        run<pp::SPH<pp::PCISPH>>(
            "pcisph", out, 1e-3F, pp::PCISPH::Configuration{4u, 2u, 0.0F, 100u, true},
            [](const pp::Domain& d, auto& p) {
                water(d, p.particles);
                fill(d, p.method.pressure_relaxation, 0.5F);
            },
            [](const pp::Domain& d, auto&, auto& pt, Random& r) {
                particle_tangent(d, pt.particles, r, false);
                random_fill(d, pt.method.pressure_relaxation, r, 0.05F);
            },
            [](auto&, auto& pt, auto&, auto& pa) {
                return dot(pt.method.pressure_relaxation, pa.method.pressure_relaxation);
            });
        run<pp::SPH<pp::IISPH>>(
            "iisph", out, 1e-3F, pp::IISPH::Configuration{4u, 2u, 0.0F, 100u, true},
            [](const pp::Domain& d, auto& p) {
                water(d, p.particles);
                fill(d, p.method.jacobi_relaxation, 0.5F);
            },
            [](const pp::Domain& d, auto&, auto& pt, Random& r) {
                particle_tangent(d, pt.particles, r, false);
                random_fill(d, pt.method.jacobi_relaxation, r, 0.05F);
            },
            [](auto&, auto& pt, auto&, auto& pa) {
                return dot(pt.method.jacobi_relaxation, pa.method.jacobi_relaxation);
            });
        run<pp::SPH<pp::WCSPH>>(
            "wcsph", out, 2e-4F, pp::WCSPH::Configuration{true},
            [](const pp::Domain& d, auto& p) {
                water(d, p.particles);
                fill(d, p.particles.surface_tensions, 0.01F);
                fill(d, p.method.speed_of_sound, 5.0F);
                fill(d, p.method.tait_exponent, 1.0F);
                fill(d, p.method.boundary_surface_tension, 0.01F);
            },
            [](const pp::Domain& d, auto&, auto& pt, Random& r) {
                particle_tangent(d, pt.particles, r, true);
                random_fill(d, pt.method.speed_of_sound, r, 0.05F);
                random_fill(d, pt.method.tait_exponent, r, 0.01F);
            },
            [](auto&, auto& pt, auto&, auto& pa) {
                return dot(pt.method.speed_of_sound, pa.method.speed_of_sound) +
                       dot(pt.method.tait_exponent, pa.method.tait_exponent) +
                       dot(pt.method.boundary_surface_tension, pa.method.boundary_surface_tension);
            });
        run<pp::PBF>(
            "pbf", out, 1.0F / 120.0F, pp::PBF::Configuration{3u, 2u},
            [](const pp::Domain& d, auto& p) {
                water(d, p.particles);
                fill(d, p.relaxation, 1e-3F);
                fill(d, p.artificial_pressure_strength, 0.0F);
                fill(d, p.artificial_pressure_exponent, 4.0F);
                fill(d, p.artificial_pressure_radius, 0.3F * 2.0F * kSpacing);
                fill(d, p.xsph_viscosity, 0.01F);
                fill(d, p.vorticity_confinement, 0.0F);
            },
            [](const pp::Domain& d, auto&, auto& pt, Random& r) {
                particle_tangent(d, pt.particles, r, false);
                random_fill(d, pt.relaxation, r, 1e-4F);
                random_fill(d, pt.xsph_viscosity, r, 1e-3F);
            },
            [](auto&, auto& pt, auto&, auto& pa) {
                return dot(pt.relaxation, pa.relaxation) + dot(pt.xsph_viscosity, pa.xsph_viscosity) +
                       dot(pt.artificial_pressure_strength, pa.artificial_pressure_strength) +
                       dot(pt.vorticity_confinement, pa.vorticity_confinement);
            });
        {  // tolerance mode
            pp::DFSPH::Configuration c{};
            c.density_error_tolerance = 1e-3F;
            c.max_iterations = 50u;
            pp::SPH<pp::DFSPH> solver(domain(1e-3F), c, pp::ExecutionMode::forward, nullptr);
            auto state = solver.allocate_state();
            auto next = solver.allocate_state();
            auto control = solver.allocate_control();
            auto parameters = solver.allocate_parameters();
            auto cache = solver.allocate_step_cache();
            solver.domain.upload(block_positions(), state.particles.positions);
            water(solver.domain, parameters.particles);
            fill(solver.domain, parameters.method.divergence_relaxation, 0.5F);
            fill(solver.domain, parameters.method.density_relaxation, 0.5F);
            solver.forward_step(state, control, parameters, next, cache);
            const auto it = solver.last_iterations();
            std::printf("tolerance iterations %u %u\n", it[0], it[1]);
        }

        {  // neighbourhood search and density with boundary samples
            pp::DomainConfiguration c = domain(1e-3F);
            for (int i = 0; i < 10; ++i)
                for (int k = 0; k < 10; ++k)
                    c.boundary_particles.push_back({{0.08F + kSpacing * i, 0.035F, 0.08F + kSpacing * k},
                                                    {0.0F, 0.0F, 0.1F},
                                                    0.8F * kSpacing * kSpacing * kSpacing});
            pp::Domain d(c, nullptr);
            pp::NeighborhoodSearch search(d);
            pp::Neighborhood nb = search.allocate();
            pp::VectorField x = d.allocate_vector_field(kCount);
            d.upload(block_positions(), x);
            search.build(3u, x, nb);
            pp::DeviceBoundary boundary(d);
            pp::ParticleParameters p = d.allocate_particle_parameters();
            water(d, p);
            pp::ScalarField rho = d.allocate_scalar_field(kCount);
            pp::density::sph_forward(d, x, x, p, nb, rho, &boundary);
            physica_amd::check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
            const std::vector<float> h = download(rho.values);
            std::ofstream f(out + "/density.bin", std::ios::binary);
            f.write(reinterpret_cast<const char*>(h.data()), static_cast<std::streamsize>(h.size() * sizeof(float)));

            // density::sph_jvp with a seeded position tangent, and sph_vjp:
            // <J t, w> = <t, J^T w> (the mass and rest-density tangents are 0)
            Random r{31};
            pp::VectorField t = d.allocate_vector_field(kCount);
            random_fill(d, t, r, 1.0F);
            pp::ParticleParameterTangent pt = d.allocate_particle_parameter_tangent();
            pp::ScalarField rho_t = d.allocate_scalar_field(kCount);
            pp::density::sph_jvp(d, x, x, t, p, pt, nb, rho_t, &boundary);
            pp::ScalarAdjointField w = d.allocate_scalar_adjoint_field(kCount);
            std::vector<double> hw(kCount);
            for (auto& e : hw) e = r.next();
            physica_amd::check_hip(hipMemcpy(w.values.data(), hw.data(), kCount * sizeof(double), hipMemcpyHostToDevice),
                                   "hipMemcpy");
            pp::VectorAdjointField xa = d.allocate_vector_adjoint_field(kCount);
            d.clear(xa);
            pp::ParticleParameterAdjoint pa = d.allocate_particle_parameter_adjoint();
            pp::density::sph_vjp(d, x, x, p, nb, w, xa, pa, &boundary);
            physica_amd::check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
            const std::vector<float> ht = download(rho_t.values);
            double left = 0.0;
            for (std::uint32_t k = 0; k < kCount; ++k) left += static_cast<double>(ht[k]) * hw[k];
            std::printf("density dot %.17g %.17g\n", left, dot(t, xa));
            std::ofstream ft(out + "/density_tangent.bin", std::ios::binary);
            ft.write(reinterpret_cast<const char*>(ht.data()), static_cast<std::streamsize>(ht.size() * sizeof(float)));
            write_positions(out + "/density_position_tangent.bin", t);
        }

        {  // Domain::accumulate
            pp::Domain d(domain(1e-3F), nullptr);
            pp::ScalarAdjointField a = d.allocate_scalar_adjoint_field(1000), b = d.allocate_scalar_adjoint_field(1000);
            std::vector<double> ha(1000), hb(1000);
            for (int k = 0; k < 1000; ++k) {
                ha[k] = 0.5 * k;
                hb[k] = 1.0 - k;
            }
            physica_amd::check_hip(hipMemcpy(a.values.data(), ha.data(), 8000, hipMemcpyHostToDevice), "hipMemcpy");
            physica_amd::check_hip(hipMemcpy(b.values.data(), hb.data(), 8000, hipMemcpyHostToDevice), "hipMemcpy");
            d.accumulate(a, b);
            const std::vector<double> r = download(b.values);
            bool ok = true;
            for (int k = 0; k < 1000; ++k) ok = ok && r[k] == 1.0 - 0.5 * k;
            std::printf("accumulate %s\n", ok ? "ok" : "wrong");
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
