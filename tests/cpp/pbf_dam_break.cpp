// The reference's only caller of the liquid path, examples/fluids/liquid/
// pbf-dam-break/{main.cpp,simulation.ixx}, written against the physica_amd
// C++ API: same lattice (16 x 28 x 12, spacing 0.025), time step 1/240,
// parameters and box. Usage: pbf_dam_break <steps> [positions.bin]
// prints the step/time lines of main.cpp:12-15 and, if a path is given,
// writes the final positions as raw float32 x[N] y[N] z[N].
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <vector>

#include "physica_amd/particle.hpp"

namespace particle = physica_amd::fluids::liquid::particle;

namespace {

constexpr std::uint32_t lattice[3] = {16u, 28u, 12u};
constexpr std::uint32_t particle_count = lattice[0] * lattice[1] * lattice[2];
constexpr float spacing = 0.025F;
constexpr float time_step = 1.0F / 240.0F;
constexpr float particle_radius = 0.010F;
constexpr float support_radius = 0.050F;

particle::DomainConfiguration create_configuration() {
    return {
        .particle_count = particle_count,
        .time_step = time_step,
        .support_radius = support_radius,
        .particle_radius = particle_radius,
        .gravity = {.x = 0.0F, .y = -9.81F, .z = 0.0F},
        .boundary = {.minimum = {0.0F, 0.0F, 0.0F}, .maximum = {0.8F, 1.1F, 0.5F}, .velocity = {}, .no_slip = true},
        .boundary_particles = {},
    };
}

void fill(particle::Domain& domain, physica_amd::device_buffer<float>& buffer, float value) {
    std::vector<float> values(particle_count, value);
    domain.upload(values, buffer);
}

}  // namespace

int main(int argc, char** argv) {
    try {
        const int steps = argc > 1 ? std::atoi(argv[1]) : 240;
        hipStream_t stream = nullptr;
        physica_amd::check_hip(hipStreamCreate(&stream), "hipStreamCreate");
        particle::PBF solver(create_configuration(), {.pressure_iterations = 5u, .checkpoint_interval = 2u},
                             particle::ExecutionMode::forward, stream);
        auto current = solver.allocate_state();
        auto next = solver.allocate_state();
        auto cache = solver.allocate_step_cache();
        auto control = solver.allocate_control();
        auto parameters = solver.allocate_parameters();
        fill(solver.domain, parameters.particles.masses, 1000.0F * spacing * spacing * spacing);
        fill(solver.domain, parameters.particles.rest_densities, 1000.0F);
        fill(solver.domain, parameters.particles.viscosities, 0.0F);
        fill(solver.domain, parameters.particles.surface_tensions, 0.0F);
        fill(solver.domain, parameters.relaxation, 1000.0F);
        fill(solver.domain, parameters.artificial_pressure_strength, 0.0F);
        fill(solver.domain, parameters.artificial_pressure_exponent, 4.0F);
        fill(solver.domain, parameters.artificial_pressure_radius, 0.3F * support_radius);
        fill(solver.domain, parameters.xsph_viscosity, 0.0F);
        fill(solver.domain, parameters.vorticity_confinement, 0.0F);
        std::vector<particle::Vector3> positions(particle_count);
        for (std::uint32_t z = 0u; z < lattice[2]; ++z)
            for (std::uint32_t y = 0u; y < lattice[1]; ++y)
                for (std::uint32_t x = 0u; x < lattice[0]; ++x)
                    positions[(z * lattice[1] + y) * lattice[0] + x] = {0.08F + static_cast<float>(x) * spacing,
                                                                        0.08F + static_cast<float>(y) * spacing,
                                                                        0.08F + static_cast<float>(z) * spacing};
        solver.domain.upload(positions, current.particles.positions);
        std::printf("PBF dam break with %u particles\n", particle_count);
        for (int step = 0; step < steps; ++step) {
            solver.forward_step(current, control, parameters, next, cache);
            std::swap(current, next);
            const auto index = current.particles.step_index;
            if (index % 60u == 0u) std::printf("step %3llu  time %.3f s\n", static_cast<unsigned long long>(index),
                                               static_cast<double>(index) * time_step);
        }
        physica_amd::check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
        if (argc > 2) {
            std::vector<float> host(3u * particle_count);
            const float* fields[3] = {current.particles.positions.x.data(), current.particles.positions.y.data(),
                                      current.particles.positions.z.data()};
            for (int a = 0; a < 3; ++a)
                physica_amd::check_hip(hipMemcpy(host.data() + a * particle_count, fields[a],
                                                 particle_count * sizeof(float), hipMemcpyDeviceToHost),
                                       "hipMemcpy");
            FILE* f = std::fopen(argv[2], "wb");
            if (f == nullptr) throw std::runtime_error("cannot open output");
            std::fwrite(host.data(), sizeof(float), host.size(), f);
            std::fclose(f);
        }
        (void)hipStreamDestroy(stream);
        return 0;
    } catch (const std::exception& error) {
        std::fprintf(stderr, "PBF dam-break example failed: %s\n", error.what());
        return 1;
    }
}
