/*
 * CPU oracle for the particle-liquid path: TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of Xayah-Graphics/physica's liquid kernels
 * (fluids/liquid/particle/) in the reference's own data layout (original
 * particle order, SoA) and summation order: 27 cells visited offset_x outer,
 * offset_y middle, offset_z inner, candidates within a cell in ascending
 * original index, fluid candidates before boundary candidates.
 *
 * The same source (oracle_impl.c) is compiled twice:
 *   *_f32 : FP32 "reference semantics" (-ffp-contract=off, IEEE +,-,*,/,sqrt)
 *   *_f64 : FP64 oracle with the same structure, used to state the error of
 *           the FP32 results.
 *
 * Only tests/, the smoke() check and bench.py's cpu_baseline leg may load
 * this library; the product path never does.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   grid build     stable-sort definition (numpy stable argsort), lower_bound
 *                  (numpy searchsorted), hand-placed known-answer cells
 *   kernel funcs   normalisation by quadrature, closed-form values, FD of W
 *   density        brute-force O(N^2) sum, closed lattice sum
 *   pressure       momentum conservation, repulsion sign on two particles
 *   lambda         closed form on two particles
 *   predict/proj.  closed form per particle
 *   G (DFSPH)      FP64 recomputation of the lattice sum
 * No golden vectors exist in the reference (it ships no tests), and the
 * reference cannot be compiled in this image (nvcc, CCCL/CUB and C++23
 * `import std` are absent), so no check against the reference's own binary
 * is possible.
 */
#ifndef PHYSICA_ORACLE_H
#define PHYSICA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* DomainConfiguration (fluids/liquid/particle/domain/domain.ixx:31-39). Always
 * FP32 as in the reference; the f64 oracle promotes the values it uses. */
typedef struct oracle_domain {
    uint32_t particle_count;
    float time_step;
    float support_radius;
    float particle_radius;
    float gravity[3];
    float box_minimum[3];
    float box_maximum[3];
    float box_velocity[3];
    uint32_t no_slip;
    uint32_t boundary_count;
    const float* boundary_position[3]; /* host arrays, boundary_count each */
    const float* boundary_velocity[3];
    const float* boundary_volume;
} oracle_domain;

/* Grid resolution exactly as neighborhood.cpp:25-30 (FP32 ceil of the box
 * extent over the support radius). */
void oracle_grid_resolution(const oracle_domain* domain, uint32_t resolution[3]);

int oracle_threads(void);
void oracle_set_threads(int threads);

#define ORACLE_DECLARE(real, suf)                                                                                      \
    /* neighborhood.cu:20-70 + neighborhood/device.cuh: keys, stable sort, lower_bound offsets */                      \
    void oracle_build_neighborhood##suf(const oracle_domain* domain, uint64_t step_index, const real* px,             \
                                        const real* py, const real* pz, uint64_t* sorted_keys,                          \
                                        uint32_t* sorted_indices, uint32_t* cell_offsets,                               \
                                        uint64_t* sorted_boundary_keys, uint32_t* sorted_boundary_indices,              \
                                        uint32_t* boundary_cell_offsets);                                               \
    /* density/device.cuh:10-88 and pbf.cu:18-71, evaluated on arrays of displacements */                              \
    void oracle_kernel_values##suf(int kind, float support_radius, uint32_t count, const real* dx, const real* dy,    \
                                   const real* dz, real* value, real* gx, real* gy, real* gz);                         \
    /* density.cu:18-40 */                                                                                              \
    void oracle_density##suf(const oracle_domain* domain, uint64_t grid_step, int pbf_kernel, const real* topology[3], \
                             const real* positions[3], const real* masses, const real* rest_densities,                 \
                             real* densities);                                                                          \
    /* density_jvp / density_vjp (density.cu:42-98); the VJP accumulates in FP64 */                                   \
    void oracle_density_jvp##suf(const oracle_domain* domain, uint64_t grid_step, int pbf_kernel,                    \
                                 const real* topology[3], const real* positions[3], const real* position_tangent[3], \
                                 const real* masses, const real* mass_tangent, const real* rest_densities,           \
                                 const real* rest_tangent, real* out);                                               \
    void oracle_density_vjp##suf(const oracle_domain* domain, uint64_t grid_step, int pbf_kernel,                    \
                                 const real* topology[3], const real* positions[3], const real* masses,              \
                                 const real* rest_densities, const double* density_adjoint,                          \
                                 double* position_adjoint[3], double* mass_adjoint, double* rest_adjoint);           \
    /* sph.cpp:37-58 */                                                                                                 \
    real oracle_reference_gradient_norm##suf(const oracle_domain* domain);                                             \
    /* SPH<DFSPH>::forward_step: sph.ixx:350-354 + DFSPH::forward sph.cpp:572-610 (fixed iterations) */                 \
    void oracle_dfsph_step##suf(const oracle_domain* domain, uint32_t divergence_iterations,                          \
                                uint32_t density_iterations, int warm_start, uint64_t step_index,                       \
                                const real* positions[3], const real* velocities[3], const real* controls[3],          \
                                const real* masses, const real* rest_densities, const real* viscosities,               \
                                const real* divergence_relaxation, const real* density_relaxation,                     \
                                const real* warm_divergence, const real* warm_density, real* next_positions[3],         \
                                real* next_velocities[3], real* next_warm_divergence, real* next_warm_density,          \
                                real* densities_out, real* non_pressure_out[3]);                                        \
    /* SPH<PCISPH|IISPH>::forward_step: sph.cpp:296-315 / 409-428 (iisph selects V = m / rho) */                        \
    void oracle_pressure_solver_step##suf(const oracle_domain* domain, int iisph, uint32_t pressure_iterations,       \
                                          uint64_t step_index, const real* positions[3], const real* velocities[3],    \
                                          const real* controls[3], const real* masses, const real* rest_densities,     \
                                          const real* viscosities, const real* relaxation, real* next_positions[3],    \
                                          real* next_velocities[3], real* densities_out, real* non_pressure_out[3],    \
                                          real* pressures_out);                                                         \
    /* PCISPH / IISPH tolerance mode (an extension): iterate until the mean                                          \
     * compression error of the prediction is within the tolerance */                                                 \
    void oracle_pressure_solver_step_tolerance##suf(                                                                  \
        const oracle_domain* domain, int iisph, double tolerance, uint32_t max_iterations, uint64_t step_index,      \
        const real* positions[3], const real* velocities[3], const real* controls[3], const real* masses,          \
        const real* rest_densities, const real* viscosities, const real* relaxation, real* next_positions[3],     \
        real* next_velocities[3], real* densities_out, real* pressures_out, uint32_t* iterations_out);                  \
    /* SPH<WCSPH>::forward_step: sph.cpp:214-226 */                                                                     \
    void oracle_wcsph_step##suf(const oracle_domain* domain, uint64_t step_index, const real* positions[3],           \
                                const real* velocities[3], const real* controls[3], const real* masses,                \
                                const real* rest_densities, const real* viscosities, const real* surface_tensions,     \
                                const real* speed_of_sound, const real* tait_exponent,                                  \
                                const real* boundary_surface_tension, real* next_positions[3],                          \
                                real* next_velocities[3], real* densities_out, real* pressures_out);                    \
    /* PBF::forward_step: pbf.cpp:167-191 */                                                                            \
    void oracle_pbf_step##suf(const oracle_domain* domain, uint32_t pressure_iterations, uint64_t step_index,         \
                              const real* positions[3], const real* velocities[3], const real* controls[3],            \
                              const real* masses, const real* rest_densities, const real* relaxation,                  \
                              const real* strength, const real* exponent, const real* radius, const real* xsph,         \
                              const real* confinement, real* next_positions[3], real* next_velocities[3],               \
                              real* densities_out, real* lambdas_out, real* predicted_out[3],                           \
                              real* vorticity_magnitudes_out);                                                          \
    /* PBF::jvp_step (pbf.cpp:193-215): tangents in the primal precision */                                           \
    void oracle_pbf_jvp##suf(const oracle_domain* domain, uint32_t pressure_iterations, uint64_t step_index,          \
                             const real* positions[3], const real* velocities[3], const real* controls[3],             \
                             const real* masses, const real* rest_densities, const real* relaxation,                   \
                             const real* strength, const real* exponent, const real* radius, const real* xsph,         \
                             const real* confinement, const real* position_tangent[3],                                 \
                             const real* velocity_tangent[3], const real* control_tangent[3],                          \
                             const real* masses_tangent, const real* rest_densities_tangent,                           \
                             const real* relaxation_tangent, const real* strength_tangent,                             \
                             const real* exponent_tangent, const real* radius_tangent, const real* xsph_tangent,        \
                             const real* confinement_tangent, real* next_position_tangent[3],                          \
                             real* next_velocity_tangent[3]);                                                           \
    /* PBF::vjp_step (pbf.cpp:217-258): FP64 adjoints, accumulated into the outputs */                                \
    void oracle_pbf_vjp##suf(const oracle_domain* domain, uint32_t pressure_iterations, uint64_t step_index,          \
                             const real* positions[3], const real* velocities[3], const real* controls[3],             \
                             const real* masses, const real* rest_densities, const real* relaxation,                   \
                             const real* strength, const real* exponent, const real* radius, const real* xsph,         \
                             const real* confinement, const double* next_position_adjoint[3],                          \
                             const double* next_velocity_adjoint[3], double* position_adjoint[3],                      \
                             double* velocity_adjoint[3], double* control_adjoint[3], double* masses_adjoint,          \
                             double* rest_densities_adjoint, double* relaxation_adjoint, double* strength_adjoint,      \
                             double* exponent_adjoint, double* radius_adjoint, double* xsph_adjoint,                    \
                             double* confinement_adjoint);                                                              \
    /* SPH<DFSPH>::forward_step in tolerance mode (an extension): each phase iterates until the FP64 mean          \
     * compression error (1/N) sum max(0, rho* - target) / rho0 of its prediction is <= tolerance or max_iterations  \
     * is reached; iterations_out receives the executed (divergence, density) counts */                               \
    void oracle_dfsph_step_tolerance##suf(const oracle_domain* domain, double tolerance, uint32_t max_iterations,     \
                                          int warm_start, uint64_t step_index, const real* positions[3],               \
                                          const real* velocities[3], const real* controls[3], const real* masses,     \
                                          const real* rest_densities, const real* viscosities,                         \
                                          const real* divergence_relaxation, const real* density_relaxation,          \
                                          const real* warm_divergence, const real* warm_density,                       \
                                          real* next_positions[3], real* next_velocities[3],                           \
                                          real* next_warm_divergence, real* next_warm_density,                         \
                                          uint32_t iterations_out[2]);                                                  \
    /* SPH<DFSPH>::jvp_step (sph.ixx:356-359, sph.cpp:612-659). parameters: masses, rest densities, viscosities,      \
     * divergence relaxation, density relaxation. volume_rest = 1 takes V = m / rho0 in the pressure update: with     \
     * divergence_iterations = 0 and no warm start the step is PCISPH's (volume_rest = 1, sph.cpp:312-331) or        \
     * IISPH's (volume_rest = 0), the density relaxation being their pressure / Jacobi relaxation */                   \
    void oracle_dfsph_jvp##suf(const oracle_domain* domain, uint32_t divergence_iterations,                           \
                               uint32_t density_iterations, int warm_start, int volume_rest, uint64_t step_index,       \
                               const real* positions[3], const real* velocities[3], const real* controls[3],           \
                               const real* const parameters[5], const real* warm_divergence,                            \
                               const real* warm_density, const real* position_tangent[3],                               \
                               const real* velocity_tangent[3], const real* control_tangent[3],                         \
                               const real* const parameter_tangent[5], const real* warm_divergence_tangent,             \
                               const real* warm_density_tangent, real* next_position_tangent[3],                        \
                               real* next_velocity_tangent[3], real* next_warm_divergence_tangent,                      \
                               real* next_warm_density_tangent);                                                        \
    /* SPH<DFSPH>::vjp_step (sph.ixx:361-365, sph.cpp:661-723): FP64 adjoints accumulated into the outputs, warm     \
     * pressure adjoints copied when warm_start */                                                                      \
    void oracle_dfsph_vjp##suf(const oracle_domain* domain, uint32_t divergence_iterations,                           \
                               uint32_t density_iterations, int warm_start, int volume_rest, uint64_t step_index,       \
                               const real* positions[3], const real* velocities[3], const real* controls[3],           \
                               const real* const parameters[5], const real* warm_divergence,                            \
                               const real* warm_density, const double* next_position_adjoint[3],                        \
                               const double* next_velocity_adjoint[3], const double* next_warm_divergence_adjoint,      \
                               const double* next_warm_density_adjoint, double* position_adjoint[3],                    \
                               double* velocity_adjoint[3], double* control_adjoint[3],                                 \
                               double* const parameter_adjoint[5], double* warm_divergence_adjoint,                     \
                               double* warm_density_adjoint);                                                           \
    /* SPH<WCSPH>::jvp_step / vjp_step (sph.ixx:356-365, sph.cpp:228-249). parameters: masses, rest densities,     \
     * viscosities, surface tensions, speed of sound, Tait exponent, boundary surface tension. The VJP accumulates   \
     * FP64 adjoints into its outputs */                                                                               \
    void oracle_wcsph_jvp##suf(const oracle_domain* domain, uint64_t step_index, const real* positions[3],            \
                               const real* velocities[3], const real* controls[3], const real* const parameters[7],    \
                               const real* position_tangent[3], const real* velocity_tangent[3],                        \
                               const real* control_tangent[3], const real* const parameter_tangent[7],                  \
                               real* next_position_tangent[3], real* next_velocity_tangent[3]);                         \
    void oracle_wcsph_vjp##suf(const oracle_domain* domain, uint64_t step_index, const real* positions[3],            \
                               const real* velocities[3], const real* controls[3], const real* const parameters[7],    \
                               const double* next_position_adjoint[3], const double* next_velocity_adjoint[3],          \
                               double* position_adjoint[3], double* velocity_adjoint[3], double* control_adjoint[3],   \
                               double* const parameter_adjoint[7]);                                                     \
    /* the primal step the two functions above differentiate: next positions and velocities */                        \
    void oracle_dfsph_record_step##suf(const oracle_domain* domain, uint32_t divergence_iterations,                   \
                                       uint32_t density_iterations, int warm_start, int volume_rest,                    \
                                       uint64_t step_index, const real* positions[3], const real* velocities[3],      \
                                       const real* controls[3], const real* const parameters[5],                       \
                                       const real* warm_divergence, const real* warm_density,                          \
                                       real* next_positions[3], real* next_velocities[3]);                             \
    /* 1: restate the reference's VJP of a no-slip collision in the SPH prediction (sph/dfsph.cu:81); 0 (default): \
     * the exact transpose of the prediction's JVP */                                                                   \
    void oracle_set_reference_no_slip_adjoint##suf(int enabled);                                                        \
    /* single kernels on an existing grid, for unit tests (sph/common.cu:14-47, 165-194) */                            \
    void oracle_non_pressure##suf(const oracle_domain* domain, uint64_t grid_step, const real* positions[3],          \
                                  const real* velocities[3], const real* controls[3], const real* masses,              \
                                  const real* viscosities, const real* densities, real* out[3]);                       \
    void oracle_pressure##suf(const oracle_domain* domain, uint64_t grid_step, const real* positions[3],              \
                              const real* masses, const real* rest_densities, const real* densities,                   \
                              const real* pressures, real* out[3]);                                                     \
    /* pbf.cu:95-127 */                                                                                                 \
    void oracle_pbf_lambda##suf(const oracle_domain* domain, uint64_t grid_step, const real* topology[3],             \
                                const real* positions[3], const real* masses, const real* rest_densities,              \
                                const real* densities, const real* relaxation, real* gradient_sums[3],                  \
                                real* denominators, real* lambdas);

ORACLE_DECLARE(float, _f32)
ORACLE_DECLARE(double, _f64)

#undef ORACLE_DECLARE

/* kernel kinds for oracle_kernel_values */
enum {
    ORACLE_POLY6 = 0,
    ORACLE_CUBIC = 1,
    ORACLE_SPIKY = 2,
    ORACLE_VISCOSITY_LAPLACIAN = 3,
};

#ifdef __cplusplus
}
#endif

#endif
