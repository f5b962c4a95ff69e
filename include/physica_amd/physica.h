/*
 * physica_amd: MI355X (gfx950) particle-liquid engine, C ABI.
 *
 * Drop-in boundary for the particle-liquid path of Xayah-Graphics/physica
 * (fluids/liquid/particle/). Plain C types only: device pointers, sizes,
 * PODs, and an opaque handle per solver. All device buffers named in the
 * argument lists are caller-owned and live on the device of the stream the
 * solver was created with. Every entry point returns a phx_status; nothing
 * throws across the ABI. Work is enqueued on the solver's stream and is
 * asynchronous unless stated otherwise.
 *
 * Reference interfaces replaced (paths relative to fluids/liquid/particle/):
 *   phx_domain_configuration   DomainConfiguration      domain/domain.ixx:12-44
 *   phx_build_neighborhood     NeighborhoodSearch::build neighborhood/neighborhood.cpp:44-89
 *                              cuda_detail::build_neighborhood neighborhood/neighborhood.cu:54-70
 *   phx_density_forward        density::sph_forward/pbf_forward density/density.cpp:12-34,
 *                              cuda_detail::density_forward density/density.cu:104-106
 *   phx_dfsph_*                SPH<DFSPH> sph/sph.ixx:273-372, DFSPH sph/sph.ixx:195-268,
 *                              DFSPH::forward sph/sph.cpp:572-610
 *   phx_pbf_*                  PBF pbf/pbf.ixx:14-141, PBF::forward_step pbf/pbf.cpp:167-191
 */
#ifndef PHYSICA_AMD_PHYSICA_H
#define PHYSICA_AMD_PHYSICA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* A hipStream_t; 0 is the null stream. */
typedef void* phx_stream;

typedef enum phx_status {
    PHX_SUCCESS = 0,
    PHX_ERROR_INVALID_ARGUMENT = 1,
    PHX_ERROR_UNSUPPORTED = 2,
    PHX_ERROR_WRONG_MODE = 3,
    PHX_ERROR_WORKSPACE_TOO_SMALL = 4,
    PHX_ERROR_OUT_OF_MEMORY = 5,
    PHX_ERROR_DEVICE = 6,
    PHX_ERROR_CAPACITY = 7
} phx_status;

/* Message of the last failing call on this host thread ("" if none). */
const char* phx_last_error(void);
const char* phx_version(void);

/* ---------------------------------------------------------------- domain */
/* Vector3, BoxBoundary, BoundaryParticle, DomainConfiguration, ExecutionMode:
 * domain/domain.ixx:12-44 */
typedef struct phx_vector3 {
    float x, y, z;
} phx_vector3;

typedef struct phx_box_boundary {
    phx_vector3 minimum;
    phx_vector3 maximum;
    phx_vector3 velocity;
    uint32_t no_slip;
} phx_box_boundary;

typedef struct phx_boundary_particle {
    phx_vector3 position;
    phx_vector3 velocity;
    float volume;
} phx_boundary_particle;

typedef struct phx_domain_configuration {
    uint32_t particle_count;
    float time_step;
    float support_radius;
    float particle_radius;
    phx_vector3 gravity;
    phx_box_boundary boundary;
    const phx_boundary_particle* boundary_particles; /* host array */
    uint32_t boundary_particle_count;
} phx_domain_configuration;

typedef enum phx_execution_mode { PHX_MODE_FORWARD = 0, PHX_MODE_DIFFERENTIABLE = 1 } phx_execution_mode;

/* Device SoA views, domain/device.h:20-36 */
typedef struct phx_vector_view {
    float* x;
    float* y;
    float* z;
} phx_vector_view;

typedef struct phx_const_vector_view {
    const float* x;
    const float* y;
    const float* z;
} phx_const_vector_view;

/* ParticleState, Control, ParticleParameters: domain/domain.ixx:70-107.
 * step_index is host-side, as in the reference. */
typedef struct phx_particle_state {
    phx_vector_view positions;
    phx_vector_view velocities;
    uint64_t step_index;
} phx_particle_state;

typedef struct phx_control {
    phx_const_vector_view external_accelerations;
} phx_control;

typedef struct phx_particle_parameters {
    const float* masses;
    const float* rest_densities;
    const float* viscosities;
    const float* surface_tensions;
} phx_particle_parameters;

/* Tangent and adjoint counterparts (domain/domain.ixx:76-113; adjoint
 * fields are FP64, domain/device.h). */
typedef struct phx_adjoint_vector_view {
    double* x;
    double* y;
    double* z;
} phx_adjoint_vector_view;

typedef struct phx_const_adjoint_vector_view {
    const double* x;
    const double* y;
    const double* z;
} phx_const_adjoint_vector_view;

typedef struct phx_control_tangent {
    phx_const_vector_view external_accelerations;
} phx_control_tangent;

typedef struct phx_particle_parameter_tangent {
    const float* masses;
    const float* rest_densities;
    const float* viscosities;
    const float* surface_tensions;
} phx_particle_parameter_tangent;

typedef struct phx_control_adjoint {
    phx_adjoint_vector_view external_accelerations;
} phx_control_adjoint;

typedef struct phx_particle_parameter_adjoint {
    double* masses;
    double* rest_densities;
    double* viscosities;
    double* surface_tensions;
} phx_particle_parameter_adjoint;

/* ParticleStateTangent / ParticleStateAdjoint (domain/domain.ixx:76-88) */
typedef struct phx_particle_state_tangent {
    phx_vector_view positions;
    phx_vector_view velocities;
} phx_particle_state_tangent;

typedef struct phx_particle_state_adjoint {
    phx_adjoint_vector_view positions;
    phx_adjoint_vector_view velocities;
} phx_particle_state_adjoint;

/* Domain::accumulate (domain/domain.cpp:189-200, cuda_detail::accumulate,
 * domain/domain.cu:8-29): destination[k] += source[k] for k < count, FP64
 * device arrays, enqueued on `stream`. */
phx_status phx_accumulate(phx_stream stream, const double* source, double* destination, size_t count);

/* Grid resolution ceil((max - min) / h) per axis (neighborhood.cpp:25-30). */
phx_status phx_grid_resolution(const phx_domain_configuration* domain, uint32_t resolution[3]);

/* Per-pass timing (no reference counterpart: the reference has no timing in
 * the liquid module). With profiling on, each pass of a step is bracketed by
 * HIP events on the solver stream; read_profile waits for them and returns,
 * per pass, the summed milliseconds, the launch count, and the pass's
 * algorithmic bytes per particle per launch (its unique per-particle inputs
 * and outputs read or written once, plus the cell-offset table). */
typedef struct phx_pass_profile {
    const char* name;
    double total_ms;
    uint64_t launches;
    double bytes_per_particle;
} phx_pass_profile;

/* ----------------------------------------------------------- neighborhood */
/* NeighborhoodView, neighborhood/device.h:8-23 (reference layout: arrays in
 * original particle order are indexed through sorted_particle_indices). */
typedef struct phx_neighborhood_view {
    const uint64_t* sorted_keys;
    const uint32_t* sorted_particle_indices;
    const uint32_t* cell_offsets;
    const uint64_t* sorted_boundary_keys;
    const uint32_t* sorted_boundary_indices;
    const uint32_t* boundary_cell_offsets;
    uint32_t particle_count;
    uint32_t boundary_count;
    uint32_t cells_x, cells_y, cells_z;
    float origin_x, origin_y, origin_z;
    float cell_size;
} phx_neighborhood_view;

/* BoundaryView, neighborhood/device.h:26-36: the boundary samples as device
 * arrays in original order; a sample's position at `time` is
 * position + time * velocity (boundary_position, neighborhood/device.cuh:16-22). */
typedef struct phx_boundary_view {
    const float* position_x;
    const float* position_y;
    const float* position_z;
    const float* velocity_x;
    const float* velocity_y;
    const float* velocity_z;
    const float* volumes;
    uint32_t count;
    float time;
} phx_boundary_view;

/* Workspace bytes phx_build_neighborhood needs (radix_sort_storage_size,
 * neighborhood.cu:48-52), boundary samples included. */
phx_status phx_neighborhood_workspace_size(const phx_domain_configuration* domain, size_t* bytes);

/* NeighborhoodSearch::build(step_index, positions, neighborhood)
 * (neighborhood/neighborhood.cpp:44-89, neighborhood.cu:54-70):
 * key = (cz*Cy + cy)*Cx + cx with cell = clamp(floor((x - min - t v_box)/h)),
 * t = step_index * time_step; stable ascending sort of (key, index);
 * cell_offsets[c] = first sorted position with key >= c for c in [0, C].
 * The domain's boundary samples get the same keys at their positions
 * x0 + t v and their own sort and offsets. Outputs (device; the sorted keys
 * and boundary arrays may be null): sorted_keys[N], sorted_particle_indices[N],
 * cell_offsets[C+1], sorted_boundary_keys[B], sorted_boundary_indices[B],
 * boundary_cell_offsets[C+1] (all zero when B = 0). The boundary samples are
 * copied from the domain's host array into the workspace, and the call waits
 * for that copy. Writes the arrays, counts, resolution and the lookup origin
 * min + t v_box into *view if view is non-null. */
phx_status phx_build_neighborhood(phx_stream stream, const phx_domain_configuration* domain, uint64_t step_index,
                                  phx_const_vector_view positions, void* workspace, size_t workspace_bytes,
                                  uint64_t* sorted_keys, uint32_t* sorted_particle_indices, uint32_t* cell_offsets,
                                  uint64_t* sorted_boundary_keys, uint32_t* sorted_boundary_indices,
                                  uint32_t* boundary_cell_offsets, phx_neighborhood_view* view);

/* Scratch bytes phx_build_neighborhood_noalloc holds per GPU (64 MiB): the
 * grid scratch of phx_build_neighborhood without boundary samples,
 * 3 round256(4N) + round256(4(C+2)) + round256(4 ceil((C+1)/2048))
 * + round256(4 (N/129 + 1)) + 256 bytes, must fit (about 4.4M particles
 * with 4M cells). */
#define PHX_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/* No counterpart in the reference. phx_build_neighborhood without a
 * workspace argument and without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same outputs as phx_build_neighborhood;
 * its scratch is a fixed device array of PHX_NOALLOC_SCRATCH_BYTES on each
 * GPU, kept until the process ends and shared by every call on that GPU, so
 * calls that overlap on one GPU race. Larger scenes return PHX_ERROR_CAPACITY;
 * boundary samples return PHX_ERROR_UNSUPPORTED (they are uploaded from host
 * memory); the stream's device must be the current device. */
phx_status phx_build_neighborhood_noalloc(phx_stream stream, const phx_domain_configuration* domain,
                                          uint64_t step_index, phx_const_vector_view positions,
                                          uint64_t* sorted_keys, uint32_t* sorted_particle_indices,
                                          uint32_t* cell_offsets, phx_neighborhood_view* view);

/* density_forward (density.cu:18-40): rho_i = sum_j m_j W(x_i - x_j) over the
 * 27 cells of topology_positions[i], j = i included, each cell's boundary
 * samples b after its particles adding rho0_i V_b W(x_i - x_b(t)); W = poly6
 * when pbf_kernel != 0, else the cubic spline. boundary.count must equal
 * neighborhood.boundary_count (0: no boundary arrays are read). */
phx_status phx_density_forward(phx_stream stream, uint32_t particle_count, float support_radius, int pbf_kernel,
                               phx_const_vector_view topology_positions, phx_const_vector_view positions,
                               phx_particle_parameters parameters, phx_neighborhood_view neighborhood,
                               phx_boundary_view boundary, float* densities);

/* density_jvp (density.cu:42-63): the FP32 density tangent for position and
 * parameter tangents (masses and rest densities are read; a null tangent
 * array counts as zero). */
phx_status phx_density_jvp(phx_stream stream, uint32_t particle_count, float support_radius, int pbf_kernel,
                           phx_const_vector_view topology_positions, phx_const_vector_view positions,
                           phx_const_vector_view position_tangent, phx_particle_parameters parameters,
                           phx_particle_parameter_tangent parameter_tangent, phx_neighborhood_view neighborhood,
                           phx_boundary_view boundary, float* density_tangent);

/* density_vjp (density.cu:65-98): the FP64 VJP of the densities for
 * density_adjoint, accumulated (+=) into position_adjoint and the mass and
 * rest-density adjoints; null output arrays are skipped. */
phx_status phx_density_vjp(phx_stream stream, uint32_t particle_count, float support_radius, int pbf_kernel,
                           phx_const_vector_view topology_positions, phx_const_vector_view positions,
                           phx_particle_parameters parameters, phx_neighborhood_view neighborhood,
                           phx_boundary_view boundary, const double* density_adjoint,
                           phx_adjoint_vector_view position_adjoint, phx_particle_parameter_adjoint parameter_adjoint);

/* ----------------------------------------------------------------- DFSPH */
/* DFSPH::Configuration, sph/sph.ixx:196-201. density_error_tolerance > 0
 * selects tolerance mode (an extension: iterate each phase until the mean
 * relative density error, reduced in FP64, is below the tolerance, at most
 * max_iterations times); 0 gives the reference's fixed iteration counts. */
typedef struct phx_dfsph_configuration {
    uint32_t divergence_iterations; /* reference default 4 */
    uint32_t density_iterations;    /* reference default 6 */
    uint32_t checkpoint_interval;   /* reference default 2 */
    uint32_t pressure_warm_start;   /* reference default 1 */
    float density_error_tolerance;  /* 0: fixed iterations */
    uint32_t max_iterations;
    /* VJP of a prediction whose no-slip box collision resets the velocity:
     * 0 (default) the reference's predict_vjp_kernel (sph/dfsph.cu:68-102,
     * line 81), which zeroes the whole velocity adjoint and with it the
     * dt * position adjoint of the components that did not collide; 1 the
     * exact transpose of the JVP (those terms kept). */
    uint32_t exact_no_slip_adjoint;
} phx_dfsph_configuration;

/* DFSPH::Parameters and DFSPH::State, sph/sph.ixx:202-217 */
typedef struct phx_dfsph_parameters {
    const float* divergence_relaxation;
    const float* density_relaxation;
} phx_dfsph_parameters;

typedef struct phx_dfsph_state {
    float* warm_divergence_pressure;
    float* warm_density_pressure;
} phx_dfsph_state;

/* SPH<DFSPH>::State / Parameters, sph/sph.ixx:275-298 */
typedef struct phx_sph_dfsph_state {
    phx_particle_state particles;
    phx_dfsph_state method;
} phx_sph_dfsph_state;

typedef struct phx_sph_dfsph_parameters {
    phx_particle_parameters particles;
    phx_dfsph_parameters method;
} phx_sph_dfsph_parameters;

/* DFSPH::StateTangent / StateAdjoint / ParameterTangent / ParameterAdjoint
 * (sph/sph.ixx:206-225) and their SPH<DFSPH> wrappers (sph/sph.ixx:279-298). */
typedef struct phx_dfsph_state_tangent {
    float* warm_divergence_pressure;
    float* warm_density_pressure;
} phx_dfsph_state_tangent;

typedef struct phx_dfsph_state_adjoint {
    double* warm_divergence_pressure;
    double* warm_density_pressure;
} phx_dfsph_state_adjoint;

typedef struct phx_dfsph_parameter_tangent {
    const float* divergence_relaxation;
    const float* density_relaxation;
} phx_dfsph_parameter_tangent;

typedef struct phx_dfsph_parameter_adjoint {
    double* divergence_relaxation;
    double* density_relaxation;
} phx_dfsph_parameter_adjoint;

typedef struct phx_sph_dfsph_state_tangent {
    phx_particle_state_tangent particles;
    phx_dfsph_state_tangent method;
} phx_sph_dfsph_state_tangent;

typedef struct phx_sph_dfsph_state_adjoint {
    phx_particle_state_adjoint particles;
    phx_dfsph_state_adjoint method;
} phx_sph_dfsph_state_adjoint;

typedef struct phx_sph_dfsph_parameter_tangent {
    phx_particle_parameter_tangent particles;
    phx_dfsph_parameter_tangent method;
} phx_sph_dfsph_parameter_tangent;

typedef struct phx_sph_dfsph_parameter_adjoint {
    phx_particle_parameter_adjoint particles;
    phx_dfsph_parameter_adjoint method;
} phx_sph_dfsph_parameter_adjoint;

typedef struct phx_dfsph phx_dfsph;
typedef struct phx_dfsph_step_cache phx_dfsph_step_cache;

/* SPH<DFSPH>(domain, configuration, mode, stream), sph/sph.ixx:309-312 */
phx_status phx_dfsph_create(const phx_domain_configuration* domain, const phx_dfsph_configuration* configuration,
                            phx_execution_mode mode, phx_stream stream, phx_dfsph** solver);
phx_status phx_dfsph_destroy(phx_dfsph* solver);
/* allocate_step_cache, sph/sph.ixx:322 */
phx_status phx_dfsph_allocate_step_cache(phx_dfsph* solver, phx_dfsph_step_cache** cache);
phx_status phx_dfsph_free_step_cache(phx_dfsph_step_cache* cache);
/* forward_step(state, control, parameters, next_state, cache), sph/sph.ixx:350-354 */
phx_status phx_dfsph_forward_step(phx_dfsph* solver, const phx_sph_dfsph_state* state, const phx_control* control,
                                  const phx_sph_dfsph_parameters* parameters, phx_sph_dfsph_state* next_state,
                                  phx_dfsph_step_cache* cache);
/* The reference's DFSPH::reference_gradient_norm (sph/sph.cpp:37-58). */
phx_status phx_dfsph_reference_gradient_norm(const phx_dfsph* solver, float* value);
/* Iterations executed by the last forward_step (divergence, density). */
phx_status phx_dfsph_last_iterations(const phx_dfsph* solver, uint32_t iterations[2]);

/* Diagnostics (no reference counterpart): the skin of the neighbour masks
 * the next step will use, in units of the support radius (adapted from the
 * predicted displacements of a completed step). */
phx_status phx_dfsph_mask_skin(const phx_dfsph* solver, float* skin);
phx_status phx_dfsph_set_profiling(phx_dfsph* solver, int enabled);
/* Fills up to capacity entries (passes never launched are omitted), sets
 * *count, and resets the totals. Synchronises with the recorded events. */
phx_status phx_dfsph_read_profile(phx_dfsph* solver, phx_pass_profile* profile, uint32_t capacity, uint32_t* count);

/* Cache inspection (StepCache fields, sph/sph.ixx:299-303), copied out in
 * original particle order into caller device buffers; enqueued on the solver
 * stream. Neighborhood in the reference layout; any output may be null. */
phx_status phx_dfsph_cache_neighborhood(const phx_dfsph_step_cache* cache, uint64_t* sorted_keys,
                                        uint32_t* sorted_particle_indices, uint32_t* cell_offsets);
phx_status phx_dfsph_cache_densities(const phx_dfsph_step_cache* cache, float* densities);
phx_status phx_dfsph_cache_non_pressure_accelerations(const phx_dfsph_step_cache* cache, phx_vector_view out);

/* jvp_step(state, control, parameters, next_state, cache, state_tangent,
 * control_tangent, parameter_tangent, next_state_tangent), sph/sph.ixx:356-359
 * and DFSPH::jvp, sph/sph.cpp:612-659. Needs a solver created in
 * PHX_MODE_DIFFERENTIABLE and the cache of that step's forward_step, which
 * records the state, control, parameters and every iterate, so they are not
 * passed again. Tangents are FP32 and overwrite next_state_tangent (its warm
 * pressure tangents may be null: not written); null parameter-tangent and
 * warm-pressure tangent arrays count as zero. */
phx_status phx_dfsph_jvp_step(phx_dfsph* solver, const phx_dfsph_step_cache* cache,
                              const phx_sph_dfsph_state_tangent* state_tangent,
                              const phx_control_tangent* control_tangent,
                              const phx_sph_dfsph_parameter_tangent* parameter_tangent,
                              phx_sph_dfsph_state_tangent* next_state_tangent);
/* vjp_step(..., next_state_adjoint, previous_state_adjoint, control_adjoint,
 * parameter_adjoint), sph/sph.ixx:361-365 and DFSPH::vjp, sph/sph.cpp:661-723:
 * the VJP of the recorded step with its neighbour graph, collision masks and
 * pressure active sets frozen, in FP64. Particle, control and parameter
 * adjoints are accumulated (+=); the warm-start pressure adjoints are written
 * (copied, sph.cpp:713) with warm start and left alone without. Null output
 * arrays are skipped; null next warm-pressure adjoints count as zero. Under a
 * no-slip collision the prediction's VJP is the exact transpose of its JVP,
 * which keeps dt * xbar of the components that did not collide (the
 * reference's sph/dfsph.cu:81 zeroes it). */
phx_status phx_dfsph_vjp_step(phx_dfsph* solver, const phx_dfsph_step_cache* cache,
                              const phx_sph_dfsph_state_adjoint* next_state_adjoint,
                              phx_sph_dfsph_state_adjoint* previous_state_adjoint,
                              phx_control_adjoint* control_adjoint,
                              phx_sph_dfsph_parameter_adjoint* parameter_adjoint);

/* ------------------------------------------------ WCSPH, PCISPH, IISPH */
/* SPH<WCSPH>, SPH<PCISPH>, SPH<IISPH> (sph/sph.ixx:39-193, 273-372): one
 * solver family selected by `method`. */
typedef enum phx_sph_method { PHX_SPH_WCSPH = 0, PHX_SPH_PCISPH = 1, PHX_SPH_IISPH = 2 } phx_sph_method;

/* WCSPH::Configuration is empty (sph/sph.ixx:40); PCISPH/IISPH::Configuration
 * {pressure_iterations, checkpoint_interval} (sph/sph.ixx:104-107, 150-153).
 * density_error_tolerance > 0 selects tolerance mode for PCISPH/IISPH (an
 * extension: iterate until the mean compression error
 * (1/N) sum max(0, rho* - rho0) / rho0 of the prediction, reduced in FP64,
 * is within the tolerance, at most max_iterations times; forward mode only);
 * 0 gives the reference's fixed pressure_iterations. */
typedef struct phx_sph_configuration {
    phx_sph_method method;
    uint32_t pressure_iterations;  /* PCISPH, IISPH; reference default 6 */
    uint32_t checkpoint_interval;  /* PCISPH, IISPH; reference default 2 */
    float density_error_tolerance; /* PCISPH, IISPH; 0: fixed iterations */
    uint32_t max_iterations;
    /* as phx_dfsph_configuration: 0 (default) the reference's VJP of a
     * no-slip collision (predict_vjp_kernel, sph/pcisph.cu:68-102 and
     * sph/iisph.cu:68-102; integrate_vjp_kernel, sph/common.cu:344-366 for
     * WCSPH), 1 the exact transpose */
    uint32_t exact_no_slip_adjoint;
} phx_sph_configuration;

/* Method parameters, one value per particle (sph/sph.ixx:45-49, 112, 158).
 * Only the arrays of the selected method are read; the others may be null. */
typedef struct phx_sph_method_parameters {
    const float* speed_of_sound;           /* WCSPH::Parameters */
    const float* tait_exponent;            /* WCSPH::Parameters */
    const float* boundary_surface_tension; /* WCSPH::Parameters */
    const float* pressure_relaxation;      /* PCISPH::Parameters */
    const float* jacobi_relaxation;        /* IISPH::Parameters */
} phx_sph_method_parameters;

/* SPH<Method>::State (the method states of these three are empty) and
 * SPH<Method>::Parameters, sph/sph.ixx:275-298. */
typedef struct phx_sph_state {
    phx_particle_state particles;
} phx_sph_state;

typedef struct phx_sph_parameters {
    phx_particle_parameters particles;
    phx_sph_method_parameters method;
} phx_sph_parameters;

typedef struct phx_sph phx_sph;
typedef struct phx_sph_step_cache phx_sph_step_cache;

/* SPH<Method>(domain, configuration, mode, stream), sph/sph.ixx:309-312 */
phx_status phx_sph_create(const phx_domain_configuration* domain, const phx_sph_configuration* configuration,
                          phx_execution_mode mode, phx_stream stream, phx_sph** solver);
phx_status phx_sph_destroy(phx_sph* solver);
/* allocate_step_cache, sph/sph.ixx:322 */
phx_status phx_sph_allocate_step_cache(phx_sph* solver, phx_sph_step_cache** cache);
phx_status phx_sph_free_step_cache(phx_sph_step_cache* cache);
/* forward_step(state, control, parameters, next_state, cache), sph/sph.ixx:350-354 */
phx_status phx_sph_forward_step(phx_sph* solver, const phx_sph_state* state, const phx_control* control,
                                const phx_sph_parameters* parameters, phx_sph_state* next_state,
                                phx_sph_step_cache* cache);
/* The reference_gradient_norm of PCISPH/IISPH (sph/sph.cpp:37-58, 252, 365). */
phx_status phx_sph_reference_gradient_norm(const phx_sph* solver, float* value);
/* Pressure iterations run by the last PCISPH/IISPH forward_step (the
 * configured count, or the count tolerance mode needed; 0 for WCSPH). */
phx_status phx_sph_last_iterations(const phx_sph* solver, uint32_t* iterations);
phx_status phx_sph_set_profiling(phx_sph* solver, int enabled);
phx_status phx_sph_read_profile(phx_sph* solver, phx_pass_profile* profile, uint32_t capacity, uint32_t* count);

/* Cache inspection in original particle order, enqueued on the solver
 * stream; any output may be null. Pressures: WCSPH::Cache::pressures, or the
 * last iteration's pressures of PCISPH/IISPH (PressureIterationCache).
 * Non-pressure accelerations: PCISPH/IISPH::Cache only. */
phx_status phx_sph_cache_neighborhood(const phx_sph_step_cache* cache, uint64_t* sorted_keys,
                                      uint32_t* sorted_particle_indices, uint32_t* cell_offsets);
phx_status phx_sph_cache_densities(const phx_sph_step_cache* cache, float* densities);
phx_status phx_sph_cache_pressures(const phx_sph_step_cache* cache, float* pressures);
phx_status phx_sph_cache_non_pressure_accelerations(const phx_sph_step_cache* cache, phx_vector_view out);

/* StateTangent / StateAdjoint of the three methods are empty and their
 * ParameterTangent / ParameterAdjoint mirror Parameters (sph/sph.ixx:42-60,
 * 108-113, 154-159); SPH<Method> wrappers, sph/sph.ixx:279-298. Only the
 * arrays of the solver's method are read or written; the others may be null. */
typedef struct phx_sph_state_tangent {
    phx_particle_state_tangent particles;
} phx_sph_state_tangent;

typedef struct phx_sph_state_adjoint {
    phx_particle_state_adjoint particles;
} phx_sph_state_adjoint;

typedef struct phx_sph_method_parameter_tangent {
    const float* speed_of_sound;
    const float* tait_exponent;
    const float* boundary_surface_tension;
    const float* pressure_relaxation;
    const float* jacobi_relaxation;
} phx_sph_method_parameter_tangent;

typedef struct phx_sph_method_parameter_adjoint {
    double* speed_of_sound;
    double* tait_exponent;
    double* boundary_surface_tension;
    double* pressure_relaxation;
    double* jacobi_relaxation;
} phx_sph_method_parameter_adjoint;

typedef struct phx_sph_parameter_tangent {
    phx_particle_parameter_tangent particles;
    phx_sph_method_parameter_tangent method;
} phx_sph_parameter_tangent;

typedef struct phx_sph_parameter_adjoint {
    phx_particle_parameter_adjoint particles;
    phx_sph_method_parameter_adjoint method;
} phx_sph_parameter_adjoint;

/* SPH<WCSPH|PCISPH|IISPH>::jvp_step (sph/sph.ixx:356-359; WCSPH::jvp
 * sph/sph.cpp:228-237, PCISPH::jvp sph/sph.cpp:312-331, IISPH::jvp
 * sph/sph.cpp:423-442). Needs a solver created in PHX_MODE_DIFFERENTIABLE and
 * the cache of that step's forward_step, which records the step. FP32
 * tangents overwrite next_state_tangent; null parameter-tangent arrays count
 * as zero. */
phx_status phx_sph_jvp_step(phx_sph* solver, const phx_sph_step_cache* cache,
                            const phx_sph_state_tangent* state_tangent, const phx_control_tangent* control_tangent,
                            const phx_sph_parameter_tangent* parameter_tangent,
                            phx_sph_state_tangent* next_state_tangent);
/* SPH<WCSPH|PCISPH|IISPH>::vjp_step (sph/sph.ixx:361-365; WCSPH::vjp
 * sph/sph.cpp:239-249, PCISPH::vjp sph/sph.cpp:333-363, IISPH::vjp
 * sph/sph.cpp:444-476): FP64 adjoints of the recorded step (neighbour graph,
 * collision masks, the compression test of every artificial-viscosity pair
 * and the active sets of the max(0, .) updates frozen), accumulated (+=) into
 * the outputs; null outputs are skipped. The integration's VJP under a no-slip
 * collision is the exact transpose of its JVP (see phx_dfsph_vjp_step). */
phx_status phx_sph_vjp_step(phx_sph* solver, const phx_sph_step_cache* cache,
                            const phx_sph_state_adjoint* next_state_adjoint,
                            phx_sph_state_adjoint* previous_state_adjoint, phx_control_adjoint* control_adjoint,
                            phx_sph_parameter_adjoint* parameter_adjoint);

/* ------------------------------------------------------------------- PBF */
/* PBF::Configuration, pbf/pbf.ixx:15-18 */
typedef struct phx_pbf_configuration {
    uint32_t pressure_iterations; /* reference default 5 */
    uint32_t checkpoint_interval; /* reference default 2 */
} phx_pbf_configuration;

/* PBF::State / Parameters, pbf/pbf.ixx:20-32 */
typedef struct phx_pbf_state {
    phx_particle_state particles;
} phx_pbf_state;

typedef struct phx_pbf_parameters {
    phx_particle_parameters particles;
    const float* relaxation;
    const float* artificial_pressure_strength;
    const float* artificial_pressure_exponent;
    const float* artificial_pressure_radius;
    const float* xsph_viscosity;
    const float* vorticity_confinement;
} phx_pbf_parameters;

/* PBF::StateTangent / ParameterTangent / StateAdjoint / ParameterAdjoint
 * (pbf/pbf.ixx:20-51). */
typedef struct phx_pbf_state_tangent {
    phx_vector_view positions;
    phx_vector_view velocities;
} phx_pbf_state_tangent;

typedef struct phx_pbf_parameter_tangent {
    phx_particle_parameter_tangent particles;
    const float* relaxation;
    const float* artificial_pressure_strength;
    const float* artificial_pressure_exponent;
    const float* artificial_pressure_radius;
    const float* xsph_viscosity;
    const float* vorticity_confinement;
} phx_pbf_parameter_tangent;

typedef struct phx_pbf_state_adjoint {
    phx_adjoint_vector_view positions;
    phx_adjoint_vector_view velocities;
} phx_pbf_state_adjoint;

typedef struct phx_pbf_parameter_adjoint {
    phx_particle_parameter_adjoint particles;
    double* relaxation;
    double* artificial_pressure_strength;
    double* artificial_pressure_exponent;
    double* artificial_pressure_radius;
    double* xsph_viscosity;
    double* vorticity_confinement;
} phx_pbf_parameter_adjoint;

typedef struct phx_pbf phx_pbf;
typedef struct phx_pbf_step_cache phx_pbf_step_cache;

/* PBF(domain, configuration, mode, stream), pbf/pbf.cpp:39-72 */
phx_status phx_pbf_create(const phx_domain_configuration* domain, const phx_pbf_configuration* configuration,
                          phx_execution_mode mode, phx_stream stream, phx_pbf** solver);
phx_status phx_pbf_destroy(phx_pbf* solver);
phx_status phx_pbf_allocate_step_cache(phx_pbf* solver, phx_pbf_step_cache** cache);
phx_status phx_pbf_free_step_cache(phx_pbf_step_cache* cache);
/* forward_step(state, control, parameters, next_state, cache), pbf/pbf.cpp:167-191 */
phx_status phx_pbf_forward_step(phx_pbf* solver, const phx_pbf_state* state, const phx_control* control,
                                const phx_pbf_parameters* parameters, phx_pbf_state* next_state,
                                phx_pbf_step_cache* cache);
/* jvp_step(state, control, parameters, next_state, cache, state_tangent,
 * control_tangent, parameter_tangent, next_state_tangent), pbf/pbf.ixx:93 and
 * pbf/pbf.cpp:193-215. Needs a solver created in PHX_MODE_DIFFERENTIABLE and
 * the cache of that step's forward_step (state, control and next state are
 * read from the cache). Tangents are FP32 and overwrite next_state_tangent;
 * null parameter-tangent arrays count as zero. */
phx_status phx_pbf_jvp_step(phx_pbf* solver, const phx_pbf_parameters* parameters, const phx_pbf_step_cache* cache,
                            const phx_pbf_state_tangent* state_tangent, const phx_control_tangent* control_tangent,
                            const phx_pbf_parameter_tangent* parameter_tangent,
                            phx_pbf_state_tangent* next_state_tangent);
/* vjp_step(..., next_state_adjoint, previous_state_adjoint, control_adjoint,
 * parameter_adjoint), pbf/pbf.ixx:94 and pbf/pbf.cpp:217-258: the exact VJP
 * of the recorded step with its neighbour graph and collision masks frozen,
 * in FP64. Every output is accumulated (+=); null adjoint arrays are
 * skipped. The adjoint of next_state positions and velocities is read only. */
phx_status phx_pbf_vjp_step(phx_pbf* solver, const phx_pbf_parameters* parameters, const phx_pbf_step_cache* cache,
                            const phx_pbf_state_adjoint* next_state_adjoint,
                            phx_pbf_state_adjoint* previous_state_adjoint, phx_control_adjoint* control_adjoint,
                            phx_pbf_parameter_adjoint* parameter_adjoint);

phx_status phx_pbf_set_profiling(phx_pbf* solver, int enabled);
phx_status phx_pbf_read_profile(phx_pbf* solver, phx_pass_profile* profile, uint32_t capacity, uint32_t* count);
/* StepCache fields (pbf/pbf.ixx:54-70), original particle order. */
phx_status phx_pbf_cache_neighborhood(const phx_pbf_step_cache* cache, uint64_t* sorted_keys,
                                      uint32_t* sorted_particle_indices, uint32_t* cell_offsets);
phx_status phx_pbf_cache_predicted_positions(const phx_pbf_step_cache* cache, phx_vector_view out);
/* Densities and lambdas of the last constraint iteration (IterationPrimal,
 * pbf/pbf.ixx:98-105). */
phx_status phx_pbf_cache_last_iteration(const phx_pbf_step_cache* cache, float* densities, float* lambdas);
/* StepCache::vorticity_magnitudes (pbf/pbf.ixx:67): |omega| of the step's
 * final positions, the field the reference's viewer provider publishes
 * (examples/fluids/liquid/pbf-dam-break/spectra/provider.ixx:90-91). */
phx_status phx_pbf_cache_vorticity_magnitudes(const phx_pbf_step_cache* cache, float* magnitudes);

/* ---------------------------------------------------------------- viewer */
/* spectra_cuda::write_vectors (examples/fluids/liquid/pbf-dam-break/spectra/
 * kernels.cu:8-18): SoA positions and velocities to interleaved x, y, z
 * (Float3) device arrays of 3 * count floats; either output may be null. */
phx_status phx_write_vectors(phx_stream stream, uint32_t count, phx_const_vector_view positions,
                             phx_const_vector_view velocities, float* positions_xyz, float* velocities_xyz);

/* ------------------------------------------------------- batched PBF scenes */
/* Independent PBF scenes of uneven particle counts stepped together. The
 * reference runs one PBF solver per scene and stream (domain/domain.ixx:
 * 121-124), each substep the fixed launch sequence of PBF::forward_step
 * (pbf/pbf.cpp:167-191). A batch runs that sequence for every scene at once:
 * each pass of a substep is one launch over all scenes, and with static
 * boxes one substep is captured in a HIP graph and replayed. Forward mode,
 * no boundary particles; the scenes share the support radius, the time step
 * and the iteration count. The scenes' particles are concatenated in scene
 * order: scene k holds particles offsets[k] .. offsets[k + 1] - 1 of every
 * state, control and parameter array, and its results equal its own PBF
 * solver's bit for bit. */
typedef struct phx_pbf_batch phx_pbf_batch;

typedef struct phx_pbf_batch_configuration {
    float mask_skin;    /* neighbour-mask skin in units of h (0: 0.3); results do not depend on it */
    uint32_t use_graph; /* replay a captured substep (only when every box is static) */
} phx_pbf_batch_configuration;

phx_status phx_pbf_batch_create(const phx_domain_configuration* domains, uint32_t scene_count,
                                const phx_pbf_configuration* configuration,
                                const phx_pbf_batch_configuration* batch_configuration, phx_stream stream,
                                phx_pbf_batch** batch);
phx_status phx_pbf_batch_destroy(phx_pbf_batch* batch);
/* offsets[0 .. scene_count]: the first particle of each scene, then the total */
phx_status phx_pbf_batch_particle_offsets(const phx_pbf_batch* batch, uint64_t* offsets);
/* `steps` forward substeps of every scene, in place on `state`, enqueued on
 * the batch's stream (ordered after and before the caller's work on it);
 * state->particles.step_index advances by `steps`. */
phx_status phx_pbf_batch_run(phx_pbf_batch* batch, phx_pbf_state* state, const phx_control* control,
                             const phx_pbf_parameters* parameters, uint32_t steps);
/* kernel launches of one substep over all scenes */
phx_status phx_pbf_batch_kernels_per_step(const phx_pbf_batch* batch, uint32_t* kernels);

#ifdef __cplusplus
}
#endif

#endif
