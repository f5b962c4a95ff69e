"""ctypes mirror of include/physica_amd/physica.h.

Loading fails loudly when libphysica_amd.so has not been built: there is no
CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_float, c_int, c_size_t, c_uint32, c_uint64, c_void_p
from pathlib import Path

LIBRARY_PATH = Path(__file__).resolve().parent / "libphysica_amd.so"

PHX_SUCCESS = 0
PHX_ERROR_INVALID_ARGUMENT = 1
PHX_ERROR_UNSUPPORTED = 2
PHX_ERROR_WRONG_MODE = 3
PHX_ERROR_WORKSPACE_TOO_SMALL = 4
PHX_ERROR_OUT_OF_MEMORY = 5
PHX_ERROR_DEVICE = 6
PHX_ERROR_CAPACITY = 7

STATUS_NAMES = {
    0: "PHX_SUCCESS",
    1: "PHX_ERROR_INVALID_ARGUMENT",
    2: "PHX_ERROR_UNSUPPORTED",
    3: "PHX_ERROR_WRONG_MODE",
    4: "PHX_ERROR_WORKSPACE_TOO_SMALL",
    5: "PHX_ERROR_OUT_OF_MEMORY",
    6: "PHX_ERROR_DEVICE",
    7: "PHX_ERROR_CAPACITY",
}


class PhxError(RuntimeError):
    """A non-success phx_status returned by the library."""

    def __init__(self, status: int, function: str, message: str):
        self.status = status
        self.function = function
        super().__init__(f"{function}: {STATUS_NAMES.get(status, status)}: {message}")


class Vector3(ctypes.Structure):
    _fields_ = [("x", c_float), ("y", c_float), ("z", c_float)]


class BoxBoundary(ctypes.Structure):
    _fields_ = [("minimum", Vector3), ("maximum", Vector3), ("velocity", Vector3), ("no_slip", c_uint32)]


class BoundaryParticle(ctypes.Structure):
    _fields_ = [("position", Vector3), ("velocity", Vector3), ("volume", c_float)]


class DomainConfiguration(ctypes.Structure):
    _fields_ = [
        ("particle_count", c_uint32),
        ("time_step", c_float),
        ("support_radius", c_float),
        ("particle_radius", c_float),
        ("gravity", Vector3),
        ("boundary", BoxBoundary),
        ("boundary_particles", POINTER(BoundaryParticle)),
        ("boundary_particle_count", c_uint32),
    ]


class VectorView(ctypes.Structure):
    _fields_ = [("x", c_void_p), ("y", c_void_p), ("z", c_void_p)]


class ParticleState(ctypes.Structure):
    _fields_ = [("positions", VectorView), ("velocities", VectorView), ("step_index", c_uint64)]


class Control(ctypes.Structure):
    _fields_ = [("external_accelerations", VectorView)]


class ParticleParameters(ctypes.Structure):
    _fields_ = [
        ("masses", c_void_p),
        ("rest_densities", c_void_p),
        ("viscosities", c_void_p),
        ("surface_tensions", c_void_p),
    ]


class PassProfile(ctypes.Structure):
    _fields_ = [("name", c_char_p), ("total_ms", ctypes.c_double), ("launches", c_uint64),
                ("bytes_per_particle", ctypes.c_double)]


class NeighborhoodView(ctypes.Structure):
    _fields_ = [
        ("sorted_keys", c_void_p),
        ("sorted_particle_indices", c_void_p),
        ("cell_offsets", c_void_p),
        ("sorted_boundary_keys", c_void_p),
        ("sorted_boundary_indices", c_void_p),
        ("boundary_cell_offsets", c_void_p),
        ("particle_count", c_uint32),
        ("boundary_count", c_uint32),
        ("cells_x", c_uint32),
        ("cells_y", c_uint32),
        ("cells_z", c_uint32),
        ("origin_x", c_float),
        ("origin_y", c_float),
        ("origin_z", c_float),
        ("cell_size", c_float),
    ]


class BoundaryView(ctypes.Structure):
    _fields_ = [
        ("position_x", c_void_p),
        ("position_y", c_void_p),
        ("position_z", c_void_p),
        ("velocity_x", c_void_p),
        ("velocity_y", c_void_p),
        ("velocity_z", c_void_p),
        ("volumes", c_void_p),
        ("count", c_uint32),
        ("time", c_float),
    ]


class DfsphConfiguration(ctypes.Structure):
    _fields_ = [
        ("divergence_iterations", c_uint32),
        ("density_iterations", c_uint32),
        ("checkpoint_interval", c_uint32),
        ("pressure_warm_start", c_uint32),
        ("density_error_tolerance", c_float),
        ("max_iterations", c_uint32),
        ("exact_no_slip_adjoint", c_uint32),
    ]


class PbfBatchConfiguration(ctypes.Structure):
    _fields_ = [("mask_skin", c_float), ("use_graph", c_uint32)]


class DfsphParameters(ctypes.Structure):
    _fields_ = [("divergence_relaxation", c_void_p), ("density_relaxation", c_void_p)]


class DfsphState(ctypes.Structure):
    _fields_ = [("warm_divergence_pressure", c_void_p), ("warm_density_pressure", c_void_p)]


class SphDfsphState(ctypes.Structure):
    _fields_ = [("particles", ParticleState), ("method", DfsphState)]


class SphDfsphParameters(ctypes.Structure):
    _fields_ = [("particles", ParticleParameters), ("method", DfsphParameters)]


class AdjointVectorView(ctypes.Structure):
    _fields_ = [("x", c_void_p), ("y", c_void_p), ("z", c_void_p)]


class ControlTangent(ctypes.Structure):
    _fields_ = [("external_accelerations", VectorView)]


class ParticleParameterTangent(ctypes.Structure):
    _fields_ = [("masses", c_void_p), ("rest_densities", c_void_p), ("viscosities", c_void_p),
                ("surface_tensions", c_void_p)]


class ControlAdjoint(ctypes.Structure):
    _fields_ = [("external_accelerations", AdjointVectorView)]


class ParticleParameterAdjoint(ctypes.Structure):
    _fields_ = [("masses", c_void_p), ("rest_densities", c_void_p), ("viscosities", c_void_p),
                ("surface_tensions", c_void_p)]


class ParticleStateTangent(ctypes.Structure):
    _fields_ = [("positions", VectorView), ("velocities", VectorView)]


class ParticleStateAdjoint(ctypes.Structure):
    _fields_ = [("positions", AdjointVectorView), ("velocities", AdjointVectorView)]


class DfsphStateTangent(ctypes.Structure):
    _fields_ = [("warm_divergence_pressure", c_void_p), ("warm_density_pressure", c_void_p)]


class DfsphStateAdjoint(ctypes.Structure):
    _fields_ = [("warm_divergence_pressure", c_void_p), ("warm_density_pressure", c_void_p)]


class DfsphParameterTangent(ctypes.Structure):
    _fields_ = [("divergence_relaxation", c_void_p), ("density_relaxation", c_void_p)]


class DfsphParameterAdjoint(ctypes.Structure):
    _fields_ = [("divergence_relaxation", c_void_p), ("density_relaxation", c_void_p)]


class SphDfsphStateTangent(ctypes.Structure):
    _fields_ = [("particles", ParticleStateTangent), ("method", DfsphStateTangent)]


class SphDfsphStateAdjoint(ctypes.Structure):
    _fields_ = [("particles", ParticleStateAdjoint), ("method", DfsphStateAdjoint)]


class SphDfsphParameterTangent(ctypes.Structure):
    _fields_ = [("particles", ParticleParameterTangent), ("method", DfsphParameterTangent)]


class SphDfsphParameterAdjoint(ctypes.Structure):
    _fields_ = [("particles", ParticleParameterAdjoint), ("method", DfsphParameterAdjoint)]


PHX_SPH_WCSPH, PHX_SPH_PCISPH, PHX_SPH_IISPH = 0, 1, 2


class SphConfiguration(ctypes.Structure):
    _fields_ = [("method", c_int), ("pressure_iterations", c_uint32), ("checkpoint_interval", c_uint32),
                ("density_error_tolerance", c_float), ("max_iterations", c_uint32),
                ("exact_no_slip_adjoint", c_uint32)]


class SphMethodParameters(ctypes.Structure):
    _fields_ = [
        ("speed_of_sound", c_void_p),
        ("tait_exponent", c_void_p),
        ("boundary_surface_tension", c_void_p),
        ("pressure_relaxation", c_void_p),
        ("jacobi_relaxation", c_void_p),
    ]


class SphState(ctypes.Structure):
    _fields_ = [("particles", ParticleState)]


class SphParameters(ctypes.Structure):
    _fields_ = [("particles", ParticleParameters), ("method", SphMethodParameters)]


class SphStateTangent(ctypes.Structure):
    _fields_ = [("particles", ParticleStateTangent)]


class SphStateAdjoint(ctypes.Structure):
    _fields_ = [("particles", ParticleStateAdjoint)]


class SphMethodParameterTangent(ctypes.Structure):
    _fields_ = SphMethodParameters._fields_


class SphMethodParameterAdjoint(ctypes.Structure):
    _fields_ = SphMethodParameters._fields_


class SphParameterTangent(ctypes.Structure):
    _fields_ = [("particles", ParticleParameterTangent), ("method", SphMethodParameterTangent)]


class SphParameterAdjoint(ctypes.Structure):
    _fields_ = [("particles", ParticleParameterAdjoint), ("method", SphMethodParameterAdjoint)]


class PbfConfiguration(ctypes.Structure):
    _fields_ = [("pressure_iterations", c_uint32), ("checkpoint_interval", c_uint32)]


class PbfState(ctypes.Structure):
    _fields_ = [("particles", ParticleState)]


class PbfParameters(ctypes.Structure):
    _fields_ = [
        ("particles", ParticleParameters),
        ("relaxation", c_void_p),
        ("artificial_pressure_strength", c_void_p),
        ("artificial_pressure_exponent", c_void_p),
        ("artificial_pressure_radius", c_void_p),
        ("xsph_viscosity", c_void_p),
        ("vorticity_confinement", c_void_p),
    ]


_PBF_FIELDS = ("relaxation", "artificial_pressure_strength", "artificial_pressure_exponent",
               "artificial_pressure_radius", "xsph_viscosity", "vorticity_confinement")


class PbfStateTangent(ctypes.Structure):
    _fields_ = [("positions", VectorView), ("velocities", VectorView)]


class PbfParameterTangent(ctypes.Structure):
    _fields_ = [("particles", ParticleParameterTangent)] + [(k, c_void_p) for k in _PBF_FIELDS]


class PbfStateAdjoint(ctypes.Structure):
    _fields_ = [("positions", AdjointVectorView), ("velocities", AdjointVectorView)]


class PbfParameterAdjoint(ctypes.Structure):
    _fields_ = [("particles", ParticleParameterAdjoint)] + [(k, c_void_p) for k in _PBF_FIELDS]


# name -> (restype, argtypes); every function declared in physica.h
PROTOTYPES = {
    "phx_last_error": (c_char_p, []),
    "phx_version": (c_char_p, []),
    "phx_grid_resolution": (c_int, [POINTER(DomainConfiguration), POINTER(c_uint32)]),
    "phx_accumulate": (c_int, [c_void_p, c_void_p, c_void_p, c_size_t]),
    "phx_write_vectors": (c_int, [c_void_p, c_uint32, VectorView, VectorView, c_void_p, c_void_p]),
    "phx_pbf_cache_vorticity_magnitudes": (c_int, [c_void_p, c_void_p]),
    "phx_neighborhood_workspace_size": (c_int, [POINTER(DomainConfiguration), POINTER(c_size_t)]),
    "phx_build_neighborhood": (
        c_int,
        [c_void_p, POINTER(DomainConfiguration), c_uint64, VectorView, c_void_p, c_size_t, c_void_p, c_void_p,
         c_void_p, c_void_p, c_void_p, c_void_p, POINTER(NeighborhoodView)],
    ),
    "phx_build_neighborhood_noalloc": (
        c_int,
        [c_void_p, POINTER(DomainConfiguration), c_uint64, VectorView, c_void_p, c_void_p, c_void_p,
         POINTER(NeighborhoodView)],
    ),
    "phx_pbf_batch_create": (
        c_int,
        [POINTER(DomainConfiguration), c_uint32, POINTER(PbfConfiguration), POINTER(PbfBatchConfiguration), c_void_p,
         POINTER(c_void_p)],
    ),
    "phx_pbf_batch_destroy": (c_int, [c_void_p]),
    "phx_pbf_batch_particle_offsets": (c_int, [c_void_p, POINTER(c_uint64)]),
    "phx_pbf_batch_run": (c_int, [c_void_p, POINTER(PbfState), POINTER(Control), POINTER(PbfParameters), c_uint32]),
    "phx_pbf_batch_kernels_per_step": (c_int, [c_void_p, POINTER(c_uint32)]),
    "phx_density_forward": (
        c_int,
        [c_void_p, c_uint32, c_float, c_int, VectorView, VectorView, ParticleParameters, NeighborhoodView,
         BoundaryView, c_void_p],
    ),
    "phx_density_jvp": (
        c_int,
        [c_void_p, c_uint32, c_float, c_int, VectorView, VectorView, VectorView, ParticleParameters,
         ParticleParameterTangent, NeighborhoodView, BoundaryView, c_void_p],
    ),
    "phx_density_vjp": (
        c_int,
        [c_void_p, c_uint32, c_float, c_int, VectorView, VectorView, ParticleParameters, NeighborhoodView,
         BoundaryView, c_void_p, AdjointVectorView, ParticleParameterAdjoint],
    ),
    "phx_dfsph_create": (
        c_int,
        [POINTER(DomainConfiguration), POINTER(DfsphConfiguration), c_int, c_void_p, POINTER(c_void_p)],
    ),
    "phx_dfsph_destroy": (c_int, [c_void_p]),
    "phx_dfsph_allocate_step_cache": (c_int, [c_void_p, POINTER(c_void_p)]),
    "phx_dfsph_free_step_cache": (c_int, [c_void_p]),
    "phx_dfsph_forward_step": (
        c_int,
        [c_void_p, POINTER(SphDfsphState), POINTER(Control), POINTER(SphDfsphParameters), POINTER(SphDfsphState),
         c_void_p],
    ),
    "phx_dfsph_reference_gradient_norm": (c_int, [c_void_p, POINTER(c_float)]),
    "phx_dfsph_last_iterations": (c_int, [c_void_p, POINTER(c_uint32)]),
    "phx_dfsph_mask_skin": (c_int, [c_void_p, POINTER(c_float)]),
    "phx_dfsph_set_profiling": (c_int, [c_void_p, c_int]),
    "phx_dfsph_read_profile": (c_int, [c_void_p, POINTER(PassProfile), c_uint32, POINTER(c_uint32)]),
    "phx_dfsph_cache_neighborhood": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "phx_dfsph_cache_densities": (c_int, [c_void_p, c_void_p]),
    "phx_dfsph_cache_non_pressure_accelerations": (c_int, [c_void_p, VectorView]),
    "phx_dfsph_jvp_step": (
        c_int,
        [c_void_p, c_void_p, POINTER(SphDfsphStateTangent), POINTER(ControlTangent),
         POINTER(SphDfsphParameterTangent), POINTER(SphDfsphStateTangent)],
    ),
    "phx_dfsph_vjp_step": (
        c_int,
        [c_void_p, c_void_p, POINTER(SphDfsphStateAdjoint), POINTER(SphDfsphStateAdjoint), POINTER(ControlAdjoint),
         POINTER(SphDfsphParameterAdjoint)],
    ),
    "phx_sph_create": (
        c_int,
        [POINTER(DomainConfiguration), POINTER(SphConfiguration), c_int, c_void_p, POINTER(c_void_p)],
    ),
    "phx_sph_destroy": (c_int, [c_void_p]),
    "phx_sph_allocate_step_cache": (c_int, [c_void_p, POINTER(c_void_p)]),
    "phx_sph_free_step_cache": (c_int, [c_void_p]),
    "phx_sph_forward_step": (
        c_int,
        [c_void_p, POINTER(SphState), POINTER(Control), POINTER(SphParameters), POINTER(SphState), c_void_p],
    ),
    "phx_sph_reference_gradient_norm": (c_int, [c_void_p, POINTER(c_float)]),
    "phx_sph_last_iterations": (c_int, [c_void_p, POINTER(c_uint32)]),
    "phx_sph_set_profiling": (c_int, [c_void_p, c_int]),
    "phx_sph_read_profile": (c_int, [c_void_p, POINTER(PassProfile), c_uint32, POINTER(c_uint32)]),
    "phx_sph_cache_neighborhood": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "phx_sph_cache_densities": (c_int, [c_void_p, c_void_p]),
    "phx_sph_cache_pressures": (c_int, [c_void_p, c_void_p]),
    "phx_sph_cache_non_pressure_accelerations": (c_int, [c_void_p, VectorView]),
    "phx_sph_jvp_step": (
        c_int,
        [c_void_p, c_void_p, POINTER(SphStateTangent), POINTER(ControlTangent), POINTER(SphParameterTangent),
         POINTER(SphStateTangent)],
    ),
    "phx_sph_vjp_step": (
        c_int,
        [c_void_p, c_void_p, POINTER(SphStateAdjoint), POINTER(SphStateAdjoint), POINTER(ControlAdjoint),
         POINTER(SphParameterAdjoint)],
    ),
    "phx_pbf_create": (
        c_int,
        [POINTER(DomainConfiguration), POINTER(PbfConfiguration), c_int, c_void_p, POINTER(c_void_p)],
    ),
    "phx_pbf_destroy": (c_int, [c_void_p]),
    "phx_pbf_allocate_step_cache": (c_int, [c_void_p, POINTER(c_void_p)]),
    "phx_pbf_free_step_cache": (c_int, [c_void_p]),
    "phx_pbf_forward_step": (
        c_int,
        [c_void_p, POINTER(PbfState), POINTER(Control), POINTER(PbfParameters), POINTER(PbfState), c_void_p],
    ),
    "phx_pbf_jvp_step": (
        c_int,
        [c_void_p, POINTER(PbfParameters), c_void_p, POINTER(PbfStateTangent), POINTER(ControlTangent),
         POINTER(PbfParameterTangent), POINTER(PbfStateTangent)],
    ),
    "phx_pbf_vjp_step": (
        c_int,
        [c_void_p, POINTER(PbfParameters), c_void_p, POINTER(PbfStateAdjoint), POINTER(PbfStateAdjoint),
         POINTER(ControlAdjoint), POINTER(PbfParameterAdjoint)],
    ),
    "phx_pbf_set_profiling": (c_int, [c_void_p, c_int]),
    "phx_pbf_read_profile": (c_int, [c_void_p, POINTER(PassProfile), c_uint32, POINTER(c_uint32)]),
    "phx_pbf_cache_neighborhood": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "phx_pbf_cache_predicted_positions": (c_int, [c_void_p, VectorView]),
    "phx_pbf_cache_last_iteration": (c_int, [c_void_p, c_void_p, c_void_p]),
}

_lib = None


def library() -> ctypes.CDLL:
    """The loaded libphysica_amd.so with prototypes set; raises if not built."""
    global _lib
    if _lib is None:
        if not LIBRARY_PATH.exists():
            raise RuntimeError(
                f"{LIBRARY_PATH} is not built; run `python -m physica_amd.build` (there is no CPU fallback)"
            )
        lib = ctypes.CDLL(str(LIBRARY_PATH))
        for name, (restype, argtypes) in PROTOTYPES.items():
            fn = getattr(lib, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _lib = lib
    return _lib


def check(status: int, function: str) -> None:
    """Raise PhxError for a non-success status."""
    if status != PHX_SUCCESS:
        message = library().phx_last_error().decode(errors="replace")
        raise PhxError(status, function, message)


def call(name: str, *args) -> None:
    """Call an entry point that returns phx_status and raise on failure."""
    check(getattr(library(), name)(*args), name)
