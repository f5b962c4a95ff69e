"""Python mirror of the reference's particle-liquid solver API.

Names and argument meaning follow physica.fluids.liquid.particle
(fluids/liquid/particle/{domain,sph,pbf}/*.ixx): a DomainConfiguration, SoA
VectorFields, ParticleState/Control/ParticleParameters, and solver objects
with allocate_* and forward_step(state, control, parameters, next_state,
cache). Device memory is held in torch tensors; every step runs through the
HIP library (libphysica_amd.so) and fails loudly if it is missing.
"""
from __future__ import annotations

import ctypes
import functools
import math
from dataclasses import dataclass, field
from enum import IntEnum
from typing import NamedTuple, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _capi


# ------------------------------------------------------------ configuration
@dataclass
class BoxBoundary:
    """BoxBoundary, domain/domain.ixx:18-23."""

    minimum: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    maximum: Tuple[float, float, float] = (1.0, 1.0, 1.0)
    velocity: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    no_slip: bool = True


@dataclass
class DomainConfiguration:
    """DomainConfiguration, domain/domain.ixx:31-39."""

    particle_count: int
    time_step: float
    support_radius: float
    particle_radius: float
    gravity: Tuple[float, float, float] = (0.0, -9.81, 0.0)
    boundary: BoxBoundary = field(default_factory=BoxBoundary)
    boundary_particles: Sequence = ()

    def to_c(self) -> _capi.DomainConfiguration:
        c = _capi.DomainConfiguration()
        c.particle_count = int(self.particle_count)
        c.time_step = float(self.time_step)
        c.support_radius = float(self.support_radius)
        c.particle_radius = float(self.particle_radius)
        c.gravity = _capi.Vector3(*map(float, self.gravity))
        c.boundary = _capi.BoxBoundary(
            _capi.Vector3(*map(float, self.boundary.minimum)),
            _capi.Vector3(*map(float, self.boundary.maximum)),
            _capi.Vector3(*map(float, self.boundary.velocity)),
            1 if self.boundary.no_slip else 0,
        )
        count = len(self.boundary_particles)
        if count:
            arr = (_capi.BoundaryParticle * count)()
            for k, (pos, vel, vol) in enumerate(self.boundary_particles):
                arr[k] = _capi.BoundaryParticle(_capi.Vector3(*pos), _capi.Vector3(*vel), float(vol))
            c.boundary_particles = ctypes.cast(arr, ctypes.POINTER(_capi.BoundaryParticle))
            c._keep = arr  # keep the host array alive with the struct
        c.boundary_particle_count = count
        return c

    def grid_resolution(self) -> Tuple[int, int, int]:
        res = (ctypes.c_uint32 * 3)()
        _capi.call("phx_grid_resolution", ctypes.byref(self.to_c()), res)
        return int(res[0]), int(res[1]), int(res[2])


class ExecutionMode(IntEnum):
    """ExecutionMode, domain/domain.ixx:41-44."""

    forward = 0
    differentiable = 1


# ------------------------------------------------------------------ fields
def _ptr(t: Optional[torch.Tensor]) -> Optional[int]:
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("device buffers must be contiguous")
    return t.data_ptr()


@dataclass
class VectorField:
    """VectorField (SoA x|y|z), domain/domain.ixx:58-62."""

    x: torch.Tensor
    y: torch.Tensor
    z: torch.Tensor

    def view(self) -> _capi.VectorView:
        return _capi.VectorView(_ptr(self.x), _ptr(self.y), _ptr(self.z))

    def numpy(self) -> np.ndarray:
        return np.stack([self.x.cpu().numpy(), self.y.cpu().numpy(), self.z.cpu().numpy()])

    def copy_(self, other: "VectorField") -> None:
        self.x.copy_(other.x)
        self.y.copy_(other.y)
        self.z.copy_(other.z)

    def upload(self, values: np.ndarray) -> None:
        """Domain::upload for a (3, N) host array (domain.cpp:206-218)."""
        values = np.asarray(values, dtype=np.float32)
        self.x.copy_(torch.from_numpy(np.ascontiguousarray(values[0])))
        self.y.copy_(torch.from_numpy(np.ascontiguousarray(values[1])))
        self.z.copy_(torch.from_numpy(np.ascontiguousarray(values[2])))

    @staticmethod
    def zeros(count: int, device) -> "VectorField":
        return VectorField(*(torch.zeros(count, dtype=torch.float32, device=device) for _ in range(3)))


@dataclass
class VectorAdjointField:
    """VectorAdjointField (FP64 SoA x|y|z), domain/domain.ixx:64-68."""

    x: torch.Tensor
    y: torch.Tensor
    z: torch.Tensor

    def view(self) -> _capi.AdjointVectorView:
        return _capi.AdjointVectorView(_ptr(self.x), _ptr(self.y), _ptr(self.z))

    def numpy(self) -> np.ndarray:
        return np.stack([self.x.cpu().numpy(), self.y.cpu().numpy(), self.z.cpu().numpy()])

    def copy_(self, other: "VectorAdjointField") -> None:
        self.x.copy_(other.x)
        self.y.copy_(other.y)
        self.z.copy_(other.z)

    def add_(self, other: "VectorAdjointField") -> None:
        """Domain::accumulate (domain/domain.cu:8-29)."""
        self.x.add_(other.x)
        self.y.add_(other.y)
        self.z.add_(other.z)

    def zero_(self) -> None:
        self.x.zero_()
        self.y.zero_()
        self.z.zero_()

    def upload(self, values: np.ndarray) -> None:
        values = np.asarray(values, dtype=np.float64)
        self.x.copy_(torch.from_numpy(np.ascontiguousarray(values[0])))
        self.y.copy_(torch.from_numpy(np.ascontiguousarray(values[1])))
        self.z.copy_(torch.from_numpy(np.ascontiguousarray(values[2])))

    @staticmethod
    def zeros(count: int, device) -> "VectorAdjointField":
        return VectorAdjointField(*(torch.zeros(count, dtype=torch.float64, device=device) for _ in range(3)))


@dataclass
class ParticleStateTangent:
    """ParticleStateTangent, domain/domain.ixx:76-79."""

    positions: VectorField
    velocities: VectorField


@dataclass
class ParticleStateAdjoint:
    """ParticleStateAdjoint, domain/domain.ixx:85-88."""

    positions: VectorAdjointField
    velocities: VectorAdjointField


@dataclass
class ControlTangent:
    external_accelerations: VectorField

    def to_c(self) -> _capi.ControlTangent:
        return _capi.ControlTangent(self.external_accelerations.view())


@dataclass
class ControlAdjoint:
    external_accelerations: VectorAdjointField

    def to_c(self) -> _capi.ControlAdjoint:
        return _capi.ControlAdjoint(self.external_accelerations.view())


@dataclass
class ParticleParameterTangent:
    """ParticleParameterTangent (FP32), domain/domain.ixx:101-106."""

    masses: torch.Tensor
    rest_densities: torch.Tensor
    viscosities: torch.Tensor
    surface_tensions: torch.Tensor

    def to_c(self) -> _capi.ParticleParameterTangent:
        return _capi.ParticleParameterTangent(_ptr(self.masses), _ptr(self.rest_densities), _ptr(self.viscosities),
                                              _ptr(self.surface_tensions))


@dataclass
class ParticleParameterAdjoint:
    """ParticleParameterAdjoint (FP64), domain/domain.ixx:108-113."""

    masses: torch.Tensor
    rest_densities: torch.Tensor
    viscosities: torch.Tensor
    surface_tensions: torch.Tensor

    def to_c(self) -> _capi.ParticleParameterAdjoint:
        return _capi.ParticleParameterAdjoint(_ptr(self.masses), _ptr(self.rest_densities), _ptr(self.viscosities),
                                              _ptr(self.surface_tensions))


@dataclass
class ParticleState:
    """ParticleState, domain/domain.ixx:70-74."""

    positions: VectorField
    velocities: VectorField
    step_index: int = 0

    def to_c(self) -> _capi.ParticleState:
        return _capi.ParticleState(self.positions.view(), self.velocities.view(), int(self.step_index))


@dataclass
class Control:
    """Control, domain/domain.ixx:76-78."""

    external_accelerations: VectorField

    def to_c(self) -> _capi.Control:
        return _capi.Control(self.external_accelerations.view())


@dataclass
class ParticleParameters:
    """ParticleParameters, domain/domain.ixx:100-105."""

    masses: torch.Tensor
    rest_densities: torch.Tensor
    viscosities: torch.Tensor
    surface_tensions: torch.Tensor

    def to_c(self) -> _capi.ParticleParameters:
        return _capi.ParticleParameters(
            _ptr(self.masses), _ptr(self.rest_densities), _ptr(self.viscosities), _ptr(self.surface_tensions)
        )


class Domain:
    """Allocation helpers of Domain (domain/domain.ixx:121-165) on a torch device."""

    def __init__(self, configuration: DomainConfiguration, device="cuda"):
        self.configuration = configuration
        self.device = torch.device(device)

    def allocate_scalar_field(self, count: Optional[int] = None) -> torch.Tensor:
        n = self.configuration.particle_count if count is None else count
        return torch.zeros(n, dtype=torch.float32, device=self.device)

    def allocate_scalar_adjoint_field(self, count: Optional[int] = None) -> torch.Tensor:
        n = self.configuration.particle_count if count is None else count
        return torch.zeros(n, dtype=torch.float64, device=self.device)

    def allocate_vector_adjoint_field(self, count: Optional[int] = None) -> VectorAdjointField:
        n = self.configuration.particle_count if count is None else count
        return VectorAdjointField.zeros(n, self.device)

    def allocate_vector_field(self, count: Optional[int] = None) -> VectorField:
        n = self.configuration.particle_count if count is None else count
        return VectorField.zeros(n, self.device)

    def allocate_particle_state(self) -> ParticleState:
        return ParticleState(self.allocate_vector_field(), self.allocate_vector_field(), 0)

    def allocate_control(self) -> Control:
        return Control(self.allocate_vector_field())

    def allocate_particle_parameters(self) -> ParticleParameters:
        return ParticleParameters(*(self.allocate_scalar_field() for _ in range(4)))

    def copy(self, source: ParticleState, destination: ParticleState) -> None:
        destination.positions.copy_(source.positions)
        destination.velocities.copy_(source.velocities)
        destination.step_index = source.step_index


def _read_profile(function: str, handle) -> list:
    entries = (_capi.PassProfile * 32)()
    count = ctypes.c_uint32()
    _capi.call(function, handle, entries, 32, ctypes.byref(count))
    return [
        {"name": e.name.decode(), "total_ms": e.total_ms, "launches": int(e.launches),
         "bytes_per_particle": e.bytes_per_particle}
        for e in entries[: count.value]
    ]


def _stream_handle(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


class _SolverStream:
    """A solver enqueues its work on the stream that was current when it was
    created; the tensors a call reads or writes are filled, read and freed on
    whatever stream is current at the call. Inside this block the solver's
    stream first waits for the caller's current stream, and on leaving it the
    current stream waits for the solver's stream, so work on the two streams
    never overlaps on shared tensors (nothing is waited for when the two are
    the same stream)."""

    def __init__(self, solver):
        self._own = solver._stream
        self._current = torch.cuda.current_stream(solver.domain.device)
        self._same = self._current.cuda_stream == self._own.cuda_stream

    def __enter__(self):
        if not self._same:
            self._own.wait_stream(self._current)
        return self

    def __exit__(self, *exc):
        if not self._same:
            self._current.wait_stream(self._own)
        return False


def _stream_ordered(method):
    """Runs a solver method inside _SolverStream (the method must enqueue no
    work of its own on the current stream)."""

    @functools.wraps(method)
    def ordered(self, *args, **kwargs):
        with _SolverStream(getattr(self, "_solver", self)):
            return method(self, *args, **kwargs)

    return ordered


def _release_handle(owner) -> None:
    """Frees owner._handle with the C entry point bound at creation
    (owner._release), once. Used as the classes' __del__ itself, so it needs no
    module global at interpreter exit, when those may already be cleared."""
    handle = getattr(owner, "_handle", None)
    release = getattr(owner, "_release", None)
    if handle is not None and handle.value and release is not None:
        owner._handle = None
        release(handle)


def _require_gpu(device: torch.device) -> None:
    if device.type != "cuda" or not torch.cuda.is_available():
        raise RuntimeError("physica_amd solvers run on the GPU only (no CPU fallback); no HIP device is available")


# -------------------------------------------------------------------- DFSPH
class DFSPH:
    """DFSPH method types, sph/sph.ixx:195-268."""

    @dataclass
    class Configuration:
        divergence_iterations: int = 4
        density_iterations: int = 6
        checkpoint_interval: int = 2
        pressure_warm_start: bool = True
        density_error_tolerance: float = 0.0
        max_iterations: int = 100
        # VJP of a no-slip box collision: False (default) the reference's
        # (sph/dfsph.cu:81 zeroes the whole velocity adjoint), True the exact
        # transpose of the JVP
        exact_no_slip_adjoint: bool = False

        def to_c(self) -> _capi.DfsphConfiguration:
            return _capi.DfsphConfiguration(
                int(self.divergence_iterations),
                int(self.density_iterations),
                int(self.checkpoint_interval),
                1 if self.pressure_warm_start else 0,
                float(self.density_error_tolerance),
                int(self.max_iterations),
                1 if self.exact_no_slip_adjoint else 0,
            )

    @dataclass
    class State:
        warm_divergence_pressure: torch.Tensor
        warm_density_pressure: torch.Tensor

        def to_c(self) -> _capi.DfsphState:
            return _capi.DfsphState(_ptr(self.warm_divergence_pressure), _ptr(self.warm_density_pressure))

    @dataclass
    class Parameters:
        divergence_relaxation: torch.Tensor
        density_relaxation: torch.Tensor

        def to_c(self) -> _capi.DfsphParameters:
            return _capi.DfsphParameters(_ptr(self.divergence_relaxation), _ptr(self.density_relaxation))

    @dataclass
    class StateTangent:
        """FP32 (sph/sph.ixx:206-209); None counts as zero on input."""

        warm_divergence_pressure: Optional[torch.Tensor]
        warm_density_pressure: Optional[torch.Tensor]

        def to_c(self) -> _capi.DfsphStateTangent:
            return _capi.DfsphStateTangent(_ptr(self.warm_divergence_pressure), _ptr(self.warm_density_pressure))

    @dataclass
    class StateAdjoint:
        """FP64 (sph/sph.ixx:210-213)."""

        warm_divergence_pressure: Optional[torch.Tensor]
        warm_density_pressure: Optional[torch.Tensor]

        def to_c(self) -> _capi.DfsphStateAdjoint:
            return _capi.DfsphStateAdjoint(_ptr(self.warm_divergence_pressure), _ptr(self.warm_density_pressure))

    @dataclass
    class ParameterTangent:
        divergence_relaxation: Optional[torch.Tensor]
        density_relaxation: Optional[torch.Tensor]

        def to_c(self) -> _capi.DfsphParameterTangent:
            return _capi.DfsphParameterTangent(_ptr(self.divergence_relaxation), _ptr(self.density_relaxation))

    @dataclass
    class ParameterAdjoint:
        divergence_relaxation: Optional[torch.Tensor]
        density_relaxation: Optional[torch.Tensor]

        def to_c(self) -> _capi.DfsphParameterAdjoint:
            return _capi.DfsphParameterAdjoint(_ptr(self.divergence_relaxation), _ptr(self.density_relaxation))


# ------------------------------------------------------ WCSPH, PCISPH, IISPH
class WCSPH:
    """WCSPH method types, sph/sph.ixx:39-101."""

    @dataclass
    class Configuration:
        # VJP of a no-slip box collision: False (default) the reference's
        # (sph/common.cu:357 zeroes the whole velocity adjoint), True the
        # exact transpose of the JVP
        exact_no_slip_adjoint: bool = False

        def to_c(self) -> _capi.SphConfiguration:
            return _capi.SphConfiguration(_capi.PHX_SPH_WCSPH, 0, 2, 0.0, 0, 1 if self.exact_no_slip_adjoint else 0)

    @dataclass
    class State:
        pass

    @dataclass
    class Parameters:
        speed_of_sound: torch.Tensor
        tait_exponent: torch.Tensor
        boundary_surface_tension: torch.Tensor

        def to_c(self) -> _capi.SphMethodParameters:
            return _capi.SphMethodParameters(
                _ptr(self.speed_of_sound), _ptr(self.tait_exponent), _ptr(self.boundary_surface_tension), None, None
            )

    @dataclass
    class StateTangent:
        """Empty (sph/sph.ixx:42)."""

    @dataclass
    class StateAdjoint:
        """Empty (sph/sph.ixx:43)."""

    @dataclass
    class ParameterTangent:
        """FP32 (sph/sph.ixx:51-55); None counts as zero."""

        speed_of_sound: Optional[torch.Tensor]
        tait_exponent: Optional[torch.Tensor]
        boundary_surface_tension: Optional[torch.Tensor]

        def to_c(self) -> _capi.SphMethodParameterTangent:
            return _capi.SphMethodParameterTangent(_ptr(self.speed_of_sound), _ptr(self.tait_exponent),
                                                   _ptr(self.boundary_surface_tension), None, None)

    @dataclass
    class ParameterAdjoint:
        """FP64 (sph/sph.ixx:57-61); None is skipped."""

        speed_of_sound: Optional[torch.Tensor]
        tait_exponent: Optional[torch.Tensor]
        boundary_surface_tension: Optional[torch.Tensor]

        def to_c(self) -> _capi.SphMethodParameterAdjoint:
            return _capi.SphMethodParameterAdjoint(_ptr(self.speed_of_sound), _ptr(self.tait_exponent),
                                                   _ptr(self.boundary_surface_tension), None, None)


class PCISPH:
    """PCISPH method types, sph/sph.ixx:103-147."""

    @dataclass
    class Configuration:
        pressure_iterations: int = 6
        checkpoint_interval: int = 2
        # tolerance mode (an extension): iterate to this mean compression
        # error, at most max_iterations times; 0: fixed pressure_iterations
        density_error_tolerance: float = 0.0
        max_iterations: int = 100
        # VJP of a no-slip box collision: False (default) the reference's
        # (sph/dfsph.cu:81 zeroes the whole velocity adjoint), True the exact
        # transpose of the JVP
        exact_no_slip_adjoint: bool = False

        def to_c(self) -> _capi.SphConfiguration:
            return _capi.SphConfiguration(
                _capi.PHX_SPH_PCISPH, int(self.pressure_iterations), int(self.checkpoint_interval),
                float(self.density_error_tolerance), int(self.max_iterations), 1 if self.exact_no_slip_adjoint else 0
            )

    @dataclass
    class State:
        pass

    @dataclass
    class Parameters:
        pressure_relaxation: torch.Tensor

        def to_c(self) -> _capi.SphMethodParameters:
            return _capi.SphMethodParameters(None, None, None, _ptr(self.pressure_relaxation), None)

    @dataclass
    class StateTangent:
        """Empty (sph/sph.ixx:109)."""

    @dataclass
    class StateAdjoint:
        """Empty (sph/sph.ixx:110)."""

    @dataclass
    class ParameterTangent:
        """FP32 (sph/sph.ixx:112); None counts as zero."""

        pressure_relaxation: Optional[torch.Tensor]

        def to_c(self) -> _capi.SphMethodParameterTangent:
            return _capi.SphMethodParameterTangent(None, None, None, _ptr(self.pressure_relaxation), None)

    @dataclass
    class ParameterAdjoint:
        """FP64 (sph/sph.ixx:113); None is skipped."""

        pressure_relaxation: Optional[torch.Tensor]

        def to_c(self) -> _capi.SphMethodParameterAdjoint:
            return _capi.SphMethodParameterAdjoint(None, None, None, _ptr(self.pressure_relaxation), None)


class IISPH:
    """IISPH method types, sph/sph.ixx:149-193."""

    @dataclass
    class Configuration:
        pressure_iterations: int = 6
        checkpoint_interval: int = 2
        # tolerance mode (an extension): iterate to this mean compression
        # error, at most max_iterations times; 0: fixed pressure_iterations
        density_error_tolerance: float = 0.0
        max_iterations: int = 100
        # VJP of a no-slip box collision: False (default) the reference's
        # (sph/dfsph.cu:81 zeroes the whole velocity adjoint), True the exact
        # transpose of the JVP
        exact_no_slip_adjoint: bool = False

        def to_c(self) -> _capi.SphConfiguration:
            return _capi.SphConfiguration(_capi.PHX_SPH_IISPH, int(self.pressure_iterations),
                                          int(self.checkpoint_interval), float(self.density_error_tolerance),
                                          int(self.max_iterations), 1 if self.exact_no_slip_adjoint else 0)

    @dataclass
    class State:
        pass

    @dataclass
    class Parameters:
        jacobi_relaxation: torch.Tensor

        def to_c(self) -> _capi.SphMethodParameters:
            return _capi.SphMethodParameters(None, None, None, None, _ptr(self.jacobi_relaxation))

    @dataclass
    class StateTangent:
        """Empty (sph/sph.ixx:155)."""

    @dataclass
    class StateAdjoint:
        """Empty (sph/sph.ixx:156)."""

    @dataclass
    class ParameterTangent:
        """FP32 (sph/sph.ixx:158); None counts as zero."""

        jacobi_relaxation: Optional[torch.Tensor]

        def to_c(self) -> _capi.SphMethodParameterTangent:
            return _capi.SphMethodParameterTangent(None, None, None, None, _ptr(self.jacobi_relaxation))

    @dataclass
    class ParameterAdjoint:
        """FP64 (sph/sph.ixx:159); None is skipped."""

        jacobi_relaxation: Optional[torch.Tensor]

        def to_c(self) -> _capi.SphMethodParameterAdjoint:
            return _capi.SphMethodParameterAdjoint(None, None, None, None, _ptr(self.jacobi_relaxation))


_METHODS = {DFSPH.Configuration: DFSPH, WCSPH.Configuration: WCSPH, PCISPH.Configuration: PCISPH,
            IISPH.Configuration: IISPH}


class SPH:
    """SPH<Method> (sph/sph.ixx:273-372) over the HIP library; the method is
    chosen by the type of the method configuration (DFSPH by default)."""

    @dataclass
    class State:
        particles: ParticleState
        method: object

        def to_c(self):
            if isinstance(self.method, DFSPH.State):
                return _capi.SphDfsphState(self.particles.to_c(), self.method.to_c())
            return _capi.SphState(self.particles.to_c())

    @dataclass
    class Parameters:
        particles: ParticleParameters
        method: object

        def to_c(self):
            if isinstance(self.method, DFSPH.Parameters):
                return _capi.SphDfsphParameters(self.particles.to_c(), self.method.to_c())
            return _capi.SphParameters(self.particles.to_c(), self.method.to_c())

    @dataclass
    class StateTangent:
        particles: ParticleStateTangent
        method: object

        def to_c(self):
            particles = _capi.ParticleStateTangent(self.particles.positions.view(), self.particles.velocities.view())
            if isinstance(self.method, DFSPH.StateTangent):
                return _capi.SphDfsphStateTangent(particles, self.method.to_c())
            return _capi.SphStateTangent(particles)

    @dataclass
    class StateAdjoint:
        particles: ParticleStateAdjoint
        method: object

        def to_c(self):
            particles = _capi.ParticleStateAdjoint(self.particles.positions.view(), self.particles.velocities.view())
            if isinstance(self.method, DFSPH.StateAdjoint):
                return _capi.SphDfsphStateAdjoint(particles, self.method.to_c())
            return _capi.SphStateAdjoint(particles)

    @dataclass
    class ParameterTangent:
        particles: ParticleParameterTangent
        method: object

        def to_c(self):
            if isinstance(self.method, DFSPH.ParameterTangent):
                return _capi.SphDfsphParameterTangent(self.particles.to_c(), self.method.to_c())
            return _capi.SphParameterTangent(self.particles.to_c(), self.method.to_c())

    @dataclass
    class ParameterAdjoint:
        particles: ParticleParameterAdjoint
        method: object

        def to_c(self):
            if isinstance(self.method, DFSPH.ParameterAdjoint):
                return _capi.SphDfsphParameterAdjoint(self.particles.to_c(), self.method.to_c())
            return _capi.SphParameterAdjoint(self.particles.to_c(), self.method.to_c())

    class StepCache:
        """StepCache (sph/sph.ixx:299-303); fields are read back in original order."""

        def __init__(self, solver: "SPH"):
            self._solver = solver
            self._prefix = solver._prefix
            handle = ctypes.c_void_p()
            _capi.call(self._prefix + "_allocate_step_cache", solver._handle, ctypes.byref(handle))
            self._handle = handle
            # bound now: module globals may already be gone when __del__ runs at exit
            self._release = getattr(_capi.library(), self._prefix + "_free_step_cache")

        __del__ = _release_handle

        def neighborhood(self):
            n = self._solver.domain.configuration.particle_count
            cells = math.prod(self._solver.domain.configuration.grid_resolution())
            dev = self._solver.domain.device
            keys = torch.empty(n, dtype=torch.int64, device=dev)
            idx = torch.empty(n, dtype=torch.int32, device=dev)
            off = torch.empty(cells + 1, dtype=torch.int32, device=dev)
            with _SolverStream(self._solver):
                _capi.call(self._prefix + "_cache_neighborhood", self._handle, _ptr(keys), _ptr(idx), _ptr(off))
            return keys, idx, off

        def densities(self) -> torch.Tensor:
            out = self._solver.domain.allocate_scalar_field()
            with _SolverStream(self._solver):
                _capi.call(self._prefix + "_cache_densities", self._handle, _ptr(out))
            return out

        def pressures(self) -> torch.Tensor:
            """WCSPH::Cache::pressures, or the last PCISPH/IISPH iteration's pressures."""
            if self._prefix != "phx_sph":
                raise AttributeError("the DFSPH step cache keeps its pressures in the method state")
            out = self._solver.domain.allocate_scalar_field()
            with _SolverStream(self._solver):
                _capi.call("phx_sph_cache_pressures", self._handle, _ptr(out))
            return out

        def non_pressure_accelerations(self) -> VectorField:
            out = self._solver.domain.allocate_vector_field()
            with _SolverStream(self._solver):
                _capi.call(self._prefix + "_cache_non_pressure_accelerations", self._handle, out.view())
            return out

    def __init__(
        self,
        domain_configuration: DomainConfiguration,
        method_configuration=None,
        mode: ExecutionMode = ExecutionMode.forward,
        device="cuda",
    ):
        self.domain = Domain(domain_configuration, device)
        self.method_configuration = method_configuration if method_configuration is not None else DFSPH.Configuration()
        self.method = _METHODS.get(type(self.method_configuration))
        if self.method is None:
            raise TypeError(f"unknown SPH method configuration {type(self.method_configuration).__name__}")
        self._prefix = "phx_dfsph" if self.method is DFSPH else "phx_sph"
        _require_gpu(self.domain.device)
        handle = ctypes.c_void_p()
        self._domain_c = domain_configuration.to_c()
        self._stream = torch.cuda.current_stream(self.domain.device)
        _capi.call(
            self._prefix + "_create",
            ctypes.byref(self._domain_c),
            ctypes.byref(self.method_configuration.to_c()),
            int(mode),
            self._stream.cuda_stream,
            ctypes.byref(handle),
        )
        self._handle = handle
        self._release = getattr(_capi.library(), self._prefix + "_destroy")

    __del__ = _release_handle

    def allocate_state(self) -> "SPH.State":
        d = self.domain
        if self.method is DFSPH:
            method = DFSPH.State(d.allocate_scalar_field(), d.allocate_scalar_field())
        else:
            method = self.method.State()
        return SPH.State(d.allocate_particle_state(), method)

    def allocate_control(self) -> Control:
        return self.domain.allocate_control()

    def allocate_parameters(self) -> "SPH.Parameters":
        d = self.domain
        if self.method is DFSPH:
            method = DFSPH.Parameters(d.allocate_scalar_field(), d.allocate_scalar_field())
        elif self.method is WCSPH:
            method = WCSPH.Parameters(d.allocate_scalar_field(), d.allocate_scalar_field(), d.allocate_scalar_field())
        elif self.method is PCISPH:
            method = PCISPH.Parameters(d.allocate_scalar_field())
        else:
            method = IISPH.Parameters(d.allocate_scalar_field())
        return SPH.Parameters(d.allocate_particle_parameters(), method)

    def allocate_step_cache(self) -> "SPH.StepCache":
        return SPH.StepCache(self)

    def copy_state(self, source: "SPH.State", destination: "SPH.State") -> None:
        self.domain.copy(source.particles, destination.particles)
        if self.method is DFSPH:
            destination.method.warm_divergence_pressure.copy_(source.method.warm_divergence_pressure)
            destination.method.warm_density_pressure.copy_(source.method.warm_density_pressure)

    @_stream_ordered
    def forward_step(self, state, control, parameters, next_state, cache) -> None:
        s = state.to_c()
        n = next_state.to_c()
        _capi.call(
            self._prefix + "_forward_step",
            self._handle,
            ctypes.byref(s),
            ctypes.byref(control.to_c()),
            ctypes.byref(parameters.to_c()),
            ctypes.byref(n),
            cache._handle,
        )
        next_state.particles.step_index = int(n.particles.step_index)

    @property
    def reference_gradient_norm(self) -> float:
        v = ctypes.c_float()
        _capi.call(self._prefix + "_reference_gradient_norm", self._handle, ctypes.byref(v))
        return float(v.value)

    def set_profiling(self, enabled: bool) -> None:
        """Bracket every pass with HIP events (per-pass timing)."""
        _capi.call(self._prefix + "_set_profiling", self._handle, 1 if enabled else 0)

    def read_profile(self) -> list:
        """Per-pass totals since profiling was enabled or last read."""
        return _read_profile(self._prefix + "_read_profile", self._handle)

    # tangents and adjoints (sph/sph.ixx:313-348), allocated zero
    def _require_derivatives(self) -> None:
        pass

    def allocate_state_tangent(self) -> "SPH.StateTangent":
        self._require_derivatives()
        d = self.domain
        n = d.configuration.particle_count
        z = lambda: torch.zeros(n, dtype=torch.float32, device=d.device)  # noqa: E731
        method = DFSPH.StateTangent(z(), z()) if self.method is DFSPH else self.method.StateTangent()
        return SPH.StateTangent(ParticleStateTangent(VectorField.zeros(n, d.device), VectorField.zeros(n, d.device)),
                                method)

    def allocate_control_tangent(self) -> ControlTangent:
        return ControlTangent(VectorField.zeros(self.domain.configuration.particle_count, self.domain.device))

    def allocate_parameter_tangent(self) -> "SPH.ParameterTangent":
        self._require_derivatives()
        d = self.domain
        z = lambda: torch.zeros(d.configuration.particle_count, dtype=torch.float32, device=d.device)  # noqa: E731
        if self.method is DFSPH:
            method = DFSPH.ParameterTangent(z(), z())
        elif self.method is WCSPH:
            method = WCSPH.ParameterTangent(z(), z(), z())
        else:
            method = self.method.ParameterTangent(z())
        return SPH.ParameterTangent(ParticleParameterTangent(z(), z(), z(), z()), method)

    def allocate_state_adjoint(self) -> "SPH.StateAdjoint":
        self._require_derivatives()
        d = self.domain
        z = d.allocate_scalar_adjoint_field
        method = DFSPH.StateAdjoint(z(), z()) if self.method is DFSPH else self.method.StateAdjoint()
        return SPH.StateAdjoint(ParticleStateAdjoint(d.allocate_vector_adjoint_field(),
                                                     d.allocate_vector_adjoint_field()), method)

    def allocate_control_adjoint(self) -> ControlAdjoint:
        return ControlAdjoint(self.domain.allocate_vector_adjoint_field())

    def allocate_parameter_adjoint(self) -> "SPH.ParameterAdjoint":
        self._require_derivatives()
        z = self.domain.allocate_scalar_adjoint_field
        if self.method is DFSPH:
            method = DFSPH.ParameterAdjoint(z(), z())
        elif self.method is WCSPH:
            method = WCSPH.ParameterAdjoint(z(), z(), z())
        else:
            method = self.method.ParameterAdjoint(z())
        return SPH.ParameterAdjoint(ParticleParameterAdjoint(z(), z(), z(), z()), method)

    def copy_state_tangent(self, source: "SPH.StateTangent", destination: "SPH.StateTangent") -> None:
        destination.particles.positions.copy_(source.particles.positions)
        destination.particles.velocities.copy_(source.particles.velocities)
        if self.method is DFSPH:
            destination.method.warm_divergence_pressure.copy_(source.method.warm_divergence_pressure)
            destination.method.warm_density_pressure.copy_(source.method.warm_density_pressure)

    def copy_state_adjoint(self, source: "SPH.StateAdjoint", destination: "SPH.StateAdjoint") -> None:
        destination.particles.positions.copy_(source.particles.positions)
        destination.particles.velocities.copy_(source.particles.velocities)
        if self.method is DFSPH:
            destination.method.warm_divergence_pressure.copy_(source.method.warm_divergence_pressure)
            destination.method.warm_density_pressure.copy_(source.method.warm_density_pressure)

    def accumulate_state_adjoint(self, source: "SPH.StateAdjoint", destination: "SPH.StateAdjoint") -> None:
        destination.particles.positions.add_(source.particles.positions)
        destination.particles.velocities.add_(source.particles.velocities)
        if self.method is DFSPH:
            destination.method.warm_divergence_pressure.add_(source.method.warm_divergence_pressure)
            destination.method.warm_density_pressure.add_(source.method.warm_density_pressure)

    @_stream_ordered
    def jvp_step(self, state, control, parameters, next_state, cache, state_tangent, control_tangent,
                 parameter_tangent, next_state_tangent) -> None:
        """SPH<Method>::jvp_step (sph/sph.ixx:356-359) of the step recorded in
        `cache`: state, control, parameters and
        next_state are those of that forward step, which the cache recorded."""
        self._require_derivatives()
        _capi.call(
            self._prefix + "_jvp_step",
            self._handle,
            cache._handle,
            ctypes.byref(state_tangent.to_c()),
            ctypes.byref(control_tangent.to_c()),
            ctypes.byref(parameter_tangent.to_c()),
            ctypes.byref(next_state_tangent.to_c()),
        )

    @_stream_ordered
    def vjp_step(self, state, control, parameters, next_state, cache, next_state_adjoint, previous_state_adjoint,
                 control_adjoint, parameter_adjoint) -> None:
        """SPH<Method>::vjp_step (sph/sph.ixx:361-365): particle, control and
        parameter adjoints accumulate; DFSPH's warm-start pressure adjoints
        are written (with warm start)."""
        self._require_derivatives()
        _capi.call(
            self._prefix + "_vjp_step",
            self._handle,
            cache._handle,
            ctypes.byref(next_state_adjoint.to_c()),
            ctypes.byref(previous_state_adjoint.to_c()),
            ctypes.byref(control_adjoint.to_c()),
            ctypes.byref(parameter_adjoint.to_c()),
        )

    def last_iterations(self) -> Tuple[int, int]:
        """Iterations of the last DFSPH step (divergence, density)."""
        if self.method is not DFSPH:
            it = ctypes.c_uint32()
            _capi.call("phx_sph_last_iterations", self._handle, ctypes.byref(it))
            return (int(it.value), 0)
        v = (ctypes.c_uint32 * 2)()
        _capi.call("phx_dfsph_last_iterations", self._handle, v)
        return int(v[0]), int(v[1])

    def mask_skin(self) -> float:
        """DFSPH diagnostics: the neighbour-mask skin of the next step, in
        units of the support radius."""
        if self.method is not DFSPH:
            raise RuntimeError("mask_skin is a DFSPH diagnostic")
        v = ctypes.c_float()
        _capi.call("phx_dfsph_mask_skin", self._handle, ctypes.byref(v))
        return float(v.value)


# ---------------------------------------------------------------------- PBF
class PBF:
    """PBF (pbf/pbf.ixx:14-141) over the HIP library."""

    @dataclass
    class Configuration:
        pressure_iterations: int = 5
        checkpoint_interval: int = 2

        def to_c(self) -> _capi.PbfConfiguration:
            return _capi.PbfConfiguration(int(self.pressure_iterations), int(self.checkpoint_interval))

    @dataclass
    class State:
        particles: ParticleState

        def to_c(self) -> _capi.PbfState:
            return _capi.PbfState(self.particles.to_c())

    @dataclass
    class Parameters:
        particles: ParticleParameters
        relaxation: torch.Tensor
        artificial_pressure_strength: torch.Tensor
        artificial_pressure_exponent: torch.Tensor
        artificial_pressure_radius: torch.Tensor
        xsph_viscosity: torch.Tensor
        vorticity_confinement: torch.Tensor

        def to_c(self) -> _capi.PbfParameters:
            return _capi.PbfParameters(
                self.particles.to_c(),
                _ptr(self.relaxation),
                _ptr(self.artificial_pressure_strength),
                _ptr(self.artificial_pressure_exponent),
                _ptr(self.artificial_pressure_radius),
                _ptr(self.xsph_viscosity),
                _ptr(self.vorticity_confinement),
            )

    @dataclass
    class StateTangent:
        particles: ParticleStateTangent

        def to_c(self) -> _capi.PbfStateTangent:
            return _capi.PbfStateTangent(self.particles.positions.view(), self.particles.velocities.view())

    @dataclass
    class StateAdjoint:
        particles: ParticleStateAdjoint

        def to_c(self) -> _capi.PbfStateAdjoint:
            return _capi.PbfStateAdjoint(self.particles.positions.view(), self.particles.velocities.view())

    @dataclass
    class ParameterTangent:
        particles: ParticleParameterTangent
        relaxation: torch.Tensor
        artificial_pressure_strength: torch.Tensor
        artificial_pressure_exponent: torch.Tensor
        artificial_pressure_radius: torch.Tensor
        xsph_viscosity: torch.Tensor
        vorticity_confinement: torch.Tensor

        def to_c(self) -> _capi.PbfParameterTangent:
            return _capi.PbfParameterTangent(self.particles.to_c(), *(_ptr(getattr(self, k)) for k in _capi._PBF_FIELDS))

    @dataclass
    class ParameterAdjoint:
        particles: ParticleParameterAdjoint
        relaxation: torch.Tensor
        artificial_pressure_strength: torch.Tensor
        artificial_pressure_exponent: torch.Tensor
        artificial_pressure_radius: torch.Tensor
        xsph_viscosity: torch.Tensor
        vorticity_confinement: torch.Tensor

        def to_c(self) -> _capi.PbfParameterAdjoint:
            return _capi.PbfParameterAdjoint(self.particles.to_c(), *(_ptr(getattr(self, k)) for k in _capi._PBF_FIELDS))

    class StepCache:
        """StepCache (pbf/pbf.ixx:54-70); fields are read back in original order."""

        def __init__(self, solver: "PBF"):
            self._solver = solver
            handle = ctypes.c_void_p()
            _capi.call("phx_pbf_allocate_step_cache", solver._handle, ctypes.byref(handle))
            self._handle = handle
            self._release = _capi.library().phx_pbf_free_step_cache

        __del__ = _release_handle

        def neighborhood(self):
            n = self._solver.domain.configuration.particle_count
            cells = math.prod(self._solver.domain.configuration.grid_resolution())
            dev = self._solver.domain.device
            keys = torch.empty(n, dtype=torch.int64, device=dev)
            idx = torch.empty(n, dtype=torch.int32, device=dev)
            off = torch.empty(cells + 1, dtype=torch.int32, device=dev)
            with _SolverStream(self._solver):
                _capi.call("phx_pbf_cache_neighborhood", self._handle, _ptr(keys), _ptr(idx), _ptr(off))
            return keys, idx, off

        def predicted_positions(self) -> VectorField:
            out = self._solver.domain.allocate_vector_field()
            with _SolverStream(self._solver):
                _capi.call("phx_pbf_cache_predicted_positions", self._handle, out.view())
            return out

        def last_iteration(self):
            rho = self._solver.domain.allocate_scalar_field()
            lam = self._solver.domain.allocate_scalar_field()
            with _SolverStream(self._solver):
                _capi.call("phx_pbf_cache_last_iteration", self._handle, _ptr(rho), _ptr(lam))
            return rho, lam

        def vorticity_magnitudes(self) -> torch.Tensor:
            """StepCache::vorticity_magnitudes (pbf/pbf.ixx:67), original order."""
            out = self._solver.domain.allocate_scalar_field()
            with _SolverStream(self._solver):
                _capi.call("phx_pbf_cache_vorticity_magnitudes", self._handle, _ptr(out))
            return out

    def __init__(
        self,
        domain_configuration: DomainConfiguration,
        configuration: Optional["PBF.Configuration"] = None,
        mode: ExecutionMode = ExecutionMode.forward,
        device="cuda",
    ):
        self.domain = Domain(domain_configuration, device)
        self.configuration = configuration or PBF.Configuration()
        _require_gpu(self.domain.device)
        handle = ctypes.c_void_p()
        self._domain_c = domain_configuration.to_c()
        self._stream = torch.cuda.current_stream(self.domain.device)
        _capi.call(
            "phx_pbf_create",
            ctypes.byref(self._domain_c),
            ctypes.byref(self.configuration.to_c()),
            int(mode),
            self._stream.cuda_stream,
            ctypes.byref(handle),
        )
        self._handle = handle
        self._release = _capi.library().phx_pbf_destroy

    __del__ = _release_handle

    def allocate_state(self) -> "PBF.State":
        return PBF.State(self.domain.allocate_particle_state())

    def allocate_control(self) -> Control:
        return self.domain.allocate_control()

    def allocate_parameters(self) -> "PBF.Parameters":
        d = self.domain
        return PBF.Parameters(d.allocate_particle_parameters(), *(d.allocate_scalar_field() for _ in range(6)))

    def allocate_step_cache(self) -> "PBF.StepCache":
        return PBF.StepCache(self)

    def copy_state(self, source: "PBF.State", destination: "PBF.State") -> None:
        self.domain.copy(source.particles, destination.particles)

    # tangents and adjoints (pbf/pbf.ixx:84-94); allocated zero
    def allocate_state_tangent(self) -> "PBF.StateTangent":
        d = self.domain
        return PBF.StateTangent(ParticleStateTangent(VectorField.zeros(d.configuration.particle_count, d.device),
                                                     VectorField.zeros(d.configuration.particle_count, d.device)))

    def allocate_control_tangent(self) -> ControlTangent:
        return ControlTangent(VectorField.zeros(self.domain.configuration.particle_count, self.domain.device))

    def allocate_parameter_tangent(self) -> "PBF.ParameterTangent":
        d = self.domain
        z = lambda: torch.zeros(d.configuration.particle_count, dtype=torch.float32, device=d.device)  # noqa: E731
        return PBF.ParameterTangent(ParticleParameterTangent(z(), z(), z(), z()), *(z() for _ in range(6)))

    def allocate_state_adjoint(self) -> "PBF.StateAdjoint":
        d = self.domain
        return PBF.StateAdjoint(ParticleStateAdjoint(d.allocate_vector_adjoint_field(), d.allocate_vector_adjoint_field()))

    def allocate_control_adjoint(self) -> ControlAdjoint:
        return ControlAdjoint(self.domain.allocate_vector_adjoint_field())

    def allocate_parameter_adjoint(self) -> "PBF.ParameterAdjoint":
        d = self.domain
        z = d.allocate_scalar_adjoint_field
        return PBF.ParameterAdjoint(ParticleParameterAdjoint(z(), z(), z(), z()), *(z() for _ in range(6)))

    def copy_state_tangent(self, source: "PBF.StateTangent", destination: "PBF.StateTangent") -> None:
        destination.particles.positions.copy_(source.particles.positions)
        destination.particles.velocities.copy_(source.particles.velocities)

    def copy_state_adjoint(self, source: "PBF.StateAdjoint", destination: "PBF.StateAdjoint") -> None:
        destination.particles.positions.copy_(source.particles.positions)
        destination.particles.velocities.copy_(source.particles.velocities)

    def accumulate_state_adjoint(self, source: "PBF.StateAdjoint", destination: "PBF.StateAdjoint") -> None:
        destination.particles.positions.add_(source.particles.positions)
        destination.particles.velocities.add_(source.particles.velocities)

    @_stream_ordered
    def jvp_step(self, state, control, parameters, next_state, cache, state_tangent, control_tangent,
                 parameter_tangent, next_state_tangent) -> None:
        """PBF::jvp_step (pbf/pbf.cpp:193-215) of the step recorded in `cache`
        (state, control and next_state are those of that forward step)."""
        _capi.call(
            "phx_pbf_jvp_step",
            self._handle,
            ctypes.byref(parameters.to_c()),
            cache._handle,
            ctypes.byref(state_tangent.to_c()),
            ctypes.byref(control_tangent.to_c()),
            ctypes.byref(parameter_tangent.to_c()),
            ctypes.byref(next_state_tangent.to_c()),
        )

    @_stream_ordered
    def vjp_step(self, state, control, parameters, next_state, cache, next_state_adjoint, previous_state_adjoint,
                 control_adjoint, parameter_adjoint) -> None:
        """PBF::vjp_step (pbf/pbf.cpp:217-258); every output is accumulated."""
        _capi.call(
            "phx_pbf_vjp_step",
            self._handle,
            ctypes.byref(parameters.to_c()),
            cache._handle,
            ctypes.byref(next_state_adjoint.to_c()),
            ctypes.byref(previous_state_adjoint.to_c()),
            ctypes.byref(control_adjoint.to_c()),
            ctypes.byref(parameter_adjoint.to_c()),
        )

    def set_profiling(self, enabled: bool) -> None:
        _capi.call("phx_pbf_set_profiling", self._handle, 1 if enabled else 0)

    def read_profile(self) -> list:
        return _read_profile("phx_pbf_read_profile", self._handle)

    @_stream_ordered
    def forward_step(self, state, control, parameters, next_state, cache) -> None:
        s = state.to_c()
        n = next_state.to_c()
        _capi.call(
            "phx_pbf_forward_step",
            self._handle,
            ctypes.byref(s),
            ctypes.byref(control.to_c()),
            ctypes.byref(parameters.to_c()),
            ctypes.byref(n),
            cache._handle,
        )
        next_state.particles.step_index = int(n.particles.step_index)


# ------------------------------------------------------------ neighborhood
def neighborhood_workspace_size(configuration: DomainConfiguration) -> int:
    size = ctypes.c_size_t()
    _capi.call("phx_neighborhood_workspace_size", ctypes.byref(configuration.to_c()), ctypes.byref(size))
    return int(size.value)


def density_jvp(
    support_radius: float,
    pbf_kernel: bool,
    topology: VectorField,
    positions: VectorField,
    position_tangent: VectorField,
    parameters: ParticleParameters,
    parameter_tangent: ParticleParameterTangent,
    view: _capi.NeighborhoodView,
    density_tangent: torch.Tensor,
    boundary: Optional[DeviceBoundary] = None,
    boundary_time: float = 0.0,
) -> None:
    """density::sph_jvp / pbf_jvp (density/density.cpp, density.cu:42-63):
    overwrites `density_tangent` (FP32)."""
    dev = density_tangent.device
    bview = boundary.view(boundary_time) if boundary is not None else _capi.BoundaryView()
    _capi.call("phx_density_jvp", _stream_handle(dev), int(view.particle_count), float(support_radius),
               1 if pbf_kernel else 0, topology.view(), positions.view(), position_tangent.view(), parameters.to_c(),
               parameter_tangent.to_c(), view, bview, _ptr(density_tangent))


def density_vjp(
    support_radius: float,
    pbf_kernel: bool,
    topology: VectorField,
    positions: VectorField,
    parameters: ParticleParameters,
    view: _capi.NeighborhoodView,
    density_adjoint: torch.Tensor,
    position_adjoint: VectorAdjointField,
    parameter_adjoint: ParticleParameterAdjoint,
    boundary: Optional[DeviceBoundary] = None,
    boundary_time: float = 0.0,
) -> None:
    """density::sph_vjp / pbf_vjp (density.cu:65-98): accumulates (+=) the
    FP64 adjoints of positions, masses and rest densities."""
    dev = density_adjoint.device
    bview = boundary.view(boundary_time) if boundary is not None else _capi.BoundaryView()
    _capi.call("phx_density_vjp", _stream_handle(dev), int(view.particle_count), float(support_radius),
               1 if pbf_kernel else 0, topology.view(), positions.view(), parameters.to_c(), view, bview,
               _ptr(density_adjoint), position_adjoint.view(), parameter_adjoint.to_c())


def write_vectors(positions: VectorField, velocities: VectorField) -> Tuple[torch.Tensor, torch.Tensor]:
    """spectra_cuda::write_vectors (examples/fluids/liquid/pbf-dam-break/
    spectra/kernels.cu:8-18): interleaved (N, 3) float32 copies of the SoA
    positions and velocities, for a viewer."""
    n = positions.x.numel()
    dev = positions.x.device
    _require_gpu(dev)
    pos = torch.empty((n, 3), dtype=torch.float32, device=dev)
    vel = torch.empty((n, 3), dtype=torch.float32, device=dev)
    _capi.call("phx_write_vectors", _stream_handle(dev), n, positions.view(), velocities.view(), _ptr(pos), _ptr(vel))
    return pos, vel


class DeviceBoundary:
    """DeviceBoundary (domain/domain.ixx:124): the configuration's boundary
    samples as device arrays in original order (positions and velocities
    (3, B), volumes (B,), FP32)."""

    def __init__(self, configuration: DomainConfiguration, device="cuda"):
        dev = torch.device(device)
        _require_gpu(dev)
        samples = list(configuration.boundary_particles)
        b = len(samples)
        pos = np.zeros((3, b), np.float32)
        vel = np.zeros((3, b), np.float32)
        vol = np.zeros(b, np.float32)
        for k, (x, v, volume) in enumerate(samples):
            pos[:, k] = x
            vel[:, k] = v
            vol[k] = volume
        self.count = b
        self.positions = torch.from_numpy(pos).to(dev)
        self.velocities = torch.from_numpy(vel).to(dev)
        self.volumes = torch.from_numpy(vol).to(dev)

    def view(self, time: float) -> _capi.BoundaryView:
        """BoundaryView at `time` (neighborhood/interop.h:29-41)."""
        if self.count == 0:
            return _capi.BoundaryView()
        return _capi.BoundaryView(
            _ptr(self.positions[0]), _ptr(self.positions[1]), _ptr(self.positions[2]), _ptr(self.velocities[0]),
            _ptr(self.velocities[1]), _ptr(self.velocities[2]), _ptr(self.volumes), self.count, float(time))


class Neighborhood(NamedTuple):
    """Neighborhood (neighborhood/neighborhood.ixx:12-23) in the reference
    layout, with its C view."""

    keys: torch.Tensor
    indices: torch.Tensor
    offsets: torch.Tensor
    view: _capi.NeighborhoodView
    boundary_keys: torch.Tensor
    boundary_indices: torch.Tensor
    boundary_offsets: torch.Tensor
    boundary_time: float


def build_neighborhood(configuration: DomainConfiguration, step_index: int, positions: VectorField,
                       device="cuda") -> Neighborhood:
    """NeighborhoodSearch::build in the reference layout: sorted keys int64[N],
    sorted particle indices int32[N], cell offsets int32[C+1], the same three
    for the boundary samples (B, B, C+1) and the boundary time step * dt."""
    dev = torch.device(device)
    _require_gpu(dev)
    n = configuration.particle_count
    b = len(configuration.boundary_particles)
    cells = math.prod(configuration.grid_resolution())
    workspace = torch.empty(neighborhood_workspace_size(configuration), dtype=torch.uint8, device=dev)
    keys = torch.empty(n, dtype=torch.int64, device=dev)
    idx = torch.empty(n, dtype=torch.int32, device=dev)
    off = torch.empty(cells + 1, dtype=torch.int32, device=dev)
    bkeys = torch.empty(b, dtype=torch.int64, device=dev)
    bidx = torch.empty(b, dtype=torch.int32, device=dev)
    boff = torch.empty(cells + 1, dtype=torch.int32, device=dev)
    view = _capi.NeighborhoodView()
    _capi.call(
        "phx_build_neighborhood",
        _stream_handle(dev),
        ctypes.byref(configuration.to_c()),
        int(step_index),
        positions.view(),
        _ptr(workspace),
        workspace.numel(),
        _ptr(keys),
        _ptr(idx),
        _ptr(off),
        _ptr(bkeys) if b else None,
        _ptr(bidx) if b else None,
        _ptr(boff),
        ctypes.byref(view),
    )
    view._keep = (workspace, keys, idx, off, bkeys, bidx, boff)
    time = float(np.float32(step_index) * np.float32(configuration.time_step))
    return Neighborhood(keys, idx, off, view, bkeys, bidx, boff, time)


def density_forward(
    support_radius: float,
    pbf_kernel: bool,
    topology: VectorField,
    positions: VectorField,
    parameters: ParticleParameters,
    view: _capi.NeighborhoodView,
    densities: torch.Tensor,
    boundary: Optional[DeviceBoundary] = None,
    boundary_time: float = 0.0,
) -> None:
    """density::sph_forward / pbf_forward (density/density.cpp:12-34), with
    the boundary samples at `boundary_time` when given."""
    dev = densities.device
    bview = boundary.view(boundary_time) if boundary is not None else _capi.BoundaryView()
    _capi.call(
        "phx_density_forward",
        _stream_handle(dev),
        int(view.particle_count),
        float(support_radius),
        1 if pbf_kernel else 0,
        topology.view(),
        positions.view(),
        parameters.to_c(),
        view,
        bview,
        _ptr(densities),
    )
