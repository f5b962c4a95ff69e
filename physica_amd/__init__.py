"""physica_amd: MI355X (gfx950) particle-liquid engine.

Drop-in for the liquid path of Xayah-Graphics/physica (WCSPH, PCISPH, IISPH,
DFSPH, PBF; fluids/liquid/particle/). The solvers run hand-written HIP
kernels in libphysica_amd.so behind the C ABI declared in
include/physica_amd/physica.h; this package is a thin harness over it.
"""
from .particle import (  # noqa: F401
    DFSPH,
    PBF,
    SPH,
    BoxBoundary,
    Control,
    DeviceBoundary,
    Domain,
    DomainConfiguration,
    ExecutionMode,
    Neighborhood,
    ParticleParameters,
    ParticleState,
    VectorField,
    build_neighborhood,
    density_forward,
    density_jvp,
    density_vjp,
    neighborhood_workspace_size,
    write_vectors,
)
from ._capi import LIBRARY_PATH, PhxError, library  # noqa: F401

__version__ = "0.1.0"
