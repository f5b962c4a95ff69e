"""Seeded synthetic scenes for the liquid path (host-side numpy).

Random numbers come from a counter-based SplitMix64 stream: value k of a
stream is splitmix64(seed + (k + 1) * golden), so every draw is a pure function
of (seed, k) and generation vectorises. A uniform float is
(u >> 40) * 2^-24 in [0, 1), exact in FP32.

Scenes follow the reference's example layout (examples/fluids/liquid/
pbf-dam-break/simulation.ixx:83-99): particle index (z * ny + y) * nx + x,
position origin + index * spacing in FP32, masses rho0 * spacing^3.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Tuple

import numpy as np

from .particle import BoxBoundary, DomainConfiguration

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """count 64-bit values of the SplitMix64 stream `seed`, starting at draw `offset`."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Uniform FP32 values in [0, 1)."""
    u = splitmix64(seed, count, offset) >> np.uint64(40)
    return (u.astype(np.float64) * (2.0**-24)).astype(np.float32)


def lattice_block(
    shape: Tuple[int, int, int],
    spacing: float,
    origin: Tuple[float, float, float],
    jitter: float,
    seed: int,
) -> np.ndarray:
    """(3, nx*ny*nz) FP32 positions of a jittered cubic lattice, x fastest."""
    nx, ny, nz = shape
    n = nx * ny * nz
    idx = np.arange(n, dtype=np.int64)
    coords = (idx % nx, (idx // nx) % ny, idx // (nx * ny))
    s = np.float32(spacing)
    out = np.empty((3, n), dtype=np.float32)
    for a in range(3):
        out[a] = np.float32(origin[a]) + coords[a].astype(np.float32) * s
        if jitter:
            u = uniform(seed, n, offset=a * n)
            out[a] += (np.float32(2.0) * u - np.float32(1.0)) * np.float32(jitter)
    return out


@dataclass
class Scene:
    """A domain plus host arrays for the state and per-particle parameters."""

    name: str
    domain: DomainConfiguration
    positions: np.ndarray
    velocities: np.ndarray
    parameters: Dict[str, np.ndarray] = field(default_factory=dict)
    solver: Dict[str, float] = field(default_factory=dict)

    @property
    def count(self) -> int:
        return self.domain.particle_count


def dfsph_dam_break(shape=(128, 64, 128), spacing=0.01, seed=20260002, box=None, jitter=None) -> Scene:
    """DFSPH 3D dam break: block in the x-corner of a box twice its x extent.

    Defaults: 128x64x128 = 1,048,576 particles, spacing 0.01, jitter
    U(+-0.001), h = 2 spacing (cubic kernel), r_p = spacing/2, dt = 1e-3,
    rest density 1000, viscosity 0.01, relaxation 0.5 for both phases, box
    2.56 x 1.28 x 1.28 m."""
    nx, ny, nz = shape
    s = float(spacing)
    h = 2.0 * s
    if box is None:
        box = (2.0 * nx * s, 2.0 * ny * s, nz * s)
    if jitter is None:
        jitter = 0.1 * s
    pos = lattice_block(shape, s, (0.5 * s, 0.5 * s, 0.5 * s), jitter, seed)
    n = pos.shape[1]
    domain = DomainConfiguration(
        particle_count=n,
        time_step=1.0e-3,
        support_radius=h,
        particle_radius=0.5 * s,
        gravity=(0.0, -9.81, 0.0),
        boundary=BoxBoundary((0.0, 0.0, 0.0), tuple(map(float, box)), (0.0, 0.0, 0.0), True),
    )
    rho0 = np.float32(1000.0)
    params = {
        "masses": np.full(n, rho0 * np.float32(s) ** 3, dtype=np.float32),
        "rest_densities": np.full(n, rho0, dtype=np.float32),
        "viscosities": np.full(n, 0.01, dtype=np.float32),
        "surface_tensions": np.zeros(n, dtype=np.float32),
        "divergence_relaxation": np.full(n, 0.5, dtype=np.float32),
        "density_relaxation": np.full(n, 0.5, dtype=np.float32),
    }
    return Scene(
        "dfsph_dam_break",
        domain,
        pos,
        np.zeros_like(pos),
        params,
        {"divergence_iterations": 4, "density_iterations": 6},
    )


def pbf_dam_break(shape=(32, 32, 32), spacing=0.05, seed=20260001, box=(3.2, 2.0, 1.6), jitter=None) -> Scene:
    """PBF dam break: 32^3 lattice at spacing 0.05 (jitter U(+-0.005)) in a
    3.2 x 2.0 x 1.6 m box; h = 0.1 (poly6 / spiky), r_p = spacing/2,
    dt = 1/120 (one substep of 1/60 with 2 substeps), rest density 1000,
    4 constraint iterations. The example's relaxation 1000 at spacing 0.025
    (simulation.ixx:69-70) is rescaled by (0.025/spacing)^2; artificial
    pressure strength 0, exponent 4, radius 0.3h; XSPH and vorticity
    confinement 0 (simulation.ixx:71-79)."""
    s = float(spacing)
    h = 2.0 * s
    if jitter is None:
        jitter = 0.1 * s
    pos = lattice_block(shape, s, (0.5 * s, 0.5 * s, 0.5 * s), jitter, seed)
    n = pos.shape[1]
    domain = DomainConfiguration(
        particle_count=n,
        time_step=1.0 / 120.0,
        support_radius=h,
        particle_radius=0.5 * s,
        gravity=(0.0, -9.81, 0.0),
        boundary=BoxBoundary((0.0, 0.0, 0.0), tuple(map(float, box)), (0.0, 0.0, 0.0), True),
    )
    rho0 = np.float32(1000.0)
    params = {
        "masses": np.full(n, rho0 * np.float32(s) ** 3, dtype=np.float32),
        "rest_densities": np.full(n, rho0, dtype=np.float32),
        "viscosities": np.zeros(n, dtype=np.float32),
        "surface_tensions": np.zeros(n, dtype=np.float32),
        "relaxation": np.full(n, 1000.0 * (0.025 / s) ** 2, dtype=np.float32),
        "artificial_pressure_strength": np.zeros(n, dtype=np.float32),
        "artificial_pressure_exponent": np.full(n, 4.0, dtype=np.float32),
        "artificial_pressure_radius": np.full(n, 0.3 * h, dtype=np.float32),
        "xsph_viscosity": np.zeros(n, dtype=np.float32),
        "vorticity_confinement": np.zeros(n, dtype=np.float32),
    }
    return Scene("pbf_dam_break", domain, pos, np.zeros_like(pos), params, {"pressure_iterations": 4})


def batched_pbf_scenes(count=64, seed=20260005, low=4096, high=200000, spacing=0.02) -> List[Scene]:
    """The uneven PBF batch: `count` independent dam breaks whose particle
    counts are N_k = floor(exp(U(ln low, ln high))), U the k-th draw of the
    SplitMix64 stream `seed`. Scene k is a cube lattice of side
    ceil(N_k^(1/3)) at `spacing` (jitter stream seed + 1 + k), truncated to
    its first N_k particles, in a box twice the block's extent, with h =
    2 spacing and pbf_dam_break's parameters."""
    u = uniform(seed, count).astype(np.float64)
    counts = np.floor(np.exp(np.log(low) + u * (np.log(high) - np.log(low)))).astype(np.int64)
    out = []
    for k, n in enumerate(counts.tolist()):
        side = max(1, int(round(n ** (1.0 / 3.0))))
        while side ** 3 < n:
            side += 1
        while side > 1 and (side - 1) ** 3 >= n:
            side -= 1
        extent = 2.0 * side * spacing
        sc = pbf_dam_break(shape=(side, side, side), spacing=spacing, seed=seed + 1 + k, box=(extent,) * 3)
        sc.name = f"pbf_batch_{k}"
        sc.positions = np.ascontiguousarray(sc.positions[:, :n])
        sc.velocities = np.ascontiguousarray(sc.velocities[:, :n])
        sc.parameters = {name: np.ascontiguousarray(v[:n]) for name, v in sc.parameters.items()}
        sc.domain.particle_count = int(n)
        out.append(sc)
    return out


def two_cubes(method="iisph", shape=(128, 128, 128), spacing=0.01, seed=20260003, box=(4.0, 4.0, 2.0),
              jitter=None) -> Scene:
    """Two falling cubes of shape[0] x shape[1] x shape[2] particles (spacing
    0.01, jitter U(+-0.001)) into an empty basin, for IISPH, PCISPH and WCSPH.

    The cubes' lower corners are at x in U[0.2, 0.6], z in U[0.2, 0.5]
    (seeded, seed + 100) and y = 1.2 and 2.6. The basin is 4.0 x 4.0 x 2.0 m:
    the upper cube reaches y = 3.88, so a 3 m tall basin would clip it.
    h = 2 spacing, r_p = spacing / 2, rest density 1000, viscosity 0.
    IISPH / PCISPH: dt = 1e-3, relaxation 0.5. WCSPH: dt = 2e-4, Tait exponent
    7, speed of sound sqrt(7) m/s (rho0 c^2 / gamma = 1000), no surface
    tension."""
    s = float(spacing)
    h = 2.0 * s
    if jitter is None:
        jitter = 0.1 * s
    offsets = uniform(seed + 100, 4)
    corners = [
        (0.2 + 0.4 * float(offsets[0]), 1.2, 0.2 + 0.3 * float(offsets[1])),
        (0.2 + 0.4 * float(offsets[2]), 2.6, 0.2 + 0.3 * float(offsets[3])),
    ]
    blocks = []
    n_block = shape[0] * shape[1] * shape[2]
    for k, corner in enumerate(corners):
        blocks.append(lattice_block(shape, s, tuple(c + 0.5 * s for c in corner), jitter, seed + k * 7919))
    pos = np.concatenate(blocks, axis=1)
    n = pos.shape[1]
    assert n == 2 * n_block
    wcsph = method == "wcsph"
    domain = DomainConfiguration(
        particle_count=n,
        time_step=2.0e-4 if wcsph else 1.0e-3,
        support_radius=h,
        particle_radius=0.5 * s,
        gravity=(0.0, -9.81, 0.0),
        boundary=BoxBoundary((0.0, 0.0, 0.0), tuple(map(float, box)), (0.0, 0.0, 0.0), True),
    )
    rho0 = np.float32(1000.0)
    params = {
        "masses": np.full(n, rho0 * np.float32(s) ** 3, dtype=np.float32),
        "rest_densities": np.full(n, rho0, dtype=np.float32),
        "viscosities": np.zeros(n, dtype=np.float32),
        "surface_tensions": np.zeros(n, dtype=np.float32),
    }
    if wcsph:
        params["speed_of_sound"] = np.full(n, np.sqrt(7.0), dtype=np.float32)
        params["tait_exponent"] = np.full(n, 7.0, dtype=np.float32)
        params["boundary_surface_tension"] = np.zeros(n, dtype=np.float32)
        solver = {}
    else:
        params["relaxation"] = np.full(n, 0.5, dtype=np.float32)
        solver = {"pressure_iterations": 6}
    return Scene(f"two_cubes_{method}", domain, pos, np.zeros_like(pos), params, solver)
