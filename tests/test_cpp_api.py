"""The C++ API (include/physica_amd/particle.hpp) driving the reference's
PBF dam-break example (examples/fluids/liquid/pbf-dam-break) through the C ABI."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from helpers import bits_equal
from physica_amd.particle import BoxBoundary, DomainConfiguration

BINARY = Path(__file__).resolve().parent / "cpp" / "_bin" / "pbf_dam_break"


def test_example_is_built():
    assert BINARY.exists(), "run python -m physica_amd.build"


def example_scene():
    lattice = (16, 28, 12)
    n = lattice[0] * lattice[1] * lattice[2]
    idx = np.arange(n)
    x, y, z = idx % 16, (idx // 16) % 28, idx // (16 * 28)
    s = np.float32(0.025)
    pos = np.stack([np.float32(0.08) + c.astype(np.float32) * s for c in (x, y, z)]).astype(np.float32)
    domain = DomainConfiguration(n, 1.0 / 240.0, 0.05, 0.01, (0.0, -9.81, 0.0),
                                 BoxBoundary((0, 0, 0), (0.8, 1.1, 0.5), (0, 0, 0), True))
    params = {
        "masses": np.full(n, np.float32(1000.0) * s * s * s, np.float32),
        "rest_densities": np.full(n, 1000.0, np.float32),
        "relaxation": np.full(n, 1000.0, np.float32),
        "artificial_pressure_strength": np.zeros(n, np.float32),
        "artificial_pressure_exponent": np.full(n, 4.0, np.float32),
        "artificial_pressure_radius": np.full(n, np.float32(0.3) * np.float32(0.05), np.float32),
        "xsph_viscosity": np.zeros(n, np.float32),
        "vorticity_confinement": np.zeros(n, np.float32),
    }
    return domain, pos, params


@pytest.mark.gpu
def test_example_matches_oracle(tmp_path, gpu_device):
    steps = 4
    out = tmp_path / "positions.bin"
    run = subprocess.run([str(BINARY), str(steps), str(out)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    assert "PBF dam break with 5376 particles" in run.stdout
    got = np.fromfile(out, dtype=np.float32).reshape(3, -1)
    domain, x, params = example_scene()
    d = oracle.make_domain(domain)
    v = np.zeros_like(x)
    for step in range(steps):
        r = oracle.pbf_step(d, step, x, v, np.zeros_like(x), params, 5)
        x, v = r["positions"], r["velocities"]
    assert bits_equal(got, x)


COVERAGE = Path(__file__).resolve().parent / "cpp" / "_bin" / "api_coverage"
SPACING = np.float32(0.02)


def coverage_domain(n, dt):
    return DomainConfiguration(n, dt, float(np.float32(2.0) * SPACING), float(np.float32(0.5) * SPACING),
                               (0.0, -9.81, 0.0), BoxBoundary((0, 0, 0), (0.4, 0.4, 0.4), (0, 0, 0), True))


def water(n):
    f = np.float32
    return {"masses": np.full(n, f(1000.0) * SPACING * SPACING * SPACING, f),
            "rest_densities": np.full(n, 1000.0, f), "viscosities": np.full(n, 0.01, f),
            "surface_tensions": np.zeros(n, f)}


def test_api_coverage_is_built():
    assert COVERAGE.exists(), "run python -m physica_amd.build"


@pytest.mark.gpu
def test_api_coverage_matches_oracle(tmp_path, gpu_device):
    """Every solver of the C++ API: forward_step bit-exact against the
    oracle, jvp_step and vjp_step transposes of each other over every field
    the API passes (dot-product identity, FP32 JVP vs FP64 VJP), tolerance
    mode's iteration counts, NeighborhoodSearch + density::sph_forward with
    moving boundary samples bit-exact, and Domain::accumulate."""
    run = subprocess.run([str(COVERAGE), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stderr
    lines = dict(line.split(" ", 1) for line in run.stdout.strip().splitlines())
    n = 216
    x0 = np.fromfile(tmp_path / "initial_positions.bin", dtype=np.float32).reshape(3, n)
    v0 = np.fromfile(tmp_path / "initial_velocities.bin", dtype=np.float32).reshape(3, n)
    u = np.zeros_like(x0)
    z = np.zeros(n, np.float32)
    f = np.float32

    def got(name):
        return np.fromfile(tmp_path / f"{name}_positions.bin", dtype=np.float32).reshape(3, n)

    p = dict(water(n), divergence_relaxation=np.full(n, 0.5, f), density_relaxation=np.full(n, 0.5, f))
    exp = oracle.dfsph_step(oracle.make_domain(coverage_domain(n, 1e-3)), 0, x0, v0, u, p, z, z, 3, 4, False)
    assert bits_equal(got("dfsph"), exp["positions"])
    for method in ("pcisph", "iisph"):
        p = dict(water(n), relaxation=np.full(n, 0.5, f))
        exp = oracle.pressure_solver_step(oracle.make_domain(coverage_domain(n, 1e-3)), 0, x0, v0, u, p, method, 4)
        assert bits_equal(got(method), exp["positions"]), method
    p = dict(water(n), speed_of_sound=np.full(n, 5.0, f), tait_exponent=np.ones(n, f),
             boundary_surface_tension=np.full(n, 0.01, f))
    p["surface_tensions"] = np.full(n, 0.01, f)
    exp = oracle.wcsph_step(oracle.make_domain(coverage_domain(n, 2e-4)), 0, x0, v0, u, p)
    assert bits_equal(got("wcsph"), exp["positions"])
    p = dict(water(n), relaxation=np.full(n, 1e-3, f), artificial_pressure_strength=z,
             artificial_pressure_exponent=np.full(n, 4.0, f),
             artificial_pressure_radius=np.full(n, f(0.3) * f(2.0) * SPACING, f),
             xsph_viscosity=np.full(n, 0.01, f), vorticity_confinement=z)
    exp = oracle.pbf_step(oracle.make_domain(coverage_domain(n, 1.0 / 120.0)), 0, x0, v0, u, p, 3)
    assert bits_equal(got("pbf"), exp["positions"])

    for name in ("dfsph", "pcisph", "iisph", "wcsph", "pbf"):
        _, left, right = lines[name].split()
        left, right = float(left), float(right)
        assert left != 0.0
        assert abs(left - right) <= 2e-4 * (abs(left) + abs(right)), (name, left, right)

    p = dict(water(n), divergence_relaxation=np.full(n, 0.5, f), density_relaxation=np.full(n, 0.5, f))
    tol = oracle.dfsph_step_tolerance(oracle.make_domain(coverage_domain(n, 1e-3)), 0, x0, np.zeros_like(x0), u, p,
                                      z, z, 1e-3, 50, True, "f32")
    assert lines["tolerance"] == "iterations %d %d" % tuple(tol["iterations"])

    cfg = coverage_domain(n, 1e-3)
    k = np.arange(100)
    bpos = np.stack([f(0.08) + SPACING * (k // 10).astype(f), np.full(100, 0.035, f),
                     f(0.08) + SPACING * (k % 10).astype(f)]).astype(f)
    bvel = np.zeros((3, 100), f)
    bvel[2] = 0.1
    bvol = np.full(100, f(0.8) * SPACING * SPACING * SPACING, f)
    rho = np.fromfile(tmp_path / "density.bin", dtype=np.float32)
    w = water(n)
    expected = oracle.density(oracle.make_domain(cfg, (bpos, bvel, bvol)), 3, False, x0, x0, w["masses"],
                              w["rest_densities"])
    free = oracle.density(oracle.make_domain(cfg), 3, False, x0, x0, w["masses"], w["rest_densities"])
    assert np.sum(expected != free) > 10
    assert bits_equal(rho, expected)
    od = oracle.make_domain(cfg, (bpos, bvel, bvol))
    tan = np.fromfile(tmp_path / "density_position_tangent.bin", dtype=np.float32).reshape(3, n)
    got_t = np.fromfile(tmp_path / "density_tangent.bin", dtype=np.float32)
    assert bits_equal(got_t, oracle.density_jvp(od, 3, False, x0, x0, tan, w["masses"], z, w["rest_densities"], z))
    _, left, right = lines["density"].split()
    assert abs(float(left) - float(right)) <= 1e-5 * (abs(float(left)) + abs(float(right)))
    assert lines["accumulate"] == "ok"
