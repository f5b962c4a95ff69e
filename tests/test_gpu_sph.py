"""SPH<PCISPH>, SPH<IISPH> and SPH<WCSPH>::forward_step on the GPU against
the FP32 oracle: bit-exact, except WCSPH's Tait pressure at a non-unit
exponent, which goes through two different powf implementations (tolerance
stated in the test)."""
import numpy as np
import pytest
import torch

import oracle
from physica_amd import scenes
from helpers import bits_equal, max_rel, sph_setup

pytestmark = pytest.mark.gpu


def scene_for(method, shape=(16, 12, 16), seed=7):
    sc = scenes.two_cubes(method, shape=shape, seed=seed)
    n = sc.count
    rng = np.random.default_rng(seed)
    sc.velocities = rng.normal(0, 0.3, size=(3, n)).astype(np.float32)
    sc.parameters["viscosities"] = rng.uniform(0.0, 0.05, n).astype(np.float32)
    if method == "wcsph":
        sc.parameters["surface_tensions"] = rng.uniform(0.0, 2.0, n).astype(np.float32)
        sc.parameters["speed_of_sound"] = rng.uniform(2.0, 3.0, n).astype(np.float32)
        # 0.97 rho0 keeps the Tait pressures of the lattice positive
        sc.parameters["rest_densities"] = np.full(n, 970.0, dtype=np.float32)
    else:
        sc.parameters["relaxation"] = rng.uniform(0.3, 0.7, n).astype(np.float32)
    return sc


def gpu_step(solver, state, nxt, control, params, cache):
    solver.forward_step(state, control, params, nxt, cache)
    torch.cuda.synchronize()
    return {"positions": nxt.particles.positions.numpy(), "velocities": nxt.particles.velocities.numpy(),
            "densities": cache.densities().cpu().numpy(), "pressures": cache.pressures().cpu().numpy()}


def oracle_step(sc, method, x, v, step, iterations=6):
    d = oracle.make_domain(sc.domain)
    if method == "wcsph":
        return oracle.wcsph_step(d, step, x, v, np.zeros_like(x), sc.parameters)
    return oracle.pressure_solver_step(d, step, x, v, np.zeros_like(x), sc.parameters, method, iterations)


@pytest.mark.parametrize("method,iterations", [("pcisph", 6), ("pcisph", 0), ("pcisph", 1), ("iisph", 6),
                                               ("iisph", 3)])
def test_pressure_solver_bit_exact(gpu_device, method, iterations):
    print(f"sph {method} iterations={iterations}", flush=True)
    sc = scene_for(method)
    solver, state, nxt, control, params, cache = sph_setup(sc, gpu_device, method, iterations)
    state.particles.step_index = 3
    got = gpu_step(solver, state, nxt, control, params, cache)
    exp = oracle_step(sc, method, sc.positions, sc.velocities, 3, iterations)
    assert nxt.particles.step_index == 4
    for name in ("positions", "velocities", "densities", "pressures"):
        assert bits_equal(got[name], exp[name]), name
    assert bits_equal(cache.non_pressure_accelerations().numpy(), exp["non_pressure_accelerations"])
    if iterations > 0:
        assert (exp["pressures"] > 0).any()


@pytest.mark.parametrize("method", ["pcisph", "iisph"])
def test_pressure_solver_rollout_bit_exact(gpu_device, method):
    """10 steps of a falling pair of 12x12x12 cubes, each step against the
    oracle fed the same input."""
    print(f"sph rollout {method}", flush=True)
    sc = scene_for(method, shape=(12, 12, 12), seed=11)
    solver, state, nxt, control, params, cache = sph_setup(sc, gpu_device, method)
    x, v = sc.positions, sc.velocities
    for step in range(10):
        got = gpu_step(solver, state, nxt, control, params, cache)
        exp = oracle_step(sc, method, x, v, step)
        for name in ("positions", "velocities"):
            assert bits_equal(got[name], exp[name]), (step, name)
        x, v = got["positions"], got["velocities"]
        solver.copy_state(nxt, state)


@pytest.mark.parametrize("method", ["pcisph", "iisph"])
def test_per_cell_skins_stay_exact_when_particles_outrun_their_bounds(gpu_device, method):
    """The forward PCISPH/IISPH step uses per-cell skins from its second step
    on (as DFSPH's). Step 1 starts at rest, so the bounds are tiny; step 2
    gets a state in which a seeded tenth of the particles move at up to
    3 m/s, far past them. Both steps equal the FP32 oracle bit for bit."""
    print(f"sph per-cell skins with outrunning particles {method}", flush=True)
    sc = scene_for(method, shape=(12, 12, 12), seed=17)
    n = sc.count
    solver, state, nxt, control, params, cache = sph_setup(sc, gpu_device, method)
    x, v = sc.positions, np.zeros((3, n), np.float32)
    state.particles.velocities.upload(v)
    got = gpu_step(solver, state, nxt, control, params, cache)
    exp = oracle_step(sc, method, x, v, 0)
    for name in ("positions", "velocities"):
        assert bits_equal(got[name], exp[name]), ("step 1", name)
    rng = np.random.default_rng(18)
    v = got["velocities"].copy()
    kicked = rng.choice(n, size=n // 10, replace=False)
    v[:, kicked] = rng.uniform(-3.0, 3.0, size=(3, kicked.size)).astype(np.float32)
    x = got["positions"]
    solver.copy_state(nxt, state)
    state.particles.velocities.upload(v)
    got = gpu_step(solver, state, nxt, control, params, cache)
    exp = oracle_step(sc, method, x, v, 1)
    for name in ("positions", "velocities"):
        assert bits_equal(got[name], exp[name]), ("step 2", name)


def test_wcsph_unit_exponent_bit_exact(gpu_device):
    """Tait exponent 1: powf(x, 1) = x exactly, so the whole step is exact."""
    sc = scene_for("wcsph")
    sc.parameters["tait_exponent"] = np.ones(sc.count, dtype=np.float32)
    solver, state, nxt, control, params, cache = sph_setup(sc, gpu_device, "wcsph")
    got = gpu_step(solver, state, nxt, control, params, cache)
    exp = oracle_step(sc, "wcsph", sc.positions, sc.velocities, 0)
    for name in ("positions", "velocities", "densities", "pressures"):
        assert bits_equal(got[name], exp[name]), name
    assert (exp["pressures"] > 0).any() and (exp["pressures"] < 0).any()


def test_wcsph_tait_within_tolerance(gpu_device):
    """Exponent 7: densities bit-exact; pressures p = B ((rho/rho0)^g - 1),
    B = rho0 c^2 / g, within 2 ulp of the power times B plus 4 ulp of p
    (powf of HIP vs the host C library, each within 1 ulp); positions within
    1e-6 of the step's displacement plus one FP32 ulp of the coordinate."""
    sc = scene_for("wcsph")
    solver, state, nxt, control, params, cache = sph_setup(sc, gpu_device, "wcsph")
    got = gpu_step(solver, state, nxt, control, params, cache)
    exp = oracle_step(sc, "wcsph", sc.positions, sc.velocities, 0)
    assert bits_equal(got["densities"], exp["densities"])
    prm = sc.parameters
    rho = exp["densities"].astype(np.float64)
    power = (rho / prm["rest_densities"]) ** prm["tait_exponent"].astype(np.float64)
    b = prm["rest_densities"].astype(np.float64) * prm["speed_of_sound"] ** 2 / prm["tait_exponent"]
    bound = 2 * np.spacing(power.astype(np.float32)) * b + 4 * np.spacing(np.abs(exp["pressures"]))
    p_err = np.abs(got["pressures"].astype(np.float64) - exp["pressures"])
    assert (p_err <= bound).all(), float((p_err / bound).max())
    identical = float(np.mean(got["pressures"].view(np.uint32) == exp["pressures"].view(np.uint32)))
    print(f"wcsph pressures bit-identical: {identical:.4f}", flush=True)
    span = np.abs(exp["positions"] - sc.positions).max()
    ulp = float(np.spacing(np.float32(np.abs(sc.positions).max())))
    assert np.abs(got["positions"] - exp["positions"]).max() <= 1e-6 * span + ulp
    assert max_rel(got["velocities"], exp["velocities"]) < 1e-4


def test_wcsph_rollout_unit_exponent_bit_exact(gpu_device):
    sc = scene_for("wcsph", shape=(12, 12, 12), seed=13)
    sc.parameters["tait_exponent"] = np.ones(sc.count, dtype=np.float32)
    solver, state, nxt, control, params, cache = sph_setup(sc, gpu_device, "wcsph")
    x, v = sc.positions, sc.velocities
    for step in range(10):
        got = gpu_step(solver, state, nxt, control, params, cache)
        exp = oracle_step(sc, "wcsph", x, v, step)
        for name in ("positions", "velocities"):
            assert bits_equal(got[name], exp[name]), (step, name)
        x, v = got["positions"], got["velocities"]
        solver.copy_state(nxt, state)


@pytest.mark.parametrize("method", ["pcisph", "iisph", "wcsph"])
def test_empty_domain(gpu_device, method):
    sc = scene_for(method, shape=(2, 2, 2))
    sc.domain.particle_count = 0
    sc.positions = sc.positions[:, :0]
    sc.velocities = sc.velocities[:, :0]
    sc.parameters = {k: v[:0] for k, v in sc.parameters.items()}
    solver, state, nxt, control, params, cache = sph_setup(sc, gpu_device, method)
    state.particles.step_index = 9
    solver.forward_step(state, control, params, nxt, cache)
    torch.cuda.synchronize()
    assert nxt.particles.step_index == 10


@pytest.mark.parametrize("method,tolerance,max_iterations", [("iisph", 1e-3, 100), ("pcisph", 1e-3, 100),
                                                             ("iisph", 1e-4, 5), ("pcisph", 1e-2, 100)])
def test_pressure_solver_tolerance_mode_bit_exact(gpu_device, method, tolerance, max_iterations):
    """Tolerance mode of PCISPH/IISPH (the IISPH scale configuration iterates
    to 0.1 % with an FP64 residual): iteration counts and states equal the
    oracle's over a rollout, a capped run stops at its cap."""
    print(f"sph tolerance {method} {tolerance} cap {max_iterations}", flush=True)
    from physica_amd.particle import IISPH, PCISPH, SPH
    sc = scene_for(method, shape=(12, 12, 12), seed=13)
    cls = IISPH if method == "iisph" else PCISPH
    solver, state, nxt, control, params, cache = sph_setup(sc, gpu_device, method)
    solver = SPH(sc.domain, cls.Configuration(density_error_tolerance=tolerance, max_iterations=max_iterations),
                 device=gpu_device)
    state = solver.allocate_state()
    nxt = solver.allocate_state()
    control = solver.allocate_control()
    cache = solver.allocate_step_cache()
    state.particles.positions.upload(sc.positions)
    state.particles.velocities.upload(sc.velocities)
    p2 = solver.allocate_parameters()
    for name in ("masses", "rest_densities", "viscosities", "surface_tensions"):
        getattr(p2.particles, name).copy_(getattr(params.particles, name))
    relax = p2.method.jacobi_relaxation if method == "iisph" else p2.method.pressure_relaxation
    relax.copy_(torch.from_numpy(sc.parameters["relaxation"]).to(gpu_device))
    d = oracle.make_domain(sc.domain)
    x, v = sc.positions, sc.velocities
    counts = []
    for step in range(4):
        solver.forward_step(state, control, p2, nxt, cache)
        torch.cuda.synchronize()
        exp = oracle.pressure_solver_step_tolerance(d, step, x, v, np.zeros_like(x), sc.parameters, method,
                                                    tolerance, max_iterations)
        assert solver.last_iterations()[0] == exp["iterations"], step
        got = {"positions": nxt.particles.positions.numpy(), "velocities": nxt.particles.velocities.numpy()}
        for name in ("positions", "velocities"):
            assert bits_equal(got[name], exp[name]), (step, name)
        assert bits_equal(cache.densities().cpu().numpy(), exp["densities"]), step
        counts.append(exp["iterations"])
        x, v = got["positions"], got["velocities"]
        solver.copy_state(nxt, state)
    print("iterations", counts, flush=True)
    assert max(counts) <= max_iterations
    if max_iterations == 5:
        assert 5 in counts
