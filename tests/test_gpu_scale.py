"""The IISPH / WCSPH scale test at its configured size and settings: two
falling 128^3 cubes (4,194,304 particles, spacing 0.01, h 0.02) in a
4 x 4 x 2 m basin. IISPH: relaxed Jacobi 0.5 to a 0.1 % mean compression
error (FP64 residual, at most 100 iterations), dt 1e-3. WCSPH: Tait equation
with c = sqrt(7), gamma = 7 (stiffness rho0 c^2 / gamma = 1000), no viscosity
or surface tension, dt 2e-4 (sph/sph.cpp:162-249, :364-476).

At full size a few steps are checked through properties that hold at any
size: finite state, every particle inside the collision box, the grid's
sorted indices a permutation with cell offsets that count every particle,
the median density within 50 % of rest, and IISPH's iteration counts within
the cap. The same scenes at 16^3 particles per cube, with the configured
parameters unchanged, are compared with the oracle: IISPH bit for bit
(iteration counts included), WCSPH within the propagated ulp bound of the
two platforms' powf at gamma = 7."""
import numpy as np
import pytest
import torch

import oracle
from physica_amd import scenes
from physica_amd.particle import IISPH, SPH, WCSPH
from helpers import bits_equal, max_rel, upload_scalar

pytestmark = pytest.mark.gpu

TOLERANCE = 1e-3
CAP = 100


def solver_for(sc, method, device):
    cfg = (IISPH.Configuration(density_error_tolerance=TOLERANCE, max_iterations=CAP) if method == "iisph"
           else WCSPH.Configuration())
    solver = SPH(sc.domain, cfg, device=device)
    state, nxt = solver.allocate_state(), solver.allocate_state()
    control, params, cache = solver.allocate_control(), solver.allocate_parameters(), solver.allocate_step_cache()
    state.particles.positions.upload(sc.positions)
    state.particles.velocities.upload(sc.velocities)
    for name in ("masses", "rest_densities", "viscosities", "surface_tensions"):
        upload_scalar(getattr(params.particles, name), sc.parameters[name])
    if method == "iisph":
        upload_scalar(params.method.jacobi_relaxation, sc.parameters["relaxation"])
    else:
        for name in ("speed_of_sound", "tait_exponent", "boundary_surface_tension"):
            upload_scalar(getattr(params.method, name), sc.parameters[name])
    return solver, state, nxt, control, params, cache


@pytest.mark.parametrize("method", ["iisph", "wcsph"])
def test_configured_scale_scene(gpu_device, method):
    sc = scenes.two_cubes(method)
    n = sc.count
    assert n == 4_194_304
    print(f"scale test {method} n={n}", flush=True)
    solver, state, nxt, control, params, cache = solver_for(sc, method, gpu_device)
    assert sc.domain.time_step == pytest.approx(1e-3 if method == "iisph" else 2e-4)
    if method == "wcsph":
        assert np.all(sc.parameters["tait_exponent"] == 7.0)
        assert np.allclose(1000.0 * sc.parameters["speed_of_sound"].astype(np.float64) ** 2 / 7.0, 1000.0,
                           rtol=1e-6)
        assert not sc.parameters["viscosities"].any() and not sc.parameters["surface_tensions"].any()
    else:
        assert np.all(sc.parameters["relaxation"] == np.float32(0.5))
    states = [state, nxt]
    rp = sc.domain.particle_radius
    lo = np.array(sc.domain.boundary.minimum, np.float32) + np.float32(rp)
    hi = np.array(sc.domain.boundary.maximum, np.float32) - np.float32(rp)
    iterations = []
    for k in range(4):
        solver.forward_step(states[k % 2], control, params, states[(k + 1) % 2], cache)
        if method == "iisph":
            iterations.append(solver.last_iterations()[0])
    torch.cuda.synchronize()
    out = states[0]
    x = out.particles.positions.numpy()
    v = out.particles.velocities.numpy()
    assert out.particles.step_index == 4
    assert np.isfinite(x).all() and np.isfinite(v).all()
    assert (x >= lo[:, None]).all() and (x <= hi[:, None]).all()
    # four steps of free fall from rest, no wall reached: the pair forces are
    # antisymmetric (equal masses), so the mean velocity is that of free fall
    fall = -9.81 * 4 * sc.domain.time_step
    mean_v = v.astype(np.float64).mean(axis=1)
    assert abs(mean_v[1] - fall) <= 1e-2 * abs(fall), mean_v
    assert abs(mean_v[0]) <= 1e-2 * abs(fall) and abs(mean_v[2]) <= 1e-2 * abs(fall), mean_v
    keys, idx, off = cache.neighborhood()
    order = idx.cpu().numpy()
    assert np.array_equal(np.sort(order), np.arange(n, dtype=order.dtype))
    offsets = off.cpu().numpy().astype(np.int64)
    assert offsets[0] == 0 and offsets[-1] == n and np.all(np.diff(offsets) >= 0)
    k = keys.cpu().numpy()
    assert np.all(np.diff(k) >= 0)
    rho = cache.densities().cpu().numpy()
    assert np.isfinite(rho).all()
    assert abs(np.median(rho) - 1000.0) <= 500.0, float(np.median(rho))
    if method == "iisph":
        print("iterations", iterations, flush=True)
        assert all(1 <= i <= CAP for i in iterations)
    else:
        p = cache.pressures().cpu().numpy()
        assert np.isfinite(p).all()


def test_reduced_iisph_configured_parameters_bit_exact(gpu_device):
    """Two 16^3 cubes, every configured IISPH setting: states, densities and
    iteration counts equal the oracle's over five steps."""
    sc = scenes.two_cubes("iisph", shape=(16, 16, 16))
    print(f"reduced iisph n={sc.count}", flush=True)
    solver, state, nxt, control, params, cache = solver_for(sc, "iisph", gpu_device)
    d = oracle.make_domain(sc.domain)
    x, v = sc.positions, sc.velocities
    for step in range(5):
        solver.forward_step(state, control, params, nxt, cache)
        torch.cuda.synchronize()
        exp = oracle.pressure_solver_step_tolerance(d, step, x, v, np.zeros_like(x), sc.parameters, "iisph",
                                                    TOLERANCE, CAP)
        assert solver.last_iterations()[0] == exp["iterations"], step
        got = {"positions": nxt.particles.positions.numpy(), "velocities": nxt.particles.velocities.numpy()}
        for name in ("positions", "velocities"):
            assert bits_equal(got[name], exp[name]), (step, name)
        assert bits_equal(cache.densities().cpu().numpy(), exp["densities"]), step
        x, v = got["positions"], got["velocities"]
        solver.copy_state(nxt, state)


def test_reduced_wcsph_configured_parameters(gpu_device):
    """Two 16^3 cubes, every configured WCSPH setting (gamma = 7): densities
    bit for bit; pressures within two ulps of (rho/rho0)^7 times the
    stiffness plus four ulps of the pressure (HIP's and the host's powf, each
    within an ulp); positions and velocities of the step within the
    propagation of that bound."""
    sc = scenes.two_cubes("wcsph", shape=(16, 16, 16))
    print(f"reduced wcsph n={sc.count}", flush=True)
    solver, state, nxt, control, params, cache = solver_for(sc, "wcsph", gpu_device)
    d = oracle.make_domain(sc.domain)
    x, v = sc.positions, sc.velocities
    for step in range(3):
        solver.forward_step(state, control, params, nxt, cache)
        torch.cuda.synchronize()
        exp = oracle.wcsph_step(d, step, x, v, np.zeros_like(x), sc.parameters)
        assert bits_equal(cache.densities().cpu().numpy(), exp["densities"]), step
        prm = sc.parameters
        rho = exp["densities"].astype(np.float64)
        power = (rho / prm["rest_densities"]) ** prm["tait_exponent"].astype(np.float64)
        b = prm["rest_densities"].astype(np.float64) * prm["speed_of_sound"] ** 2 / prm["tait_exponent"]
        bound = 2 * np.spacing(power.astype(np.float32)) * b + 4 * np.spacing(np.abs(exp["pressures"]))
        p_err = np.abs(cache.pressures().cpu().numpy().astype(np.float64) - exp["pressures"])
        assert (p_err <= bound).all(), (step, float((p_err / bound).max()))
        got_x = nxt.particles.positions.numpy()
        got_v = nxt.particles.velocities.numpy()
        span = np.abs(exp["positions"] - x).max()
        ulp = float(np.spacing(np.float32(np.abs(x).max())))
        assert np.abs(got_x - exp["positions"]).max() <= 1e-6 * span + ulp, step
        assert max_rel(got_v, exp["velocities"]) < 1e-4, step
        # continue both from the GPU state so each step is compared alone
        x, v = got_x, got_v
        solver.copy_state(nxt, state)
