"""Pins of the oracle's boundary-particle terms (density.cu:34-37,
common.cu:38-43 and :185-190, with boundary_position of
neighborhood/device.cuh:16-22).

A boundary sample b of volume V_b at x_b(t) = x0_b + t v_b acts on fluid
particle i exactly as a fluid neighbour j at x_b(t) would, with

- density:   m_j = rho0 V_b (uniform rest density rho0),
- pressure:  m_j = rho0 V_b and p_j = 0 (only p_i / rho_i^2 enters),
- viscosity: mu_j = mu_i, m_j / rho_j = V_b, v_j = v_b.

So the oracle's sums over particles + samples must equal, in FP64 and up to
summation order, its sums over particles + the samples turned into such
particles; a wrong sign, a missing factor, an unmoved sample or a skipped
cell fails this. The density term is also linear in rho0_i, the particle's
own rest density (not a constant)."""
import numpy as np

import oracle
from physica_amd import scenes


def setup(moving):
    sc = scenes.dfsph_dam_break(shape=(6, 5, 6), seed=7, box=(0.2, 0.2, 0.2))
    n = sc.count
    rng = np.random.default_rng(11)
    x = sc.positions.astype(np.float64)
    # a floor of samples just below the block and a wall beside it
    s = 0.01
    g = np.arange(-1, 9) * s
    gx, gz = np.meshgrid(g + x[0].min(), g + x[2].min(), indexing="ij")
    floor = np.stack([gx.ravel(), np.full(gx.size, x[1].min() - 0.009), gz.ravel()])
    gy, gz2 = np.meshgrid(g + x[1].min(), g + x[2].min(), indexing="ij")
    wall = np.stack([np.full(gy.size, x[0].max() + 0.011), gy.ravel(), gz2.ravel()])
    x0 = np.concatenate([floor, wall], axis=1)
    b = x0.shape[1]
    v = rng.uniform(-0.3, 0.3, size=(3, b)) if moving else np.zeros((3, b))
    step = 3 if moving else 0
    t = np.float64(step) * np.float64(np.float32(sc.domain.time_step))
    x0 = (x0 - t * v).astype(np.float32)  # the samples reach x0 + t v at the step
    v = v.astype(np.float32)
    vol = rng.uniform(0.5, 1.5, b).astype(np.float32) * np.float32(s ** 3)
    xb = x0.astype(np.float64) + t * v.astype(np.float64)
    return sc, n, (x0, v, vol), xb, step


def merged(sc, fluid, extra):
    return np.concatenate([fluid, extra], axis=-1)


def test_density_term_equals_mirrored_fluid_neighbours():
    for moving in (False, True):
        sc, n, samples, xb, step = setup(moving)
        x0, v, vol = samples
        b = vol.shape[0]
        rho0 = 1000.0
        m = sc.parameters["masses"].astype(np.float64)
        x = sc.positions.astype(np.float64)
        with_b = oracle.density(oracle.make_domain(sc.domain, samples), step, False, x, x, m, np.full(n, rho0), "f64")
        sc.domain.particle_count = n + b
        d = oracle.make_domain(sc.domain)
        xa = merged(sc, x, xb)
        ma = merged(sc, m, rho0 * vol.astype(np.float64))
        as_fluid = oracle.density(d, step, False, xa, xa, ma, np.full(n + b, rho0), "f64")[:n]
        sc.domain.particle_count = n
        diff = np.abs(with_b - as_fluid).max() / np.abs(as_fluid).max()
        assert diff < 1e-13, (moving, diff)
        without = oracle.density(oracle.make_domain(sc.domain), step, False, x, x, m, np.full(n, rho0), "f64")
        assert np.sum(np.abs(with_b - without) > 1e-6 * rho0) > n // 4  # the samples reach many particles


def test_density_term_is_linear_in_the_particle_rest_density():
    sc, n, samples, _, step = setup(True)
    d = oracle.make_domain(sc.domain, samples)
    x = sc.positions.astype(np.float64)
    m = sc.parameters["masses"].astype(np.float64)
    rng = np.random.default_rng(3)
    r0 = rng.uniform(900.0, 1100.0, n)
    base = oracle.density(oracle.make_domain(sc.domain), step, False, x, x, m, r0, "f64")
    one = oracle.density(d, step, False, x, x, m, r0, "f64") - base
    two = oracle.density(d, step, False, x, x, m, 2.0 * r0, "f64") - base
    assert np.abs(one).max() > 0.0
    assert np.allclose(two, 2.0 * one, rtol=1e-12, atol=1e-9)


def test_pressure_term_equals_pressure_free_fluid_neighbours():
    sc, n, samples, xb, step = setup(True)
    x0, v, vol = samples
    b = vol.shape[0]
    rng = np.random.default_rng(5)
    rho0 = 1000.0
    m = sc.parameters["masses"].astype(np.float64)
    x = sc.positions.astype(np.float64)
    rho = rng.uniform(990.0, 1030.0, n)
    p = rng.uniform(0.0, 300.0, n)
    with_b = oracle.pressure(oracle.make_domain(sc.domain, samples), step, x, m, np.full(n, rho0), rho, p, "f64")
    sc.domain.particle_count = n + b
    d = oracle.make_domain(sc.domain)
    as_fluid = oracle.pressure(d, step, merged(sc, x, xb), merged(sc, m, rho0 * vol.astype(np.float64)),
                               np.full(n + b, rho0), merged(sc, rho, np.full(b, rho0)), merged(sc, p, np.zeros(b)),
                               "f64")[:, :n]
    sc.domain.particle_count = n
    without = oracle.pressure(oracle.make_domain(sc.domain), step, x, m, np.full(n, rho0), rho, p, "f64")
    scale = np.abs(as_fluid).max()
    assert np.abs(with_b - as_fluid).max() <= 1e-12 * scale
    assert np.abs(with_b - without).max() > 1e-3 * scale


def test_viscosity_term_equals_matched_fluid_neighbours():
    sc, n, samples, xb, step = setup(True)
    x0, v, vol = samples
    b = vol.shape[0]
    rng = np.random.default_rng(9)
    mu = 0.02
    m = sc.parameters["masses"].astype(np.float64)
    x = sc.positions.astype(np.float64)
    vel = rng.normal(0.0, 0.5, size=(3, n))
    rho = rng.uniform(990.0, 1030.0, n)
    u = np.zeros((3, n))
    with_b = oracle.non_pressure(oracle.make_domain(sc.domain, samples), step, x, vel, u, m, np.full(n, mu), rho,
                                 "f64")
    sc.domain.particle_count = n + b
    d = oracle.make_domain(sc.domain)
    rho_b = np.full(b, 1000.0)
    as_fluid = oracle.non_pressure(d, step, merged(sc, x, xb), merged(sc, vel, v.astype(np.float64)),
                                   np.zeros((3, n + b)), merged(sc, m, rho_b * vol.astype(np.float64)),
                                   np.full(n + b, mu), merged(sc, rho, rho_b), "f64")[:, :n]
    sc.domain.particle_count = n
    without = oracle.non_pressure(oracle.make_domain(sc.domain), step, x, vel, u, m, np.full(n, mu), rho, "f64")
    gravity = np.abs(without).max()
    assert np.abs(with_b - as_fluid).max() <= 1e-12 * gravity
    assert np.abs(with_b - without).max() > 1e-4 * gravity
