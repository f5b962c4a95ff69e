"""Pins of the oracle's WCSPH JVP and VJP (SPH<WCSPH>::jvp_step/vjp_step,
sph/sph.cpp:228-249): the dot-product identity between the two in FP64,
central finite differences of the FP64 forward step along random directions
and along one input group at a time, and FP32 tracking FP64.

The scene is the compressed, swirling block of the DFSPH derivative tests
with WCSPH parameters drawn per particle (speed of sound, Tait exponent,
artificial viscosity, cohesion), so every branch of the step is taken:
approaching and separating pairs (the artificial viscosity acts only on the
former), cohesion pairs closer than a particle diameter (clamped weight) and
farther, and, in the variants, a plane of boundary particles and a wall that
clamps the integration."""
import numpy as np
import pytest

import oracle
from test_oracle_dfsph_derivatives import case

PARAMS = oracle.WCSPH_PARAMETERS
STATE = ("positions", "velocities", "controls")


def setup(seed=3, shape=(6, 6, 5), wall=False, boundary=False):
    cs = case(shape=shape, seed=seed, wall=wall, boundary=boundary)
    n = cs.sc.count
    if wall:
        # the Tait pressure is negative (tension) at the block's surface and
        # lifts the lowest layer; a faster fall still drives it into the floor
        cs.sc.velocities[1] -= np.float32(1.0)
    rng = np.random.default_rng(seed + 100)
    p = cs.sc.parameters
    p["speed_of_sound"] = rng.uniform(4.0, 6.0, n).astype(np.float32)
    p["tait_exponent"] = rng.uniform(1.5, 7.0, n).astype(np.float32)
    p["viscosities"] = rng.uniform(0.05, 0.2, n).astype(np.float32)
    # cohesion of the order of the pressure forces, not above them
    p["surface_tensions"] = rng.uniform(0.005, 0.02, n).astype(np.float32)
    p["boundary_surface_tension"] = rng.uniform(0.005, 0.02, n).astype(np.float32)
    return cs


def inputs(cs, precision="f64"):
    dt = np.float64 if precision == "f64" else np.float32
    sc = cs.sc
    x = sc.positions.astype(dt)
    v = sc.velocities.astype(dt)
    u = np.zeros_like(x)
    p = {k: sc.parameters[k].astype(dt) for k in PARAMS}
    return x, v, u, p


def forward(cs, x, v, u, p, precision="f64"):
    return oracle.wcsph_step(cs.d, 0, x, v, u, p, precision)


def directions(cs, rng):
    n = cs.sc.count
    t = {k: rng.normal(0, 1, size=(3, n)) for k in STATE}
    t["positions"] *= 1e-4
    t["velocities"] *= 1e-2
    for k in PARAMS:
        t[k] = rng.normal(0, 1, n) * 0.05 * np.abs(cs.sc.parameters[k].astype(np.float64)).mean()
    w = {"positions": rng.normal(0, 1, size=(3, n)), "velocities": rng.normal(0, 1, size=(3, n))}
    return t, w


def pairing(jt, w):
    return sum(np.sum(jt[k] * w[k]) for k in ("positions", "velocities"))


def test_scene_takes_every_branch():
    cs = setup()
    x, v, _, _ = inputs(cs)
    d = x[:, :, None] - x[:, None, :]
    dist = np.sqrt((d ** 2).sum(0))
    dv = v[:, :, None] - v[:, None, :]
    compression = (dv * d).sum(0)
    h = cs.sc.domain.support_radius
    near = (dist < h) & (dist > 0)
    diameter = 2 * cs.sc.domain.particle_radius
    assert np.sum(near & (compression < 0)) > 50 and np.sum(near & (compression > 0)) > 50
    assert np.sum(near & (dist < diameter)) > 20 and np.sum(near & (dist > diameter)) > 50


@pytest.mark.parametrize("variant", ["free", "wall", "boundary"])
def test_dot_product_identity(variant):
    cs = setup(wall=variant == "wall", boundary=variant == "boundary")
    x, v, u, p = inputs(cs)
    if variant == "wall":
        nxt = forward(cs, x, v, u, p)
        assert np.sum(nxt["positions"][1] <= cs.sc.domain.particle_radius + 1e-12) > 5
    rng = np.random.default_rng(11)
    t, w = directions(cs, rng)
    jt = oracle.wcsph_jvp(cs.d, 0, x, v, u, p, t, "f64")
    jw = oracle.wcsph_vjp(cs.d, 0, x, v, u, p, w, "f64")
    left = pairing(jt, w)
    right = sum(np.sum(t[k] * jw[k]) for k in STATE + PARAMS)
    assert abs(left - right) <= 1e-11 * (abs(left) + abs(right)), (left, right)


@pytest.mark.parametrize("boundary", [False, True])
def test_jvp_matches_central_differences(boundary):
    cs = setup(boundary=boundary)
    x, v, u, p = inputs(cs)
    rng = np.random.default_rng(21)
    t, _ = directions(cs, rng)
    jt = oracle.wcsph_jvp(cs.d, 0, x, v, u, p, t, "f64")

    def central(eps):
        def shifted(sign):
            q = {k: p[k] + sign * eps * t[k] for k in PARAMS}
            return forward(cs, x + sign * eps * t["positions"], v + sign * eps * t["velocities"],
                           u + sign * eps * t["controls"], q)

        a, b = shifted(1.0), shifted(-1.0)
        return {k: (a[k] - b[k]) / (2 * eps) for k in ("positions", "velocities")}

    c1, c2 = central(2e-6), central(1e-6)
    for k in ("positions", "velocities"):
        fd = (4 * c2[k] - c1[k]) / 3
        err = np.abs(jt[k] - fd).max() / np.abs(fd).max()
        assert err < 2e-6, (k, err)


@pytest.mark.parametrize("boundary", [False, True])
def test_vjp_of_each_input_matches_finite_differences(boundary):
    cs = setup(seed=5, shape=(5, 5, 4), boundary=boundary)
    x, v, u, p = inputs(cs)
    rng = np.random.default_rng(31)
    t_all, w = directions(cs, rng)
    jw = oracle.wcsph_vjp(cs.d, 0, x, v, u, p, w, "f64")
    for group in STATE + PARAMS:
        t = t_all[group]

        def value(sign, eps):
            xs, vs, us, q = x, v, u, dict(p)
            if group == "positions":
                xs = x + sign * eps * t
            elif group == "velocities":
                vs = v + sign * eps * t
            elif group == "controls":
                us = u + sign * eps * t
            else:
                q[group] = p[group] + sign * eps * t
            return pairing(forward(cs, xs, vs, us, q), w)

        def central(eps):
            return (value(1.0, eps) - value(-1.0, eps)) / (2 * eps)

        # small steps: an approaching pair whose compression is near zero
        # switches the artificial viscosity off under a larger perturbation
        fd = (4 * central(2e-4) - central(4e-4)) / 3
        got = np.sum(t * jw[group])
        if group == "boundary_surface_tension" and not boundary:
            assert got == 0.0 and abs(fd) < 1e-12
            continue
        assert abs(got - fd) <= 3e-3 * abs(fd) + 1e-7, (group, got, fd)


def test_f32_tracks_f64():
    cs = setup()
    rng = np.random.default_rng(41)
    t, w = directions(cs, rng)
    x64, v64, u64, p64 = inputs(cs, "f64")
    x32, v32, u32, p32 = inputs(cs, "f32")
    j64 = oracle.wcsph_jvp(cs.d, 0, x64, v64, u64, p64, t, "f64")
    j32 = oracle.wcsph_jvp(cs.d, 0, x32, v32, u32, p32, t, "f32")
    for k in ("positions", "velocities"):
        assert np.abs(j32[k] - j64[k]).max() <= 2e-3 * np.abs(j64[k]).max(), k
    a64 = oracle.wcsph_vjp(cs.d, 0, x64, v64, u64, p64, w, "f64")
    a32 = oracle.wcsph_vjp(cs.d, 0, x32, v32, u32, p32, w, "f32")
    for k in STATE + PARAMS:
        scale = np.abs(a64[k]).max()
        if scale == 0.0:
            assert np.abs(a32[k]).max() == 0.0, k
            continue
        assert np.abs(a32[k] - a64[k]).max() <= 2e-3 * scale, k
