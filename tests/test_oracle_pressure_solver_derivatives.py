"""Pins of the oracle's PCISPH and IISPH JVP and VJP (SPH<PCISPH|IISPH>::
jvp_step/vjp_step, sph/sph.cpp:312-363, 423-476).

The oracle differentiates these steps as the DFSPH step without a divergence
phase or warm start, with the pressure (Jacobi) relaxation in the density
phase and, for PCISPH, V = m / rho0 in the pressure update. The first test
pins that this primal is the PCISPH / IISPH forward step bit for bit; the
others check the derivatives themselves: the dot-product identity between the
JVP and the VJP in FP64, central finite differences of the FP64 forward step
along random directions, the VJP of each input group against differences, and
FP32 tracking FP64. The scene is the compressed, swirling block of the DFSPH
derivative tests."""
import numpy as np
import pytest

import oracle
from test_oracle_dfsph_derivatives import case

PARAMS = oracle.PRESSURE_SOLVER_PARAMETERS
STATE = ("positions", "velocities", "controls")
METHODS = ("pcisph", "iisph")


def setup(method, seed=3, shape=(6, 6, 5), boundary=False, wall=False):
    cs = case(shape=shape, seed=seed, boundary=boundary, wall=wall)
    p = cs.sc.parameters
    p["relaxation"] = p["density_relaxation"]
    return cs


def inputs(cs, precision="f64"):
    dt = np.float64 if precision == "f64" else np.float32
    sc = cs.sc
    x = sc.positions.astype(dt)
    v = sc.velocities.astype(dt)
    u = np.zeros_like(x)
    p = {k: sc.parameters[k].astype(dt) for k in PARAMS}
    return x, v, u, p


def forward(cs, method, x, v, u, p, iterations, precision="f64"):
    return oracle.pressure_solver_step(cs.d, 0, x, v, u, p, method, iterations, precision)


def directions(cs, rng):
    n = cs.sc.count
    t = {k: rng.normal(0, 1, size=(3, n)) for k in STATE}
    t["positions"] *= 1e-4
    t["velocities"] *= 1e-2
    for k in PARAMS:
        t[k] = rng.normal(0, 1, n) * 0.1 * np.abs(cs.sc.parameters[k].astype(np.float64)).mean()
    w = {"positions": rng.normal(0, 1, size=(3, n)), "velocities": rng.normal(0, 1, size=(3, n))}
    return t, w


def pairing(jt, w):
    return sum(np.sum(jt[k] * w[k]) for k in ("positions", "velocities"))


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_recorded_primal_is_the_forward_step(method, precision):
    cs = setup(method)
    x, v, u, p = inputs(cs, precision)
    n = cs.sc.count
    zero = np.zeros(n)
    for iterations in (1, 4):
        ref = forward(cs, method, x, v, u, p, iterations, precision)
        rec = oracle.dfsph_record_step(cs.d, 0, x, v, u, oracle._as_dfsph(p, n), zero, zero, 0, iterations, False,
                                       method == "pcisph", precision)
        for k in ("positions", "velocities"):
            assert np.array_equal(rec[k], ref[k]), (iterations, k)
    # the volume choice matters: the other method's volume gives another step
    other = oracle.dfsph_record_step(cs.d, 0, x, v, u, oracle._as_dfsph(p, n), zero, zero, 0, 4, False,
                                     method != "pcisph", precision)
    assert not np.array_equal(other["positions"], forward(cs, method, x, v, u, p, 4, precision)["positions"])


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("iterations,boundary", [(1, False), (4, False), (3, True)])
def test_dot_product_identity(method, iterations, boundary):
    cs = setup(method, boundary=boundary)
    x, v, u, p = inputs(cs)
    rng = np.random.default_rng(40 + iterations)
    t, w = directions(cs, rng)
    jt = oracle.pressure_solver_jvp(cs.d, 0, x, v, u, p, t, method, iterations, "f64")
    jw = oracle.pressure_solver_vjp(cs.d, 0, x, v, u, p, w, method, iterations, "f64")
    left = pairing(jt, w)
    right = sum(np.sum(t[k] * jw[k]) for k in STATE + PARAMS)
    assert abs(left - right) <= 1e-11 * (abs(left) + abs(right)), (left, right)


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("boundary", [False, True])
def test_jvp_matches_central_differences(method, boundary):
    cs = setup(method, boundary=boundary)
    iterations = 3
    x, v, u, p = inputs(cs)
    rng = np.random.default_rng(50)
    t, _ = directions(cs, rng)
    jt = oracle.pressure_solver_jvp(cs.d, 0, x, v, u, p, t, method, iterations, "f64")

    def central(eps):
        def shifted(sign):
            q = {k: p[k] + sign * eps * t[k] for k in PARAMS}
            return forward(cs, method, x + sign * eps * t["positions"], v + sign * eps * t["velocities"],
                           u + sign * eps * t["controls"], q, iterations)

        a, b = shifted(1.0), shifted(-1.0)
        return {k: (a[k] - b[k]) / (2 * eps) for k in ("positions", "velocities")}

    c1, c2 = central(2e-6), central(1e-6)
    for k in ("positions", "velocities"):
        fd = (4 * c2[k] - c1[k]) / 3
        err = np.abs(jt[k] - fd).max() / np.abs(fd).max()
        assert err < 2e-6, (k, err)


@pytest.mark.parametrize("method", METHODS)
def test_vjp_of_each_input_matches_finite_differences(method):
    cs = setup(method, seed=5, shape=(5, 5, 4))
    iterations = 3
    x, v, u, p = inputs(cs)
    rng = np.random.default_rng(60)
    t_all, w = directions(cs, rng)
    jw = oracle.pressure_solver_vjp(cs.d, 0, x, v, u, p, w, method, iterations, "f64")
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
            return pairing(forward(cs, method, xs, vs, us, q, iterations), w)

        def central(eps):
            return (value(1.0, eps) - value(-1.0, eps)) / (2 * eps)

        fd = (4 * central(1.5e-3) - central(3e-3)) / 3
        got = np.sum(t * jw[group])
        assert abs(got - fd) <= 3e-3 * abs(fd) + 1e-6, (group, got, fd)


@pytest.mark.parametrize("method", METHODS)
def test_pcisph_and_iisph_derivatives_differ_only_through_the_volume(method):
    # a mass or rest-density direction separates V = m / rho0 from m / rho:
    # the two methods' JVPs along the rest densities must differ
    cs = setup(method)
    x, v, u, p = inputs(cs)
    rng = np.random.default_rng(70)
    t, _ = directions(cs, rng)
    only = {k: np.zeros_like(t[k]) for k in t}
    only["rest_densities"] = t["rest_densities"]
    a = oracle.pressure_solver_jvp(cs.d, 0, x, v, u, p, only, "pcisph", 3, "f64")
    b = oracle.pressure_solver_jvp(cs.d, 0, x, v, u, p, only, "iisph", 3, "f64")
    assert not np.allclose(a["positions"], b["positions"], rtol=1e-6, atol=0)


@pytest.mark.parametrize("method", METHODS)
def test_f32_tracks_f64(method):
    cs = setup(method)
    rng = np.random.default_rng(80)
    t, w = directions(cs, rng)
    x64, v64, u64, p64 = inputs(cs, "f64")
    x32, v32, u32, p32 = inputs(cs, "f32")
    j64 = oracle.pressure_solver_jvp(cs.d, 0, x64, v64, u64, p64, t, method, 3, "f64")
    j32 = oracle.pressure_solver_jvp(cs.d, 0, x32, v32, u32, p32, t, method, 3, "f32")
    for k in ("positions", "velocities"):
        assert np.abs(j32[k] - j64[k]).max() <= 2e-3 * np.abs(j64[k]).max(), k
    a64 = oracle.pressure_solver_vjp(cs.d, 0, x64, v64, u64, p64, w, method, 3, "f64")
    a32 = oracle.pressure_solver_vjp(cs.d, 0, x32, v32, u32, p32, w, method, 3, "f32")
    for k in STATE + PARAMS:
        assert np.abs(a32[k] - a64[k]).max() <= 2e-3 * np.abs(a64[k]).max(), k
