"""The differentiable PBF rollout's gradient in the FP64 oracle, on the
construction of the 262,144-particle rollout (lattice spacing 0.02, h 0.04,
dt 1/120, 4 iterations, loss the mean squared distance of the final positions
to a seeded target cloud, gradient with respect to the initial velocities)
at 512 particles.

The chained VJPs (pbf/pbf.cpp:217-258 step after step) are the exact
derivative: central differences converge to them as the step shrinks, at 5
and at 20 steps. The same numbers show why a central difference at
eps = 1e-3 cannot check a 120-step gradient: the dam break's derivative
grows by about 3x per step (the scene clumps, as PBF without artificial
pressure does), so at eps = 1e-3 the difference quotient already leaves the
linear regime after a few steps."""
import numpy as np
import pytest

import oracle
from physica_amd import scenes

SEED = 20260004


def setup(side=8):
    sc = scenes.pbf_dam_break(shape=(side, side, side), spacing=0.02, seed=SEED)
    n = sc.count
    rng = np.random.default_rng(SEED + 100)
    target = sc.positions + np.array([[0.1], [-0.05], [0.08]]) + rng.normal(0, 0.01, size=(3, n))
    params = {k: v.astype(np.float64) for k, v in sc.parameters.items()}
    return sc, oracle.make_domain(sc.domain), target, params


def rollout(sc, d, params, v0, steps, keep=False):
    x, v = sc.positions.astype(np.float64), v0
    zero = np.zeros_like(x)
    states = []
    for k in range(steps):
        if keep:
            states.append((x, v))
        r = oracle.pbf_step(d, k, x, v, zero, params, 4, "f64")
        x, v = r["positions"], r["velocities"]
    return x, states


def loss(x, target):
    return float(np.mean(np.sum((x - target) ** 2, axis=0)))


def gradient(sc, d, params, target, steps):
    """dL/dv0 by the chained FP64 VJPs."""
    v0 = sc.velocities.astype(np.float64)
    x_final, states = rollout(sc, d, params, v0, steps, keep=True)
    n = sc.count
    xa = 2.0 * (x_final - target) / n
    va = np.zeros_like(xa)
    zero = np.zeros_like(xa)
    for k in reversed(range(steps)):
        x, v = states[k]
        r = oracle.pbf_vjp(d, k, x, v, zero, params, xa, va, 4, "f64")
        xa, va = r["positions"], r["velocities"]
    return va


def directional_fd(sc, d, params, target, steps, u, eps):
    v0 = sc.velocities.astype(np.float64)
    plus = loss(rollout(sc, d, params, v0 + eps * u, steps)[0], target)
    minus = loss(rollout(sc, d, params, v0 - eps * u, steps)[0], target)
    return (plus - minus) / (2.0 * eps)


@pytest.mark.parametrize("steps,eps", [(5, 1e-7), (20, 1e-10)])
def test_chained_vjp_is_the_derivative(steps, eps):
    sc, d, target, params = setup()
    g = gradient(sc, d, params, target, steps)
    rng = np.random.default_rng(7)
    for _ in range(3):
        u = rng.normal(0, 1, size=g.shape)
        u /= np.sqrt(np.mean(u * u))
        fd = directional_fd(sc, d, params, target, steps, u, eps)
        ad = float(np.sum(g * u))
        assert abs(fd - ad) <= 1e-3 * abs(ad), (steps, fd, ad)


def test_derivative_grows_past_what_eps_1e3_resolves():
    """|dL/dv0| grows by orders of magnitude from 5 to 20 steps, and at 20
    steps the eps = 1e-3 difference quotient no longer resembles the
    derivative, while the eps = 1e-10 one matches it (previous test)."""
    sc, d, target, params = setup()
    norms = {steps: float(np.linalg.norm(gradient(sc, d, params, target, steps))) for steps in (5, 20)}
    assert norms[20] > 1e3 * norms[5], norms
    g = gradient(sc, d, params, target, 20)
    rng = np.random.default_rng(8)
    u = rng.normal(0, 1, size=g.shape)
    u /= np.sqrt(np.mean(u * u))
    coarse = directional_fd(sc, d, params, target, 20, u, 1e-3)
    assert abs(coarse - float(np.sum(g * u))) > 0.5 * abs(float(np.sum(g * u)))
