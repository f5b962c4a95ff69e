"""The differentiable PBF rollout's gradient on the GPU, on the construction
of the 262,144-particle rollout (lattice spacing 0.02, h 0.04, dt 1/120, 4
iterations, loss the mean squared distance of the final positions to a
seeded target cloud, gradient with respect to the initial velocities) at 512
particles, where the oracle is fast.

- 120 steps, a state checkpoint every 10 steps with segment recompute and
  the unkept steps on a forward-mode solver: the final state equals the FP32
  oracle's bit for bit, and the gradient equals the oracle's chained VJPs
  (FP64 adjoints of the FP32 primal, pbf/pbf.cpp:217-258 step after step)
  to 1e-6 along each of 64 seeded directions.
- 2 steps: the gradient agrees with central differences of the FP64 oracle
  (eps = 1e-7, in the linear regime) to 1e-3 along each of 16 directions.
  The GPU differentiates the FP32 trajectory and the differences the FP64
  one; in this clumping dam break their derivatives part by up to 4 % along
  single directions from the third step on (median 4e-4), while the FP64
  oracle's own gradient keeps matching its differences (the oracle's pins).
At 120 steps a difference quotient cannot check the gradient: the exact
derivative grows about 3x per step (the oracle's own pins show it), far
beyond what eps = 1e-3 resolves."""
import numpy as np
import pytest
import torch

import oracle
from physica_amd import scenes
from physica_amd.particle import ExecutionMode, PBF
from physica_amd.rollout import PbfRollout
from helpers import bits_equal, pbf_setup

pytestmark = pytest.mark.gpu

SEED = 20260004


def scene():
    sc = scenes.pbf_dam_break(shape=(8, 8, 8), spacing=0.02, seed=SEED)
    rng = np.random.default_rng(SEED + 100)
    target = sc.positions + np.array([[0.1], [-0.05], [0.08]]) + rng.normal(0, 0.01, size=(3, sc.count))
    return sc, target


def gpu_gradient(sc, target, steps, device):
    cfg = PBF.Configuration(pressure_iterations=4)
    solver, state, _, control, params, _ = pbf_setup(sc, device, cfg, ExecutionMode.differentiable)
    forward_solver = pbf_setup(sc, device, cfg, ExecutionMode.forward)[0]
    rollout = PbfRollout(solver, steps, checkpoint_every=10, parameter_gradients=False,
                         forward_solver=forward_solver)
    final = rollout.forward(state, control, params)
    x_final = final.particles.positions.numpy()
    seed = solver.allocate_state_adjoint()
    seed.particles.positions.upload(2.0 * (x_final.astype(np.float64) - target) / sc.count)
    initial, _, _ = rollout.backward(seed)
    torch.cuda.synchronize()
    return x_final, initial.particles.velocities.numpy()


def oracle_gradient(sc, target, steps):
    """FP32 primal rollout and FP64 chained VJPs in the oracle."""
    d = oracle.make_domain(sc.domain)
    zero = np.zeros_like(sc.positions)
    x, v = sc.positions, sc.velocities
    states = []
    for k in range(steps):
        states.append((x, v))
        r = oracle.pbf_step(d, k, x, v, zero, sc.parameters, 4)
        x, v = r["positions"], r["velocities"]
    xa = 2.0 * (x.astype(np.float64) - target) / sc.count
    va = np.zeros_like(xa)
    for k in reversed(range(steps)):
        r = oracle.pbf_vjp(d, k, states[k][0], states[k][1], zero, sc.parameters, xa, va, 4)
        xa, va = r["positions"], r["velocities"]
    return x, va


def directions(count, shape, seed):
    rng = np.random.default_rng(seed)
    for _ in range(count):
        u = rng.normal(0, 1, size=shape)
        yield u / np.sqrt(np.mean(u * u))


def test_120_step_gradient_matches_the_chained_oracle_vjps(gpu_device):
    sc, target = scene()
    print(f"rollout gradient n={sc.count} steps=120", flush=True)
    x_gpu, g_gpu = gpu_gradient(sc, target, 120, gpu_device)
    x_ref, g_ref = oracle_gradient(sc, target, 120)
    assert bits_equal(x_gpu, x_ref)
    assert np.isfinite(g_gpu).all()
    errors = []
    for u in directions(64, g_ref.shape, 20260104):
        ref = float(np.sum(g_ref * u))
        got = float(np.sum(g_gpu * u))
        errors.append(abs(got - ref) / abs(ref))
    print(f"  |grad| {np.linalg.norm(g_ref):.3e}, max per-direction relative error {max(errors):.2e}", flush=True)
    assert max(errors) <= 1e-6


def test_short_horizon_gradient_matches_central_differences(gpu_device):
    sc, target = scene()
    steps, eps = 2, 1e-7
    print(f"rollout gradient vs FD n={sc.count} steps={steps}", flush=True)
    _, g_gpu = gpu_gradient(sc, target, steps, gpu_device)
    d = oracle.make_domain(sc.domain)
    params = {k: v.astype(np.float64) for k, v in sc.parameters.items()}
    zero = np.zeros((3, sc.count))

    def loss(v0):
        x, v = sc.positions.astype(np.float64), v0
        for k in range(steps):
            r = oracle.pbf_step(d, k, x, v, zero, params, 4, "f64")
            x, v = r["positions"], r["velocities"]
        return float(np.mean(np.sum((x - target) ** 2, axis=0)))

    v0 = sc.velocities.astype(np.float64)
    errors = []
    for u in directions(16, g_gpu.shape, 20260204):
        fd = (loss(v0 + eps * u) - loss(v0 - eps * u)) / (2.0 * eps)
        ad = float(np.sum(g_gpu * u))
        errors.append(abs(ad - fd) / abs(fd))
    print(f"  max per-direction relative error {max(errors):.2e}", flush=True)
    assert max(errors) <= 1e-3
