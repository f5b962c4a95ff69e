"""PBF::forward_step with boundary particles on the GPU against the FP32
oracle: a floor of samples under the dam-break block and a wall beside it
moving toward the fluid. The samples enter the density (rho0_i V_b W), the
gradient sum of every lambda (V_b grad W) and every correction
(V_b (lambda_i + s_corr) grad W), each cell's samples after its particles.
Bit-exact with the reference parameters (s_corr vanishes); with artificial
pressure, where powf enters, within the 1e-5 relative tolerance of the
boundary-free case."""
import numpy as np
import pytest
import torch

import oracle
from physica_amd import scenes
from physica_amd.particle import PBF, ExecutionMode
from helpers import add_boundary_samples, bits_equal, max_rel, pbf_setup

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("moving", [False, True])
def test_rollout_with_boundary_bit_exact(gpu_device, moving):
    print(f"pbf boundary moving={moving}", flush=True)
    sc = scenes.pbf_dam_break(shape=(16, 12, 16))
    samples = add_boundary_samples(sc, moving)
    solver, state, nxt, control, params, cache = pbf_setup(sc, gpu_device)
    d = oracle.make_domain(sc.domain, samples)
    x, v = sc.positions, sc.velocities
    u = np.zeros_like(x)
    for step in range(4):
        solver.forward_step(state, control, params, nxt, cache)
        torch.cuda.synchronize()
        exp = oracle.pbf_step(d, step, x, v, u, sc.parameters, 4)
        if step == 0:
            free = oracle.pbf_step(oracle.make_domain(sc.domain), step, x, v, u, sc.parameters, 4)
            assert np.sum(free["positions"] != exp["positions"]) > sc.count // 10
        gx, gv = nxt.particles.positions.numpy(), nxt.particles.velocities.numpy()
        assert bits_equal(gx, exp["positions"]), step
        assert bits_equal(gv, exp["velocities"]), step
        rho, lam = cache.last_iteration()
        assert bits_equal(rho.cpu().numpy(), exp["densities"]), step
        assert bits_equal(lam.cpu().numpy(), exp["lambdas"]), step
        x, v = gx, gv
        solver.copy_state(nxt, state)


def test_artificial_pressure_with_boundary_within_tolerance(gpu_device):
    sc = scenes.pbf_dam_break(shape=(16, 12, 16))
    n = sc.count
    samples = add_boundary_samples(sc, True)
    sc.parameters["artificial_pressure_strength"] = np.full(n, 1e-4, np.float32)
    solver, state, nxt, control, params, cache = pbf_setup(sc, gpu_device)
    solver.forward_step(state, control, params, nxt, cache)
    torch.cuda.synchronize()
    exp = oracle.pbf_step(oracle.make_domain(sc.domain, samples), 0, sc.positions, sc.velocities,
                          np.zeros((3, n)), sc.parameters, 4)
    assert max_rel(nxt.particles.positions.numpy(), exp["positions"]) < 1e-5
    assert max_rel(nxt.particles.velocities.numpy(), exp["velocities"]) < 1e-5
