"""Solvers created under a side stream and used from another stream, with no
torch.cuda.synchronize() between the calls: the Python mirror orders the
solver's stream after the caller's (zero-fills, uploads) and the caller's
after the solver's (read-backs, frees). Every result equals the same run on
one stream bit for bit: PBF steps, the step cache's accessors and a
checkpointed rollout's gradient; DFSPH steps and cached densities."""
import numpy as np
import pytest
import torch

from physica_amd import scenes
from physica_amd.particle import DFSPH, PBF, SPH, ExecutionMode
from physica_amd.rollout import PbfRollout
from helpers import PBF_NAMES, bits_equal, dfsph_setup, upload_scalar

pytestmark = pytest.mark.gpu


def _pbf_run(solver, sc, steps=3):
    """Steps, cache read-backs and a checkpointed rollout gradient, all
    enqueued from the current stream without a synchronize in between."""
    state, nxt = solver.allocate_state(), solver.allocate_state()
    control, params, cache = solver.allocate_control(), solver.allocate_parameters(), solver.allocate_step_cache()
    state.particles.positions.upload(sc.positions)
    state.particles.velocities.upload(sc.velocities)
    for name in ("masses", "rest_densities", "viscosities", "surface_tensions"):
        upload_scalar(getattr(params.particles, name), sc.parameters[name])
    for name in PBF_NAMES:
        upload_scalar(getattr(params, name), sc.parameters[name])
    states = [state, nxt]
    for k in range(steps):
        solver.forward_step(states[k % 2], control, params, states[(k + 1) % 2], cache)
    out = {"positions": states[steps % 2].particles.positions.numpy(),
           "velocities": states[steps % 2].particles.velocities.numpy(),
           "vorticity": cache.vorticity_magnitudes().cpu().numpy(),
           "predicted": cache.predicted_positions().numpy()}
    rho, lam = cache.last_iteration()
    out["rho"], out["lambda"] = rho.cpu().numpy(), lam.cpu().numpy()
    rollout = PbfRollout(solver, 6, checkpoint_every=3)
    final = rollout.forward(state, control, params)
    seed = solver.allocate_state_adjoint()
    seed.particles.positions.upload(np.ones((3, sc.count)) / sc.count)
    initial, _, parameter_adjoint = rollout.backward(seed)
    out["final"] = final.particles.positions.numpy()
    out["grad_v"] = initial.particles.velocities.numpy()
    out["grad_relaxation"] = parameter_adjoint.relaxation.cpu().numpy()
    return out


def test_pbf_solver_created_on_side_stream(gpu_device):
    print("pbf side-stream solver", flush=True)
    sc = scenes.pbf_dam_break(shape=(16, 16, 16), seed=20260004)
    config = PBF.Configuration(pressure_iterations=3)
    reference = _pbf_run(PBF(sc.domain, config, mode=ExecutionMode.differentiable, device=gpu_device), sc)
    torch.cuda.synchronize()
    side = torch.cuda.Stream(gpu_device)
    with torch.cuda.stream(side):
        solver = PBF(sc.domain, config, mode=ExecutionMode.differentiable, device=gpu_device)
    assert solver._stream.cuda_stream == side.cuda_stream
    got = _pbf_run(solver, sc)  # issued from the default stream
    for key, value in reference.items():
        if value.dtype == np.float32:
            assert bits_equal(got[key], value), key
        else:
            assert np.array_equal(got[key], value), key


def test_dfsph_solver_created_on_side_stream(gpu_device):
    print("dfsph side-stream solver", flush=True)
    sc = scenes.dfsph_dam_break(shape=(16, 8, 16))
    config = DFSPH.Configuration(divergence_iterations=2, density_iterations=3)

    def run(solver):
        state, nxt = solver.allocate_state(), solver.allocate_state()
        control, params, cache = solver.allocate_control(), solver.allocate_parameters(), solver.allocate_step_cache()
        state.particles.positions.upload(sc.positions)
        state.particles.velocities.upload(sc.velocities)
        for name in ("masses", "rest_densities", "viscosities", "surface_tensions"):
            upload_scalar(getattr(params.particles, name), sc.parameters[name])
        upload_scalar(params.method.divergence_relaxation, sc.parameters["divergence_relaxation"])
        upload_scalar(params.method.density_relaxation, sc.parameters["density_relaxation"])
        states = [state, nxt]
        for k in range(4):
            solver.forward_step(states[k % 2], control, params, states[(k + 1) % 2], cache)
        return (states[0].particles.positions.numpy(), states[0].particles.velocities.numpy(),
                cache.densities().cpu().numpy(), cache.non_pressure_accelerations().numpy())

    reference = run(dfsph_setup(sc, gpu_device, config)[0])
    torch.cuda.synchronize()
    side = torch.cuda.Stream(gpu_device)
    with torch.cuda.stream(side):
        solver = SPH(sc.domain, config, device=gpu_device)
    got = run(solver)
    for a, b in zip(got, reference):
        assert bits_equal(a, b)
