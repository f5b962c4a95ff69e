"""The PBF rollout driver and its torch.autograd view: store-all and
checkpointed rollouts give the same gradient, and torch.autograd returns the
chained VJP."""
import numpy as np
import pytest
import torch

from physica_amd import scenes
from physica_amd.autograd import pbf_rollout
from physica_amd.particle import ExecutionMode, PBF
from physica_amd.rollout import PbfRollout
from helpers import pbf_setup

pytestmark = pytest.mark.gpu


def setup(device, mode=ExecutionMode.differentiable):
    sc = scenes.pbf_dam_break(shape=(8, 8, 8), seed=20260004)
    solver, state, _, control, params, _ = pbf_setup(sc, device, PBF.Configuration(pressure_iterations=3), mode)
    return sc, solver, state, control, params


def test_checkpointed_rollout_matches_store_all(gpu_device):
    """Store-all, checkpoints every 4 or 3 steps (the last segment recorded by
    the forward pass, the others recomputed), on one stream or overlapped over
    two solvers on two streams, with the unkept steps on the differentiable
    solver or on a forward-mode solver: identical states and gradients. A
    second backward pass of the same rollout (whose recorded segment the first
    pass overwrote) gives them again."""
    print("rollout checkpoint 1 vs 4 and 3, one and two streams", flush=True)
    sc, solver, state, control, params = setup(gpu_device)
    forward_solver = setup(gpu_device, ExecutionMode.forward)[1]
    n = sc.count
    rng = np.random.default_rng(1)
    w = rng.normal(0, 1, size=(3, n))
    grads = []
    finals = []
    variants = ((1, None, True), (4, None, True), (4, None, False), (4, forward_solver, True),
                (3, forward_solver, True), (3, forward_solver, False))
    for every, stepper, overlap in variants:
        rollout = PbfRollout(solver, 10, checkpoint_every=every, forward_solver=stepper, overlap=overlap)
        final = rollout.forward(state, control, params)
        finals.append(final.particles.positions.numpy())
        for _ in range(2 if every == 3 else 1):
            seed = solver.allocate_state_adjoint()
            seed.particles.positions.upload(w)
            initial, control_adjoint, parameter_adjoint = rollout.backward(seed)
            torch.cuda.synchronize()
            grads.append((initial.particles.velocities.numpy(), parameter_adjoint.relaxation.cpu().numpy(),
                          control_adjoint.external_accelerations.numpy()))
    for k in range(1, len(finals)):
        np.testing.assert_array_equal(finals[0], finals[k])
    for k in range(1, len(grads)):
        for a, b in zip(grads[0], grads[k]):
            np.testing.assert_array_equal(a, b)
    # the state passed in is not modified
    np.testing.assert_array_equal(state.particles.positions.numpy(), sc.positions)


def test_autograd_function_returns_chained_vjp(gpu_device):
    print("rollout autograd", flush=True)
    sc, solver, state, control, params = setup(gpu_device)
    n = sc.count
    x0 = torch.from_numpy(sc.positions).to(gpu_device).requires_grad_(True)
    v0 = torch.from_numpy(sc.velocities).to(gpu_device).requires_grad_(True)
    target = torch.from_numpy(sc.positions + 0.01).to(gpu_device)
    rollout = PbfRollout(solver, 6)
    xT, vT = pbf_rollout(x0, v0, rollout, control, params)
    loss = ((xT - target) ** 2).sum(dim=0).mean() + 0.1 * (vT**2).sum()
    loss.backward()
    # the same through the driver directly
    ref = PbfRollout(solver, 6)
    final = ref.forward(state, control, params)
    seed = solver.allocate_state_adjoint()
    xT_ref = final.particles.positions.numpy().astype(np.float64)
    seed.particles.positions.upload(2.0 * (xT_ref - (sc.positions + 0.01)) / n)
    seed.particles.velocities.upload(0.2 * final.particles.velocities.numpy().astype(np.float64))
    initial, _, _ = ref.backward(seed)
    torch.cuda.synchronize()
    # torch forms the loss gradient in FP32, the direct seed in FP64
    for got, ref_field in ((v0.grad, initial.particles.velocities), (x0.grad, initial.particles.positions)):
        expected = ref_field.numpy()
        assert np.abs(got.cpu().numpy() - expected).max() <= 1e-5 * np.abs(expected).max()
    assert rollout.last_parameter_adjoint is not None
