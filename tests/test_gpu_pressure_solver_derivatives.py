"""SPH<PCISPH|IISPH>::jvp_step and vjp_step on the GPU against the oracle.

In differentiable mode the solver records the step on the projection engine
(no divergence phase, no warm start, V = m / rho0 for PCISPH); its forward
step must equal the oracle's PCISPH / IISPH step and the forward-mode solver
bit for bit, cache fields included. The JVP (FP32) agrees with the oracle's
FP32 JVP to 1e-6 of the largest entry and the VJP (FP64 adjoints of the FP32
primal) to 1e-10, as for DFSPH; the GPU's own JVP and VJP satisfy the
dot-product identity. The scene is the compressed swirling block of the
oracle's derivative pins, also over a moving plane of boundary samples, and
pushed into a wall for the VJP of a no-slip collision: the exact transpose
in most tests, the reference's (the solver's default) against the oracle's
restatement of it."""
import numpy as np
import pytest
import torch

import oracle
from physica_amd.particle import ExecutionMode
from helpers import bits_equal, reference_no_slip_oracle, sph_setup
from test_oracle_pressure_solver_derivatives import PARAMS, STATE, directions, inputs, setup

pytestmark = pytest.mark.gpu

METHODS = ("pcisph", "iisph")


def method_field(method):
    return "pressure_relaxation" if method == "pcisph" else "jacobi_relaxation"


def recorded(cs, method, iterations, device, exact=True):
    solver, state, nxt, control, params, cache = sph_setup(cs.sc, device, method, iterations,
                                                           mode=ExecutionMode.differentiable,
                                                           exact_no_slip_adjoint=exact)
    solver.forward_step(state, control, params, nxt, cache)
    torch.cuda.synchronize()
    x, v, u, p = inputs(cs, "f32")
    ref = oracle.pressure_solver_step(cs.d, 0, x, v, u, p, method, iterations, "f32")
    assert bits_equal(nxt.particles.positions.numpy(), ref["positions"])
    assert bits_equal(nxt.particles.velocities.numpy(), ref["velocities"])
    assert nxt.particles.step_index == 1
    return solver, state, nxt, control, params, cache


def tensor(values, device, dtype):
    return torch.from_numpy(np.ascontiguousarray(values, dtype=dtype)).to(device)


def run_jvp(method, solver, state, nxt, control, params, cache, t, device):
    st = solver.allocate_state_tangent()
    st.particles.positions.upload(t["positions"])
    st.particles.velocities.upload(t["velocities"])
    ct = solver.allocate_control_tangent()
    ct.external_accelerations.upload(t["controls"])
    pt = solver.allocate_parameter_tangent()
    for k in ("masses", "rest_densities", "viscosities"):
        getattr(pt.particles, k).copy_(tensor(t[k], device, np.float32))
    getattr(pt.method, method_field(method)).copy_(tensor(t["relaxation"], device, np.float32))
    nt = solver.allocate_state_tangent()
    solver.jvp_step(state, control, params, nxt, cache, st, ct, pt, nt)
    torch.cuda.synchronize()
    return {"positions": nt.particles.positions.numpy(), "velocities": nt.particles.velocities.numpy()}


def run_vjp(method, solver, state, nxt, control, params, cache, w):
    next_a = solver.allocate_state_adjoint()
    next_a.particles.positions.upload(w["positions"])
    next_a.particles.velocities.upload(w["velocities"])
    prev_a = solver.allocate_state_adjoint()
    ctl_a = solver.allocate_control_adjoint()
    par_a = solver.allocate_parameter_adjoint()
    solver.vjp_step(state, control, params, nxt, cache, next_a, prev_a, ctl_a, par_a)
    torch.cuda.synchronize()
    got = {"positions": prev_a.particles.positions.numpy(), "velocities": prev_a.particles.velocities.numpy(),
           "controls": ctl_a.external_accelerations.numpy()}
    for k in ("masses", "rest_densities", "viscosities"):
        got[k] = getattr(par_a.particles, k).cpu().numpy()
    got["relaxation"] = getattr(par_a.method, method_field(method)).cpu().numpy()
    return got


def close(got, ref, tol):
    scale = np.abs(ref).max()
    return np.abs(got - ref).max() <= tol * max(scale, 1e-30)


@pytest.mark.parametrize("method", METHODS)
def test_differentiable_forward_equals_forward_mode_and_its_cache(method):
    dev = torch.device("cuda:0")
    cs = setup(method)
    a = recorded(cs, method, 4, dev)
    b = sph_setup(cs.sc, dev, method, 4)
    solver, state, nxt, control, params, cache = b
    solver.forward_step(state, control, params, nxt, cache)
    torch.cuda.synchronize()
    assert bits_equal(a[2].particles.positions.numpy(), b[2].particles.positions.numpy())
    assert bits_equal(a[2].particles.velocities.numpy(), b[2].particles.velocities.numpy())
    ca, cb = a[5], b[5]
    assert bits_equal(ca.densities().cpu().numpy(), cb.densities().cpu().numpy())
    assert bits_equal(ca.pressures().cpu().numpy(), cb.pressures().cpu().numpy())
    assert bits_equal(ca.non_pressure_accelerations().numpy(), cb.non_pressure_accelerations().numpy())
    for x, y in zip(ca.neighborhood(), cb.neighborhood()):
        assert torch.equal(x, y)


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("iterations,boundary", [(1, False), (4, False), (3, True)])
def test_jvp_and_vjp_match_the_oracle(method, iterations, boundary):
    print(f"{method} derivatives iterations={iterations} boundary={boundary}", flush=True)
    dev = torch.device("cuda:0")
    cs = setup(method, boundary=boundary)
    run = recorded(cs, method, iterations, dev)
    rng = np.random.default_rng(90 + iterations)
    t, w = directions(cs, rng)
    x, v, u, p = inputs(cs, "f32")
    jt = run_jvp(method, *run, t, dev)
    rj = oracle.pressure_solver_jvp(cs.d, 0, x, v, u, p, t, method, iterations, "f32")
    for k in ("positions", "velocities"):
        assert close(jt[k], rj[k], 1e-6), k
    jw = run_vjp(method, *run, w)
    rw = oracle.pressure_solver_vjp(cs.d, 0, x, v, u, p, w, method, iterations, "f32")
    for k in STATE + PARAMS:
        assert close(jw[k], rw[k], 1e-10), k


@pytest.mark.parametrize("method", METHODS)
def test_gpu_dot_product_identity_and_accumulation(method):
    dev = torch.device("cuda:0")
    cs = setup(method, seed=5, shape=(5, 5, 4))
    run = recorded(cs, method, 3, dev)
    rng = np.random.default_rng(95)
    t, w = directions(cs, rng)
    jt = run_jvp(method, *run, t, dev)
    jw = run_vjp(method, *run, w)
    left = sum(np.sum(jt[k].astype(np.float64) * w[k]) for k in ("positions", "velocities"))
    right = sum(np.sum(t[k] * jw[k]) for k in STATE + PARAMS)
    # FP32 JVP against the FP64 VJP: the identity holds to the JVP's rounding
    assert abs(left - right) <= 1e-4 * (abs(left) + abs(right)), (left, right)
    # adjoints accumulate (+=): a second VJP into the same outputs doubles them
    solver, state, nxt, control, params, cache = run
    next_a = solver.allocate_state_adjoint()
    next_a.particles.positions.upload(w["positions"])
    next_a.particles.velocities.upload(w["velocities"])
    prev_a = solver.allocate_state_adjoint()
    ctl_a = solver.allocate_control_adjoint()
    par_a = solver.allocate_parameter_adjoint()
    for _ in range(2):
        solver.vjp_step(state, control, params, nxt, cache, next_a, prev_a, ctl_a, par_a)
    torch.cuda.synchronize()
    assert np.allclose(prev_a.particles.positions.numpy(), 2 * jw["positions"], rtol=1e-12, atol=0)
    assert np.allclose(getattr(par_a.method, method_field(method)).cpu().numpy(), 2 * jw["relaxation"],
                       rtol=1e-12, atol=0)


@pytest.mark.parametrize("method", METHODS)
def test_reference_no_slip_vjp_matches_oracle(method):
    """Default configuration: the reference's VJP of a no-slip collision
    (sph/pcisph.cu:81, sph/iisph.cu:81) against the oracle's restatement, in
    the wall scene where predictions collide."""
    print(f"{method} reference no-slip vjp", flush=True)
    dev = torch.device("cuda:0")
    cs = setup(method, wall=True)
    run = recorded(cs, method, 3, dev, exact=False)
    rng = np.random.default_rng(97)
    _, w = directions(cs, rng)
    x, v, u, p = inputs(cs, "f32")
    jw = run_vjp(method, *run, w)
    with reference_no_slip_oracle():
        rw = oracle.pressure_solver_vjp(cs.d, 0, x, v, u, p, w, method, 3, "f32")
    exact = oracle.pressure_solver_vjp(cs.d, 0, x, v, u, p, w, method, 3, "f32")
    assert not np.array_equal(rw["velocities"], exact["velocities"])
    for k in STATE + PARAMS:
        assert close(jw[k], rw[k], 1e-10), k


def test_forward_mode_refuses_derivatives():
    dev = torch.device("cuda:0")
    cs = setup("pcisph")
    solver, state, nxt, control, params, cache = sph_setup(cs.sc, dev, "iisph", 2)
    solver.forward_step(state, control, params, nxt, cache)
    st, ct, pt, nt = (solver.allocate_state_tangent(), solver.allocate_control_tangent(),
                      solver.allocate_parameter_tangent(), solver.allocate_state_tangent())
    with pytest.raises(RuntimeError, match="differentiable mode"):
        solver.jvp_step(state, control, params, nxt, cache, st, ct, pt, nt)
