"""SPH<DFSPH>::jvp_step and vjp_step on the GPU against the oracle.

The recorded forward step is bit-exact against the FP32 oracle. The VJP
(FP32 primal, FP64 adjoints) agrees with the oracle's FP32-primal VJP to
1e-10 of the largest entry of each output: the same per-particle sums, with
only the order in which a few accumulators add their parts differing. The
JVP (FP32, as the reference's) runs the oracle's operations in the same
order and agrees to 1e-6. The VJP of a no-slip collision is the exact
transpose in most tests; with the solver's default (the reference's, which
zeroes the whole velocity adjoint of a colliding particle) it is compared
with the oracle's restatement of that variant. The scenes are those of the
oracle's own pins
(a compressed swirling block, the same block pushed into a wall so that
every prediction clamps some particles, and the block over a moving plane of
boundary samples), plus a 16,384-particle dam break at rest in its corner."""
import numpy as np
import pytest
import torch

import oracle
from physica_amd import scenes
from physica_amd.particle import DFSPH, ExecutionMode
from helpers import bits_equal, dfsph_setup, dfsph_state_numpy, reference_no_slip_oracle
from test_oracle_dfsph_derivatives import PARAMS, STATE, WARM, Case, case, directions, inputs

pytestmark = pytest.mark.gpu


def recorded(cs, device, nd, ne, warm, exact=True):
    cfg = DFSPH.Configuration(divergence_iterations=nd, density_iterations=ne, pressure_warm_start=warm,
                              exact_no_slip_adjoint=exact)
    solver, state, nxt, control, params, cache = dfsph_setup(cs.sc, device, cfg, warm=cs.warm,
                                                             mode=ExecutionMode.differentiable)
    solver.forward_step(state, control, params, nxt, cache)
    torch.cuda.synchronize()
    x, v, u, p, wd, we = inputs(cs, "f32")
    ref = oracle.dfsph_step(cs.d, 0, x, v, u, p, wd, we, nd, ne, warm)
    got = dfsph_state_numpy(nxt)
    for k in ("positions", "velocities") + WARM:
        assert bits_equal(got[k], ref[k]), k
    return solver, state, nxt, control, params, cache


def tensor(values, device, dtype):
    return torch.from_numpy(np.ascontiguousarray(values, dtype=dtype)).to(device)


def run_vjp(solver, state, nxt, control, params, cache, w, device):
    next_a = solver.allocate_state_adjoint()
    next_a.particles.positions.upload(w["positions"])
    next_a.particles.velocities.upload(w["velocities"])
    next_a.method.warm_divergence_pressure.copy_(tensor(w[WARM[0]], device, np.float64))
    next_a.method.warm_density_pressure.copy_(tensor(w[WARM[1]], device, np.float64))
    prev_a = solver.allocate_state_adjoint()
    ctl_a = solver.allocate_control_adjoint()
    par_a = solver.allocate_parameter_adjoint()
    solver.vjp_step(state, control, params, nxt, cache, next_a, prev_a, ctl_a, par_a)
    torch.cuda.synchronize()
    got = {"positions": prev_a.particles.positions.numpy(), "velocities": prev_a.particles.velocities.numpy(),
           "controls": ctl_a.external_accelerations.numpy()}
    for k in ("masses", "rest_densities", "viscosities"):
        got[k] = getattr(par_a.particles, k).cpu().numpy()
    for k in ("divergence_relaxation", "density_relaxation"):
        got[k] = getattr(par_a.method, k).cpu().numpy()
    got[WARM[0]] = prev_a.method.warm_divergence_pressure.cpu().numpy()
    got[WARM[1]] = prev_a.method.warm_density_pressure.cpu().numpy()
    return got


def run_jvp(solver, state, nxt, control, params, cache, t, device):
    st = solver.allocate_state_tangent()
    st.particles.positions.upload(t["positions"])
    st.particles.velocities.upload(t["velocities"])
    st.method.warm_divergence_pressure.copy_(tensor(t[WARM[0]], device, np.float32))
    st.method.warm_density_pressure.copy_(tensor(t[WARM[1]], device, np.float32))
    ct = solver.allocate_control_tangent()
    ct.external_accelerations.upload(t["controls"])
    pt = solver.allocate_parameter_tangent()
    for k in ("masses", "rest_densities", "viscosities"):
        getattr(pt.particles, k).copy_(tensor(t[k], device, np.float32))
    for k in ("divergence_relaxation", "density_relaxation"):
        getattr(pt.method, k).copy_(tensor(t[k], device, np.float32))
    nt = solver.allocate_state_tangent()
    solver.jvp_step(state, control, params, nxt, cache, st, ct, pt, nt)
    torch.cuda.synchronize()
    return {"positions": nt.particles.positions.numpy(), "velocities": nt.particles.velocities.numpy(),
            WARM[0]: nt.method.warm_divergence_pressure.cpu().numpy(),
            WARM[1]: nt.method.warm_density_pressure.cpu().numpy()}


CASES = [(0, 0, True, "free"), (1, 1, False, "free"), (2, 3, True, "free"), (4, 6, True, "free"),
         (2, 3, True, "wall"), (2, 2, True, "boundary")]


@pytest.mark.parametrize("nd,ne,warm,variant", CASES)
def test_vjp_matches_oracle(gpu_device, nd, ne, warm, variant):
    print(f"dfsph vjp nd={nd} ne={ne} warm={warm} {variant}", flush=True)
    cs = case(wall=variant == "wall", boundary=variant == "boundary")
    run = recorded(cs, gpu_device, nd, ne, warm)
    rng = np.random.default_rng(70 + nd + 7 * ne)
    _, w = directions(cs, rng)
    got = run_vjp(*run, w, gpu_device)
    x, v, u, p, wd, we = inputs(cs, "f32")
    exp = oracle.dfsph_vjp(cs.d, 0, x, v, u, p, wd, we, w, nd, ne, warm, "f32")
    for k in STATE + PARAMS + WARM:
        ref = exp[k]
        if not warm and k in WARM:
            assert not np.any(got[k]), k
            continue
        err = np.abs(got[k] - ref).max() / max(np.abs(ref).max(), 1e-300)
        assert err < 1e-10, (k, err)


@pytest.mark.parametrize("nd,ne,warm,variant", CASES)
def test_jvp_matches_oracle(gpu_device, nd, ne, warm, variant):
    print(f"dfsph jvp nd={nd} ne={ne} warm={warm} {variant}", flush=True)
    cs = case(wall=variant == "wall", boundary=variant == "boundary")
    run = recorded(cs, gpu_device, nd, ne, warm)
    rng = np.random.default_rng(80 + nd + 7 * ne)
    t, _ = directions(cs, rng)
    got = run_jvp(*run, t, gpu_device)
    x, v, u, p, wd, we = inputs(cs, "f32")
    exp = oracle.dfsph_jvp(cs.d, 0, x, v, u, p, wd, we, t, nd, ne, warm, "f32")
    for k in ("positions", "velocities") + WARM:
        err = np.abs(got[k] - exp[k]).max() / max(np.abs(exp[k]).max(), 1e-30)
        same = np.mean(got[k].view(np.uint32) == exp[k].view(np.uint32))
        print(f"  jvp {k}: max rel {err:.2e}, bit-identical {same:.3f}")
        assert err <= 1e-6, (k, err)


@pytest.mark.parametrize("boundary", [False, True])
def test_dot_product_identity_on_gpu(gpu_device, boundary):
    """<J t, w> from the FP32 JVP against <t, J^T w> from the FP64 VJP."""
    print(f"dfsph dot-product identity boundary={boundary}", flush=True)
    cs = case(wall=True, boundary=boundary)
    run = recorded(cs, gpu_device, 2, 3, True)
    rng = np.random.default_rng(90)
    t, w = directions(cs, rng)
    t = {k: np.asarray(t[k], dtype=np.float32).astype(np.float64) for k in t}
    jt = run_jvp(*run, t, gpu_device)
    jw = run_vjp(*run, w, gpu_device)
    left = sum(np.sum(jt[k].astype(np.float64) * w[k]) for k in ("positions", "velocities") + WARM)
    right = sum(np.sum(t[k] * jw[k]) for k in STATE + PARAMS + WARM)
    assert abs(left - right) <= 1e-5 * (abs(left) + abs(right)), (left, right)


def test_dam_break_vjp_matches_oracle(gpu_device):
    """16,384 particles at rest in the box corner: wall collisions in every
    prediction and full cells."""
    sc = scenes.dfsph_dam_break(shape=(32, 16, 32))
    d = oracle.make_domain(sc.domain)
    z = np.zeros(sc.count, dtype=np.float32)
    first = oracle.dfsph_step(d, 0, sc.positions, sc.velocities, np.zeros_like(sc.positions), sc.parameters, z, z, 2, 3)
    cs = Case(sc, d, (first[WARM[0]], first[WARM[1]]))
    print(f"dfsph dam break vjp n={sc.count}", flush=True)
    run = recorded(cs, gpu_device, 2, 3, True)
    rng = np.random.default_rng(100)
    _, w = directions(cs, rng)
    got = run_vjp(*run, w, gpu_device)
    x, v, u, p, wd, we = inputs(cs, "f32")
    exp = oracle.dfsph_vjp(cs.d, 0, x, v, u, p, wd, we, w, 2, 3, True, "f32")
    for k in STATE + PARAMS + WARM:
        err = np.abs(got[k] - exp[k]).max() / max(np.abs(exp[k]).max(), 1e-300)
        assert err < 1e-10, (k, err)


@pytest.mark.parametrize("scene", ["wall", "dam_break"])
def test_reference_no_slip_vjp_matches_oracle(gpu_device, scene):
    """The default configuration: the reference's VJP of a no-slip collision
    (sph/dfsph.cu:81) against the oracle's restatement of it, in scenes where
    predictions collide with the box; it differs from the exact transpose
    there."""
    print(f"dfsph reference no-slip vjp {scene}", flush=True)
    if scene == "wall":
        cs = case(wall=True)
    else:
        sc = scenes.dfsph_dam_break(shape=(32, 16, 32))
        d = oracle.make_domain(sc.domain)
        z = np.zeros(sc.count, dtype=np.float32)
        first = oracle.dfsph_step(d, 0, sc.positions, sc.velocities, np.zeros_like(sc.positions), sc.parameters, z,
                                  z, 2, 3)
        cs = Case(sc, d, (first[WARM[0]], first[WARM[1]]))
    assert DFSPH.Configuration().exact_no_slip_adjoint is False
    run = recorded(cs, gpu_device, 2, 3, True, exact=False)
    rng = np.random.default_rng(110)
    _, w = directions(cs, rng)
    got = run_vjp(*run, w, gpu_device)
    x, v, u, p, wd, we = inputs(cs, "f32")
    with reference_no_slip_oracle():
        exp = oracle.dfsph_vjp(cs.d, 0, x, v, u, p, wd, we, w, 2, 3, True, "f32")
    exact = oracle.dfsph_vjp(cs.d, 0, x, v, u, p, wd, we, w, 2, 3, True, "f32")
    assert not np.array_equal(exp["velocities"], exact["velocities"])
    for k in STATE + PARAMS + WARM:
        err = np.abs(got[k] - exp[k]).max() / max(np.abs(exp[k]).max(), 1e-300)
        assert err < 1e-10, (k, err)


def test_forward_mode_rejects_derivatives(gpu_device):
    cs = case(shape=(4, 4, 4))
    solver, state, nxt, control, params, cache = dfsph_setup(cs.sc, gpu_device, DFSPH.Configuration(
        divergence_iterations=1, density_iterations=1), warm=cs.warm)
    solver.forward_step(state, control, params, nxt, cache)
    with pytest.raises(Exception, match="differentiable"):
        solver.vjp_step(state, control, params, nxt, cache, solver.allocate_state_adjoint(),
                        solver.allocate_state_adjoint(), solver.allocate_control_adjoint(),
                        solver.allocate_parameter_adjoint())
