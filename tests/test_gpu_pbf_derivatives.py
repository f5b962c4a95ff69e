"""PBF::jvp_step and PBF::vjp_step on the GPU against the oracle.

With the artificial pressure off, the recorded forward step is bit-exact
against the FP32 oracle, and the VJP (FP32 primal, FP64 adjoints) agrees with
the oracle's to 1e-10 of the largest entry of each output (the same
per-particle sums; only the order in which the final accumulation adds the
parts differs), except the artificial-pressure parameter adjoints, whose
derivatives go through powf / logf (1e-6). The JVP (FP32, as the reference's) agrees to 2e-4 (the
strength derivative goes through powf). With the artificial pressure on, the
forward step itself differs by powf ulps, so the VJP is compared within
1e-5. Rollout gradients: the chained GPU VJPs against the oracle's chain
(1e-9), and against central differences of the FP64 oracle rollout on a
compressed block, whose constraints keep the rollout smooth (1e-3)."""
import numpy as np
import pytest
import torch

import oracle
from physica_amd import scenes
from physica_amd.particle import ExecutionMode, PBF
from helpers import bits_equal, pbf_setup

pytestmark = pytest.mark.gpu

PARAMS = oracle.PBF_PARAMETERS


def swirl_scene(shape=(8, 8, 6), seed=3, strength=0.0):
    """Vorticity confinement and XSPH active, particles clear of the walls."""
    sc = scenes.pbf_dam_break(shape=shape, seed=seed)
    n = sc.count
    rng = np.random.default_rng(seed)
    sc.positions = (sc.positions + np.array([[1.2], [0.7], [0.5]], dtype=np.float32)).astype(np.float32)
    r = sc.positions - sc.positions.mean(axis=1, keepdims=True)
    sc.velocities = (np.stack([-r[2], 0.2 * r[0], r[0]]) * 2.0 + rng.normal(0, 0.05, size=(3, n))).astype(np.float32)
    p = sc.parameters
    p["artificial_pressure_strength"] = np.full(n, strength, dtype=np.float32)
    p["xsph_viscosity"] = rng.uniform(0.005, 0.02, n).astype(np.float32)
    p["vorticity_confinement"] = rng.uniform(0.5, 2.0, n).astype(np.float32)
    p["masses"] = (p["masses"] * rng.uniform(0.95, 1.05, n)).astype(np.float32)
    p["rest_densities"] = (p["rest_densities"] * rng.uniform(0.97, 1.03, n)).astype(np.float32)
    return sc


def forward_recorded(sc, device, iterations):
    solver, state, nxt, control, params, cache = pbf_setup(sc, device, PBF.Configuration(pressure_iterations=iterations),
                                                           ExecutionMode.differentiable)
    solver.forward_step(state, control, params, nxt, cache)
    torch.cuda.synchronize()
    return solver, state, nxt, control, params, cache


def to_param_tensor(values, device, dtype):
    return torch.from_numpy(np.ascontiguousarray(values, dtype=dtype)).to(device)


@pytest.mark.parametrize("iterations,strength,bound", [(0, 0.0, 1e-10), (1, 0.0, 1e-10), (4, 0.0, 1e-10),
                                                        (1, 0.05, 1e-5)])
def test_vjp_matches_oracle(gpu_device, iterations, strength, bound):
    print(f"pbf vjp iterations={iterations} strength={strength}", flush=True)
    sc = swirl_scene(strength=strength)
    n = sc.count
    solver, state, nxt, control, params, cache = forward_recorded(sc, gpu_device, iterations)
    d = oracle.make_domain(sc.domain)
    zeros = np.zeros_like(sc.positions)
    fwd = oracle.pbf_step(d, 0, sc.positions, sc.velocities, zeros, sc.parameters, iterations)
    if strength == 0.0:
        assert bits_equal(nxt.particles.positions.numpy(), fwd["positions"])
        assert bits_equal(nxt.particles.velocities.numpy(), fwd["velocities"])
    else:
        assert np.abs(nxt.particles.positions.numpy() - fwd["positions"]).max() < 1e-6
    rng = np.random.default_rng(7 + iterations)
    wx, wv = rng.normal(0, 1, size=(3, n)), rng.normal(0, 1, size=(3, n))
    next_a = solver.allocate_state_adjoint()
    next_a.particles.positions.upload(wx)
    next_a.particles.velocities.upload(wv)
    prev_a = solver.allocate_state_adjoint()
    ctl_a = solver.allocate_control_adjoint()
    par_a = solver.allocate_parameter_adjoint()
    solver.vjp_step(state, control, params, nxt, cache, next_a, prev_a, ctl_a, par_a)
    torch.cuda.synchronize()
    exp = oracle.pbf_vjp(d, 0, sc.positions, sc.velocities, zeros, sc.parameters, wx, wv, iterations)
    got = {"positions": prev_a.particles.positions.numpy(), "velocities": prev_a.particles.velocities.numpy(),
           "controls": ctl_a.external_accelerations.numpy(), "masses": par_a.particles.masses.cpu().numpy(),
           "rest_densities": par_a.particles.rest_densities.cpu().numpy()}
    for k in PARAMS[2:]:
        got[k] = getattr(par_a, k).cpu().numpy()
    # the artificial-pressure parameter derivatives go through powf / logf
    powf_terms = ("artificial_pressure_strength", "artificial_pressure_exponent", "artificial_pressure_radius")
    for k, ref in exp.items():
        scale = max(np.abs(ref).max(), 1e-300)
        err = np.abs(got[k] - ref).max() / scale
        assert err < (max(bound, 1e-6) if k in powf_terms else bound), (k, err)
    # outputs accumulate
    solver.vjp_step(state, control, params, nxt, cache, next_a, prev_a, ctl_a, par_a)
    torch.cuda.synchronize()
    np.testing.assert_allclose(prev_a.particles.velocities.numpy(), 2 * got["velocities"], rtol=1e-12, atol=0)


@pytest.mark.parametrize("iterations", [1, 4])
def test_jvp_matches_oracle(gpu_device, iterations):
    print(f"pbf jvp iterations={iterations}", flush=True)
    sc = swirl_scene()
    n = sc.count
    solver, state, nxt, control, params, cache = forward_recorded(sc, gpu_device, iterations)
    rng = np.random.default_rng(11 + iterations)
    t = {"positions": rng.normal(0, 1e-3, size=(3, n)), "velocities": rng.normal(0, 1e-2, size=(3, n)),
         "controls": rng.normal(0, 1, size=(3, n))}
    for k in PARAMS:
        t[k] = rng.normal(0, 1, n) * 0.1 * np.abs(sc.parameters[k].astype(np.float64)).mean()
    st = solver.allocate_state_tangent()
    st.particles.positions.upload(t["positions"])
    st.particles.velocities.upload(t["velocities"])
    ct = solver.allocate_control_tangent()
    ct.external_accelerations.upload(t["controls"])
    pt = solver.allocate_parameter_tangent()
    pt.particles.masses.copy_(to_param_tensor(t["masses"], gpu_device, np.float32))
    pt.particles.rest_densities.copy_(to_param_tensor(t["rest_densities"], gpu_device, np.float32))
    for k in PARAMS[2:]:
        getattr(pt, k).copy_(to_param_tensor(t[k], gpu_device, np.float32))
    out = solver.allocate_state_tangent()
    solver.jvp_step(state, control, params, nxt, cache, st, ct, pt, out)
    torch.cuda.synchronize()
    d = oracle.make_domain(sc.domain)
    t32 = {k: np.asarray(v, dtype=np.float32) for k, v in t.items()}
    exp = oracle.pbf_jvp(d, 0, sc.positions, sc.velocities, np.zeros_like(sc.positions), sc.parameters, t32,
                         iterations)
    for k, got in (("positions", out.particles.positions.numpy()), ("velocities", out.particles.velocities.numpy())):
        err = np.abs(got - exp[k]).max() / np.abs(exp[k]).max()
        same = float(np.mean(got.view(np.uint32) == exp[k].view(np.uint32)))
        print(f"  jvp {k}: max rel {err:.2e}, bit-identical {same:.3f}", flush=True)
        assert err < 2e-4, (k, err)


def test_dot_product_identity_on_gpu(gpu_device):
    sc = swirl_scene(shape=(6, 6, 6))
    n = sc.count
    solver, state, nxt, control, params, cache = forward_recorded(sc, gpu_device, 3)
    rng = np.random.default_rng(5)
    tx, tv = rng.normal(0, 1e-3, size=(3, n)), rng.normal(0, 1e-2, size=(3, n))
    wx, wv = rng.normal(0, 1, size=(3, n)), rng.normal(0, 1, size=(3, n))
    st = solver.allocate_state_tangent()
    st.particles.positions.upload(tx)
    st.particles.velocities.upload(tv)
    out = solver.allocate_state_tangent()
    solver.jvp_step(state, control, params, nxt, cache, st, solver.allocate_control_tangent(),
                    solver.allocate_parameter_tangent(), out)
    next_a = solver.allocate_state_adjoint()
    next_a.particles.positions.upload(wx)
    next_a.particles.velocities.upload(wv)
    prev_a = solver.allocate_state_adjoint()
    solver.vjp_step(state, control, params, nxt, cache, next_a, prev_a, solver.allocate_control_adjoint(),
                    solver.allocate_parameter_adjoint())
    torch.cuda.synchronize()
    left = np.sum(out.particles.positions.numpy() * wx) + np.sum(out.particles.velocities.numpy() * wv)
    right = np.sum(np.float32(tx).astype(np.float64) * prev_a.particles.positions.numpy()) + \
        np.sum(np.float32(tv).astype(np.float64) * prev_a.particles.velocities.numpy())
    assert abs(left - right) <= 1e-4 * (abs(left) + abs(right))


def rollout_gpu(sc, device, steps, target):
    """Forward `steps` recorded steps, then chain the VJPs from
    L = mean |x_T - y|^2; returns (adjoint of v0, adjoint of x0)."""
    n = sc.count
    solver, state, nxt, control, params, _ = pbf_setup(sc, device, PBF.Configuration(pressure_iterations=4),
                                                      ExecutionMode.differentiable)
    states, caches = [state], []
    for _ in range(steps):
        cache = solver.allocate_step_cache()
        nxt = solver.allocate_state()
        solver.forward_step(states[-1], control, params, nxt, cache)
        states.append(nxt)
        caches.append(cache)
    torch.cuda.synchronize()
    xT = states[-1].particles.positions.numpy().astype(np.float64)
    adj = solver.allocate_state_adjoint()
    adj.particles.positions.upload(2.0 * (xT - target) / n)
    for k in reversed(range(steps)):
        prev = solver.allocate_state_adjoint()
        solver.vjp_step(states[k], control, params, states[k + 1], caches[k], adj, prev,
                        solver.allocate_control_adjoint(), solver.allocate_parameter_adjoint())
        adj = prev
    torch.cuda.synchronize()
    return adj.particles.velocities.numpy(), adj.particles.positions.numpy()


def rollout_oracle(sc, steps, target, precision):
    n = sc.count
    d = oracle.make_domain(sc.domain)
    zeros = np.zeros((3, n))
    dt = np.float32 if precision == "f32" else np.float64
    p = {k: v.astype(dt) for k, v in sc.parameters.items()}
    xs, vs = [sc.positions.astype(dt)], [sc.velocities.astype(dt)]
    for k in range(steps):
        r = oracle.pbf_step(d, k, xs[-1], vs[-1], zeros, p, 4, precision)
        xs.append(r["positions"])
        vs.append(r["velocities"])
    ax, av = 2.0 * (xs[-1].astype(np.float64) - target) / n, np.zeros((3, n))
    for k in reversed(range(steps)):
        a = oracle.pbf_vjp(d, k, xs[k], vs[k], zeros, p, ax, av, 4, precision)
        ax, av = a["positions"], a["velocities"]
    return av, ax


def test_rollout_adjoint_matches_oracle_chain(gpu_device):
    """10 steps of a 6^3 dam-break block: bit-exact forward, so the chained
    FP64 adjoints agree with the oracle's chain to 1e-9."""
    sc = scenes.pbf_dam_break(shape=(6, 6, 6), seed=20260004)
    n = sc.count
    rng = np.random.default_rng(20260104)
    target = (sc.positions + np.array([[0.05], [-0.02], [0.04]]) + rng.normal(0, 0.01, size=(3, n))).astype(np.float64)
    gv, gx = rollout_gpu(sc, gpu_device, 10, target)
    ov, ox = rollout_oracle(sc, 10, target, "f32")
    for got, ref in ((gv, ov), (gx, ox)):
        assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()


def test_rollout_gradient_matches_finite_differences(gpu_device):
    """Loss L = mean |x_T - y|^2 after 5 steps of a 6^3 block compressed to 0.9
    of its rest spacing (the constraints push it apart smoothly); dL/dv0 from
    the chained GPU VJPs against central differences (eps 1e-6) of the FP64
    oracle rollout along 4 seeded unit-RMS directions. Error relative to the
    RMS of the four directional derivatives <= 1e-3 (a single projection can
    nearly cancel; the FP32 primal alone contributes ~5e-4)."""
    sc = scenes.pbf_dam_break(shape=(6, 6, 6), spacing=0.045, seed=20260004)
    sc.parameters["masses"][:] = np.float32(1000.0 * 0.05**3)
    sc.positions = (sc.positions + np.array([[0.8], [0.4], [0.5]], dtype=np.float32)).astype(np.float32)
    n = sc.count
    steps = 5
    rng = np.random.default_rng(20260104)
    target = (sc.positions + np.array([[0.05], [-0.02], [0.04]]) + rng.normal(0, 0.01, size=(3, n))).astype(np.float64)
    grad_v, _ = rollout_gpu(sc, gpu_device, steps, target)
    d = oracle.make_domain(sc.domain)
    zeros = np.zeros((3, n))
    p64 = {k: v.astype(np.float64) for k, v in sc.parameters.items()}

    def loss(v0):
        x, v = sc.positions.astype(np.float64), v0
        for k in range(steps):
            r = oracle.pbf_step(d, k, x, v, zeros, p64, 4, "f64")
            x, v = r["positions"], r["velocities"]
        return float(np.mean(np.sum((x - target) ** 2, axis=0)))

    v0 = sc.velocities.astype(np.float64)
    pairs = []
    for trial in range(4):
        direction = rng.normal(0, 1, size=(3, n))
        direction /= np.sqrt(np.mean(direction**2))
        eps = 1e-6
        fd = (loss(v0 + eps * direction) - loss(v0 - eps * direction)) / (2 * eps)
        got = float(np.sum(grad_v * direction))
        print(f"  rollout direction {trial}: adjoint {got:.6e} fd {fd:.6e}", flush=True)
        pairs.append((got, fd))
    rms = np.sqrt(np.mean([fd * fd for _, fd in pairs]))
    worst = max(abs(got - fd) for got, fd in pairs) / rms
    assert worst <= 1e-3, worst
