"""PBF with crowded cells on the GPU against the oracle.

A particle whose 27 stencil cells include one of more than 16 particles walks
the full stencil in every pass. A wave whose crowded particles all sit in one
cell walks that stencil together: each lane loads one candidate of a run of 64
and the candidates reach every lane from scalar registers, while each lane
still adds its own terms in the reference order. The scene clumps a few
hundred particles into single cells, so whole waves, waves that mix crowded
and uncrowded lanes, and waves whose crowded lanes span two cells all occur.
Forward steps are bit-exact against the FP32 oracle (forward and
differentiable mode, and in the batched engine), the VJP is within 1e-10 of
the oracle's."""
import numpy as np
import pytest
import torch

import oracle
from physica_amd import scenes
from physica_amd.batch import PbfBatch
from physica_amd.particle import ExecutionMode, PBF
from helpers import bits_equal, pbf_setup

pytestmark = pytest.mark.gpu

PARAMS = oracle.PBF_PARAMETERS
ITERATIONS = 3
# (cell, particles moved into it): one cell above 256 particles (whole
# waves), one around two waves, one just above the mask limit
CLUMPS = (((3, 2, 4), 300), ((1, 4, 1), 150), ((4, 1, 2), 40))


def clumped_scene(seed=5):
    """A 12^3 PBF block (h 0.1, eight particles a cell) with three clumps,
    random velocities, XSPH and vorticity confinement on, artificial pressure
    off (the forward step is then bit-exact)."""
    sc = scenes.pbf_dam_break(shape=(12, 12, 12), seed=seed)
    n = sc.count
    rng = np.random.default_rng(seed)
    pos = sc.positions.copy()
    moved = rng.permutation(n)
    start = 0
    h = sc.domain.support_radius
    for cell, count in CLUMPS:
        chosen = moved[start:start + count]
        start += count
        lo = np.array(cell, dtype=np.float64) * h + 0.05 * h
        pos[:, chosen] = (lo[:, None] + rng.uniform(0.0, 0.9 * h, size=(3, count))).astype(np.float32)
    sc.positions = pos
    sc.velocities = rng.normal(0, 0.3, size=(3, n)).astype(np.float32)
    p = sc.parameters
    p["xsph_viscosity"] = rng.uniform(0.005, 0.02, n).astype(np.float32)
    p["vorticity_confinement"] = rng.uniform(0.5, 2.0, n).astype(np.float32)
    p["masses"] = (p["masses"] * rng.uniform(0.95, 1.05, n)).astype(np.float32)
    return sc


def test_clumps_fill_whole_waves():
    """The scene has the cases the wave-shared walk distinguishes."""
    sc = clumped_scene()
    h = sc.domain.support_radius
    cells = np.floor(sc.positions / h).astype(np.int64)
    key = (cells[2] * 64 + cells[1]) * 64 + cells[0]
    _, counts = np.unique(key, return_counts=True)
    assert counts.max() >= 256 and (counts > 16).sum() >= 3


@pytest.mark.parametrize("mode", [ExecutionMode.forward, ExecutionMode.differentiable])
def test_crowded_steps_bit_exact(gpu_device, mode):
    print(f"pbf crowded forward mode={mode}", flush=True)
    sc = clumped_scene()
    solver, state, nxt, control, params, cache = pbf_setup(sc, gpu_device, PBF.Configuration(ITERATIONS), mode)
    d = oracle.make_domain(sc.domain)
    x, v = sc.positions, sc.velocities
    for step in range(3):
        solver.forward_step(state, control, params, nxt, cache)
        torch.cuda.synchronize()
        exp = oracle.pbf_step(d, step, x, v, np.zeros_like(x), sc.parameters, ITERATIONS)
        gx, gv = nxt.particles.positions.numpy(), nxt.particles.velocities.numpy()
        assert bits_equal(gx, exp["positions"]), step
        assert bits_equal(gv, exp["velocities"]), step
        rho, lam = cache.last_iteration()
        assert bits_equal(rho.cpu().numpy(), exp["densities"]), step
        assert bits_equal(lam.cpu().numpy(), exp["lambdas"]), step
        assert bits_equal(cache.vorticity_magnitudes().cpu().numpy(), exp["vorticity_magnitudes"]), step
        x, v = gx, gv
        solver.copy_state(nxt, state)


def test_crowded_vjp_matches_oracle(gpu_device):
    print("pbf crowded vjp", flush=True)
    sc = clumped_scene(seed=6)
    n = sc.count
    solver, state, nxt, control, params, cache = pbf_setup(sc, gpu_device, PBF.Configuration(ITERATIONS),
                                                           ExecutionMode.differentiable)
    solver.forward_step(state, control, params, nxt, cache)
    torch.cuda.synchronize()
    d = oracle.make_domain(sc.domain)
    zeros = np.zeros_like(sc.positions)
    fwd = oracle.pbf_step(d, 0, sc.positions, sc.velocities, zeros, sc.parameters, ITERATIONS)
    assert bits_equal(nxt.particles.positions.numpy(), fwd["positions"])
    rng = np.random.default_rng(17)
    wx, wv = rng.normal(0, 1, size=(3, n)), rng.normal(0, 1, size=(3, n))
    next_a = solver.allocate_state_adjoint()
    next_a.particles.positions.upload(wx)
    next_a.particles.velocities.upload(wv)
    prev_a = solver.allocate_state_adjoint()
    ctl_a = solver.allocate_control_adjoint()
    par_a = solver.allocate_parameter_adjoint()
    solver.vjp_step(state, control, params, nxt, cache, next_a, prev_a, ctl_a, par_a)
    torch.cuda.synchronize()
    exp = oracle.pbf_vjp(d, 0, sc.positions, sc.velocities, zeros, sc.parameters, wx, wv, ITERATIONS)
    got = {"positions": prev_a.particles.positions.numpy(), "velocities": prev_a.particles.velocities.numpy(),
           "controls": ctl_a.external_accelerations.numpy(), "masses": par_a.particles.masses.cpu().numpy(),
           "rest_densities": par_a.particles.rest_densities.cpu().numpy()}
    for k in PARAMS[2:]:
        got[k] = getattr(par_a, k).cpu().numpy()
    powf_terms = ("artificial_pressure_strength", "artificial_pressure_exponent", "artificial_pressure_radius")
    for k, ref in exp.items():
        scale = max(np.abs(ref).max(), 1e-300)
        err = np.abs(got[k] - ref).max() / scale
        assert err < (1e-6 if k in powf_terms else 1e-10), (k, err)


def test_crowded_batch_equals_the_oracle(gpu_device):
    print("pbf crowded batch", flush=True)
    batch = [clumped_scene(seed=5), clumped_scene(seed=8)]
    engine = PbfBatch([sc.domain for sc in batch], PBF.Configuration(pressure_iterations=ITERATIONS),
                      use_graph=True, device=gpu_device)
    state, control, params = engine.allocate_state(), engine.allocate_control(), engine.allocate_parameters()
    engine.upload(batch, state, params)
    engine.run(state, control, params, 2)
    torch.cuda.synchronize()
    x_all, v_all = state.particles.positions.numpy(), state.particles.velocities.numpy()
    for k, sc in enumerate(batch):
        d = oracle.make_domain(sc.domain)
        x, v = sc.positions, sc.velocities
        for step in range(2):
            r = oracle.pbf_step(d, step, x, v, np.zeros_like(x), sc.parameters, ITERATIONS)
            x, v = r["positions"], r["velocities"]
        assert bits_equal(x_all[:, engine.scene_slice(k)], x), k
        assert bits_equal(v_all[:, engine.scene_slice(k)], v), k
