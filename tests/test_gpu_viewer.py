"""The viewer export of the reference's PBF example (spectra provider,
examples/fluids/liquid/pbf-dam-break/spectra/provider.ixx:74-96): SoA to
interleaved Float3 positions and velocities (write_vectors), and the step
cache's vorticity magnitudes, against numpy and the oracle, bit for bit."""
import numpy as np
import pytest
import torch

import oracle
from physica_amd import scenes, write_vectors
from physica_amd.particle import PBF, ExecutionMode, VectorField
from helpers import bits_equal, pbf_setup

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 255, 256, 1_000_003])
def test_write_vectors_interleaves(gpu_device, n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=(3, n)).astype(np.float32)
    v = rng.normal(size=(3, n)).astype(np.float32)
    xf, vf = VectorField.zeros(n, gpu_device), VectorField.zeros(n, gpu_device)
    xf.upload(x)
    vf.upload(v)
    pos, vel = write_vectors(xf, vf)
    torch.cuda.synchronize()
    assert bits_equal(pos.cpu().numpy(), np.ascontiguousarray(x.T))
    assert bits_equal(vel.cpu().numpy(), np.ascontiguousarray(v.T))


@pytest.mark.parametrize("mode", [ExecutionMode.forward, ExecutionMode.differentiable])
def test_vorticity_magnitudes_match_oracle(gpu_device, mode):
    sc = scenes.pbf_dam_break(shape=(16, 16, 16))
    n = sc.count
    rng = np.random.default_rng(5)
    sc.velocities = rng.normal(0, 0.5, size=(3, n)).astype(np.float32)
    sc.parameters["vorticity_confinement"] = rng.uniform(0, 0.5, n).astype(np.float32)
    solver, state, nxt, control, params, cache = pbf_setup(sc, gpu_device, PBF.Configuration(3), mode=mode)
    solver.forward_step(state, control, params, nxt, cache)
    got = cache.vorticity_magnitudes()
    torch.cuda.synchronize()
    exp = oracle.pbf_step(oracle.make_domain(sc.domain), 0, sc.positions, sc.velocities, np.zeros((3, n)),
                          sc.parameters, 3)
    assert np.abs(exp["vorticity_magnitudes"]).max() > 0.0
    assert bits_equal(got.cpu().numpy(), exp["vorticity_magnitudes"])
