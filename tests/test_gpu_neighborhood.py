"""GPU neighbour search and reference-layout density vs the oracle: keys,
sorted order and cell offsets bit-exact, densities bit-exact."""
import numpy as np
import pytest
import torch

import oracle
from physica_amd import build_neighborhood, density_forward
from physica_amd.particle import BoxBoundary, DomainConfiguration, ParticleParameters, VectorField
from helpers import bits_equal

pytestmark = pytest.mark.gpu


def domain(n, h=0.05, box=(1.0, 0.6, 0.8), vel=(0.0, 0.0, 0.0), dt=1e-2):
    return DomainConfiguration(n, dt, h, 0.25 * h, (0.0, -9.81, 0.0), BoxBoundary((0, 0, 0), box, vel, True))


def positions(kind, n, rng, box):
    b = np.array(box)[:, None]
    if kind == "uniform":
        return (rng.uniform(0.0, 1.0, size=(3, n)) * b).astype(np.float32)
    if kind == "outside":  # a third of the particles lie outside the box and are clamped into edge cells
        return (rng.uniform(-0.5, 1.5, size=(3, n)) * b).astype(np.float32)
    if kind == "one_cell":  # every particle in the same cell
        return (np.full((3, n), 0.3) + rng.uniform(0, 0.01, size=(3, n))).astype(np.float32)
    if kind == "cell_faces":  # exactly on cell boundaries
        k = rng.integers(0, 12, size=(3, n))
        return (k * np.float32(0.05)).astype(np.float32)
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["uniform", "outside", "one_cell", "cell_faces"])
@pytest.mark.parametrize("n", [0, 1, 777, 65536])
def test_build_neighborhood_bit_exact(gpu_device, kind, n):
    print(f"neighbourhood case kind={kind} n={n}", flush=True)
    if kind == "one_cell" and n > 4096:
        n = 4096  # the in-cell rank is O(k) per particle
    cfg = domain(n)
    rng = np.random.default_rng(n + len(kind))
    pos = positions(kind, n, rng, (1.0, 0.6, 0.8))
    field = VectorField.zeros(n, gpu_device)
    field.upload(pos)
    keys, idx, off, view = build_neighborhood(cfg, 3, field, gpu_device)
    torch.cuda.synchronize()
    ok, oi, oo = oracle.build_neighborhood(oracle.make_domain(cfg), 3, pos)
    assert np.array_equal(keys.cpu().numpy().astype(np.uint64), ok)
    assert np.array_equal(idx.cpu().numpy().astype(np.uint32), oi)
    assert np.array_equal(off.cpu().numpy().astype(np.uint32), oo)
    assert (view.cells_x, view.cells_y, view.cells_z) == cfg.grid_resolution()


def test_moving_box(gpu_device):
    n = 5000
    cfg = domain(n, vel=(0.3, -0.2, 0.1), dt=0.05)
    rng = np.random.default_rng(5)
    pos = positions("outside", n, rng, (1.0, 0.6, 0.8))
    field = VectorField.zeros(n, gpu_device)
    field.upload(pos)
    for step in (0, 7):
        keys, idx, off, view = build_neighborhood(cfg, step, field, gpu_device)
        ok, oi, oo = oracle.build_neighborhood(oracle.make_domain(cfg), step, pos)
        assert np.array_equal(idx.cpu().numpy().astype(np.uint32), oi)
        assert np.array_equal(off.cpu().numpy().astype(np.uint32), oo)
        assert np.array_equal(keys.cpu().numpy().astype(np.uint64), ok)


@pytest.mark.parametrize("pbf_kernel", [False, True])
def test_reference_layout_density_bit_exact(gpu_device, pbf_kernel):
    n = 20000
    cfg = domain(n)
    rng = np.random.default_rng(9)
    topo = positions("uniform", n, rng, (1.0, 0.6, 0.8))
    pos = (topo + rng.uniform(-0.003, 0.003, size=topo.shape)).astype(np.float32)
    m = rng.uniform(0.5, 1.5, size=n).astype(np.float32)
    tf, pf = VectorField.zeros(n, gpu_device), VectorField.zeros(n, gpu_device)
    tf.upload(topo)
    pf.upload(pos)
    z = torch.zeros(n, device=gpu_device)
    params = ParticleParameters(torch.from_numpy(m).to(gpu_device), z + 1000.0, z, z)
    keys, idx, off, view = build_neighborhood(cfg, 0, tf, gpu_device)
    rho = torch.empty(n, device=gpu_device)
    density_forward(cfg.support_radius, pbf_kernel, tf, pf, params, view, rho)
    expected = oracle.density(oracle.make_domain(cfg), 0, pbf_kernel, topo, pos, m, np.full(n, 1000.0, np.float32))
    assert bits_equal(rho.cpu().numpy(), expected)
