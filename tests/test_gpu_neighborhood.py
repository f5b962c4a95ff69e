"""GPU neighbour search and reference-layout density vs the oracle: keys,
sorted order and cell offsets bit-exact, for the particles and for static
and moving boundary samples; densities bit-exact, boundary terms included."""
import numpy as np
import pytest
import torch

import oracle
from physica_amd import DeviceBoundary, build_neighborhood, density_forward
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
    cfg = domain(n)
    rng = np.random.default_rng(n + len(kind))
    pos = positions(kind, n, rng, (1.0, 0.6, 0.8))
    field = VectorField.zeros(n, gpu_device)
    field.upload(pos)
    keys, idx, off, view = build_neighborhood(cfg, 3, field, gpu_device)[:4]
    torch.cuda.synchronize()
    ok, oi, oo = oracle.build_neighborhood(oracle.make_domain(cfg), 3, pos)
    assert np.array_equal(keys.cpu().numpy().astype(np.uint64), ok)
    assert np.array_equal(idx.cpu().numpy().astype(np.uint32), oi)
    assert np.array_equal(off.cpu().numpy().astype(np.uint32), oo)
    assert (view.cells_x, view.cells_y, view.cells_z) == cfg.grid_resolution()


@pytest.mark.parametrize("n", [129, 5000, 1_000_000])
def test_every_particle_in_one_cell(gpu_device, n):
    """Cells with more than 128 particles are ordered by a radix sort of
    their indices (one workgroup per cell): up to every particle of a
    1,000,000-particle scene in one cell, next to ordinary cells, the order
    is still the stable sort of the keys and the offsets count every cell."""
    print(f"neighbourhood one cell n={n}", flush=True)
    rng = np.random.default_rng(n)
    pos = (np.full((3, n), 0.3) + rng.uniform(0, 0.01, size=(3, n))).astype(np.float32)
    # a tenth spread over the box, so small and large cells mix
    spread = n // 10
    pos[:, :spread] = (rng.uniform(0.0, 1.0, size=(3, spread)) * np.array([[1.0], [0.6], [0.8]])).astype(np.float32)
    cfg = domain(n)
    field = VectorField.zeros(n, gpu_device)
    field.upload(pos)
    keys, idx, off, view = build_neighborhood(cfg, 0, field, gpu_device)[:4]
    torch.cuda.synchronize()
    ok, oi, oo = oracle.build_neighborhood(oracle.make_domain(cfg), 0, pos)
    assert np.array_equal(keys.cpu().numpy().astype(np.uint64), ok)
    assert np.array_equal(idx.cpu().numpy().astype(np.uint32), oi)
    assert np.array_equal(off.cpu().numpy().astype(np.uint32), oo)
    assert np.diff(oo).max() >= n - spread - 1


def test_moving_box(gpu_device):
    n = 5000
    cfg = domain(n, vel=(0.3, -0.2, 0.1), dt=0.05)
    rng = np.random.default_rng(5)
    pos = positions("outside", n, rng, (1.0, 0.6, 0.8))
    field = VectorField.zeros(n, gpu_device)
    field.upload(pos)
    for step in (0, 7):
        keys, idx, off, view = build_neighborhood(cfg, step, field, gpu_device)[:4]
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
    keys, idx, off, view = build_neighborhood(cfg, 0, tf, gpu_device)[:4]
    rho = torch.empty(n, device=gpu_device)
    density_forward(cfg.support_radius, pbf_kernel, tf, pf, params, view, rho)
    expected = oracle.density(oracle.make_domain(cfg), 0, pbf_kernel, topo, pos, m, np.full(n, 1000.0, np.float32))
    assert bits_equal(rho.cpu().numpy(), expected)


def boundary_samples(rng, count, box=(1.0, 0.6, 0.8), moving=False):
    """A floor plane and a wall plane of samples (ascending original index
    mixed over cells), plus a few outside the box; volumes ~ spacing^3."""
    k = count // 2
    floor = np.stack([rng.uniform(0, box[0], k), np.zeros(k), rng.uniform(0, box[2], k)])
    wall = np.stack([np.full(count - k, box[0]), rng.uniform(0, box[1], count - k), rng.uniform(0, box[2], count - k)])
    pos = np.concatenate([floor, wall], axis=1)
    outside = min(7, count)
    pos[:, :outside] = rng.uniform(-0.2, 1.2, size=(3, outside)) * np.array(box)[:, None]
    order = rng.permutation(count)
    pos = pos[:, order].astype(np.float32)
    vel = np.zeros((3, count), np.float32)
    if moving:
        vel[:] = rng.uniform(-0.5, 0.5, size=(3, count)).astype(np.float32)
    vol = rng.uniform(0.5, 1.5, count).astype(np.float32) * np.float32(0.02 ** 3)
    return pos, vel, vol


def with_boundary(cfg, samples):
    pos, vel, vol = samples
    cfg.boundary_particles = [(tuple(pos[:, k]), tuple(vel[:, k]), float(vol[k])) for k in range(vol.shape[0])]
    return cfg


@pytest.mark.parametrize("moving", [False, True])
@pytest.mark.parametrize("count", [1, 3000])
def test_boundary_neighborhood_bit_exact(gpu_device, moving, count):
    print(f"boundary neighbourhood moving={moving} count={count}", flush=True)
    n = 4000
    rng = np.random.default_rng(31 + count + moving)
    samples = boundary_samples(rng, count, moving=moving)
    cfg = with_boundary(domain(n, vel=(0.2, 0.0, -0.1) if moving else (0.0, 0.0, 0.0), dt=0.05), samples)
    pos = positions("uniform", n, rng, (1.0, 0.6, 0.8))
    field = VectorField.zeros(n, gpu_device)
    field.upload(pos)
    od = oracle.make_domain(cfg, samples)
    for step in (0, 5):
        nb = build_neighborhood(cfg, step, field, gpu_device)
        torch.cuda.synchronize()
        ok, oi, oo, bk, bi, bo = oracle.build_neighborhood(od, step, pos, boundary=True)
        assert np.array_equal(nb.indices.cpu().numpy().astype(np.uint32), oi)
        assert np.array_equal(nb.offsets.cpu().numpy().astype(np.uint32), oo)
        assert np.array_equal(nb.boundary_keys.cpu().numpy().astype(np.uint64), bk)
        assert np.array_equal(nb.boundary_indices.cpu().numpy().astype(np.uint32), bi)
        assert np.array_equal(nb.boundary_offsets.cpu().numpy().astype(np.uint32), bo)
        assert nb.view.boundary_count == count
        assert np.float32(nb.boundary_time) == np.float32(step) * np.float32(cfg.time_step)


@pytest.mark.parametrize("pbf_kernel", [False, True])
def test_reference_layout_density_with_boundary_bit_exact(gpu_device, pbf_kernel):
    n = 20000
    rng = np.random.default_rng(19)
    samples = boundary_samples(rng, 5000, moving=True)
    cfg = with_boundary(domain(n, dt=0.02), samples)
    topo = positions("uniform", n, rng, (1.0, 0.6, 0.8))
    pos = (topo + rng.uniform(-0.003, 0.003, size=topo.shape)).astype(np.float32)
    m = rng.uniform(0.5, 1.5, size=n).astype(np.float32)
    r0 = rng.uniform(900.0, 1100.0, size=n).astype(np.float32)
    tf, pf = VectorField.zeros(n, gpu_device), VectorField.zeros(n, gpu_device)
    tf.upload(topo)
    pf.upload(pos)
    z = torch.zeros(n, device=gpu_device)
    params = ParticleParameters(torch.from_numpy(m).to(gpu_device), torch.from_numpy(r0).to(gpu_device), z, z)
    step = 4
    nb = build_neighborhood(cfg, step, tf, gpu_device)
    boundary = DeviceBoundary(cfg, gpu_device)
    rho = torch.empty(n, device=gpu_device)
    density_forward(cfg.support_radius, pbf_kernel, tf, pf, params, nb.view, rho, boundary, nb.boundary_time)
    expected = oracle.density(oracle.make_domain(cfg, samples), step, pbf_kernel, topo, pos, m, r0)
    without = oracle.density(oracle.make_domain(cfg), step, pbf_kernel, topo, pos, m, r0)
    assert np.sum(expected != without) > 1000  # the boundary terms are exercised
    assert bits_equal(rho.cpu().numpy(), expected)


@pytest.mark.parametrize("pbf_kernel", [False, True])
def test_reference_layout_density_jvp_vjp_bit_exact(gpu_device, pbf_kernel):
    """density::{sph,pbf}_{jvp,vjp} in the reference layout with moving
    boundary samples: the FP32 tangent and the FP64 adjoints (accumulated
    onto non-zero outputs) equal the oracle's bit for bit."""
    from physica_amd import density_jvp, density_vjp
    from physica_amd.particle import ParticleParameterAdjoint, ParticleParameterTangent, VectorAdjointField
    n = 20000
    rng = np.random.default_rng(23 + pbf_kernel)
    samples = boundary_samples(rng, 3000, moving=True)
    cfg = with_boundary(domain(n, dt=0.02), samples)
    x = positions("uniform", n, rng, (1.0, 0.6, 0.8))
    xt = rng.normal(size=(3, n)).astype(np.float32)
    m = rng.uniform(0.5, 1.5, size=n).astype(np.float32)
    r0 = rng.uniform(900.0, 1100.0, size=n).astype(np.float32)
    mt = rng.normal(size=n).astype(np.float32)
    rt = rng.normal(size=n).astype(np.float32)
    w = rng.normal(size=n)
    xf, tf = VectorField.zeros(n, gpu_device), VectorField.zeros(n, gpu_device)
    xf.upload(x)
    tf.upload(xt)
    z = torch.zeros(n, device=gpu_device)
    dev = lambda a, dt=torch.float32: torch.from_numpy(np.ascontiguousarray(a)).to(device=gpu_device, dtype=dt)
    params = ParticleParameters(dev(m), dev(r0), z, z)
    step = 3
    nb = build_neighborhood(cfg, step, xf, gpu_device)
    boundary = DeviceBoundary(cfg, gpu_device)
    rho_t = torch.empty(n, device=gpu_device)
    density_jvp(cfg.support_radius, pbf_kernel, xf, xf, tf, params, ParticleParameterTangent(dev(mt), dev(rt), z, z),
                nb.view, rho_t, boundary, nb.boundary_time)
    start = rng.normal(size=(5, n))
    pos_a = VectorAdjointField(*(dev(start[k], torch.float64) for k in range(3)))
    zd = torch.zeros(n, device=gpu_device, dtype=torch.float64)
    par_a = ParticleParameterAdjoint(dev(start[3], torch.float64), dev(start[4], torch.float64), zd, zd)
    density_vjp(cfg.support_radius, pbf_kernel, xf, xf, params, nb.view, dev(w, torch.float64), pos_a, par_a,
                boundary, nb.boundary_time)
    torch.cuda.synchronize()
    od = oracle.make_domain(cfg, samples)
    exp_t = oracle.density_jvp(od, step, pbf_kernel, x, x, xt, m, mt, r0, rt)
    assert bits_equal(rho_t.cpu().numpy(), exp_t)
    ea, em, er = oracle.density_vjp(od, step, pbf_kernel, x, x, m, r0, w)
    assert np.abs(er).max() > 0
    assert bits_equal(pos_a.numpy(), start[:3] + ea)
    assert bits_equal(par_a.masses.cpu().numpy(), start[3] + em)
    assert bits_equal(par_a.rest_densities.cpu().numpy(), start[4] + er)
