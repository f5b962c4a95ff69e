"""The state dump format (physica_amd.state_io): a round trip of particle
state, method state and extra fields, metadata, and refusal of wrong sizes
and foreign files; CPU tensors stand in for device ones."""
import dataclasses
import types

import numpy as np
import pytest
import torch

from physica_amd import scenes
from physica_amd.particle import ParticleState, VectorField
from physica_amd.state_io import load_state, read_state, save_state


@dataclasses.dataclass
class WarmPressures:
    warm_divergence_pressure: torch.Tensor
    warm_density_pressure: torch.Tensor


def state(n, seed):
    g = torch.Generator().manual_seed(seed)
    vec = lambda: VectorField(*(torch.randn(n, generator=g) for _ in range(3)))
    return types.SimpleNamespace(particles=ParticleState(vec(), vec(), 17 + seed),
                                 method=WarmPressures(torch.randn(n, generator=g), torch.randn(n, generator=g)))


def test_round_trip(tmp_path):
    sc = scenes.dfsph_dam_break(shape=(4, 3, 2))
    a = state(sc.count, 1)
    mags = torch.rand(sc.count)
    save_state(tmp_path / "s.npz", a, sc.domain, fields={"vorticity": mags, "velocity_copy": a.particles.velocities})
    b = state(sc.count, 2)
    meta = load_state(tmp_path / "s.npz", b)
    for name in ("positions", "velocities"):
        for axis in "xyz":
            assert torch.equal(getattr(getattr(a.particles, name), axis), getattr(getattr(b.particles, name), axis))
    assert b.particles.step_index == a.particles.step_index
    assert torch.equal(a.method.warm_density_pressure, b.method.warm_density_pressure)
    assert meta["domain"]["particle_count"] == sc.count and meta["domain"]["time_step"] == pytest.approx(1e-3)
    data = read_state(tmp_path / "s.npz")
    assert np.array_equal(data["field/vorticity"], mags.numpy())
    assert data["field/velocity_copy"].shape == (3, sc.count)


def test_wrong_size_and_foreign_files_are_refused(tmp_path):
    save_state(tmp_path / "s.npz", state(10, 1))
    with pytest.raises(ValueError, match="positions"):
        load_state(tmp_path / "s.npz", state(11, 1))
    np.savez(tmp_path / "other.npz", meta=np.array('{"format": "else"}'), positions=np.zeros((3, 2)))
    with pytest.raises(ValueError, match="physica_amd.state"):
        read_state(tmp_path / "other.npz")


def test_malformed_fields_are_refused_before_anything_is_copied(tmp_path):
    """A file whose velocities or method field would broadcast into the state
    (a length-1 field, a (3, 1) vector) or whose domain metadata names another
    particle count raises ValueError and leaves the state untouched."""
    sc = scenes.dfsph_dam_break(shape=(4, 3, 2))
    n = sc.count
    good = state(n, 1)
    save_state(tmp_path / "good.npz", good, sc.domain)
    data = dict(np.load(tmp_path / "good.npz", allow_pickle=False))
    cases = {"velocities": np.zeros((3, 1), np.float32),
             "method/warm_density_pressure": np.zeros(1, np.float32),
             "meta": np.array(str(data["meta"]).replace(f'"particle_count": {n}', '"particle_count": 5'))}
    for key, bad in cases.items():
        arrays = dict(data)
        arrays[key] = bad
        np.savez(tmp_path / "bad.npz", **arrays)
        target = state(n, 2)
        before = target.particles.positions.x.clone()
        with pytest.raises(ValueError):
            load_state(tmp_path / "bad.npz", target)
        assert torch.equal(target.particles.positions.x, before), key
