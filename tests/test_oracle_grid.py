"""Pins of the oracle's neighbourhood build (keys, stable sort, offsets).

The key is the reference's closed form (neighborhood.cu:20-39); the sort and
the offsets are checked against numpy's stable argsort and searchsorted, and
hand-placed particles give known keys.
"""
import numpy as np
import pytest

import oracle
from physica_amd.particle import BoxBoundary, DomainConfiguration


def make(n, h=0.1, box=(1.0, 0.8, 0.6), vel=(0.0, 0.0, 0.0), dt=0.01):
    return DomainConfiguration(n, dt, h, 0.5 * h * 0.5, (0.0, -9.81, 0.0), BoxBoundary((0, 0, 0), box, vel, True))


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_hand_placed_keys(precision):
    cfg = make(7)
    d = oracle.make_domain(cfg)
    cx, cy, cz = oracle.grid_resolution(d)
    assert (cx, cy, cz) == (10, 8, 6)
    h = 0.1
    pos = np.array(
        [
            [0.05, 0.15, 0.05, 0.05, -1.0, 5.0, 0.95],
            [0.05, 0.05, 0.15, 0.05, -1.0, 5.0, 0.75],
            [0.05, 0.05, 0.05, 0.15, -1.0, 5.0, 0.55],
        ]
    )
    keys, idx, off = oracle.build_neighborhood(d, 0, pos, precision)
    expected = {0: 0, 1: 1, 2: cx, 3: cx * cy, 4: 0, 5: cx * cy * cz - 1, 6: cx * cy * cz - 1}
    got = dict(zip(idx.tolist(), keys.tolist()))
    assert got == expected
    # equal keys keep input order (stable): 0 then 4 in cell 0; 5 then 6 in the last cell
    assert idx.tolist() == [0, 4, 1, 2, 3, 5, 6]
    assert off[0] == 0 and off[1] == 2 and off[-1] == 7
    assert int(off[cx * cy * cz - 1]) == 5


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("n", [0, 1, 500, 20000])
def test_sort_and_offsets_match_numpy(precision, n):
    cfg = make(n)
    d = oracle.make_domain(cfg)
    res = oracle.grid_resolution(d)
    cells = int(np.prod(res))
    rng = np.random.default_rng(n + 7)
    pos = rng.uniform(-0.1, 1.1, size=(3, n)) * np.array([[1.0], [0.8], [0.6]])
    dtype = np.float32 if precision == "f32" else np.float64
    pos = pos.astype(dtype)
    keys, idx, off = oracle.build_neighborhood(d, 0, pos, precision)
    # every particle's key from its own cell coordinates
    h = dtype(0.1)
    c = [np.clip(np.floor((pos[a] - dtype(0.0)) / h), 0, res[a] - 1).astype(np.int64) for a in range(3)]
    unsorted = (c[2] * res[1] + c[1]) * res[0] + c[0]
    order = np.argsort(unsorted, kind="stable")
    np.testing.assert_array_equal(idx, order)
    np.testing.assert_array_equal(keys, unsorted[order])
    np.testing.assert_array_equal(off, np.searchsorted(unsorted[order], np.arange(cells + 1), side="left"))


def test_moving_box_shifts_keys():
    cfg = make(3, vel=(0.5, 0.0, 0.0), dt=0.1)
    d = oracle.make_domain(cfg)
    pos = np.array([[0.06, 0.16, 0.26], [0.05, 0.05, 0.05], [0.05, 0.05, 0.05]], dtype=np.float32)
    k0, _, _ = oracle.build_neighborhood(d, 0, pos)
    k2, i2, _ = oracle.build_neighborhood(d, 2, pos)  # box moved by t v = 0.1 in x
    assert k0.tolist() == [0, 1, 2]
    # x - 0.1 -> cells -1 (clamped to 0), 0, 1
    assert dict(zip(i2.tolist(), k2.tolist())) == {0: 0, 1: 0, 2: 1}
