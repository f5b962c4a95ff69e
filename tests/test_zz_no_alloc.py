"""phx_build_neighborhood_noalloc through the C ABI: the same sorted keys,
indices, offsets and view as phx_build_neighborhood on two calls in a row,
the second with fewer particles and a smaller grid (the fixed device scratch
then still holds the first call's contents), and the host-side refusals of
scenes above its scratch and of boundary samples."""
import ctypes
import math

import numpy as np
import pytest
import torch

from physica_amd import _capi, build_neighborhood
from physica_amd.particle import BoxBoundary, DomainConfiguration, VectorField

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _domain(n, box, h):
    return DomainConfiguration(n, 1e-2, h, 0.25 * h, (0.0, -9.81, 0.0), BoxBoundary((0, 0, 0), box, (0, 0, 0), True))


def _noalloc(cfg, step, field, device):
    n = cfg.particle_count
    cells = math.prod(cfg.grid_resolution())
    keys = torch.full((n,), -1, dtype=torch.int64, device=device)
    idx = torch.full((n,), -1, dtype=torch.int32, device=device)
    off = torch.full((cells + 1,), -1, dtype=torch.int32, device=device)
    view = _capi.NeighborhoodView()
    _capi.call("phx_build_neighborhood_noalloc", torch.cuda.current_stream(device).cuda_stream,
               ctypes.byref(cfg.to_c()), int(step), field.view(), _ptr(keys), _ptr(idx), _ptr(off),
               ctypes.byref(view))
    return keys, idx, off, view


def test_noalloc_matches_workspace_entry_point_twice(gpu_device):
    rng = np.random.default_rng(20260007)
    cases = [(200_000, (1.0, 0.6, 0.8), 0.02, 5), (9_001, (0.5, 0.4, 0.3), 0.05, 2)]
    for n, box, h, step in cases:
        print(f"noalloc case n={n} box={box} h={h}", flush=True)
        cfg = _domain(n, box, h)
        pos = (rng.uniform(-0.1, 1.1, size=(3, n)) * np.array(box)[:, None]).astype(np.float32)
        # a dense clump so some cells hold many particles
        pos[:, : n // 10] = (0.3 * np.array(box)[:, None] + rng.uniform(0, 0.5 * h, size=(3, n // 10))).astype(
            np.float32)
        field = VectorField.zeros(n, gpu_device)
        field.upload(pos)
        keys, idx, off, view = _noalloc(cfg, step, field, gpu_device)
        ref_keys, ref_idx, ref_off, ref_view = build_neighborhood(cfg, step, field, gpu_device)[:4]
        torch.cuda.synchronize()
        assert torch.equal(keys, ref_keys)
        assert torch.equal(idx, ref_idx)
        assert torch.equal(off, ref_off)
        for name in ("particle_count", "boundary_count", "cells_x", "cells_y", "cells_z", "origin_x", "origin_y",
                     "origin_z", "cell_size"):
            assert getattr(view, name) == getattr(ref_view, name), name
        # the order is the stable sort of the particles' keys
        sorted_keys, order = keys.cpu().numpy(), idx.cpu().numpy()
        per_particle = np.empty(n, dtype=np.int64)
        per_particle[order] = sorted_keys
        assert np.array_equal(order, np.argsort(per_particle, kind="stable").astype(np.int32))
        assert np.all(np.diff(sorted_keys) >= 0)


def test_noalloc_refuses_what_does_not_fit(gpu_device):
    # 20M particles need 3 x 80 MB of per-particle scratch: above the fixed 64 MiB
    big = _domain(20_000_000, (1.0, 1.0, 1.0), 0.05)
    with pytest.raises(_capi.PhxError) as err:
        _capi.call("phx_build_neighborhood_noalloc", torch.cuda.current_stream(gpu_device).cuda_stream,
                   ctypes.byref(big.to_c()), 0, _capi.VectorView(1, 1, 1), ctypes.c_void_p(1), ctypes.c_void_p(1),
                   ctypes.c_void_p(1), None)
    assert err.value.status == _capi.PHX_ERROR_CAPACITY
    with_samples = _domain(10, (1.0, 1.0, 1.0), 0.05)
    with_samples.boundary_particles = [((0.5, 0.5, 0.5), (0.0, 0.0, 0.0), 1e-6)]
    with pytest.raises(_capi.PhxError) as err:
        _capi.call("phx_build_neighborhood_noalloc", torch.cuda.current_stream(gpu_device).cuda_stream,
                   ctypes.byref(with_samples.to_c()), 0, _capi.VectorView(1, 1, 1), ctypes.c_void_p(1),
                   ctypes.c_void_p(1), ctypes.c_void_p(1), None)
    assert err.value.status == _capi.PHX_ERROR_UNSUPPORTED
