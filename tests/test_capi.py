"""CPU checks of the C ABI: the library loads, exports every function the
header declares, and rejects invalid input on the host before touching the
device."""
import ctypes
import re
from pathlib import Path

import pytest

from physica_amd import _capi
from physica_amd.particle import PCISPH, BoxBoundary, DFSPH, PBF, DomainConfiguration

HEADER = Path(__file__).resolve().parent.parent / "include" / "physica_amd" / "physica.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(phx_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "phx_dfsph_forward_step" in names and "phx_pbf_forward_step" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = _capi.library()
    for name in declared_functions():
        assert hasattr(lib, name), name
        assert name in _capi.PROTOTYPES, f"{name} has no ctypes prototype"
    assert set(_capi.PROTOTYPES) == set(declared_functions())


def test_version_and_clean_error_state():
    lib = _capi.library()
    assert lib.phx_version().startswith(b"physica_amd")


def cfg(**kw):
    base = dict(
        particle_count=1000,
        time_step=1e-3,
        support_radius=0.02,
        particle_radius=0.005,
        boundary=BoxBoundary((0, 0, 0), (1.0, 0.5, 0.5), (0, 0, 0), True),
    )
    base.update(kw)
    return DomainConfiguration(**base)


def create_dfsph(domain, config=None, mode=0):
    handle = ctypes.c_void_p()
    status = _capi.library().phx_dfsph_create(
        ctypes.byref(domain.to_c()), ctypes.byref((config or DFSPH.Configuration()).to_c()), mode, None,
        ctypes.byref(handle))
    return status, handle


def test_grid_resolution_host_only():
    assert cfg().grid_resolution() == (50, 25, 25)


@pytest.mark.parametrize(
    "domain,status,fragment",
    [
        (cfg(support_radius=0.0), _capi.PHX_ERROR_INVALID_ARGUMENT, "support_radius"),
        (cfg(time_step=-1.0), _capi.PHX_ERROR_INVALID_ARGUMENT, "time_step"),
        (cfg(boundary=BoxBoundary((0, 0, 0), (0.0, 1, 1))), _capi.PHX_ERROR_INVALID_ARGUMENT, "extent"),
        (cfg(boundary_particles=[((0.1, float("nan"), 0.1), (0, 0, 0), 1e-6)]), _capi.PHX_ERROR_INVALID_ARGUMENT,
         "non-finite boundary"),
        (cfg(support_radius=1e-5), _capi.PHX_ERROR_CAPACITY, "cells"),
    ],
)
def test_invalid_domain_rejected_before_device_work(domain, status, fragment):
    got, handle = create_dfsph(domain)
    assert got == status
    assert handle.value is None
    assert fragment in _capi.library().phx_last_error().decode()


def test_boundary_particles_with_a_null_array_rejected():
    # a non-zero count with a null host array
    c = cfg().to_c()
    c.boundary_particle_count = 3
    size = ctypes.c_size_t()
    assert _capi.library().phx_neighborhood_workspace_size(ctypes.byref(c), ctypes.byref(size)) == \
        _capi.PHX_ERROR_INVALID_ARGUMENT


def test_sph_tolerance_configuration_rejected_where_invalid():
    lib = _capi.library()
    h = ctypes.c_void_p()

    def create(configuration, mode=0):
        return lib.phx_sph_create(ctypes.byref(cfg().to_c()), ctypes.byref(configuration), mode, None,
                                  ctypes.byref(h))

    c = PCISPH.Configuration(density_error_tolerance=1e-3).to_c()
    assert create(c, mode=1) == _capi.PHX_ERROR_UNSUPPORTED
    assert "forward mode only" in lib.phx_last_error().decode()
    c = PCISPH.Configuration(density_error_tolerance=1e-3, max_iterations=0).to_c()
    assert create(c) == _capi.PHX_ERROR_INVALID_ARGUMENT
    c = PCISPH.Configuration(density_error_tolerance=float("nan")).to_c()
    assert create(c) == _capi.PHX_ERROR_INVALID_ARGUMENT
    c = _capi.SphConfiguration(_capi.PHX_SPH_WCSPH, 0, 2, 1e-3, 10)
    assert create(c) == _capi.PHX_ERROR_INVALID_ARGUMENT
    assert "PCISPH and IISPH" in lib.phx_last_error().decode()
    assert h.value is None


def test_neighborhood_workspace_grows_with_boundary_samples():
    lib = _capi.library()
    plain, with_b = ctypes.c_size_t(), ctypes.c_size_t()
    assert lib.phx_neighborhood_workspace_size(ctypes.byref(cfg().to_c()), ctypes.byref(plain)) == 0
    samples = [((0.01 * k, 0.0, 0.1), (0, 0, 0), 1e-6) for k in range(500)]
    assert lib.phx_neighborhood_workspace_size(ctypes.byref(cfg(boundary_particles=samples).to_c()),
                                               ctypes.byref(with_b)) == 0
    # keys, slots, ranks, the seven packed arrays and the permutation
    assert with_b.value >= plain.value + 500 * 4 * 11


def test_invalid_method_configuration_rejected():
    got, _ = create_dfsph(cfg(), DFSPH.Configuration(checkpoint_interval=0))
    assert got == _capi.PHX_ERROR_INVALID_ARGUMENT
    # tolerance mode: forward only, with 1 <= max_iterations, a finite non-negative tolerance
    got, _ = create_dfsph(cfg(), DFSPH.Configuration(density_error_tolerance=1e-3), mode=1)
    assert got == _capi.PHX_ERROR_UNSUPPORTED
    got, _ = create_dfsph(cfg(), DFSPH.Configuration(density_error_tolerance=1e-3, max_iterations=0))
    assert got == _capi.PHX_ERROR_INVALID_ARGUMENT
    got, _ = create_dfsph(cfg(), DFSPH.Configuration(density_error_tolerance=-1e-3))
    assert got == _capi.PHX_ERROR_INVALID_ARGUMENT
    handle = ctypes.c_void_p()
    got = _capi.library().phx_pbf_create(ctypes.byref(cfg().to_c()), None, 0, None, ctypes.byref(handle))
    assert got == _capi.PHX_ERROR_INVALID_ARGUMENT
    got = _capi.library().phx_pbf_create(ctypes.byref(cfg().to_c()), ctypes.byref(PBF.Configuration().to_c()), 7,
                                         None, ctypes.byref(handle))
    assert got == _capi.PHX_ERROR_INVALID_ARGUMENT
    assert "mode" in _capi.library().phx_last_error().decode()


def test_null_arguments_rejected():
    lib = _capi.library()
    assert lib.phx_dfsph_forward_step(None, None, None, None, None, None) == _capi.PHX_ERROR_INVALID_ARGUMENT
    assert lib.phx_pbf_forward_step(None, None, None, None, None, None) == _capi.PHX_ERROR_INVALID_ARGUMENT
    assert lib.phx_pbf_jvp_step(None, None, None, None, None, None, None) == _capi.PHX_ERROR_INVALID_ARGUMENT
    assert lib.phx_pbf_vjp_step(None, None, None, None, None, None, None) == _capi.PHX_ERROR_INVALID_ARGUMENT
    assert lib.phx_dfsph_jvp_step(None, None, None, None, None, None) == _capi.PHX_ERROR_INVALID_ARGUMENT
    assert lib.phx_dfsph_vjp_step(None, None, None, None, None, None) == _capi.PHX_ERROR_INVALID_ARGUMENT
    assert lib.phx_sph_jvp_step(None, None, None, None, None, None) == _capi.PHX_ERROR_INVALID_ARGUMENT
    assert lib.phx_sph_vjp_step(None, None, None, None, None, None) == _capi.PHX_ERROR_INVALID_ARGUMENT
    assert lib.phx_grid_resolution(None, None) == _capi.PHX_ERROR_INVALID_ARGUMENT
    assert lib.phx_dfsph_cache_densities(None, None) == _capi.PHX_ERROR_INVALID_ARGUMENT
    size = ctypes.c_size_t()
    assert lib.phx_neighborhood_workspace_size(ctypes.byref(cfg().to_c()), ctypes.byref(size)) == 0
    assert size.value >= 3 * 4 * 1000 + 4 * (50 * 25 * 25 + 1)


def test_python_api_refuses_cpu_device():
    with pytest.raises(RuntimeError, match="GPU"):
        from physica_amd.particle import SPH

        SPH(cfg(), device="cpu")
