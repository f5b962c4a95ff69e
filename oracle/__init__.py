"""CPU oracle for the particle-liquid path: TEST INFRASTRUCTURE ONLY.

ctypes wrapper over oracle/_build/liboracle.so, the plain-C restatement of the
reference's kernels (see oracle/oracle.h for the restated files and the checks
that pin each function). Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this package.
"""
from __future__ import annotations

import ctypes
import subprocess
from ctypes import POINTER, c_double, c_float, c_int, c_uint32, c_uint64, c_void_p
from pathlib import Path
from typing import Dict, Optional

import numpy as np

HERE = Path(__file__).resolve().parent
LIBRARY_PATH = HERE / "_build" / "liboracle.so"

POLY6, CUBIC, SPIKY, VISCOSITY_LAPLACIAN = 0, 1, 2, 3


class OracleDomain(ctypes.Structure):
    _fields_ = [
        ("particle_count", c_uint32),
        ("time_step", c_float),
        ("support_radius", c_float),
        ("particle_radius", c_float),
        ("gravity", c_float * 3),
        ("box_minimum", c_float * 3),
        ("box_maximum", c_float * 3),
        ("box_velocity", c_float * 3),
        ("no_slip", c_uint32),
        ("boundary_count", c_uint32),
        ("boundary_position", c_void_p * 3),
        ("boundary_velocity", c_void_p * 3),
        ("boundary_volume", c_void_p),
    ]


_lib = None


def build() -> Path:
    subprocess.run(["make", "-C", str(HERE), "-j4"], check=True, stdout=subprocess.DEVNULL)
    return LIBRARY_PATH


def library() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIBRARY_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(LIBRARY_PATH))
        for suf in ("_f32", "_f64"):
            getattr(_lib, "oracle_reference_gradient_norm" + suf).restype = c_float if suf == "_f32" else c_double
            getattr(_lib, "oracle_reference_gradient_norm" + suf).argtypes = [POINTER(OracleDomain)]
    return _lib


def set_threads(threads: int) -> None:
    library().oracle_set_threads(c_int(int(threads)))


def threads() -> int:
    return int(library().oracle_threads())


def make_domain(configuration, boundary_particles=None) -> OracleDomain:
    """OracleDomain from a physica_amd.particle.DomainConfiguration-like object."""
    d = OracleDomain()
    d.particle_count = int(configuration.particle_count)
    d.time_step = float(configuration.time_step)
    d.support_radius = float(configuration.support_radius)
    d.particle_radius = float(configuration.particle_radius)
    d.gravity[:] = list(map(float, configuration.gravity))
    b = configuration.boundary
    d.box_minimum[:] = list(map(float, b.minimum))
    d.box_maximum[:] = list(map(float, b.maximum))
    d.box_velocity[:] = list(map(float, b.velocity))
    d.no_slip = 1 if b.no_slip else 0
    d.boundary_count = 0
    if boundary_particles is not None:
        pos, vel, vol = boundary_particles
        pos = np.ascontiguousarray(pos, dtype=np.float32)
        vel = np.ascontiguousarray(vel, dtype=np.float32)
        vol = np.ascontiguousarray(vol, dtype=np.float32)
        d.boundary_count = int(vol.shape[0])
        for a in range(3):
            d.boundary_position[a] = pos[a].ctypes.data
            d.boundary_velocity[a] = vel[a].ctypes.data
        d.boundary_volume = vol.ctypes.data
        d._keep = (pos, vel, vol)
    return d


def grid_resolution(domain: OracleDomain):
    res = (c_uint32 * 3)()
    library().oracle_grid_resolution(ctypes.byref(domain), res)
    return int(res[0]), int(res[1]), int(res[2])


def _dtype(precision: str):
    return np.float32 if precision == "f32" else np.float64


def _arr(a, precision):
    return np.ascontiguousarray(a, dtype=_dtype(precision))


def _vec(a, precision):
    a = _arr(a, precision)
    ptrs = (c_void_p * 3)(*(a[k].ctypes.data for k in range(3)))
    return a, ptrs


def _out_vec(n, precision):
    a = np.zeros((3, n), dtype=_dtype(precision))
    ptrs = (c_void_p * 3)(*(a[k].ctypes.data for k in range(3)))
    return a, ptrs


def build_neighborhood(domain: OracleDomain, step_index: int, positions, precision="f32", boundary=False):
    """(sorted_keys u64[N], sorted_indices u32[N], cell_offsets u32[C+1]);
    with boundary=True also (sorted boundary keys u64[B], sorted boundary
    indices u32[B], boundary cell offsets u32[C+1])."""
    n = domain.particle_count
    b = domain.boundary_count
    cells = int(np.prod(grid_resolution(domain)))
    pos = _arr(positions, precision)
    keys = np.zeros(n, dtype=np.uint64)
    idx = np.zeros(n, dtype=np.uint32)
    off = np.zeros(cells + 1, dtype=np.uint32)
    bkeys = np.zeros(max(b, 1), dtype=np.uint64)
    bidx = np.zeros(max(b, 1), dtype=np.uint32)
    boff = np.zeros(cells + 1, dtype=np.uint32)
    getattr(library(), "oracle_build_neighborhood_" + precision)(
        ctypes.byref(domain), c_uint64(step_index), c_void_p(pos[0].ctypes.data), c_void_p(pos[1].ctypes.data),
        c_void_p(pos[2].ctypes.data), c_void_p(keys.ctypes.data), c_void_p(idx.ctypes.data),
        c_void_p(off.ctypes.data), c_void_p(bkeys.ctypes.data), c_void_p(bidx.ctypes.data),
        c_void_p(boff.ctypes.data))
    if boundary:
        return keys, idx, off, bkeys[:b], bidx[:b], boff
    return keys, idx, off


def kernel_values(kind: int, support_radius: float, displacements, precision="f32"):
    """(value[N], gradient[3, N]) of poly6 / cubic / spiky / viscosity Laplacian."""
    d = _arr(displacements, precision)
    n = d.shape[1]
    value = np.zeros(n, dtype=_dtype(precision))
    grad = np.zeros((3, n), dtype=_dtype(precision))
    getattr(library(), "oracle_kernel_values_" + precision)(
        c_int(kind), c_float(support_radius), c_uint32(n), c_void_p(d[0].ctypes.data), c_void_p(d[1].ctypes.data),
        c_void_p(d[2].ctypes.data), c_void_p(value.ctypes.data), c_void_p(grad[0].ctypes.data),
        c_void_p(grad[1].ctypes.data), c_void_p(grad[2].ctypes.data))
    return value, grad


def density(domain, grid_step, pbf_kernel, topology, positions, masses, rest_densities, precision="f32"):
    n = domain.particle_count
    topo, tp = _vec(topology, precision)
    pos, pp = _vec(positions, precision)
    m = _arr(masses, precision)
    r0 = _arr(rest_densities, precision)
    out = np.zeros(n, dtype=_dtype(precision))
    getattr(library(), "oracle_density_" + precision)(
        ctypes.byref(domain), c_uint64(grid_step), c_int(1 if pbf_kernel else 0), tp, pp, c_void_p(m.ctypes.data),
        c_void_p(r0.ctypes.data), c_void_p(out.ctypes.data))
    return out


def density_jvp(domain, grid_step, pbf_kernel, topology, positions, position_tangent, masses, mass_tangent,
                rest_densities, rest_tangent, precision="f32"):
    """density_jvp (density.cu:42-63): the density tangent, primal precision."""
    n = domain.particle_count
    topo, tp = _vec(topology, precision)
    pos, pp = _vec(positions, precision)
    tan, tnp = _vec(position_tangent, precision)
    arrs = [_arr(a, precision) for a in (masses, mass_tangent, rest_densities, rest_tangent)]
    out = np.zeros(n, dtype=_dtype(precision))
    getattr(library(), "oracle_density_jvp_" + precision)(
        ctypes.byref(domain), c_uint64(grid_step), c_int(1 if pbf_kernel else 0), tp, pp, tnp,
        *(c_void_p(a.ctypes.data) for a in arrs), c_void_p(out.ctypes.data))
    return out


def density_vjp(domain, grid_step, pbf_kernel, topology, positions, masses, rest_densities, density_adjoint,
                precision="f32"):
    """density_vjp (density.cu:65-98) in FP64: (position adjoint [3, N],
    mass adjoint [N], rest-density adjoint [N]) from zero."""
    n = domain.particle_count
    topo, tp = _vec(topology, precision)
    pos, pp = _vec(positions, precision)
    m = _arr(masses, precision)
    r0 = _arr(rest_densities, precision)
    rho_a = np.ascontiguousarray(density_adjoint, dtype=np.float64)
    pos_a = np.zeros((3, n))
    m_a = np.zeros(n)
    r_a = np.zeros(n)
    ptrs = (c_void_p * 3)(*(c_void_p(pos_a[k].ctypes.data) for k in range(3)))
    getattr(library(), "oracle_density_vjp_" + precision)(
        ctypes.byref(domain), c_uint64(grid_step), c_int(1 if pbf_kernel else 0), tp, pp, c_void_p(m.ctypes.data),
        c_void_p(r0.ctypes.data), c_void_p(rho_a.ctypes.data), ptrs, c_void_p(m_a.ctypes.data),
        c_void_p(r_a.ctypes.data))
    return pos_a, m_a, r_a


def reference_gradient_norm(domain, precision="f32") -> float:
    return float(getattr(library(), "oracle_reference_gradient_norm_" + precision)(ctypes.byref(domain)))


def non_pressure(domain, grid_step, positions, velocities, controls, masses, viscosities, densities, precision="f32"):
    n = domain.particle_count
    pos, pp = _vec(positions, precision)
    vel, vp = _vec(velocities, precision)
    ctl, cp = _vec(controls, precision)
    m = _arr(masses, precision)
    mu = _arr(viscosities, precision)
    rho = _arr(densities, precision)
    out, op = _out_vec(n, precision)
    getattr(library(), "oracle_non_pressure_" + precision)(
        ctypes.byref(domain), c_uint64(grid_step), pp, vp, cp, c_void_p(m.ctypes.data), c_void_p(mu.ctypes.data),
        c_void_p(rho.ctypes.data), op)
    return out


def pressure(domain, grid_step, positions, masses, rest_densities, densities, pressures, precision="f32"):
    n = domain.particle_count
    pos, pp = _vec(positions, precision)
    m = _arr(masses, precision)
    r0 = _arr(rest_densities, precision)
    rho = _arr(densities, precision)
    p = _arr(pressures, precision)
    out, op = _out_vec(n, precision)
    getattr(library(), "oracle_pressure_" + precision)(
        ctypes.byref(domain), c_uint64(grid_step), pp, c_void_p(m.ctypes.data), c_void_p(r0.ctypes.data),
        c_void_p(rho.ctypes.data), c_void_p(p.ctypes.data), op)
    return out


def pbf_lambda(domain, grid_step, topology, positions, masses, rest_densities, densities, relaxation, precision="f32"):
    n = domain.particle_count
    topo, tp = _vec(topology, precision)
    pos, pp = _vec(positions, precision)
    m = _arr(masses, precision)
    r0 = _arr(rest_densities, precision)
    rho = _arr(densities, precision)
    eps = _arr(relaxation, precision)
    gs, gp = _out_vec(n, precision)
    den = np.zeros(n, dtype=_dtype(precision))
    lam = np.zeros(n, dtype=_dtype(precision))
    getattr(library(), "oracle_pbf_lambda_" + precision)(
        ctypes.byref(domain), c_uint64(grid_step), tp, pp, c_void_p(m.ctypes.data), c_void_p(r0.ctypes.data),
        c_void_p(rho.ctypes.data), c_void_p(eps.ctypes.data), gp, c_void_p(den.ctypes.data),
        c_void_p(lam.ctypes.data))
    return gs, den, lam


def dfsph_step(domain, step_index, positions, velocities, controls, params: Dict[str, np.ndarray],
               warm_divergence, warm_density, divergence_iterations=4, density_iterations=6, warm_start=True,
               precision="f32"):
    """One SPH<DFSPH>::forward_step. Returns a dict of next state and cache fields."""
    n = domain.particle_count
    pos, pp = _vec(positions, precision)
    vel, vp = _vec(velocities, precision)
    ctl, cp = _vec(controls, precision)
    arrs = {k: _arr(params[k], precision) for k in
            ("masses", "rest_densities", "viscosities", "divergence_relaxation", "density_relaxation")}
    wd = _arr(warm_divergence, precision)
    we = _arr(warm_density, precision)
    nx, nxp = _out_vec(n, precision)
    nv, nvp = _out_vec(n, precision)
    nwd = np.zeros(n, dtype=_dtype(precision))
    nwe = np.zeros(n, dtype=_dtype(precision))
    rho = np.zeros(n, dtype=_dtype(precision))
    anp, anpp = _out_vec(n, precision)
    getattr(library(), "oracle_dfsph_step_" + precision)(
        ctypes.byref(domain), c_uint32(divergence_iterations), c_uint32(density_iterations),
        c_int(1 if warm_start else 0), c_uint64(step_index), pp, vp, cp,
        *(c_void_p(arrs[k].ctypes.data) for k in
          ("masses", "rest_densities", "viscosities", "divergence_relaxation", "density_relaxation")),
        c_void_p(wd.ctypes.data), c_void_p(we.ctypes.data), nxp, nvp, c_void_p(nwd.ctypes.data),
        c_void_p(nwe.ctypes.data), c_void_p(rho.ctypes.data), anpp)
    return {"positions": nx, "velocities": nv, "warm_divergence_pressure": nwd, "warm_density_pressure": nwe,
            "densities": rho, "non_pressure_accelerations": anp}


def dfsph_step_tolerance(domain, step_index, positions, velocities, controls, params: Dict[str, np.ndarray],
                         warm_divergence, warm_density, tolerance=1e-3, max_iterations=100, warm_start=True,
                         precision="f32"):
    """SPH<DFSPH>::forward_step in tolerance mode (an extension of the
    reference's fixed counts). Returns the next state and "iterations", the
    executed (divergence, density) counts."""
    n = domain.particle_count
    pos, pp = _vec(positions, precision)
    vel, vp = _vec(velocities, precision)
    ctl, cp = _vec(controls, precision)
    arrs = {k: _arr(params[k], precision) for k in
            ("masses", "rest_densities", "viscosities", "divergence_relaxation", "density_relaxation")}
    wd = _arr(warm_divergence, precision)
    we = _arr(warm_density, precision)
    nx, nxp = _out_vec(n, precision)
    nv, nvp = _out_vec(n, precision)
    nwd = np.zeros(n, dtype=_dtype(precision))
    nwe = np.zeros(n, dtype=_dtype(precision))
    its = (ctypes.c_uint32 * 2)()
    getattr(library(), "oracle_dfsph_step_tolerance_" + precision)(
        ctypes.byref(domain), ctypes.c_double(tolerance), c_uint32(max_iterations), c_int(1 if warm_start else 0),
        c_uint64(step_index), pp, vp, cp,
        *(c_void_p(arrs[k].ctypes.data) for k in ("masses", "rest_densities", "viscosities",
                                                   "divergence_relaxation", "density_relaxation")),
        c_void_p(wd.ctypes.data), c_void_p(we.ctypes.data), nxp, nvp, c_void_p(nwd.ctypes.data),
        c_void_p(nwe.ctypes.data), its)
    return {"positions": nx, "velocities": nv, "warm_divergence_pressure": nwd, "warm_density_pressure": nwe,
            "iterations": (int(its[0]), int(its[1]))}


def pressure_solver_step(domain, step_index, positions, velocities, controls, params: Dict[str, np.ndarray],
                         method="pcisph", pressure_iterations=6, precision="f32"):
    """One SPH<PCISPH> or SPH<IISPH>::forward_step (params["relaxation"] is the
    pressure or Jacobi relaxation). Returns next state and cache fields."""
    if method not in ("pcisph", "iisph"):
        raise ValueError(method)
    n = domain.particle_count
    pos, pp = _vec(positions, precision)
    vel, vp = _vec(velocities, precision)
    ctl, cp = _vec(controls, precision)
    names = ("masses", "rest_densities", "viscosities", "relaxation")
    arrs = {k: _arr(params[k], precision) for k in names}
    nx, nxp = _out_vec(n, precision)
    nv, nvp = _out_vec(n, precision)
    rho = np.zeros(n, dtype=_dtype(precision))
    prs = np.zeros(n, dtype=_dtype(precision))
    anp, anpp = _out_vec(n, precision)
    getattr(library(), "oracle_pressure_solver_step_" + precision)(
        ctypes.byref(domain), c_int(1 if method == "iisph" else 0), c_uint32(pressure_iterations),
        c_uint64(step_index), pp, vp, cp, *(c_void_p(arrs[k].ctypes.data) for k in names), nxp, nvp,
        c_void_p(rho.ctypes.data), anpp, c_void_p(prs.ctypes.data))
    return {"positions": nx, "velocities": nv, "densities": rho, "non_pressure_accelerations": anp,
            "pressures": prs}


def pressure_solver_step_tolerance(domain, step_index, positions, velocities, controls, params: Dict[str, np.ndarray],
                                   method="iisph", tolerance=1e-3, max_iterations=100, precision="f32"):
    """SPH<PCISPH|IISPH>::forward_step in tolerance mode (an extension: the
    reference runs fixed counts, sph.cpp:302, 415): iterate until the mean
    compression error (1/N) sum max(0, rho* - rho0) / rho0 of the prediction
    (FP64) is within the tolerance, at most max_iterations; the update of the
    iteration that meets it still applies. Adds "iterations" to the result."""
    if method not in ("pcisph", "iisph"):
        raise ValueError(method)
    n = domain.particle_count
    pos, pp = _vec(positions, precision)
    vel, vp = _vec(velocities, precision)
    ctl, cp = _vec(controls, precision)
    names = ("masses", "rest_densities", "viscosities", "relaxation")
    arrs = {k: _arr(params[k], precision) for k in names}
    nx, nxp = _out_vec(n, precision)
    nv, nvp = _out_vec(n, precision)
    rho = np.zeros(n, dtype=_dtype(precision))
    prs = np.zeros(n, dtype=_dtype(precision))
    it = c_uint32(0)
    getattr(library(), "oracle_pressure_solver_step_tolerance_" + precision)(
        ctypes.byref(domain), c_int(1 if method == "iisph" else 0), ctypes.c_double(tolerance),
        c_uint32(max_iterations), c_uint64(step_index), pp, vp, cp, *(c_void_p(arrs[k].ctypes.data) for k in names),
        nxp, nvp, c_void_p(rho.ctypes.data), c_void_p(prs.ctypes.data), ctypes.byref(it))
    return {"positions": nx, "velocities": nv, "densities": rho, "pressures": prs, "iterations": int(it.value)}


def wcsph_step(domain, step_index, positions, velocities, controls, params: Dict[str, np.ndarray],
               precision="f32"):
    """One SPH<WCSPH>::forward_step. Returns next state, densities and pressures."""
    n = domain.particle_count
    pos, pp = _vec(positions, precision)
    vel, vp = _vec(velocities, precision)
    ctl, cp = _vec(controls, precision)
    names = ("masses", "rest_densities", "viscosities", "surface_tensions", "speed_of_sound", "tait_exponent",
             "boundary_surface_tension")
    arrs = {k: _arr(params[k], precision) for k in names}
    nx, nxp = _out_vec(n, precision)
    nv, nvp = _out_vec(n, precision)
    rho = np.zeros(n, dtype=_dtype(precision))
    prs = np.zeros(n, dtype=_dtype(precision))
    getattr(library(), "oracle_wcsph_step_" + precision)(
        ctypes.byref(domain), c_uint64(step_index), pp, vp, cp, *(c_void_p(arrs[k].ctypes.data) for k in names),
        nxp, nvp, c_void_p(rho.ctypes.data), c_void_p(prs.ctypes.data))
    return {"positions": nx, "velocities": nv, "densities": rho, "pressures": prs}


def pbf_step(domain, step_index, positions, velocities, controls, params: Dict[str, np.ndarray],
             pressure_iterations=4, precision="f32"):
    """One PBF::forward_step. Returns a dict of next state and cache fields."""
    n = domain.particle_count
    pos, pp = _vec(positions, precision)
    vel, vp = _vec(velocities, precision)
    ctl, cp = _vec(controls, precision)
    names = ("masses", "rest_densities", "relaxation", "artificial_pressure_strength", "artificial_pressure_exponent",
             "artificial_pressure_radius", "xsph_viscosity", "vorticity_confinement")
    arrs = {k: _arr(params[k], precision) for k in names}
    nx, nxp = _out_vec(n, precision)
    nv, nvp = _out_vec(n, precision)
    rho = np.zeros(n, dtype=_dtype(precision))
    lam = np.zeros(n, dtype=_dtype(precision))
    pred, predp = _out_vec(n, precision)
    mag = np.zeros(n, dtype=_dtype(precision))
    getattr(library(), "oracle_pbf_step_" + precision)(
        ctypes.byref(domain), c_uint32(pressure_iterations), c_uint64(step_index), pp, vp, cp,
        *(c_void_p(arrs[k].ctypes.data) for k in names), nxp, nvp, c_void_p(rho.ctypes.data),
        c_void_p(lam.ctypes.data), predp, c_void_p(mag.ctypes.data))
    return {"positions": nx, "velocities": nv, "densities": rho, "lambdas": lam, "predicted_positions": pred,
            "vorticity_magnitudes": mag}


PBF_PARAMETERS = ("masses", "rest_densities", "relaxation", "artificial_pressure_strength",
                  "artificial_pressure_exponent", "artificial_pressure_radius", "xsph_viscosity",
                  "vorticity_confinement")


def pbf_jvp(domain, step_index, positions, velocities, controls, params: Dict[str, np.ndarray],
            tangents: Dict[str, np.ndarray], pressure_iterations=4, precision="f32"):
    """PBF::jvp_step: tangents of the next positions and velocities. `tangents`
    holds "positions", "velocities", "controls" (3, N) and any of
    PBF_PARAMETERS (missing ones are zero)."""
    n = domain.particle_count
    dt = _dtype(precision)
    pos, pp = _vec(positions, precision)
    vel, vp = _vec(velocities, precision)
    ctl, cp = _vec(controls, precision)
    arrs = {k: _arr(params[k], precision) for k in PBF_PARAMETERS}
    zero3 = np.zeros((3, n))
    tv = [_vec(tangents.get(k, zero3), precision) for k in ("positions", "velocities", "controls")]
    ts = {k: _arr(tangents.get(k, np.zeros(n)), precision) for k in PBF_PARAMETERS}
    nx, nxp = _out_vec(n, precision)
    nv, nvp = _out_vec(n, precision)
    getattr(library(), "oracle_pbf_jvp_" + precision)(
        ctypes.byref(domain), c_uint32(pressure_iterations), c_uint64(step_index), pp, vp, cp,
        *(c_void_p(arrs[k].ctypes.data) for k in PBF_PARAMETERS), *(t[1] for t in tv),
        *(c_void_p(ts[k].ctypes.data) for k in PBF_PARAMETERS), nxp, nvp)
    return {"positions": nx, "velocities": nv}


def pbf_vjp(domain, step_index, positions, velocities, controls, params: Dict[str, np.ndarray],
            next_position_adjoint, next_velocity_adjoint, pressure_iterations=4, precision="f32"):
    """PBF::vjp_step with FP64 adjoints. Returns the adjoints of the state
    ("positions", "velocities"), the control and every PBF_PARAMETERS entry."""
    n = domain.particle_count
    pos, pp = _vec(positions, precision)
    vel, vp = _vec(velocities, precision)
    ctl, cp = _vec(controls, precision)
    arrs = {k: _arr(params[k], precision) for k in PBF_PARAMETERS}
    nxa, nxap = _vec(next_position_adjoint, "f64")
    nva, nvap = _vec(next_velocity_adjoint, "f64")
    outs = {k: _out_vec(n, "f64") for k in ("positions", "velocities", "controls")}
    scal = {k: np.zeros(n, dtype=np.float64) for k in PBF_PARAMETERS}
    getattr(library(), "oracle_pbf_vjp_" + precision)(
        ctypes.byref(domain), c_uint32(pressure_iterations), c_uint64(step_index), pp, vp, cp,
        *(c_void_p(arrs[k].ctypes.data) for k in PBF_PARAMETERS), nxap, nvap,
        *(outs[k][1] for k in ("positions", "velocities", "controls")),
        *(c_void_p(scal[k].ctypes.data) for k in PBF_PARAMETERS))
    result = {k: outs[k][0] for k in outs}
    result.update(scal)
    return result


DFSPH_PARAMETERS = ("masses", "rest_densities", "viscosities", "divergence_relaxation", "density_relaxation")
DFSPH_WARM = ("warm_divergence_pressure", "warm_density_pressure")


def _pointers(arrays):
    return (c_void_p * len(arrays))(*(a.ctypes.data for a in arrays))


def dfsph_record_step(domain, step_index, positions, velocities, controls, params: Dict[str, np.ndarray],
                      warm_divergence, warm_density, divergence_iterations=4, density_iterations=6, warm_start=True,
                      volume_rest=False, precision="f32"):
    """The primal step that dfsph_jvp / dfsph_vjp differentiate (next
    positions and velocities), with the same options."""
    n = domain.particle_count
    pos, pp = _vec(positions, precision)
    vel, vp = _vec(velocities, precision)
    ctl, cp = _vec(controls, precision)
    prm = [_arr(params[k], precision) for k in DFSPH_PARAMETERS]
    wd, we = _arr(warm_divergence, precision), _arr(warm_density, precision)
    nx, nxp = _out_vec(n, precision)
    nv, nvp = _out_vec(n, precision)
    getattr(library(), "oracle_dfsph_record_step_" + precision)(
        ctypes.byref(domain), c_uint32(divergence_iterations), c_uint32(density_iterations),
        c_int(1 if warm_start else 0), c_int(1 if volume_rest else 0), c_uint64(step_index), pp, vp, cp,
        _pointers(prm), c_void_p(wd.ctypes.data), c_void_p(we.ctypes.data), nxp, nvp)
    return {"positions": nx, "velocities": nv}


def dfsph_jvp(domain, step_index, positions, velocities, controls, params: Dict[str, np.ndarray],
              warm_divergence, warm_density, tangents: Dict[str, np.ndarray], divergence_iterations=4,
              density_iterations=6, warm_start=True, precision="f32", volume_rest=False):
    """SPH<DFSPH>::jvp_step: tangents of the next positions, velocities and
    warm-start pressures. `tangents` holds "positions", "velocities",
    "controls" (3, N), any of DFSPH_PARAMETERS and DFSPH_WARM (missing ones
    are zero). volume_rest takes V = m / rho0 in the pressure update (the
    PCISPH step, see pressure_solver_jvp)."""
    n = domain.particle_count
    pos, pp = _vec(positions, precision)
    vel, vp = _vec(velocities, precision)
    ctl, cp = _vec(controls, precision)
    prm = [_arr(params[k], precision) for k in DFSPH_PARAMETERS]
    wd, we = _arr(warm_divergence, precision), _arr(warm_density, precision)
    zero3 = np.zeros((3, n))
    tv = [_vec(tangents.get(k, zero3), precision) for k in ("positions", "velocities", "controls")]
    prm_t = [_arr(tangents.get(k, np.zeros(n)), precision) for k in DFSPH_PARAMETERS]
    wt = [_arr(tangents.get(k, np.zeros(n)), precision) for k in DFSPH_WARM]
    nx, nxp = _out_vec(n, precision)
    nv, nvp = _out_vec(n, precision)
    nw = [np.zeros(n, dtype=_dtype(precision)) for _ in DFSPH_WARM]
    getattr(library(), "oracle_dfsph_jvp_" + precision)(
        ctypes.byref(domain), c_uint32(divergence_iterations), c_uint32(density_iterations),
        c_int(1 if warm_start else 0), c_int(1 if volume_rest else 0), c_uint64(step_index), pp, vp, cp,
        _pointers(prm),
        c_void_p(wd.ctypes.data), c_void_p(we.ctypes.data), *(t[1] for t in tv), _pointers(prm_t),
        c_void_p(wt[0].ctypes.data), c_void_p(wt[1].ctypes.data), nxp, nvp, c_void_p(nw[0].ctypes.data),
        c_void_p(nw[1].ctypes.data))
    return {"positions": nx, "velocities": nv, DFSPH_WARM[0]: nw[0], DFSPH_WARM[1]: nw[1]}


def dfsph_vjp(domain, step_index, positions, velocities, controls, params: Dict[str, np.ndarray],
              warm_divergence, warm_density, next_adjoints: Dict[str, np.ndarray], divergence_iterations=4,
              density_iterations=6, warm_start=True, precision="f32", volume_rest=False):
    """SPH<DFSPH>::vjp_step with FP64 adjoints. `next_adjoints` holds
    "positions", "velocities" (3, N) and optionally DFSPH_WARM entries.
    Returns the adjoints of the state, the control, every DFSPH_PARAMETERS
    entry and, with warm_start, the warm-start pressures (zero otherwise)."""
    n = domain.particle_count
    pos, pp = _vec(positions, precision)
    vel, vp = _vec(velocities, precision)
    ctl, cp = _vec(controls, precision)
    prm = [_arr(params[k], precision) for k in DFSPH_PARAMETERS]
    wd, we = _arr(warm_divergence, precision), _arr(warm_density, precision)
    nxa, nxap = _vec(next_adjoints["positions"], "f64")
    nva, nvap = _vec(next_adjoints["velocities"], "f64")
    nwa = [_arr(next_adjoints.get(k, np.zeros(n)), "f64") for k in DFSPH_WARM]
    outs = {k: _out_vec(n, "f64") for k in ("positions", "velocities", "controls")}
    scal = {k: np.zeros(n, dtype=np.float64) for k in DFSPH_PARAMETERS + DFSPH_WARM}
    getattr(library(), "oracle_dfsph_vjp_" + precision)(
        ctypes.byref(domain), c_uint32(divergence_iterations), c_uint32(density_iterations),
        c_int(1 if warm_start else 0), c_int(1 if volume_rest else 0), c_uint64(step_index), pp, vp, cp,
        _pointers(prm),
        c_void_p(wd.ctypes.data), c_void_p(we.ctypes.data), nxap, nvap, c_void_p(nwa[0].ctypes.data),
        c_void_p(nwa[1].ctypes.data), *(outs[k][1] for k in ("positions", "velocities", "controls")),
        _pointers([scal[k] for k in DFSPH_PARAMETERS]), c_void_p(scal[DFSPH_WARM[0]].ctypes.data),
        c_void_p(scal[DFSPH_WARM[1]].ctypes.data))
    result = {k: outs[k][0] for k in outs}
    result.update(scal)
    return result


# SPH<PCISPH> / SPH<IISPH> parameters (sph/sph.ixx:112, 158) in the order of
# their tangents and adjoints; "relaxation" is the pressure (PCISPH) or Jacobi
# (IISPH) relaxation.
PRESSURE_SOLVER_PARAMETERS = ("masses", "rest_densities", "viscosities", "relaxation")


def _as_dfsph(params, n):
    """PCISPH/IISPH parameters as the DFSPH parameters of the same step: no
    divergence phase, the relaxation in the density phase."""
    zero = np.zeros(n)
    return {"masses": params.get("masses", zero), "rest_densities": params.get("rest_densities", zero),
            "viscosities": params.get("viscosities", zero), "divergence_relaxation": zero,
            "density_relaxation": params.get("relaxation", zero)}


def pressure_solver_jvp(domain, step_index, positions, velocities, controls, params, tangents, method="pcisph",
                        pressure_iterations=6, precision="f32"):
    """SPH<PCISPH> / SPH<IISPH>::jvp_step (sph/sph.cpp:312-331, 423-442): the
    DFSPH derivative with no divergence phase and no warm start, and V = m /
    rho0 for PCISPH. `tangents` holds "positions", "velocities", "controls" and
    any of PRESSURE_SOLVER_PARAMETERS."""
    if method not in ("pcisph", "iisph"):
        raise ValueError(method)
    n = domain.particle_count
    zero = np.zeros(n)
    t = dict(_as_dfsph({k: tangents[k] for k in PRESSURE_SOLVER_PARAMETERS if k in tangents}, n))
    t.update({k: tangents[k] for k in ("positions", "velocities", "controls") if k in tangents})
    r = dfsph_jvp(domain, step_index, positions, velocities, controls, _as_dfsph(params, n), zero, zero, t,
                  divergence_iterations=0, density_iterations=pressure_iterations, warm_start=False,
                  precision=precision, volume_rest=method == "pcisph")
    return {"positions": r["positions"], "velocities": r["velocities"]}


def pressure_solver_vjp(domain, step_index, positions, velocities, controls, params, next_adjoints,
                        method="pcisph", pressure_iterations=6, precision="f32"):
    """SPH<PCISPH> / SPH<IISPH>::vjp_step (sph/sph.cpp:333-363, 444-476) with
    FP64 adjoints: the state, control and PRESSURE_SOLVER_PARAMETERS adjoints."""
    if method not in ("pcisph", "iisph"):
        raise ValueError(method)
    n = domain.particle_count
    zero = np.zeros(n)
    r = dfsph_vjp(domain, step_index, positions, velocities, controls, _as_dfsph(params, n), zero, zero,
                  next_adjoints, divergence_iterations=0, density_iterations=pressure_iterations, warm_start=False,
                  precision=precision, volume_rest=method == "pcisph")
    out = {k: r[k] for k in ("positions", "velocities", "controls", "masses", "rest_densities", "viscosities")}
    out["relaxation"] = r["density_relaxation"]
    return out


# SPH<WCSPH> particle and method parameters in the order of its tangents and
# adjoints (sph/sph.ixx:45-60 and domain/domain.ixx:58-63).
WCSPH_PARAMETERS = ("masses", "rest_densities", "viscosities", "surface_tensions", "speed_of_sound",
                    "tait_exponent", "boundary_surface_tension")


def wcsph_jvp(domain, step_index, positions, velocities, controls, params, tangents, precision="f32"):
    """SPH<WCSPH>::jvp_step (sph/sph.cpp:228-237): tangents of the next
    positions and velocities. `tangents` holds "positions", "velocities",
    "controls" and any of WCSPH_PARAMETERS (missing ones are zero)."""
    n = domain.particle_count
    pos, pp = _vec(positions, precision)
    vel, vp = _vec(velocities, precision)
    ctl, cp = _vec(controls, precision)
    prm = [_arr(params[k], precision) for k in WCSPH_PARAMETERS]
    zero3 = np.zeros((3, n))
    tv = [_vec(tangents.get(k, zero3), precision) for k in ("positions", "velocities", "controls")]
    prm_t = [_arr(tangents.get(k, np.zeros(n)), precision) for k in WCSPH_PARAMETERS]
    nx, nxp = _out_vec(n, precision)
    nv, nvp = _out_vec(n, precision)
    getattr(library(), "oracle_wcsph_jvp_" + precision)(
        ctypes.byref(domain), c_uint64(step_index), pp, vp, cp, _pointers(prm), *(t[1] for t in tv),
        _pointers(prm_t), nxp, nvp)
    return {"positions": nx, "velocities": nv}


def wcsph_vjp(domain, step_index, positions, velocities, controls, params, next_adjoints, precision="f32"):
    """SPH<WCSPH>::vjp_step (sph/sph.cpp:239-249) with FP64 adjoints of the
    state, the control and every WCSPH_PARAMETERS entry."""
    n = domain.particle_count
    pos, pp = _vec(positions, precision)
    vel, vp = _vec(velocities, precision)
    ctl, cp = _vec(controls, precision)
    prm = [_arr(params[k], precision) for k in WCSPH_PARAMETERS]
    nxa, nxap = _vec(next_adjoints["positions"], "f64")
    nva, nvap = _vec(next_adjoints["velocities"], "f64")
    outs = {k: _out_vec(n, "f64") for k in ("positions", "velocities", "controls")}
    scal = {k: np.zeros(n, dtype=np.float64) for k in WCSPH_PARAMETERS}
    getattr(library(), "oracle_wcsph_vjp_" + precision)(
        ctypes.byref(domain), c_uint64(step_index), pp, vp, cp, _pointers(prm), nxap, nvap,
        *(outs[k][1] for k in ("positions", "velocities", "controls")), _pointers([scal[k] for k in WCSPH_PARAMETERS]))
    result = {k: outs[k][0] for k in outs}
    result.update(scal)
    return result


def set_reference_no_slip_adjoint(enabled: bool) -> None:
    """Switches the SPH prediction's VJP between the exact transpose (default)
    and the reference's variant, which drops dt * xbar of the components that
    did not collide when a no-slip collision resets the velocity."""
    for precision in ("f32", "f64"):
        getattr(library(), "oracle_set_reference_no_slip_adjoint_" + precision)(c_int(1 if enabled else 0))
