"""Benchmark: SPH<DFSPH> dam break, 1,048,576 particles, one MI355X.

Prints one JSON line (rank 0). A step is one SPH<DFSPH>::forward_step of the
whole scene (grid build, density, 4 divergence + 6 density iterations) with
every input resident in HBM. `value` is particle-steps/s over the timed steps;
`roofline` is the dominant pass, its algorithmic bytes per launch over its
average launch time (HIP events on the solver stream during the timed steps);
`cpu_baseline` is the FP32 oracle (OpenMP over particles, the same
reference-order arithmetic) timed on this host over one step of the same scene.

`fwd_adjoint_ratio` is the differentiable PBF rollout: 262,144 particles (64^3
lattice, spacing 0.02, h 0.04), 120 steps at dt 1/120, 4 iterations, loss the
mean squared distance of the final positions to a seeded target cloud,
gradient with respect to the initial velocities (no parameter gradients). Ratio = (recording forward
rollout + chained VJPs) / plain forward rollout, with every step cache kept in
HBM; the same with a state checkpoint every 10 steps and recompute is reported
beside it.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s (1M DFSPH, 1 GPU, %HBM peak); fwd+adjoint/fwd time ratio"


def parse_args(argv=None):
    p = argparse.ArgumentParser(description=__doc__)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--workload", choices=["dfsph", "pbf"], default="dfsph")
    p.add_argument("--no-cpu-baseline", action="store_true", help="skip the CPU oracle timing")
    p.add_argument("--no-pass-timing", action="store_true", help="time without per-pass events")
    p.add_argument("--no-adjoint", action="store_true", help="skip the differentiable PBF rollout ratio")
    p.add_argument("--adjoint-steps", type=int, default=120,
                   help="steps of the differentiable PBF rollout (120: the configured rollout)")
    return p.parse_args(argv)


def dfsph_bytes_per_step(iterations, particles, cells):
    """Minimal-round-trip algorithmic bytes per particle per DFSPH step:
    384 + 124 I_div + 140 I_den + 8 C / N."""
    nd, ne = iterations
    return 384.0 + 124.0 * nd + 140.0 * ne + 8.0 * cells / particles


def pbf_bytes_per_step(iterations, particles, cells):
    """PBF: 252 + 108 I + 8 C / N."""
    return 252.0 + 108.0 * iterations + 8.0 * cells / particles


def traffic_for(kernel: str):
    """Measured HBM bytes per launch of `kernel` from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), if any."""
    path = ROOT / "profiles" / "traffic.json"
    if not path.exists():
        return None
    try:
        data = json.loads(path.read_text())
    except ValueError:
        return None
    entry = data.get(kernel)
    return None if entry is None else float(entry["bytes_per_launch"])


def setup_dfsph(device):
    import torch

    from physica_amd import scenes
    from physica_amd.particle import DFSPH, SPH

    sc = scenes.dfsph_dam_break()
    cfg = DFSPH.Configuration(divergence_iterations=4, density_iterations=6)
    solver = SPH(sc.domain, cfg, device=device)
    a, b = solver.allocate_state(), solver.allocate_state()
    control = solver.allocate_control()
    params = solver.allocate_parameters()
    cache = solver.allocate_step_cache()
    a.particles.positions.upload(sc.positions)
    a.particles.velocities.upload(sc.velocities)
    for name in ("masses", "rest_densities", "viscosities", "surface_tensions"):
        getattr(params.particles, name).copy_(torch.from_numpy(sc.parameters[name]))
    params.method.divergence_relaxation.copy_(torch.from_numpy(sc.parameters["divergence_relaxation"]))
    params.method.density_relaxation.copy_(torch.from_numpy(sc.parameters["density_relaxation"]))
    return sc, solver, [a, b], control, params, cache


def setup_pbf(device):
    import torch

    from physica_amd import scenes
    from physica_amd.particle import PBF

    sc = scenes.pbf_dam_break()
    solver = PBF(sc.domain, PBF.Configuration(pressure_iterations=4), device=device)
    a, b = solver.allocate_state(), solver.allocate_state()
    control = solver.allocate_control()
    params = solver.allocate_parameters()
    cache = solver.allocate_step_cache()
    a.particles.positions.upload(sc.positions)
    a.particles.velocities.upload(sc.velocities)
    for name in ("masses", "rest_densities", "viscosities", "surface_tensions"):
        getattr(params.particles, name).copy_(torch.from_numpy(sc.parameters[name]))
    for name in ("relaxation", "artificial_pressure_strength", "artificial_pressure_exponent",
                 "artificial_pressure_radius", "xsph_viscosity", "vorticity_confinement"):
        getattr(params, name).copy_(torch.from_numpy(sc.parameters[name]))
    return sc, solver, [a, b], control, params, cache


def cpu_baseline(workload, sc):
    """The FP32 oracle on this host: one step of the same scene, all threads."""
    import oracle

    d = oracle.make_domain(sc.domain)
    zeros = np.zeros_like(sc.positions)
    t0 = time.perf_counter()
    if workload == "dfsph":
        z = np.zeros(sc.count, np.float32)
        oracle.dfsph_step(d, 0, sc.positions, sc.velocities, zeros, sc.parameters, z, z, 4, 6)
        sample = "1 step of the 1,048,576-particle DFSPH dam break (FP32 oracle, OpenMP over particles)"
    else:
        oracle.pbf_step(d, 0, sc.positions, sc.velocities, zeros, sc.parameters, 4)
        sample = "1 step of the 32,768-particle PBF dam break (FP32 oracle, OpenMP over particles)"
    elapsed = time.perf_counter() - t0
    return {"value": sc.count / elapsed, "unit": "particle-steps/s", "cores": oracle.threads(), "kind": "port",
            "sample": sample, "seconds": elapsed}


def adjoint_ratio(device, steps=120):
    """Config of the differentiable PBF rollout; returns the timing record."""
    import torch

    from physica_amd import scenes
    from physica_amd.particle import PBF, ExecutionMode
    from physica_amd.rollout import PbfRollout

    sc = scenes.pbf_dam_break(shape=(64, 64, 64), spacing=0.02, seed=20260004)
    n = sc.count
    rng = np.random.default_rng(20260104)
    target = (sc.positions + np.array([[0.1], [-0.05], [0.08]]) + rng.normal(0, 0.01, size=(3, n))).astype(np.float64)

    def make(mode):
        solver = PBF(sc.domain, PBF.Configuration(pressure_iterations=4), mode=mode, device=device)
        state = solver.allocate_state()
        state.particles.positions.upload(sc.positions)
        state.particles.velocities.upload(sc.velocities)
        params = solver.allocate_parameters()
        for name in ("masses", "rest_densities", "viscosities", "surface_tensions"):
            getattr(params.particles, name).copy_(torch.from_numpy(sc.parameters[name]))
        for name in ("relaxation", "artificial_pressure_strength", "artificial_pressure_exponent",
                     "artificial_pressure_radius", "xsph_viscosity", "vorticity_confinement"):
            getattr(params, name).copy_(torch.from_numpy(sc.parameters[name]))
        return solver, state, solver.allocate_control(), params

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return out, (time.perf_counter() - t0) * 1e3

    # plain forward rollout (forward mode, one cache reused)
    solver, state, control, params = make(ExecutionMode.forward)
    cache = solver.allocate_step_cache()
    states = [solver.allocate_state(), solver.allocate_state()]

    def plain():
        solver.copy_state(state, states[0])
        states[0].particles.step_index = 0
        for k in range(steps):
            solver.forward_step(states[k % 2], control, params, states[(k + 1) % 2], cache)

    plain()  # warm-up (allocations, code loading)
    _, plain_ms = timed(plain)
    forward_solver = solver  # steps whose caches a checkpointed rollout does not keep
    del states, cache

    record = {"workload": "differentiable PBF rollout: 262,144 particles (64^3 lattice, spacing 0.02, h 0.04), "
                          f"{steps} steps, dt 1/120, 4 iterations, gradient of the final-position loss w.r.t. v0",
              "plain_forward_ms": round(plain_ms, 2)}
    dsolver, dstate, dcontrol, dparams = make(ExecutionMode.differentiable)
    for every in (1, 10):
        rollout = PbfRollout(dsolver, steps, checkpoint_every=every, parameter_gradients=False,
                             forward_solver=forward_solver if every > 1 else None)
        final, fwd_ms = timed(lambda: rollout.forward(dstate, dcontrol, dparams))
        xT = final.particles.positions.numpy().astype(np.float64)
        seed = dsolver.allocate_state_adjoint()
        seed.particles.positions.upload(2.0 * (xT - target) / n)
        (initial, _, _), bwd_ms = timed(lambda: rollout.backward(seed))
        grad = initial.particles.velocities.numpy()
        if not np.isfinite(grad).all():
            raise RuntimeError("non-finite rollout gradient")
        key = "store_all" if every == 1 else f"checkpoint_{every}"
        record[key] = {"forward_ms": round(fwd_ms, 2), "backward_ms": round(bwd_ms, 2),
                       "ratio": round((fwd_ms + bwd_ms) / plain_ms, 3)}
        rollout.release()
        del rollout, final, initial
    record["ratio"] = record["store_all"]["ratio"]
    return record


def main(argv=None) -> int:
    args = parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(multi-GPU slab decomposition is a design, not part of this build)", file=sys.stderr)
        return 2
    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        return 2
    device = torch.device("cuda:0")
    torch.cuda.set_device(device)
    if args.workload == "dfsph":
        sc, solver, states, control, params, cache = setup_dfsph(device)
    else:
        sc, solver, states, control, params, cache = setup_pbf(device)
    n = sc.count
    cells = math.prod(sc.domain.grid_resolution())

    def step(k):
        solver.forward_step(states[k % 2], control, params, states[(k + 1) % 2], cache)

    for k in range(args.warmup):
        step(k)
    torch.cuda.synchronize()
    if not args.no_pass_timing:
        solver.set_profiling(True)
    stream = torch.cuda.current_stream(device)
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for k in range(args.warmup, args.warmup + args.steps):
        step(k)
    stop.record(stream)
    torch.cuda.synchronize()
    elapsed_ms = start.elapsed_time(stop)
    passes = [] if args.no_pass_timing else solver.read_profile()
    solver.set_profiling(False)
    ms_per_step = elapsed_ms / args.steps
    value = n * args.steps / (elapsed_ms * 1e-3)

    final = states[(args.warmup + args.steps) % 2]
    x = final.particles.positions.numpy()
    if not np.isfinite(x).all():
        print("non-finite positions after the timed steps", file=sys.stderr)
        return 1

    if args.workload == "dfsph":
        iterations = solver.last_iterations()
        bytes_step = dfsph_bytes_per_step(iterations, n, cells)
        workload = ("DFSPH 3D dam break: 1,048,576 particles (128x64x128 block, spacing 0.01, jitter 0.001), "
                    "h 0.02 cubic spline, dt 1e-3, 4 divergence + 6 density iterations (fixed counts, the "
                    "reference's DFSPH), viscosity 0.01, static box 2.56x1.28x1.28")
        config = {"workload": workload, "particles": n, "grid_cells": cells,
                  "divergence_iterations": iterations[0], "density_iterations": iterations[1]}
    else:
        bytes_step = pbf_bytes_per_step(4, n, cells)
        workload = "PBF dam break: 32,768 particles (32^3, spacing 0.05), h 0.1, dt 1/120, 4 iterations"
        config = {"workload": workload, "particles": n, "grid_cells": cells, "pressure_iterations": 4}

    roofline = None
    if passes:
        dom = max(passes, key=lambda p: p["total_ms"])
        avg_ms = dom["total_ms"] / dom["launches"]
        achieved = dom["bytes_per_particle"] * n / (avg_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic_for(dom["name"]),
                    "kernel": dom["name"], "avg_launch_us": round(avg_ms * 1e3, 2),
                    "algorithmic_bytes_per_launch": dom["bytes_per_particle"] * n}
    step_achieved = bytes_step * n * args.steps / (elapsed_ms * 1e-3) / 1e9

    baseline = None
    if not args.no_cpu_baseline:
        baseline = cpu_baseline(args.workload, sc)
    adjoint = None
    if not args.no_adjoint:
        del states, cache
        adjoint = adjoint_ratio(device, steps=args.adjoint_steps)

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": config,
        "roofline": roofline,
        "cpu_baseline": baseline,
        "hbm_fraction_step": round(step_achieved / HBM_PEAK_GBS, 5),
        "algorithmic_bytes_per_particle_step": round(bytes_step, 2),
        "fwd_adjoint_ratio": None if adjoint is None else adjoint["ratio"],
        "fwd_adjoint": adjoint,
        "passes": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in p.items()} for p in passes],
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
