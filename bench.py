"""Benchmark: SPH<DFSPH> dam break, 1,048,576 particles, one MI355X.

Prints one JSON line (rank 0). A step is one SPH<DFSPH>::forward_step of the
whole scene (grid build, density, the divergence and density solvers) with
every input resident in HBM. The configured workload runs both solvers to a
0.1 % mean compression error (FP64 residual, at most 100 iterations each):
the W warm-up steps are the scenario's first steps (untimed, their iteration
counts reported), and `value` is particle-steps/s over the K steps that
follow them. Timed steps that end a solver at the iteration cap without
reaching the tolerance are counted in `config.unconverged_steps`.
`fixed_iterations` gives the reference's own fixed counts (4 divergence + 6
density iterations, no tolerance) measured over steady steps; `roofline` is
its dominant pass, the algorithmic bytes per launch over the average launch
time (HIP events on the solver stream during those timed steps).
`cpu_baseline` is the FP32 oracle (OpenMP over particles, the same
reference-order arithmetic) timed on this host over one fixed-count step of
the same scene with every host thread; `cpu_baseline_1core` the same on one
thread over a 131,072-particle block of the same spacing and kernel.

`pbf_32k` is the PBF dam break (32,768 particles, h 0.1, 4 iterations):
its whole configured run of 200 forward steps at dt 1/120 (100 frames of
two substeps). `iisph_4m` and `wcsph_4m` are the scale test: two falling
128^3 cubes (4,194,304 particles, h 0.02) for 100 steps each from rest, IISPH
with relaxed Jacobi (0.5) to a 0.1 % mean compression error (FP64 residual,
at most 100 iterations) at dt 1e-3, WCSPH with the Tait equation (c = sqrt 7,
gamma 7) at dt 2e-4.

`fwd_adjoint_ratio` is the differentiable PBF rollout: 262,144 particles (64^3
lattice, spacing 0.02, h 0.04), 120 steps at dt 1/120, 4 iterations, loss the
mean squared distance of the final positions to a seeded target cloud,
gradient with respect to the initial velocities (no parameter gradients).
Ratio = (forward rollout + chained VJPs) / plain forward rollout, with a state
checkpoint every 10 steps: the forward records the last segment, and the
backward pass recomputes each other segment on one of two solvers (two
streams) while the other solver differentiates the segment after it. The same
on one stream, and with every step cache kept in HBM (no recompute), are
reported beside it.

`batched_pbf` is the uneven batch: 64 independent PBF dam breaks (particle
counts log-uniform in [4,096, 200,000] from seed 20260005, 3.79 M in all),
100 steps each, every scene its own solver on one of 32 streams; then forward
plus adjoint on the first 16 scenes against their plain forward.
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
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--mode", choices=["tolerance", "fixed"], default="tolerance",
                   help="DFSPH solver stopping: 0.1 %% tolerance (the configured workload) or fixed 4 + 6 only")
    p.add_argument("--fixed-steps", type=int, default=20,
                   help="timed steps of the fixed-count measurement (0: skip it, and the kernel roofline with it)")
    p.add_argument("--workload", choices=["dfsph", "pbf"], default="dfsph")
    p.add_argument("--no-cpu-baseline", action="store_true", help="skip the CPU oracle timing")
    p.add_argument("--no-pass-timing", action="store_true", help="time without per-pass events")
    p.add_argument("--no-adjoint", action="store_true", help="skip the differentiable PBF rollout ratio")
    p.add_argument("--no-batched", action="store_true", help="skip the 64-scene uneven PBF batch")
    p.add_argument("--no-pbf", action="store_true", help="skip the 32,768-particle PBF dam break")
    p.add_argument("--no-scale", action="store_true", help="skip the 4,194,304-particle IISPH and WCSPH scale test")
    p.add_argument("--batch-streams", type=int, default=32, help="HIP streams the PBF batch's scenes share")
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


def iisph_bytes_per_step(iterations, particles, cells):
    """IISPH: 188 + 128 I + 8 C / N."""
    return 188.0 + 128.0 * iterations + 8.0 * cells / particles


def wcsph_bytes_per_step(particles, cells):
    """WCSPH: 172 + 8 C / N."""
    return 172.0 + 8.0 * cells / particles


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


TOLERANCE = 1e-3  # 0.1 % mean density error
MAX_ITERATIONS = 100


def setup_dfsph(device, tolerance=0.0):
    import torch

    from physica_amd import scenes
    from physica_amd.particle import DFSPH, SPH

    sc = scenes.dfsph_dam_break()
    if tolerance > 0.0:
        cfg = DFSPH.Configuration(density_error_tolerance=tolerance, max_iterations=MAX_ITERATIONS)
    else:
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


def cpu_baseline(workload, sc, threads=None):
    """The FP32 oracle on this host: one step of `sc`, on `threads` OpenMP
    threads (all host threads when None)."""
    import oracle

    previous = oracle.threads()
    if threads is not None:
        oracle.set_threads(threads)
    try:
        used = oracle.threads()
        d = oracle.make_domain(sc.domain)
        zeros = np.zeros_like(sc.positions)
        t0 = time.perf_counter()
        if workload == "dfsph":
            z = np.zeros(sc.count, np.float32)
            oracle.dfsph_step(d, 0, sc.positions, sc.velocities, zeros, sc.parameters, z, z, 4, 6)
            sample = (f"1 step of the {sc.count:,}-particle DFSPH dam break (spacing 0.01, h 0.02), fixed 4 + 6 "
                      f"iterations (FP32 oracle, OpenMP over particles, {used} thread{'s' if used > 1 else ''})")
        else:
            oracle.pbf_step(d, 0, sc.positions, sc.velocities, zeros, sc.parameters, 4)
            sample = (f"1 step of the {sc.count:,}-particle PBF dam break (FP32 oracle, OpenMP over particles, "
                      f"{used} thread{'s' if used > 1 else ''})")
        elapsed = time.perf_counter() - t0
    finally:
        oracle.set_threads(previous)
    return {"value": sc.count / elapsed, "unit": "particle-steps/s", "cores": used, "kind": "port",
            "sample": sample, "seconds": elapsed}


def setup_two_cubes(device, method):
    """The scale test: two falling 128^3 cubes (4,194,304 particles), IISPH to
    a 0.1 % mean compression error or WCSPH with the Tait equation."""
    import torch

    from physica_amd import scenes
    from physica_amd.particle import IISPH, SPH, WCSPH

    sc = scenes.two_cubes(method)
    if method == "iisph":
        cfg = IISPH.Configuration(density_error_tolerance=TOLERANCE, max_iterations=MAX_ITERATIONS)
    else:
        cfg = WCSPH.Configuration()
    solver = SPH(sc.domain, cfg, device=device)
    a, b = solver.allocate_state(), solver.allocate_state()
    control, params, cache = solver.allocate_control(), solver.allocate_parameters(), solver.allocate_step_cache()
    a.particles.positions.upload(sc.positions)
    a.particles.velocities.upload(sc.velocities)
    for name in ("masses", "rest_densities", "viscosities", "surface_tensions"):
        getattr(params.particles, name).copy_(torch.from_numpy(sc.parameters[name]))
    if method == "iisph":
        params.method.jacobi_relaxation.copy_(torch.from_numpy(sc.parameters["relaxation"]))
    else:
        for name in ("speed_of_sound", "tait_exponent", "boundary_surface_tension"):
            getattr(params.method, name).copy_(torch.from_numpy(sc.parameters[name]))
    return sc, solver, [a, b], control, params, cache


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
        return pbf_solver_for(sc, mode, device)

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
    for every, overlap in ((1, False), (10, False), (10, True)):
        rollout = PbfRollout(dsolver, steps, checkpoint_every=every, parameter_gradients=False,
                             forward_solver=forward_solver if every > 1 else None, overlap=overlap)
        final, fwd_ms = timed(lambda: rollout.forward(dstate, dcontrol, dparams))
        xT = final.particles.positions.numpy().astype(np.float64)
        seed = dsolver.allocate_state_adjoint()
        seed.particles.positions.upload(2.0 * (xT - target) / n)
        (initial, _, _), bwd_ms = timed(lambda: rollout.backward(seed))
        grad = initial.particles.velocities.numpy()
        if not np.isfinite(grad).all():
            raise RuntimeError("non-finite rollout gradient")
        if every == 1:
            store_all_grad = grad
        key = "store_all" if every == 1 else f"checkpoint_{every}" + ("" if overlap else "_one_stream")
        record[key] = {"forward_ms": round(fwd_ms, 2), "backward_ms": round(bwd_ms, 2),
                       "ratio": round((fwd_ms + bwd_ms) / plain_ms, 3)}
        if every > 1:
            # the same gradient bit for bit as the store-all rollout
            record[key]["gradient_equals_store_all"] = bool(np.array_equal(grad, store_all_grad))
        rollout.release()
        del rollout, final, initial
    # BASELINE's definition: forward + backward with checkpoint-every-10 recompute
    record["ratio"] = record["checkpoint_10"]["ratio"]
    return record


def pbf_solver_for(scene, mode, device):
    """A PBF solver with `scene` uploaded: (solver, state, control, parameters)."""
    import torch

    from physica_amd.particle import PBF

    solver = PBF(scene.domain, PBF.Configuration(pressure_iterations=4), mode=mode, device=device)
    state = solver.allocate_state()
    state.particles.positions.upload(scene.positions)
    state.particles.velocities.upload(scene.velocities)
    params = solver.allocate_parameters()
    for name in ("masses", "rest_densities", "viscosities", "surface_tensions"):
        getattr(params.particles, name).copy_(torch.from_numpy(scene.parameters[name]))
    for name in ("relaxation", "artificial_pressure_strength", "artificial_pressure_exponent",
                 "artificial_pressure_radius", "xsph_viscosity", "vorticity_confinement"):
        getattr(params, name).copy_(torch.from_numpy(scene.parameters[name]))
    return solver, state, solver.allocate_control(), params


def batched_pbf(device, steps=100, adjoint_scenes=16, streams=32):
    """The uneven PBF batch: 64 independent dam breaks (4,100 to 192,450
    particles, 3.79 M in all), 100 steps. The batched engine runs every pass
    of a substep as one launch over all scenes and replays a captured substep
    (settings from the autotuner's JSON when present). Beside it: the same
    batch as 64 solvers spread over `streams` streams, issued step by step.
    Then forward + adjoint on the first `adjoint_scenes` scenes: store-all
    rollouts (every step cache kept) of one solver per scene on its own
    stream, issued step by step, forward and then the chained VJPs, against
    their plain forward issued the same way."""
    import torch

    from physica_amd import scenes
    from physica_amd.batch import PbfBatch
    from physica_amd.particle import ExecutionMode, PBF

    batch = scenes.batched_pbf_scenes()
    total = sum(sc.count for sc in batch)

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3

    # batched engine
    engine = PbfBatch([sc.domain for sc in batch], PBF.Configuration(pressure_iterations=4), device=device)
    state, control, params = engine.allocate_state(), engine.allocate_control(), engine.allocate_parameters()
    engine.upload(batch, state, params)
    initial = engine.allocate_state()
    initial.particles.positions.copy_(state.particles.positions)
    initial.particles.velocities.copy_(state.particles.velocities)
    engine.run(state, control, params, 2)  # warm-up: code loading and the graph capture
    state.particles.positions.copy_(initial.particles.positions)
    state.particles.velocities.copy_(initial.particles.velocities)
    state.particles.step_index = 0
    engine_ms = timed(lambda: engine.run(state, control, params, steps))
    if not np.isfinite(state.particles.positions.numpy()).all():
        raise RuntimeError("non-finite positions after the batched run")
    record = {"workload": f"{len(batch)} uneven PBF dam breaks, {total} particles "
                          f"({min(sc.count for sc in batch)} to {max(sc.count for sc in batch)}), "
                          f"{steps} steps, dt 1/120, 4 iterations",
              "particles": total, "steps": steps, "forward_ms": round(engine_ms, 2),
              "particle_steps_per_s": round(total * steps / (engine_ms * 1e-3), 1),
              "engine": {"launches_per_pass": 1, "kernels_per_step": engine.kernels_per_step,
                         "graph_replay": engine.use_graph, "mask_skin": engine.mask_skin}}
    del engine, state, control, params, initial

    # the same batch as one solver per scene on a pool of streams
    pool = [torch.cuda.Stream(device) for _ in range(streams)]
    runs = []
    for k, sc in enumerate(batch):
        with torch.cuda.stream(pool[k % streams]):
            solver, state, control, params = pbf_solver_for(sc, ExecutionMode.forward, device)
            runs.append((pool[k % streams], solver, state, [solver.allocate_state(), solver.allocate_state()],
                         control, params, solver.allocate_step_cache()))
    torch.cuda.synchronize()

    def forward(selected):
        for stream, solver, initial, states, _, _, _ in selected:
            with torch.cuda.stream(stream):
                solver.copy_state(initial, states[0])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            for stream, solver, _, states, control, params, cache in selected:
                with torch.cuda.stream(stream):
                    solver.forward_step(states[k % 2], control, params, states[(k + 1) % 2], cache)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3

    forward(runs)  # warm-up
    streams_ms = forward(runs)
    record["per_solver_streams"] = {"streams": streams, "forward_ms": round(streams_ms, 2),
                                    "particle_steps_per_s": round(total * steps / (streams_ms * 1e-3), 1)}
    del runs

    # forward + adjoint on the first scenes, one stream per scene
    chosen = batch[:adjoint_scenes]
    lanes = []
    for k, sc in enumerate(chosen):
        stream = torch.cuda.Stream(device)
        with torch.cuda.stream(stream):
            solver, state, control, params = pbf_solver_for(sc, ExecutionMode.differentiable, device)
            plain, pstate, _, _ = pbf_solver_for(sc, ExecutionMode.forward, device)
            lanes.append({"stream": stream, "solver": solver, "state": state, "control": control, "params": params,
                          "plain": plain, "plain_states": [plain.allocate_state(), plain.allocate_state()],
                          "plain_cache": plain.allocate_step_cache(), "scene": sc})
    torch.cuda.synchronize()

    def plain_forward():
        for lane in lanes:
            with torch.cuda.stream(lane["stream"]):
                lane["plain"].copy_state(lane["state"], lane["plain_states"][0])
        for k in range(steps):
            for lane in lanes:
                with torch.cuda.stream(lane["stream"]):
                    st = lane["plain_states"]
                    lane["plain"].forward_step(st[k % 2], lane["control"], lane["params"], st[(k + 1) % 2],
                                               lane["plain_cache"])

    def recorded_forward():
        for lane in lanes:
            with torch.cuda.stream(lane["stream"]):
                lane["states"] = [lane["state"]]
                lane["caches"] = []
        for k in range(steps):
            for lane in lanes:
                with torch.cuda.stream(lane["stream"]):
                    solver = lane["solver"]
                    nxt, cache = solver.allocate_state(), solver.allocate_step_cache()
                    solver.forward_step(lane["states"][-1], lane["control"], lane["params"], nxt, cache)
                    lane["states"].append(nxt)
                    lane["caches"].append(cache)

    def backward():
        from physica_amd.particle import ParticleParameterAdjoint

        for lane in lanes:
            with torch.cuda.stream(lane["stream"]):
                solver, sc = lane["solver"], lane["scene"]
                target = sc.positions.astype(np.float64) + np.array([[0.1], [-0.05], [0.08]])
                seed = solver.allocate_state_adjoint()
                final = lane["states"][-1].particles.positions.numpy().astype(np.float64)
                seed.particles.positions.upload(2.0 * (final - target) / sc.count)
                lane["adjoint"] = seed
                lane["control_adjoint"] = solver.allocate_control_adjoint()
                lane["parameter_adjoint"] = PBF.ParameterAdjoint(ParticleParameterAdjoint(None, None, None, None),
                                                                 None, None, None, None, None, None)
        for k in reversed(range(steps)):
            for lane in lanes:
                with torch.cuda.stream(lane["stream"]):
                    solver = lane["solver"]
                    previous = solver.allocate_state_adjoint()
                    solver.vjp_step(lane["states"][k], lane["control"], lane["params"], lane["states"][k + 1],
                                    lane["caches"][k], lane["adjoint"], previous, lane["control_adjoint"],
                                    lane["parameter_adjoint"])
                    lane["adjoint"] = previous

    plain_forward()  # warm-up
    plain_ms = timed(plain_forward)
    fwd_ms = timed(recorded_forward)
    bwd_ms = timed(backward)
    for lane in lanes:
        if not np.isfinite(lane["adjoint"].particles.velocities.numpy()).all():
            raise RuntimeError("non-finite batch gradient")
    record["adjoint"] = {"scenes": adjoint_scenes, "particles": sum(sc.count for sc in chosen),
                         "streams": adjoint_scenes, "plain_forward_ms": round(plain_ms, 2),
                         "forward_ms": round(fwd_ms, 2), "backward_ms": round(bwd_ms, 2),
                         "ratio": round((fwd_ms + bwd_ms) / plain_ms, 3)}
    del lanes
    torch.cuda.synchronize()
    return record


def progress(message: str) -> None:
    """One line per finished leg on stderr (the JSON line stays alone on stdout)."""
    print(f"bench: {message}", file=sys.stderr, flush=True)


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

    def measure(setup, steps, warmup, from_start, profile):
        """Times `steps` forward steps; with from_start the warm-up steps run
        on a copy and the timed steps start from the initial state, otherwise
        the timed steps continue the scenario after the warm-up steps.
        Returns (scene, elapsed ms, per-pass profile, iteration counts per
        timed step, solver, states, cache, iteration counts per warm-up step)."""
        sc, solver, states, control, params, cache = setup()
        initial = solver.allocate_state() if from_start else None
        if from_start:
            solver.copy_state(states[0], initial)

        def step(k):
            solver.forward_step(states[k % 2], control, params, states[(k + 1) % 2], cache)
            return solver.last_iterations() if hasattr(solver, "last_iterations") else None

        warm_counts = [step(k) for k in range(warmup)]
        if from_start:
            solver.copy_state(initial, states[0])
            states[0].particles.step_index = 0
            first = 0
        else:
            first = warmup
        torch.cuda.synchronize()
        if profile:
            solver.set_profiling(True)
        stream = torch.cuda.current_stream(device)
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        start.record(stream)
        counts = [step(k) for k in range(first, first + steps)]
        stop.record(stream)
        torch.cuda.synchronize()
        elapsed_ms = start.elapsed_time(stop)
        passes = solver.read_profile() if profile else []
        solver.set_profiling(False)
        final = states[(first + steps) % 2]
        if not np.isfinite(final.particles.positions.numpy()).all():
            raise RuntimeError("non-finite positions after the timed steps")
        return sc, elapsed_ms, passes, counts, solver, states, cache, warm_counts

    def roofline_of(passes, n):
        if not passes:
            return None
        dom = max(passes, key=lambda p: p["total_ms"])
        avg_ms = dom["total_ms"] / dom["launches"]
        achieved = dom["bytes_per_particle"] * n / (avg_ms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic_for(dom["name"]),
                "kernel": dom["name"], "avg_launch_us": round(avg_ms * 1e3, 2),
                "algorithmic_bytes_per_launch": dom["bytes_per_particle"] * n}

    fixed = None
    roofline = None
    if args.workload == "dfsph":
        if args.mode == "fixed" and args.fixed_steps < 1:
            print("--mode fixed needs --fixed-steps >= 1", file=sys.stderr)
            return 2
        if args.fixed_steps > 0:
            progress("DFSPH fixed 4 + 6 iterations")
            # the reference's fixed counts over steady steps: per-pass timing and the kernel roofline
            sc, f_ms, f_passes, f_counts, solver, states, cache, _ = measure(
                lambda: setup_dfsph(device), args.fixed_steps, args.warmup, False, not args.no_pass_timing)
            n = sc.count
            cells = math.prod(sc.domain.grid_resolution())
            f_bytes = dfsph_bytes_per_step((4, 6), n, cells)
            fixed = {"iterations": [4, 6], "steps": args.fixed_steps, "warmup": args.warmup,
                     "ms_per_step": round(f_ms / args.fixed_steps, 4),
                     "value": round(n * args.fixed_steps / (f_ms * 1e-3), 1),
                     "hbm_fraction_step": round(
                         f_bytes * n * args.fixed_steps / (f_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
                     "algorithmic_bytes_per_particle_step": round(f_bytes, 2),
                     "passes": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in p.items()}
                                for p in f_passes]}
            roofline = roofline_of(f_passes, n)
            del solver, states, cache
        if args.mode == "tolerance":
            progress("DFSPH 0.1 % tolerance")
            sc, elapsed_ms, passes, counts, solver, states, cache, warm_counts = measure(
                lambda: setup_dfsph(device, TOLERANCE), args.steps, args.warmup, False, False)
            n = sc.count
            cells = math.prod(sc.domain.grid_resolution())
            bytes_total = sum(dfsph_bytes_per_step(c, n, cells) for c in counts)
            bytes_step = bytes_total / len(counts)
            first, last = args.warmup, args.warmup + args.steps - 1
            workload = ("DFSPH 3D dam break: 1,048,576 particles (128x64x128 block, spacing 0.01, jitter 0.001), "
                        "h 0.02 cubic spline, dt 1e-3, divergence and density solvers to a 0.1 % mean density "
                        "error (FP64 residual, max 100 iterations each), viscosity 0.01, static box 2.56x1.28x1.28; "
                        f"steps 0..{first - 1} untimed warm-up, timed steps {first}..{last} of the scenario")
            nd = [c[0] for c in counts]
            ne = [c[1] for c in counts]
            unconverged = sum(1 for a, b in counts if a >= MAX_ITERATIONS or b >= MAX_ITERATIONS)
            config = {"workload": workload, "particles": n, "grid_cells": cells, "tolerance": TOLERANCE,
                      "max_iterations": MAX_ITERATIONS,
                      "divergence_iterations": {"total": sum(nd), "max": max(nd), "mean": round(sum(nd) / len(nd), 3)},
                      "density_iterations": {"total": sum(ne), "max": max(ne), "mean": round(sum(ne) / len(ne), 3)},
                      "unconverged_steps": unconverged,
                      "warmup_iterations": [list(c) for c in warm_counts]}
            steps_timed, warmup_used = args.steps, args.warmup
        else:
            elapsed_ms, passes, steps_timed, warmup_used = f_ms, f_passes, args.fixed_steps, args.warmup
            bytes_step = f_bytes
            workload = ("DFSPH 3D dam break: 1,048,576 particles (128x64x128 block, spacing 0.01, jitter 0.001), "
                        "h 0.02 cubic spline, dt 1e-3, 4 divergence + 6 density iterations (fixed counts, the "
                        "reference's DFSPH), viscosity 0.01, static box 2.56x1.28x1.28")
            config = {"workload": workload, "particles": n, "grid_cells": cells, "divergence_iterations": 4,
                      "density_iterations": 6}
            fixed = None
    else:
        sc, elapsed_ms, passes, _, solver, states, cache, _ = measure(
            lambda: setup_pbf(device), args.steps, args.warmup, False, not args.no_pass_timing)
        n = sc.count
        cells = math.prod(sc.domain.grid_resolution())
        bytes_step = pbf_bytes_per_step(4, n, cells)
        workload = "PBF dam break: 32,768 particles (32^3, spacing 0.05), h 0.1, dt 1/120, 4 iterations"
        config = {"workload": workload, "particles": n, "grid_cells": cells, "pressure_iterations": 4}
        roofline = roofline_of(passes, n)
        steps_timed, warmup_used = args.steps, args.warmup
    ms_per_step = elapsed_ms / steps_timed
    value = n * steps_timed / (elapsed_ms * 1e-3)
    step_achieved = bytes_step * n * steps_timed / (elapsed_ms * 1e-3) / 1e9

    solver = states = cache = None  # release the headline solver's buffers
    progress("headline done")
    pbf_leg = None
    if not args.no_pbf and args.workload == "dfsph":
        # configuration 1: the whole 200-step PBF dam break from the initial state
        psc, p_ms, _, _, psolver, pstates, pcache, _ = measure(lambda: setup_pbf(device), 200, 3, True, False)
        pn = psc.count
        pcells = math.prod(psc.domain.grid_resolution())
        pbytes = pbf_bytes_per_step(4, pn, pcells)
        pbf_leg = {"workload": "PBF dam break: 32,768 particles (32^3 lattice, spacing 0.05, jitter 0.005), h 0.1, "
                               "rest density 1000, 4 iterations, dt 1/120 (two substeps of 1/60), box 3.2x2.0x1.6; "
                               "forward steps 0..199 from the initial state",
                   "particles": pn, "steps": 200, "ms_per_step": round(p_ms / 200, 4),
                   "value": round(pn * 200 / (p_ms * 1e-3), 1),
                   "hbm_fraction_step": round(pbytes * pn * 200 / (p_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
                   "algorithmic_bytes_per_particle_step": round(pbytes, 2)}
        del psolver, pstates, pcache
    progress("PBF 32,768 done")
    scale = None
    if not args.no_scale and args.workload == "dfsph":
        scale = {}
        for method in ("iisph", "wcsph"):
            ssc, s_ms, _, s_counts, ssolver, sstates, scache, _ = measure(
                lambda: setup_two_cubes(device, method), 100, 1, True, False)
            sn = ssc.count
            scells = math.prod(ssc.domain.grid_resolution())
            rec = {"particles": sn, "grid_cells": scells, "steps": 100, "ms_per_step": round(s_ms / 100, 4),
                   "value": round(sn * 100 / (s_ms * 1e-3), 1)}
            if method == "iisph":
                its = [c[0] for c in s_counts]
                sbytes = sum(iisph_bytes_per_step(i, sn, scells) for i in its) / len(its)
                rec["workload"] = ("IISPH scale test: two falling 128^3 cubes (4,194,304 particles, spacing 0.01, "
                                   "jitter 0.001) in a 4x4x2 basin, h 0.02, relaxed Jacobi 0.5 to a 0.1 % mean "
                                   "compression error (FP64 residual, max 100 iterations), dt 1e-3, steps 0..99")
                rec["pressure_iterations"] = {"total": sum(its), "max": max(its), "first_step": its[0],
                                              "mean": round(sum(its) / len(its), 3)}
                rec["unconverged_steps"] = sum(1 for i in its if i >= MAX_ITERATIONS)
            else:
                sbytes = wcsph_bytes_per_step(sn, scells)
                rec["workload"] = ("WCSPH scale test: two falling 128^3 cubes (4,194,304 particles, spacing 0.01, "
                                   "jitter 0.001) in a 4x4x2 basin, h 0.02, Tait equation c = sqrt(7) gamma 7 "
                                   "(stiffness 1000), no viscosity or surface tension, dt 2e-4, steps 0..99")
            rec["hbm_fraction_step"] = round(sbytes * sn * 100 / (s_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)
            rec["algorithmic_bytes_per_particle_step"] = round(sbytes, 2)
            scale[f"{method}_4m"] = rec
            del ssolver, sstates, scache

    progress("scale test done")
    baseline = baseline_1core = None
    if not args.no_cpu_baseline:
        baseline = cpu_baseline(args.workload, sc)
        if args.workload == "dfsph":
            from physica_amd import scenes

            baseline_1core = cpu_baseline("dfsph", scenes.dfsph_dam_break(shape=(64, 32, 64)), threads=1)
    progress("CPU baselines done")
    adjoint = None
    if not args.no_adjoint:
        adjoint = adjoint_ratio(device, steps=args.adjoint_steps)
    progress("differentiable rollout done")
    batched = None
    if not args.no_batched:
        batched = batched_pbf(device, streams=args.batch_streams)

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps_timed,
        "warmup": warmup_used,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": config,
        "roofline": roofline,
        "cpu_baseline": baseline,
        "cpu_baseline_1core": baseline_1core,
        "hbm_fraction_step": round(step_achieved / HBM_PEAK_GBS, 5),
        "algorithmic_bytes_per_particle_step": round(bytes_step, 2),
        "fwd_adjoint_ratio": None if adjoint is None else adjoint["ratio"],
        "fwd_adjoint": adjoint,
        "batched_pbf": batched,
        "pbf_32k": pbf_leg,
        "iisph_4m": None if scale is None else scale["iisph_4m"],
        "wcsph_4m": None if scale is None else scale["wcsph_4m"],
        "fixed_iterations": fixed,
        "passes": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in p.items()} for p in passes],
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
