"""Differentiable PBF rollouts: forward steps recorded in step caches, then
PBF::vjp_step chained backwards (pbf/pbf.cpp:217-258).

`checkpoint_every=1` keeps every step cache resident (about 260 bytes per
particle per step at 4 constraint iterations, so a 262,144-particle,
120-step rollout holds roughly 8 GB of HBM and the backward pass recomputes
nothing). `checkpoint_every=k` keeps only every k-th state and recomputes each
k-step segment's caches before differentiating it, into a segment workspace
(states and caches) that the segments reuse. The steps whose caches are not
kept can run on `forward_solver`, a solver of the same scene created in
forward mode: it skips the recording and gives the same states bit for bit.

Checkpointed rollouts record the last segment during the forward pass, so
the backward pass starts differentiating at once. With `overlap` (the
default) the backward pass alternates the segments between the solver and a
second differentiable solver of the same scene on a stream of its own, each
with its own segment workspace: while one solver differentiates segment j,
the other recomputes segment j - 1. The adjoint passes from one stream to the
other through stream waits, so every vjp_step still runs in the order of the
single-stream schedule and the results are the same bit for bit.
"""
from __future__ import annotations

from typing import Dict, List, Optional, Tuple

import torch

from .particle import PBF, Control, ExecutionMode, ParticleParameterAdjoint


class PbfRollout:
    """A fixed-length rollout of one PBF solver created in differentiable mode."""

    def __init__(self, solver: PBF, steps: int, checkpoint_every: int = 1, parameter_gradients: bool = True,
                 forward_solver: Optional[PBF] = None, overlap: bool = True):
        if steps < 0 or checkpoint_every < 1:
            raise ValueError("steps must be >= 0 and checkpoint_every >= 1")
        self.solver = solver
        self.forward_solver = forward_solver
        self.steps = steps
        self.checkpoint_every = checkpoint_every
        # without parameter gradients the VJPs receive null parameter adjoints
        # and skip the terms that only feed them
        self.parameter_gradients = parameter_gradients
        self.overlap = overlap and checkpoint_every > 1
        self._states: List[Optional[PBF.State]] = []
        self._caches: List[Optional[PBF.StepCache]] = []
        self._control: Optional[Control] = None
        self._parameters: Optional[PBF.Parameters] = None
        # checkpointed rollouts: per solver, its segment workspace (states and
        # caches); the segment recorded by the forward pass
        self._work: Dict[int, Tuple[List[PBF.State], List[PBF.StepCache]]] = {}
        self._recorded: Optional[Tuple[int, int]] = None
        self._scratch_states: List[PBF.State] = []
        self._scratch_cache: Optional[PBF.StepCache] = None
        self._partner: Optional[PBF] = None

    def _workspace(self, solver: PBF, count: int):
        states, caches = self._work.setdefault(id(solver), ([], []))
        while len(caches) < count:
            caches.append(solver.allocate_step_cache())
            states.append(solver.allocate_state())
        return states, caches

    def _segments(self) -> List[Tuple[int, int]]:
        """The checkpointed segments [first, last), last one first."""
        segments = []
        k = self.steps
        while k > 0:
            first = ((k - 1) // self.checkpoint_every) * self.checkpoint_every
            segments.append((first, k))
            k = first
        return segments

    def _solvers(self) -> List[PBF]:
        """The solvers the backward pass alternates between: the rollout's own,
        and with `overlap` a second differentiable solver of the same scene
        created on a new stream (at the default priority)."""
        if not self.overlap:
            return [self.solver]
        if self._partner is None:
            device = self.solver.domain.device
            stream = torch.cuda.Stream(device)
            with torch.cuda.stream(stream):
                self._partner = PBF(self.solver.domain.configuration, self.solver.configuration,
                                    mode=ExecutionMode.differentiable, device=device)
        return [self.solver, self._partner]

    def forward(self, state: PBF.State, control: Control, parameters: PBF.Parameters) -> PBF.State:
        """Runs the rollout from `state` (not modified) and returns the final state."""
        self._control, self._parameters = control, parameters
        solver = self.solver
        start = solver.allocate_state()
        solver.copy_state(state, start)
        start.particles.step_index = state.particles.step_index
        self._states = [start]
        self._caches = []
        self._recorded = None
        keep_all = self.checkpoint_every == 1
        if not keep_all and not self._scratch_states:
            self._scratch_states = [solver.allocate_state(), solver.allocate_state()]
        # checkpointed: the last segment is recorded now, into the workspace of
        # the solver that differentiates it first
        last_first = self._segments()[0][0] if not keep_all and self.steps > 0 else self.steps
        if last_first < self.steps:
            seg_states, seg_caches = self._workspace(solver, self.steps - last_first)
            self._recorded = (last_first, self.steps)
        current = start
        for k in range(self.steps):
            checkpoint = keep_all or (k + 1) % self.checkpoint_every == 0 or k + 1 == self.steps
            if keep_all:
                nxt = solver.allocate_state()
                cache = solver.allocate_step_cache()
                solver.forward_step(current, control, parameters, nxt, cache)
                self._caches.append(cache)
            elif k >= last_first:
                nxt = solver.allocate_state() if checkpoint else seg_states[k - last_first]
                solver.forward_step(current, control, parameters, nxt, seg_caches[k - last_first])
                self._caches.append(None)
            else:
                nxt = solver.allocate_state() if checkpoint else self._scratch_states[k % 2]
                stepper = self.forward_solver or solver
                if self._scratch_cache is None:
                    self._scratch_cache = stepper.allocate_step_cache()
                stepper.forward_step(current, control, parameters, nxt, self._scratch_cache)
                self._caches.append(None)
            self._states.append(nxt if checkpoint else None)
            current = nxt
        return current

    def _segment(self, solver: PBF, first: int, last: int):
        """The states and step caches of steps [first, last) in `solver`'s
        workspace: recomputed from the checkpoint at `first`, unless the
        forward pass recorded them there."""
        count = last - first
        states, caches = self._workspace(solver, count)
        chain = [self._states[first]] + states[:count - 1] + [self._states[last]]
        if self._recorded != (first, last) or solver is not self.solver:
            if solver is self.solver:
                self._recorded = None  # the recorded segment is overwritten
            targets = states[:count]  # the checkpoint at `last` is read by other segments: not rewritten
            source = self._states[first]
            for local in range(count):
                solver.forward_step(source, self._control, self._parameters, targets[local], caches[local])
                source = targets[local]
        return chain, caches[:count]

    def backward(self, final_adjoint: PBF.StateAdjoint):
        """Chains the VJPs from the adjoint of the final state. Returns
        (adjoint of the initial state, control adjoint, parameter adjoint),
        all freshly allocated and accumulated over every step (the parameter
        adjoint is None without parameter gradients)."""
        solver = self.solver
        control_adjoint = solver.allocate_control_adjoint()
        if self.parameter_gradients:
            parameter_adjoint = solver.allocate_parameter_adjoint()
        else:
            parameter_adjoint = PBF.ParameterAdjoint(ParticleParameterAdjoint(None, None, None, None),
                                                     None, None, None, None, None, None)
        adjoint = solver.allocate_state_adjoint()
        solver.copy_state_adjoint(final_adjoint, adjoint)
        if self.checkpoint_every == 1:
            for k in reversed(range(self.steps)):
                previous = solver.allocate_state_adjoint()
                solver.vjp_step(self._states[k], self._control, self._parameters, self._states[k + 1],
                                self._caches[k], adjoint, previous, control_adjoint, parameter_adjoint)
                adjoint = previous
            return adjoint, control_adjoint, parameter_adjoint if self.parameter_gradients else None

        # checkpointed: segment i (from the end) runs on solvers[i % 2], each
        # solver on its own stream with its own workspace. A stream recomputes
        # its next segment while the other differentiates; each segment's
        # VJPs wait for the previous segment's.
        solvers = self._solvers()
        current = torch.cuda.current_stream(solver.domain.device)
        for s in solvers:
            s._stream.wait_stream(current)
        handoff = [adjoint]  # adjoints read on another stream: kept until every stream is done
        last_vjp: Optional[torch.cuda.Event] = None
        for index, (first, last) in enumerate(self._segments()):
            owner = solvers[index % len(solvers)]
            with torch.cuda.stream(owner._stream):
                chain, caches = self._segment(owner, first, last)
                if last_vjp is not None:
                    owner._stream.wait_event(last_vjp)
                for local in reversed(range(last - first)):
                    previous = owner.allocate_state_adjoint()
                    owner.vjp_step(chain[local], self._control, self._parameters, chain[local + 1], caches[local],
                                   adjoint, previous, control_adjoint, parameter_adjoint)
                    adjoint = previous
                last_vjp = torch.cuda.Event()
                last_vjp.record(owner._stream)
            handoff.append(adjoint)
        for s in solvers:
            current.wait_stream(s._stream)
        del handoff
        return adjoint, control_adjoint, parameter_adjoint if self.parameter_gradients else None

    def release(self) -> None:
        """Frees the recorded states, caches and segment workspaces."""
        torch.cuda.synchronize()
        self._states, self._caches = [], []
        self._work, self._recorded = {}, None
        self._scratch_states, self._scratch_cache = [], None
