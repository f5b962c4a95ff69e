"""Differentiable PBF rollouts: forward steps recorded in step caches, then
PBF::vjp_step chained backwards (pbf/pbf.cpp:217-258).

`checkpoint_every=1` keeps every step cache resident (about 260 bytes per
particle per step at 4 constraint iterations, so a 262,144-particle,
120-step rollout holds roughly 8 GB of HBM and the backward pass recomputes
nothing). `checkpoint_every=k` keeps only every k-th state and recomputes each
k-step segment's caches before differentiating it; the segment's states and
caches are one workspace reused by every segment. The steps whose caches are
not kept can run on `forward_solver`, a solver of the same scene created in
forward mode: it skips the recording and gives the same states bit for bit.
"""
from __future__ import annotations

from typing import List, Optional

import torch

from .particle import PBF, Control, ParticleParameterAdjoint


class PbfRollout:
    """A fixed-length rollout of one PBF solver created in differentiable mode."""

    def __init__(self, solver: PBF, steps: int, checkpoint_every: int = 1, parameter_gradients: bool = True,
                 forward_solver: Optional[PBF] = None):
        if steps < 0 or checkpoint_every < 1:
            raise ValueError("steps must be >= 0 and checkpoint_every >= 1")
        self.solver = solver
        self.forward_solver = forward_solver
        self.steps = steps
        self.checkpoint_every = checkpoint_every
        # without parameter gradients the VJPs receive null parameter adjoints
        # and skip the terms that only feed them
        self.parameter_gradients = parameter_gradients
        self._states: List[Optional[PBF.State]] = []
        self._caches: List[Optional[PBF.StepCache]] = []
        self._control: Optional[Control] = None
        self._parameters: Optional[PBF.Parameters] = None
        # segment workspace (checkpointed rollouts)
        self._segment_states: List[PBF.State] = []
        self._segment_caches: List[PBF.StepCache] = []
        self._scratch_states: List[PBF.State] = []
        self._scratch_cache: Optional[PBF.StepCache] = None

    def forward(self, state: PBF.State, control: Control, parameters: PBF.Parameters) -> PBF.State:
        """Runs the rollout from `state` (not modified) and returns the final state."""
        self._control, self._parameters = control, parameters
        solver = self.solver
        start = solver.allocate_state()
        solver.copy_state(state, start)
        start.particles.step_index = state.particles.step_index
        self._states = [start]
        self._caches = []
        keep_all = self.checkpoint_every == 1
        if not keep_all and not self._scratch_states:
            self._scratch_states = [solver.allocate_state(), solver.allocate_state()]
        current = start
        for k in range(self.steps):
            checkpoint = keep_all or (k + 1) % self.checkpoint_every == 0 or k + 1 == self.steps
            nxt = solver.allocate_state() if checkpoint else self._scratch_states[k % 2]
            if keep_all:
                cache = solver.allocate_step_cache()
                solver.forward_step(current, control, parameters, nxt, cache)
                self._caches.append(cache)
            else:
                stepper = self.forward_solver or solver
                if self._scratch_cache is None:
                    self._scratch_cache = stepper.allocate_step_cache()
                stepper.forward_step(current, control, parameters, nxt, self._scratch_cache)
                self._caches.append(None)
            self._states.append(nxt if checkpoint else None)
            current = nxt
        return current

    def _segment(self, first: int, last: int):
        """Recomputes the states and caches of steps [first, last) from the
        checkpoint at `first` into the segment workspace."""
        solver = self.solver
        count = last - first
        while len(self._segment_caches) < count:
            self._segment_caches.append(solver.allocate_step_cache())
            self._segment_states.append(solver.allocate_state())
        states = [self._states[first]] + self._segment_states[:count]
        for local in range(count):
            solver.forward_step(states[local], self._control, self._parameters, states[local + 1],
                                self._segment_caches[local])
        return states, self._segment_caches[:count]

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
        k = self.steps
        while k > 0:
            if self.checkpoint_every == 1:
                first = k - 1
                states, caches = [self._states[k - 1], self._states[k]], [self._caches[k - 1]]
            else:
                first = ((k - 1) // self.checkpoint_every) * self.checkpoint_every
                states, caches = self._segment(first, k)
            for local in reversed(range(len(caches))):
                previous = solver.allocate_state_adjoint()
                solver.vjp_step(states[local], self._control, self._parameters, states[local + 1], caches[local],
                                adjoint, previous, control_adjoint, parameter_adjoint)
                adjoint = previous
            k = first
        return adjoint, control_adjoint, parameter_adjoint if self.parameter_gradients else None

    def release(self) -> None:
        """Frees the recorded states, caches and segment workspace."""
        self._states, self._caches = [], []
        self._segment_states, self._segment_caches = [], []
        self._scratch_states, self._scratch_cache = [], None
        torch.cuda.synchronize()
